"""Mat assembly on the device (mpk_mat_fold, mpk_dense_to_csr) against the
host fold and numpy's nonzero, bit for bit: the CSR arrays of logs of
setValue / setValues writes (inserts and adds, repeated entries, -0.0 and
NaN values, writes on top of an assembled matrix, empty rows), and of dense
arrays.  MPK_FOLD_DEVICE_MIN / MPK_DENSE_DEVICE_MIN select the path."""

import importlib

import numpy as np
import pytest

from _helpers import same_bits

pytestmark = pytest.mark.gpu


def _fill(A, rng, m, n, writes, base=None):
    if base is not None:
        A.setValuesCSR(*base)
    for _ in range(writes):
        kind = rng.integers(0, 3)
        add = bool(rng.integers(0, 2))
        mode = True if add else None  # any addv adds (the reference's Mat.setValue)
        if kind == 0:
            r, c = int(rng.integers(0, m)), int(rng.integers(0, n))
            v = rng.choice([rng.normal(), -0.0, 0.0, np.nan, 1e300, -1e-300])
            A.setValue(r, c, v, addv=mode)
        else:
            rows = rng.integers(0, m, int(rng.integers(1, 6)))
            cols = rng.integers(0, n, int(rng.integers(1, 6)))
            vals = rng.normal(size=len(rows) * len(cols))
            vals[rng.random(len(vals)) < 0.1] = -0.0
            A.setValues(rows, cols, vals, addv=mode)
    A.assemble()


def _csr(A):
    ip, ix, va = A.getValuesCSR()
    return np.asarray(ip, np.int64), np.asarray(ix, np.int64), np.asarray(va)


@pytest.mark.parametrize("seed,m,n,writes,with_base", [(0, 50, 40, 300, False), (1, 200, 200, 2000, True),
                                                       (2, 7, 300, 500, True), (3, 1000, 3, 800, False)])
def test_device_fold_matches_host(native, monkeypatch, seed, m, n, writes, with_base):
    from mojopetsc4py import PETSc
    base = None
    if with_base:
        rng = np.random.default_rng(seed + 100)
        dense = np.where(rng.random((m, n)) < 0.05, rng.normal(size=(m, n)), 0.0)
        rr, cc = np.nonzero(dense)
        ip = np.zeros(m + 1, np.int64)
        np.cumsum(np.bincount(rr, minlength=m), out=ip[1:])
        base = (ip, cc.astype(np.int64), dense[rr, cc])
    got = []
    for threshold in ("1", str(10 ** 9)):
        monkeypatch.setenv("MPK_FOLD_DEVICE_MIN", threshold)
        A = PETSc.Mat().createSeqAIJ((m, n), nnz=3)
        _fill(A, np.random.default_rng(seed), m, n, writes, base)
        got.append((_csr(A), A))
    (ip_d, ix_d, va_d), Ad = got[0]
    (ip_h, ix_h, va_h), Ah = got[1]
    assert np.array_equal(ip_d, ip_h) and np.array_equal(ix_d, ix_h)
    assert same_bits(va_d, va_h)
    assert Ad._devc is not None and Ah._devc is None  # the device fold left its copy on the device
    # the adopted device copy multiplies as an uploaded one does
    x = PETSc.Vec().createWithArray(np.random.default_rng(seed).normal(size=n))
    ys = []
    for A in (Ad, Ah):
        y = PETSc.Vec().createSeq(m)
        A.mult(x, y)
        ys.append(y.getArray().copy())
    assert same_bits(ys[0], ys[1])


def test_device_fold_empty_and_sparse_rows(native, monkeypatch):
    """No writes at all, and writes to a few rows of many (empty rows between)."""
    from mojopetsc4py import PETSc
    monkeypatch.setenv("MPK_FOLD_DEVICE_MIN", "0")
    A = PETSc.Mat().createSeqAIJ((10, 10), nnz=2)
    A.assemble()
    ip, ix, va = _csr(A)
    assert np.array_equal(ip, np.zeros(11)) and len(ix) == 0
    B = PETSc.Mat().createSeqAIJ((1000, 1000), nnz=1)
    B.setValue(999, 0, 2.5)
    B.setValue(3, 998, -1.0)
    B.setValue(3, 2, 4.0, addv=True)
    B.assemble()
    ip, ix, va = _csr(B)
    assert ip[3] == 0 and ip[4] == 2 and ip[999] == 2 and ip[1000] == 3
    assert ix.tolist() == [2, 998, 0] and va.tolist() == [4.0, -1.0, 2.5]


@pytest.mark.parametrize("shape", [(1, 1), (3, 700), (257, 65), (1000, 1000)])
def test_device_dense_to_csr(native, monkeypatch, shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.normal(size=shape)
    a[rng.random(shape) < 0.7] = 0.0
    a[rng.random(shape) < 0.05] = -0.0
    a[rng.random(shape) < 0.01] = np.nan
    mats = []
    for threshold in ("1", str(10 ** 12)):
        monkeypatch.setenv("MPK_DENSE_DEVICE_MIN", threshold)
        mats.append(importlib.import_module("mojopetsc4py.PETSc.Mat").from_dense(a))
    (ip_d, ix_d, va_d), (ip_h, ix_h, va_h) = (_csr(M) for M in mats)
    assert np.array_equal(ip_d, ip_h) and np.array_equal(ix_d, ix_h)
    assert same_bits(va_d, va_h)
    rr, cc = np.nonzero(a)
    assert np.array_equal(ix_h, cc) and len(va_h) == len(rr)
