"""Host-side Mat assembly (no device involved): the folded write log against
a per-entry dictionary, which is how the reference's Mat defines the result
of setValue / setValues (python/mojopetsc4py/PETSc/Mat.py:383-439): an
insert replaces the entry, an add accumulates onto its value so far (0.0 if
it has none), the entries present when writing starts are kept (the last of
duplicate columns wins), rows come out with columns ascending and every
entry ever written stays stored, zeros included."""

import numpy as np
import pytest

from mojopetsc4py import PETSc


class DictModel:
    def __init__(self, m, n):
        self.m, self.n, self.rows = m, n, {}

    def load_csr(self, ip, ix, va):
        self.rows = {}
        for i in range(self.m):
            for j in range(ip[i], ip[i + 1]):
                self.rows.setdefault(i, {})[int(ix[j])] = float(va[j])

    def write(self, r, c, v, add):
        cols = self.rows.setdefault(r, {})
        cols[c] = cols.get(c, 0.0) + v if add else v

    def csr(self):
        ip, ix, va = [0], [], []
        for i in range(self.m):
            cols = self.rows.get(i, {})
            for c in sorted(cols):
                ix.append(c)
                va.append(cols[c])
            ip.append(len(ix))
        return np.array(ip), np.array(ix, dtype=np.int64), np.array(va)


def _same(A, model):
    ip, ix, va = A.getValuesCSR()
    mi, mx, mv = model.csr()
    assert np.array_equal(ip, mi) and np.array_equal(ix, mx)
    assert np.array_equal(va.view(np.int64), mv.view(np.int64))


@pytest.mark.parametrize("seed", range(6))
def test_random_write_sequences(seed):
    rng = np.random.default_rng(seed)
    m, n = 7, 5
    A = PETSc.Mat().createSeqAIJ((m, n), nnz=3)
    model = DictModel(m, n)
    if seed % 2:  # start from installed CSR arrays, duplicates included
        ip = np.array([0, 2, 2, 5, 5, 6, 6, 8])
        ix = np.array([1, 3, 0, 0, 4, 2, 1, 1])
        va = rng.normal(size=8)
        A.setValuesCSR(ip, ix, va)
        model.load_csr(ip, ix, va)
    for step in range(40):
        kind = rng.integers(0, 3)
        add = bool(rng.integers(0, 2))
        if kind == 0:
            r, c, v = int(rng.integers(0, m)), int(rng.integers(0, n)), float(rng.normal())
            A.setValue(r, c, v, addv=True if add else None)
            model.write(r, c, v, add)
        elif kind == 1:
            rows = rng.choice(m, size=2, replace=False)
            cols = rng.choice(n, size=3, replace=False)
            vals = rng.normal(size=6) * 10.0 ** rng.integers(-8, 9, size=6)
            A.setValues(rows, cols, vals, addv=True if add else None)
            for i, r in enumerate(rows):
                for j, c in enumerate(cols):
                    model.write(int(r), int(c), float(vals[i * 3 + j]), add)
        else:
            r, c = int(rng.integers(0, m)), int(rng.integers(0, n))
            cols = model.rows.get(r, {})
            assert A.getValue(r, c) == cols.get(c, 0.0)
        if step % 13 == 12:
            _same(A, model)  # assemble part way; later writes start from these arrays
    _same(A, model)


def test_adds_fold_in_call_order():
    """Several adds to one entry sum left to right: 1e16 + 1 + 1 stays 1e16
    (each 1 rounds away), while 1 + 1 + 1e16 would not."""
    A = PETSc.Mat().createSeqAIJ((2, 2), nnz=1)
    for v in (1e16, 1.0, 1.0):
        A.setValue(0, 0, v, addv=True)
    for v in (1.0, 1.0, 1e16):
        A.setValue(1, 1, v, addv=True)
    assert A.getValue(0, 0) == 1e16
    assert A.getValue(1, 1) == 1e16 + 2.0
    _, _, va = A.getValuesCSR()
    assert va.tolist() == [1e16, 1e16 + 2.0]


def test_insert_after_add_and_zero_entries_stay():
    A = PETSc.Mat().createSeqAIJ((2, 3), nnz=2)
    A.setValue(0, 2, 5.0, addv=True)
    A.setValue(0, 2, 7.0)
    A.setValue(1, 0, 0.0)
    ip, ix, va = A.getValuesCSR()
    assert ip.tolist() == [0, 1, 2] and ix.tolist() == [2, 0] and va.tolist() == [7.0, 0.0]
    assert A.getNNZ() == 2


def test_created_matrix_is_empty_and_room_counts():
    A = PETSc.Mat().createSeqAIJ((4, 4), nnz=3)
    assert A.getNNZ() == 0 and A.getInfo()["nz_allocated"] == 1
    B = PETSc.Mat().createSeqAIJ((4, 4), nnz=3)
    B.setValuesCSR([0, 1, 2, 3, 4], [0, 1, 2, 3], [1.0, 2.0, 3.0, 4.0])
    assert B.getInfo() == {"nz_allocated": 12, "nz_used": 4, "rcond": 0.0}
    assert len(B.icol) == 12 and B.icol.dtype == np.int64 and B.irow.dtype == np.int64


def test_int32_csr_is_kept_narrow():
    ip = np.array([0, 2, 3], dtype=np.int32)
    ix = np.array([0, 1, 1], dtype=np.int32)
    A = PETSc.Mat().createSeqAIJ((2, 2), nnz=2)
    A.setValuesCSR(ip, ix, np.array([1.0, 2.0, 3.0]))
    h = A._host()
    assert h.indptr.dtype == np.int32 and h.indices.dtype == np.int32
    ix[0] = 1  # the matrix holds its own copy
    assert A.getValue(0, 0) == 1.0
    assert A.irow.dtype == np.int64 and A.icol.tolist()[:3] == [0, 1, 1]


def test_out_of_range_keeps_earlier_writes():
    A = PETSc.Mat().createSeqAIJ((2, 2), nnz=2)
    with pytest.raises(ValueError):
        A.setValues([0, 5], [0, 1], [1.0, 2.0, 3.0, 4.0])
    assert A.getValue(0, 0) == 1.0 and A.getValue(0, 1) == 2.0
    with pytest.raises(ValueError):
        A.setValues([0], [0, 1], [1.0])  # fewer values than entries


def test_copy_with_different_pattern():
    """The target keeps its pattern and takes the source's values there,
    0.0 where the source has none (MatCopy, DIFFERENT_NONZERO_PATTERN)."""
    S = PETSc.Mat().createSeqAIJ((3, 3), nnz=2)
    S.setValuesCSR([0, 2, 3, 4], [0, 2, 1, 0], [1.0, 2.0, 3.0, 4.0])
    T = PETSc.Mat().createSeqAIJ((3, 3), nnz=2)
    T.setValuesCSR([0, 1, 3, 4], [2, 0, 1, 2], [9.0, 9.0, 9.0, 9.0])
    S.copy(T, PETSc.Mat.Structure.DIFFERENT_NONZERO_PATTERN)
    ip, ix, va = T.getValuesCSR()
    assert ix.tolist() == [2, 0, 1, 2] and va.tolist() == [2.0, 0.0, 3.0, 0.0]
    with pytest.raises(ValueError):
        S.copy(PETSc.Mat().createSeqAIJ((2, 3)), PETSc.Mat.Structure.DIFFERENT_NONZERO_PATTERN)


def test_dense_symmetry_transpose_and_products():
    rng = np.random.default_rng(4)
    D = rng.normal(size=(5, 4)) * (rng.random((5, 4)) < 0.5)
    A = PETSc.Mat().createSeqAIJ((5, 4), nnz=4)
    for i, j in zip(*np.nonzero(D)):
        A.setValue(int(i), int(j), float(D[i, j]))
    assert np.array_equal(A._dense(), D)
    T = A.transpose()
    assert T.getSize() == (4, 5) and np.array_equal(T._dense(), D.T)
    assert T.isTranspose(A) and A.isTranspose(T) and not A.isSymmetric()
    S = PETSc.Mat().createSeqAIJ((3, 3), nnz=3)
    S.setValues([0, 1, 2], [0, 1, 2], np.array([[2.0, 1.0, 0.0], [1.0, 3.0, 0.5], [0.0, 0.5, 1.0]]))
    assert S.isSymmetric() and S.isTranspose(S)
    P = A * T
    assert np.allclose(P._dense(), D @ D.T)
    N = -A
    assert np.array_equal(N._dense(), -D)
    assert A.equal(A.duplicate(copy=True)) and not A.equal(N)
