"""The rule the device IC(0) analysis sorts each column's visits by, checked
on the CPU against the product's replay of the up-looking pass's row chains
(mpk_icc_visits, which follows MatICCFactorNumeric_SeqAIJ,
ported.mojo:1040-1124), and the IC(0) pattern against mp_icc_symbolic at
level 0, on symmetric and nonsymmetric patterns, with and without diagonals.

The rule: entry (i, k) of the factor's strict part gets the key "row i's
columns before k, last first, then i, then an end mark"; column k's visits
come in key order, comparing element by element, larger first at even
depths and smaller first at odd ones, the end mark last at even depths and
first at odd ones."""

import functools

import numpy as np
import pytest

from mpk import _native
from mpk import workloads as W


@pytest.fixture(scope="module")
def L():
    return _native.load(False)


def _a(x):
    return x.ctypes.data


def _ic0_pattern(ip, ix, n):
    bi = np.zeros(n + 1, np.int64)
    bj = []
    for k in range(n):
        bj += sorted(int(c) for c in ix[ip[k]:ip[k + 1]] if c > k) + [k]
        bi[k + 1] = len(bj)
    return bi, np.array(bj, np.int64)


def _visits_by_rule(bi, bj, n):
    cols = [[] for _ in range(n)]
    for i in range(n):
        for p in range(bi[i], bi[i + 1] - 1):
            cols[bj[p]].append((i, p))

    def elem(i, p, d):
        q = p - 1 - d
        if q >= bi[i]:
            return int(bj[q])
        return i if q == bi[i] - 1 else None

    def order(a, b):
        d = 0
        while True:
            ea, eb = elem(*a, d), elem(*b, d)
            if ea != eb:
                if d % 2 == 0:
                    first = eb is None or (ea is not None and ea > eb)
                else:
                    first = ea is None or (eb is not None and ea < eb)
                return -1 if first else 1
            d += 1

    rows, pos = [], []
    for k in range(n):
        for i, p in sorted(cols[k], key=functools.cmp_to_key(order)):
            rows.append(i)
            pos.append(p)
    return np.array(rows, np.int64), np.array(pos, np.int64)


def _patterns():
    out = [("lap", W.laplacian_2d(9)), ("diff", W.diffusion_3d(5)[0]), ("tri", W.tridiagonal(17)),
           ("cd", W.convection_diffusion_3d(4)[0])]
    rng = np.random.default_rng(11)
    for s in range(24):
        n = int(rng.integers(4, 60))
        M = rng.random((n, n)) < rng.uniform(0.03, 0.4)
        if s % 2 == 0:
            M |= M.T
        np.fill_diagonal(M, s % 3 != 0)
        ip = np.concatenate([[0], np.cumsum(M.sum(1))]).astype(np.int64)
        ix = np.nonzero(M)[1].astype(np.int64)
        out.append((f"rand{s}", (ip, ix, np.ones(len(ix)))))
    return out


@pytest.mark.parametrize("name,csr", _patterns(), ids=[p[0] for p in _patterns()])
def test_visits_follow_the_rule(L, name, csr):
    ip, ix = np.asarray(csr[0], np.int64), np.asarray(csr[1], np.int64)
    n = len(ip) - 1
    bi, bj = _ic0_pattern(ip, ix, n)
    cap = int(ip[-1]) + n
    ui, uj, ud, lv = (np.zeros(k, np.int64) for k in (n + 1, cap, n, cap))
    sc = np.zeros(6 * (n + 1) + 2 * cap, np.int64)
    r = np.arange(n, dtype=np.int64)
    got = L.mp_icc_symbolic(_a(ui), _a(uj), _a(lv), _a(ud), _a(ip), _a(ix), _a(r), _a(r), 0, cap, _a(sc), n,
                            int(sc.size))
    assert got == bi[n]
    assert np.array_equal(ui, bi) and np.array_equal(uj[:got], bj) and np.array_equal(ud, bi[1:] - 1)
    m = int(bi[n])
    vptr, vrow, vpos, lrow = (np.zeros(k, np.int32) for k in (n + 1, m, m, n))
    lptr = np.zeros(n + 1, np.int64)
    nlev = L.mpk_icc_visits(_a(bi), _a(bj), _a(ip), _a(ix), _a(r), _a(r), n, _a(vptr), _a(vrow), _a(vpos),
                            _a(lrow), _a(lptr))
    assert nlev > 0
    rows, pos = _visits_by_rule(bi, bj, n)
    nv = int(vptr[n])
    assert np.array_equal(vrow[:nv], rows) and np.array_equal(vpos[:nv], pos)
    # vptr: each column's count of strict entries
    assert np.array_equal(np.diff(vptr), np.bincount(bj[np.concatenate(
        [np.arange(bi[i], bi[i + 1] - 1) for i in range(n)]).astype(np.int64)], minlength=n))
