"""The product's host-side factor set up (libmpk's symbolic and numeric
routines; no GPU needed) against the oracle, entry for entry and bit for bit,
over levels of fill, orderings and matrices."""

import numpy as np
import pytest

from mpk import workloads as W


def a_(x):
    return x.ctypes.data


@pytest.fixture(scope="module")
def L():
    from mpk import _native
    return _native.load(init=False)


def product_factor(L, O, A, kind, levels, ordering):
    n = A.n
    r, ic = O.permutations(A, ordering)
    adiag = np.zeros(n, np.int64)
    O.lib().or_mat_get_diagonal_markers(a_(A.irow), a_(A.icol), a_(adiag), n)
    scratch = np.zeros(7 * n + 8, np.int64)
    cap = max(4 * n, 4 * int(A.irow[-1]) + 4 * n, 16)
    if kind == "ilu":
        while True:
            bi, bj = np.zeros(n + 1, np.int64), np.zeros(cap, np.int64)
            bjlvl, bdiag = np.zeros(cap, np.int64), np.zeros(n + 1, np.int64)
            if levels == 0 and ordering == "natural":
                nnz = L.mp_ilu_symbolic_ilu0(a_(bi), a_(bj), a_(bdiag), a_(adiag), a_(A.irow), a_(A.icol), n)
                cont = False
            else:
                bjsrc = np.zeros(cap, np.int64)
                nnz = L.mp_ilu_symbolic(a_(bi), a_(bjsrc), a_(bjlvl), a_(bdiag), a_(A.irow), a_(A.icol), a_(r),
                                        a_(ic), levels, cap, a_(scratch), n)
                cont = True
            if nnz >= 0:
                if cont:
                    L.mp_ilu_contiguous_lu(a_(bjsrc), a_(bi), a_(bdiag), a_(bj), n)
                break
            cap = -nnz + n
        ba, rtmp = np.zeros(max(int(bdiag[0]) + 1, 1)), np.zeros(n)
        L.mp_pc_ilu_numeric(a_(ba), a_(rtmp), a_(A.irow), a_(A.icol), a_(A.data), a_(bi), a_(bj), a_(bdiag),
                            a_(r), a_(ic), n)
        return bi, bj, bdiag, ba, int(bdiag[0]) + 1
    rinv = np.empty(n, np.int64)
    rinv[r] = np.arange(n)
    while True:
        ui, uj = np.zeros(n + 1, np.int64), np.zeros(cap, np.int64)
        ujlvl, udiag = np.zeros(cap, np.int64), np.zeros(n, np.int64)
        need = 6 * (n + 1) + 2 * cap
        if scratch.size < need:
            scratch = np.zeros(need, np.int64)
        nnz = L.mp_icc_symbolic(a_(ui), a_(uj), a_(ujlvl), a_(udiag), a_(A.irow), a_(A.icol), a_(r), a_(rinv),
                                levels, cap, a_(scratch), n, scratch.size)
        if nnz >= 0:
            break
        cap = -nnz
    bi, bdiag = np.zeros(n + 1, np.int64), np.zeros(n + 1, np.int64)
    bj = np.zeros(max(cap, int(ui[-1]) + 1), np.int64)
    bdiag[:n] = udiag
    L.mp_icc_contiguous_cholesky(a_(uj), a_(ui), a_(bdiag), a_(bi), a_(bdiag), a_(bj), n)
    ua, rtmp = np.zeros(max(int(ui[-1]), 1)), np.zeros(n)
    il, c2r = np.zeros(n, np.int64), np.zeros(n, np.int64)
    L.mp_pc_icc_numeric(a_(ua), a_(rtmp), a_(il), a_(c2r), a_(A.irow), a_(A.icol), a_(A.data), a_(bi), a_(bj),
                        a_(bdiag), a_(r), a_(rinv), n)
    return bi, bj, bdiag, ua, int(bi[-1])


MATS = {
    "lap12": lambda: W.laplacian_2d(12),
    "lap20": lambda: W.laplacian_2d(20),
    "tri30": lambda: W.tridiagonal(30),
    "diff6": lambda: W.diffusion_3d(6)[0],
    "cd5": lambda: W.convection_diffusion_3d(5)[0],
}


@pytest.mark.parametrize("name", sorted(MATS))
@pytest.mark.parametrize("kind", ["ilu", "icc"])
@pytest.mark.parametrize("levels", [0, 1, 2, 3])
@pytest.mark.parametrize("ordering", ["natural", "rcm"])
def test_factor_set_up_bitwise(L, oracle, name, kind, levels, ordering):
    A = oracle.CSR(*MATS[name]())
    f = oracle.build_factor(A, kind, levels, ordering)
    bi, bj, bdiag, ba, nnz = product_factor(L, oracle, A, kind, levels, ordering)
    assert np.array_equal(bi, f["bi"]) and np.array_equal(bdiag, f["bdiag"])
    assert np.array_equal(bj[:nnz], f["bj"][:nnz])
    assert np.array_equal(ba[:nnz].view(np.int64), f["ba"][:nnz].view(np.int64))


def test_rcm_bitwise(L, oracle):
    A = oracle.CSR(*W.convection_diffusion_3d(6)[0])
    n = A.n
    bufs = [np.zeros(n, np.int64) for _ in range(4)]
    L.mp_ordering_rcm(*[a_(b) for b in bufs], a_(A.irow), a_(A.icol), n)
    r, _ = oracle.permutations(A, "rcm")
    assert np.array_equal(bufs[0], r)
