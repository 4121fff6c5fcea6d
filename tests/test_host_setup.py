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


@pytest.mark.parametrize("name", sorted(MATS))
@pytest.mark.parametrize("levels,ordering", [(0, "natural"), (2, "natural"), (1, "rcm")])
def test_ilu_levels(L, oracle, name, levels, ordering):
    """mpk_ilu_levels: every row's level is one past its deepest L pivot's
    (brute force here), rows ascending inside a level, all rows once."""
    A = oracle.CSR(*MATS[name]())
    n = A.n
    f = oracle.build_factor(A, "ilu", levels, ordering)
    bi, bj = f["bi"], f["bj"]
    want = np.zeros(n, np.int64)
    for i in range(n):
        piv = bj[bi[i]:bi[i + 1]]
        assert np.all(piv < i)
        want[i] = 1 + want[piv].max() if len(piv) else 0
    lrow = np.zeros(n, np.int32)
    lptr = np.zeros(n + 1, np.int64)
    nlev = L.mpk_ilu_levels(a_(bi), a_(bj), a_(f["bdiag"]), n, a_(lrow), a_(lptr))
    assert nlev == want.max() + 1 and lptr[0] == 0 and lptr[nlev] == n
    for lev in range(nlev):
        rows = lrow[lptr[lev]:lptr[lev + 1]]
        assert np.all(np.diff(rows) > 0) and np.all(want[rows] == lev)
    assert np.array_equal(np.sort(lrow), np.arange(n))


def _random_nonsym(n, per_row, seed):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n):
        cs = set(rng.integers(0, n, per_row).tolist()) | {i}
        for c in sorted(cs):
            rows.append(i)
            cols.append(c)
    vals = rng.normal(size=len(rows))
    return W._csr_from_coo(n, np.array(rows), np.array(cols), vals)


def test_ilu_levels_refuses_rows_without_a_diagonal_slot(L, oracle):
    """Level-of-fill on this random nonsymmetric pattern leaves rows that
    share a later row's diagonal slot; the level listing refuses the layout
    (the serial host pass factors it), and accepts the ILU(0) one."""
    A = oracle.CSR(*_random_nonsym(300, 14, 11))
    n = A.n
    lrow, lptr = np.zeros(n, np.int32), np.zeros(n + 1, np.int64)
    f = oracle.build_factor(A, "ilu", 1, "natural")
    bd = f["bdiag"]
    assert any(bd[i] <= bd[i + 1] for i in range(n))
    assert L.mpk_ilu_levels(a_(f["bi"]), a_(f["bj"]), a_(bd), n, a_(lrow), a_(lptr)) == -2
    f0 = oracle.build_factor(A, "ilu", 0, "natural")
    assert L.mpk_ilu_levels(a_(f0["bi"]), a_(f0["bj"]), a_(f0["bdiag"]), n, a_(lrow), a_(lptr)) > 0
    # plant a mistake: two L columns of a row out of order
    bj = f0["bj"].copy()
    i = next(i for i in range(n) if f0["bi"][i + 1] - f0["bi"][i] >= 2)
    j = int(f0["bi"][i])
    bj[j], bj[j + 1] = bj[j + 1], bj[j]
    assert L.mpk_ilu_levels(a_(f0["bi"]), a_(bj), a_(f0["bdiag"]), n, a_(lrow), a_(lptr)) == -2


def _missing_diagonals(n, seed):
    """A nonsymmetric pattern where every fifth row has no diagonal entry."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n):
        cs = set(rng.integers(0, n, 6).tolist())
        if i % 5 == 0:
            cs.discard(i)
        else:
            cs.add(i)
        for c in sorted(cs):
            rows.append(i)
            cols.append(c)
    return W._csr_from_coo(n, np.array(rows), np.array(cols), rng.normal(size=len(rows)) + 3.0)


ODD = {
    # level-of-fill outgrows the first capacity guess: the retry runs on the
    # scratch the failed attempt left behind (PC.py:393-428), and some rows
    # come out without their diagonal
    "nonsym300": lambda: _random_nonsym(300, 14, 11),
    "nodiag120": lambda: _missing_diagonals(120, 4),
}


@pytest.mark.parametrize("name", sorted(ODD))
@pytest.mark.parametrize("levels,ordering", [(0, "natural"), (1, "natural"), (2, "natural"), (0, "rcm"),
                                             (1, "rcm")])
def test_ilu_set_up_on_degenerate_patterns(L, oracle, name, levels, ordering):
    """Rows without a diagonal (in A, or left so by a retried symbolic pass)
    give rows whose diagonal slot is the next row's; the product's host set up
    still follows the reference's copy loops (which copy nothing for a
    negative count) and matches the oracle entry for entry and bit for bit."""
    A = oracle.CSR(*ODD[name]())
    f = oracle.build_factor(A, "ilu", levels, ordering)
    bi, bj, bdiag, ba, nnz = product_factor(L, oracle, A, "ilu", levels, ordering)
    assert np.array_equal(bi, f["bi"]) and np.array_equal(bdiag, f["bdiag"])
    assert np.array_equal(bj[:nnz], f["bj"][:nnz])
    assert np.array_equal(ba[:nnz].view(np.int64), f["ba"][:nnz].view(np.int64))


def test_degenerate_patterns_have_shared_diagonal_slots(oracle):
    """The cases above do reach the degenerate layout."""
    for name, lv in (("nonsym300", 1), ("nodiag120", 0)):
        f = oracle.build_factor(oracle.CSR(*ODD[name]()), "ilu", lv, "natural")
        bd = f["bdiag"]
        assert any(bd[i] <= bd[i + 1] for i in range(len(bd) - 1)), name


def _fma(a, b, c):
    """fma(a, b, c) rounded once (exact product and sum, then to double)."""
    from fractions import Fraction
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _icc_numeric_levels(A, f, vptr, vrow, vpos, lrow, lptr, nlev):
    """The device IC numeric pass (k_icc_numeric) restated, level by level."""
    n = A.n
    bi, bj, r, rinv = f["bi"], f["bj"], f["r"], f["rinv"]
    ua = np.zeros(max(int(bi[n]), 1))
    u0 = np.zeros_like(ua)

    def cell(k, c):
        s0, dg = bi[k], bi[k + 1] - 1
        hit = np.nonzero(bj[s0:dg] == c)[0]
        if len(hit):
            return s0 + hit[0]
        if c == bj[dg]:
            return dg
        return -2 if c == k else -1

    for lev in range(nlev):
        for k in lrow[lptr[lev]:lptr[lev + 1]]:
            s0, dg = bi[k], bi[k + 1] - 1
            ua[s0:dg + 1] = 0.0
            wk = 0.0
            ra = r[k]
            for j in range(A.irow[ra], A.irow[ra + 1]):
                c = rinv[A.icol[j]]
                if c >= k:
                    q = cell(k, c)
                    if q >= 0:
                        ua[q] = A.data[j]
                    elif q == -2:
                        wk = A.data[j]
            qk = cell(k, k)
            dk = ua[qk] if qk >= 0 else wk
            for v in range(vptr[k], vptr[k + 1]):
                i, fp = vrow[v], vpos[v]
                uf = u0[fp]
                di = bi[i + 1] - 1
                uikdi = -uf * u0[di]
                dk = _fma(uikdi, uf, dk)
                ua[fp] = uikdi
                for j in range(fp + 1, di + 1):
                    q = cell(k, bj[j])
                    if q >= 0:
                        ua[q] = _fma(uikdi, u0[j], ua[q])
            u0[s0:dg] = ua[s0:dg]
            u0[dg] = ua[dg] = 1.0 / dk
    return ua


ICC_MATS = {
    "lap12": lambda: W.laplacian_2d(12),
    "tri30": lambda: W.tridiagonal(30),
    "diff5": lambda: W.diffusion_3d(5)[0],
    "diff7": lambda: W.diffusion_3d(7)[0],
    "lap20": lambda: W.laplacian_2d(20),
    "cd4sym": lambda: _symmetrized(W.convection_diffusion_3d(4)[0]),
}


def _symmetrized(csr):
    """A + A^T of a nonsymmetric operator (an SPD-shaped pattern with uneven
    values)."""
    import scipy.sparse as sp
    ip, ix, va = csr
    n = len(ip) - 1
    M = sp.csr_matrix((va, ix, ip), shape=(n, n))
    S = (M + M.T).tocsr()
    S.sort_indices()
    return S.indptr.astype(np.int64), S.indices.astype(np.int64), S.data.astype(np.float64)


@pytest.mark.parametrize("name", sorted(ICC_MATS))
@pytest.mark.parametrize("levels,ordering", [(0, "natural"), (1, "natural"), (2, "natural"), (0, "rcm"),
                                             (2, "rcm")])
def test_icc_visits_reproduce_the_oracle(L, oracle, name, levels, ordering):
    """mpk_icc_visits replays the up-looking pass's row chains: the device
    pass restated over its visits and level sets gives the oracle's IC factor
    (MatICCFactorNumeric_SeqAIJ restated) bit for bit, and every row's
    updaters sit in earlier levels."""
    A = oracle.CSR(*ICC_MATS[name]())
    n = A.n
    f = oracle.build_factor(A, "icc", levels, ordering)
    bi, bj = f["bi"], f["bj"]
    m = max(int(bi[n]), 1)
    vptr, vrow, vpos = np.zeros(n + 1, np.int32), np.zeros(m, np.int32), np.zeros(m, np.int32)
    lrow, lptr = np.zeros(n, np.int32), np.zeros(n + 1, np.int64)
    nlev = L.mpk_icc_visits(a_(bi), a_(bj), a_(A.irow), a_(A.icol), a_(f["r"]), a_(f["rinv"]), n, a_(vptr),
                            a_(vrow), a_(vpos), a_(lrow), a_(lptr))
    assert nlev > 0
    lev = np.zeros(n, np.int64)
    for l in range(nlev):
        lev[lrow[lptr[l]:lptr[l + 1]]] = l
    seen = set()
    for k in range(n):
        for v in range(vptr[k], vptr[k + 1]):
            # the updater is finished first; no entry is visited twice (a
            # visit scales its entry).  With columns below the diagonal the
            # chains can hand a row over at another column's step.
            assert lev[vrow[v]] < lev[k] and vpos[v] not in seen
            seen.add(vpos[v])
    got = _icc_numeric_levels(A, f, vptr, vrow, vpos, lrow, lptr, nlev)
    nnz = int(bi[n])
    assert np.array_equal(got[:nnz].view(np.int64), f["ba"][:nnz].view(np.int64))
