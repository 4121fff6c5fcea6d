"""Pins for the oracle's kernels that do not rely on the oracle's own
formulas: exact rational arithmetic (fractions) for the elementwise fused
multiply-adds, math.fsum for reductions, dense numpy for products and
set ups, the defining property of an incomplete factorisation, brute-force
level of fill for the symbolic phase, and closed-form solves."""

import math
from fractions import Fraction as Fr

import numpy as np
import pytest

from mpk import workloads as W


def fma_exact(a, b, c):
    """The correctly rounded a*b + c."""
    return float(Fr(a) * Fr(b) + Fr(c))


def a_(x):
    return x.ctypes.data


@pytest.fixture(params=["ref", "dev"])
def O(oracle, request):
    oracle.set_mode(request.param)
    yield oracle
    oracle.set_mode("ref")


def test_elementwise_fma_placement(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(0)
    n = 37
    x, y, z = rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)
    al, be, ga = 0.3711, -1.237, 0.913
    h = y.copy()
    L.or_vec_axpy(a_(h), al, a_(x), n)
    assert all(h[i] == fma_exact(al, x[i], y[i]) for i in range(n))
    h = y.copy()
    L.or_vec_aypx(a_(h), al, a_(x), n)
    assert all(h[i] == fma_exact(al, y[i], x[i]) for i in range(n))
    h = y.copy()
    L.or_vec_axpby(a_(h), al, be, a_(x), n)
    assert all(h[i] == fma_exact(al, x[i], be * y[i]) for i in range(n))
    h = z.copy()
    L.or_vec_axpbypcz(a_(h), al, be, ga, a_(x), a_(y), n)
    assert all(h[i] == fma_exact(ga, z[i], fma_exact(al, x[i], be * y[i])) for i in range(n))
    w = np.zeros(n)
    L.or_vec_waxpy(a_(w), al, a_(x), a_(y), n)
    assert all(w[i] == fma_exact(al, x[i], y[i]) for i in range(n))


def test_aypx_axpby_special_branches(oracle):
    """ported.mojo:285-308: alpha 0 copies, alpha 1 is axpy; beta 0 scales x."""
    L = oracle.lib()
    x = np.array([1.5, -2.0, np.inf])
    y = np.array([np.nan, 3.0, 1.0])
    h = y.copy()
    L.or_vec_aypx(a_(h), 0.0, a_(x), 3)
    assert np.array_equal(h, x)
    h = y.copy()
    L.or_vec_axpby(a_(h), 2.0, 0.0, a_(x), 3)
    assert np.array_equal(h, 2.0 * x)


def test_spmv_row_order_fma(oracle):
    ip, ix, va = W.diffusion_3d(4)[0]
    A = oracle.CSR(ip, ix, va)
    x = np.random.default_rng(1).normal(size=A.n)
    y = A.mult(x)
    for i in range(A.n):
        s = 0.0
        for j in range(ip[i], ip[i + 1]):
            s = fma_exact(va[j], x[ix[j]], s)
        assert y[i] == s
    np.testing.assert_allclose(y, A.dense() @ x, rtol=1e-13, atol=1e-13)


def test_transpose_and_diagonal(oracle):
    L = oracle.lib()
    ip, ix, va = W.convection_diffusion_3d(4)[0]
    A = oracle.CSR(ip, ix, va)
    x = np.random.default_rng(2).normal(size=A.n)
    y = np.zeros(A.n)
    L.or_mat_mult_transpose(a_(x), a_(y), a_(A.irow), a_(A.icol), a_(A.data), A.n, A.n)
    np.testing.assert_allclose(y, A.dense().T @ x, rtol=1e-13, atol=1e-13)
    d = np.zeros(A.n)
    L.or_mat_get_diagonal(a_(d), a_(A.irow), a_(A.icol), a_(A.data), A.n)
    assert np.array_equal(d, np.diag(A.dense()))


LENGTHS = [0, 1, 3, 4, 5, 15, 16, 17, 31, 33, 63, 64, 65, 2047, 2048, 2049, 4096, 4097, 16385]


@pytest.mark.parametrize("n", LENGTHS)
def test_reductions_against_fsum(O, n):
    L = O.lib()
    rng = np.random.default_rng(n)
    x, y = rng.normal(size=n), rng.normal(size=n)
    exact_dot = math.fsum(float(Fr(p) * Fr(q)) for p, q in zip(x, y)) if n else 0.0
    scale = max(1.0, math.fsum(abs(p * q) for p, q in zip(x, y)))
    assert abs(L.or_red_dot(a_(x), a_(y), n) - exact_dot) <= 1e-13 * scale
    assert abs(L.or_red_total(a_(x), n) - math.fsum(x)) <= 1e-13 * max(1.0, math.fsum(abs(x)))
    assert abs(L.or_red_asum(a_(x), n) - math.fsum(abs(x))) <= 1e-13 * max(1.0, math.fsum(abs(x)))
    assert abs(L.or_red_sumsq(a_(x), n) - math.fsum(x * x)) <= 1e-13 * max(1.0, math.fsum(x * x))
    z = np.zeros(2)
    L.or_vec_norm(a_(x), 3, a_(z), n)
    assert z[0] == (np.max(np.abs(x)) if n else 0.0)


def test_norm_infinity_nan_rule(oracle):
    x = np.array([1.0, np.nan, 5.0])
    z = np.zeros(2)
    oracle.lib().or_vec_norm(a_(x), 3, a_(z), 3)
    assert np.isnan(z[0])


def test_minmax_first_index(oracle):
    L = oracle.lib()
    x = np.array([2.0, 7.0, -1.0, 7.0, -1.0])
    i, v = np.zeros(1, np.int64), np.zeros(1)
    L.or_vec_max(a_(x), a_(i), a_(v), 5)
    assert (i[0], v[0]) == (1, 7.0)
    L.or_vec_min(a_(x), a_(i), a_(v), 5)
    assert (i[0], v[0]) == (2, -1.0)


def test_reciprocal_guard(oracle):
    x = np.array([0.0, 1e-20, -4.0, 5e-324])
    oracle.lib().or_vec_reciprocal(a_(x), 4)
    assert x[0] == 0.0 and x[1] == 1e20 and x[2] == -0.25 and np.isinf(x[3])


@pytest.mark.parametrize("jtype", [0, 1, 2, 3])
def test_jacobi_closed_forms(oracle, jtype):
    L = oracle.lib()
    ip, ix, va = W.convection_diffusion_3d(3)[0]
    A = oracle.CSR(ip, ix, va)
    D = A.dense()
    n = A.n
    fill = [np.diag(D), np.max(np.abs(D), axis=1), np.sum(np.abs(D), axis=1), np.sum(D, axis=1)][jtype]
    d, dx = np.zeros(n), np.zeros(n)
    L.or_pc_jacobi_setup(a_(d), jtype, 0, 1, 1.0, a_(A.irow), a_(A.icol), a_(A.data), n, a_(dx))
    want = np.where(fill == 0, 1.0, 1.0 / np.where(fill == 0, 1.0, fill))
    np.testing.assert_allclose(d, want, rtol=1e-14)
    ds = np.zeros(n)
    L.or_pc_jacobi_setup_symmetric(a_(ds), jtype, a_(A.irow), a_(A.icol), a_(A.data), n)
    np.testing.assert_allclose(ds, np.where(fill == 0, 1.0, 1.0 / np.sqrt(np.abs(np.where(fill == 0, 1, fill)))),
                               rtol=1e-14)


def _sor_numpy(D, b, omega, symmetric):
    """Symmetric (or forward) SOR with a zero initial guess, written from the
    textbook sweeps."""
    n = len(b)
    y = np.zeros(n)
    t = np.zeros(n)
    for i in range(n):
        t[i] = b[i] - D[i, :i] @ y[:i]
        y[i] = omega * t[i] / D[i, i]
    if not symmetric:
        return y
    for i in range(n - 1, -1, -1):
        s = t[i] - D[i, i + 1:] @ y[i + 1:]
        y[i] = (1 - omega) * y[i] + omega * s / D[i, i]
    return y


@pytest.mark.parametrize("omega,symmetric", [(1.0, True), (1.3, True), (0.7, False)])
def test_sor_against_textbook(oracle, omega, symmetric):
    ip, ix, va = W.convection_diffusion_3d(4)[0]
    A = oracle.CSR(ip, ix, va)
    b = np.random.default_rng(3).normal(size=A.n)
    pc = oracle.PCSetup(A, "sor", omega=omega, symmetric=symmetric)
    np.testing.assert_allclose(pc.apply(b), _sor_numpy(A.dense(), b, omega, symmetric), rtol=1e-12, atol=1e-13)


def _ilu_pattern_bruteforce(D, levels):
    """Textbook level of fill, IKJ: lev(i,j) = min over k < min(i,j) of
    lev(i,k) + lev(k,j) + 1, keeping entries with lev <= levels."""
    n = D.shape[0]
    lev = np.where(D != 0, 0, 10 ** 9)
    for i in range(n):
        for k in range(i):
            if lev[i, k] <= levels:
                for j in range(k + 1, n):
                    if lev[k, j] <= levels:
                        lev[i, j] = min(lev[i, j], lev[i, k] + lev[k, j] + 1)
    return lev <= levels


def _ilu_pattern_oracle(f, n):
    P = np.zeros((n, n), bool)
    bi, bj, bd = f["bi"], f["bj"], f["bdiag"]
    for i in range(n):
        for j in range(bi[i], bi[i + 1]):
            P[i, bj[j]] = True
        for j in range(bd[i + 1] + 1, bd[i] + 1):
            P[i, bj[j]] = True
    return P


@pytest.mark.parametrize("name,levels", [("lap8", 0), ("lap8", 1), ("lap8", 2), ("cd3", 1), ("cd3", 2),
                                         ("tri20", 1)])
def test_ilu_pattern_is_level_of_fill(oracle, name, levels):
    csr = {"lap8": lambda: W.laplacian_2d(8), "cd3": lambda: W.convection_diffusion_3d(3)[0],
           "tri20": lambda: W.tridiagonal(20)}[name]()
    A = oracle.CSR(*csr)
    f = oracle.symbolic(A, "ilu", levels)
    assert np.array_equal(_ilu_pattern_oracle(f, A.n), _ilu_pattern_bruteforce(A.dense(), levels))


@pytest.mark.parametrize("name,levels", [("lap8", 0), ("lap8", 1), ("cd3", 0), ("cd3", 2), ("diff3", 1)])
def test_ilu_numeric_defining_property(oracle, name, levels):
    """(L U)_ij = a_ij on the factor's pattern (the definition of ILU)."""
    csr = {"lap8": lambda: W.laplacian_2d(8), "cd3": lambda: W.convection_diffusion_3d(3)[0],
           "diff3": lambda: W.diffusion_3d(3)[0]}[name]()
    A = oracle.CSR(*csr)
    n = A.n
    f = oracle.build_factor(A, "ilu", levels)
    Lm, Um = np.eye(n), np.zeros((n, n))
    bi, bj, bd, ba = f["bi"], f["bj"], f["bdiag"], f["ba"]
    for i in range(n):
        for j in range(bi[i], bi[i + 1]):
            Lm[i, bj[j]] = ba[j]
        for j in range(bd[i + 1] + 1, bd[i]):
            Um[i, bj[j]] = ba[j]
        Um[i, i] = 1.0 / ba[bd[i]]
    P = _ilu_pattern_oracle(f, n)
    np.testing.assert_allclose((Lm @ Um)[P], A.dense()[P], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name,levels", [("lap8", 0), ("tri20", 0), ("tri20", 2), ("diff3", 0)])
def test_icc_numeric_defining_property(oracle, name, levels):
    """U^T D U = A on the upper pattern of A (IC(0) on a symmetric matrix)."""
    csr = {"lap8": lambda: W.laplacian_2d(8), "tri20": lambda: W.tridiagonal(20),
           "diff3": lambda: W.diffusion_3d(3)[0]}[name]()
    A = oracle.CSR(*csr)
    n = A.n
    f = oracle.build_factor(A, "icc", levels)
    bi, bj, bd, ba = f["bi"], f["bj"], f["bdiag"], f["ba"]
    U = np.eye(n)
    d = np.zeros(n)
    Pt = np.zeros((n, n), bool)
    for i in range(n):
        for j in range(bi[i], bd[i]):
            U[i, bj[j]] = -ba[j]  # the stored slots hold -U_ij
            Pt[i, bj[j]] = True
        d[i] = 1.0 / ba[bd[i]]
        Pt[i, i] = True
    M = U.T @ np.diag(d) @ U
    np.testing.assert_allclose(M[Pt], A.dense()[Pt], rtol=1e-12, atol=1e-12)


def test_factor_exact_on_tridiagonal(oracle):
    A = oracle.CSR(*W.tridiagonal(50))
    x = np.random.default_rng(4).normal(size=50)
    for kind in ("ilu", "icc"):
        pc = oracle.PCSetup(A, kind, 0)
        np.testing.assert_allclose(pc.apply(A.mult(x)), x, rtol=1e-12, atol=1e-13)


def test_icc_layout_invariants(oracle):
    """test_factor_layout.py:82-156 of the reference: strict columns
    ascending, above the diagonal at level 0 (the level-of-fill merges can add
    columns on either side, as that test notes), the diagonal last, pattern
    containing the upper triangle of A."""
    for k, levels in ((4, 0), (6, 1), (8, 2)):
        A = oracle.CSR(*W.laplacian_2d(k))
        f = oracle.symbolic(A, "icc", levels)
        bi, bj, bd = f["bi"], f["bj"], f["bdiag"]
        D = A.dense()
        for i in range(A.n):
            strict = bj[bi[i]:bd[i]]
            assert bd[i] == bi[i + 1] - 1 and bj[bd[i]] == i
            assert np.all(np.diff(strict) > 0)
            if levels == 0:
                assert np.all(strict > i)
            upper = set(np.nonzero(D[i, i + 1:])[0] + i + 1)
            assert upper <= set(strict.tolist())


def test_rcm_is_permutation(oracle):
    A = oracle.CSR(*W.laplacian_2d(9))
    r, ic = oracle.permutations(A, "rcm")
    assert sorted(r.tolist()) == list(range(A.n))
    assert np.array_equal(ic[r], np.arange(A.n))


@pytest.mark.parametrize("ksp_type,pc_type", [("cg", "jacobi"), ("cg", "icc"), ("gmres", "ilu"), ("fgmres", "sor"),
                                              ("bcgs", "ilu"), ("gmres", "none")])
def test_solves_against_numpy(O, ksp_type, pc_type):
    symmetric = ksp_type == "cg"
    csr, b = (W.diffusion_3d(6) if symmetric else W.convection_diffusion_3d(6))
    A = O.CSR(*csr)
    r = O.ksp_solve(A, b, ksp_type, O.PCSetup(A, pc_type, 0), rtol=1e-12, max_it=2000)
    assert r.reason > 0
    np.testing.assert_allclose(r.x, np.linalg.solve(A.dense(), b), rtol=1e-8, atol=1e-10)


def test_failure_reasons(oracle):
    """The host-forced failures: NaN right-hand side -> -9 at its 0, an
    indefinite matrix in CG -> -8 or -10, max_it reached -> +4 (CONVERGED_ITS,
    KSP.py:223-225), tiny divtol -> -4, zero pivot in ILU -> -9."""
    O = oracle
    csr, b = W.diffusion_3d(5)
    A = O.CSR(*csr)
    bn = b.copy()
    bn[0] = np.nan
    r = O.ksp_solve(A, bn, "cg", O.PCSetup(A, "jacobi"), rtol=1e-10, max_it=50)
    assert r.reason == -9 and r.its == 0
    Ai = O.CSR(np.array([0, 1, 2]), np.array([0, 1]), np.array([1.0, -1.0]))
    r = O.ksp_solve(Ai, np.array([1.0, 1.0]), "cg", O.PCSetup(Ai, "none"), rtol=1e-10, max_it=50)
    assert r.reason in (-8, -10)
    r = O.ksp_solve(A, b, "cg", O.PCSetup(A, "jacobi"), rtol=1e-14, max_it=1)
    assert r.reason == 4 and r.its == 1
    r = O.ksp_solve(A, b, "gmres", O.PCSetup(A, "none"), rtol=1e-10, divtol=1e-30, max_it=50)
    assert r.reason == -4
    Az = O.CSR(np.array([0, 2, 4]), np.array([0, 1, 0, 1]), np.array([0.0, 1.0, 1.0, 1.0]))
    r = O.ksp_solve(Az, np.array([1.0, 2.0]), "gmres", O.PCSetup(Az, "ilu", 0), rtol=1e-10, max_it=10)
    assert r.reason == -9


def test_convergence_test_branches(oracle):
    """KSPConvergedDefault (ksp.mojo:96-152) on hand-made inputs."""
    L = oracle.lib()

    def run(n, rnorm, cfg_over=None, cfgi_over=None, seed=True):
        cfg = np.zeros(32)
        cfg[0], cfg[1], cfg[2] = 1e-5, 1e-50, 1e4
        cfgi = np.zeros(16, np.int64)
        cfgi[0], cfgi[4], cfgi[6] = 2, -1, 100
        for k, v in (cfg_over or {}).items():
            cfg[k] = v
        for k, v in (cfgi_over or {}).items():
            cfgi[k] = v
        if seed:
            L.or_ksp_converged_default(0, 10.0, a_(cfg), a_(cfgi))
        return L.or_ksp_converged_default(n, rnorm, a_(cfg), a_(cfgi))

    assert run(3, 1.0) == 0
    assert run(3, 1e-5) == 2                       # rnorm <= rtol * rnorm0
    assert run(3, 1e-60, {1: 1e-55}) == 3          # below atol
    assert run(3, 1e6) == -4                       # divergence
    assert run(3, float("nan")) == -9
    assert run(3, float("inf")) == -4              # MAX_REAL test misses inf (ksp.mojo:85,138)
    assert run(3, 1.0, cfgi_over={5: 1, 6: 3}) == 4  # convmaxits
    assert run(3, 1e-9, cfgi_over={3: 5}) == 0     # below min_it
    assert run(3, 1e-9, cfgi_over={4: 3}) == 0     # chknorm defers the test
    assert run(3, 1.0, cfgi_over={0: 1}) == 0      # NORM_NONE never converges


# ---------------------------------------------------------------- MAXPY, MDot, normalize, MultAdd, RCM

def test_maxpy_sequential_fma(oracle):
    """VecMAXPY (simdvec.mojo:503-518, ported.mojo:329-335): y_i takes
    y_i + a_k v_k,i for k = 0, 1, ... one correctly rounded multiply-add at a
    time.  Summing the terms first, or another k order, changes some bits
    here (checked: the alternatives differ from the sequential result)."""
    L = oracle.lib()
    rng = np.random.default_rng(7)
    n, nv = 41, 9
    V = [rng.normal(size=n) * 10.0 ** rng.integers(-3, 4, size=n) for _ in range(nv)]
    alpha = rng.normal(size=nv)
    y0 = rng.normal(size=n)
    y = y0.copy()
    ptrs = np.array([v.ctypes.data for v in V], dtype=np.int64)
    L.or_vec_maxpy(a_(y), nv, a_(alpha), a_(ptrs), n)
    want = y0.copy()
    for k in range(nv):
        want = np.array([fma_exact(alpha[k], V[k][i], want[i]) for i in range(n)])
    assert np.array_equal(y.view(np.int64), want.view(np.int64))
    terms_first = np.array([float(Fr(y0[i]) + sum(Fr(alpha[k]) * Fr(V[k][i]) for k in range(nv))) for i in range(n)])
    reversed_k = y0.copy()
    for k in reversed(range(nv)):
        reversed_k = np.array([fma_exact(alpha[k], V[k][i], reversed_k[i]) for i in range(n)])
    assert not np.array_equal(want, terms_first) and not np.array_equal(want, reversed_k)


@pytest.mark.parametrize("n", [0, 1, 15, 16, 17, 31, 33, 100, 2049])
@pytest.mark.parametrize("nv", [1, 3, 4, 5, 8, 9])
def test_mdot_against_fsum(O, n, nv):
    """_mdot_basis (ksp.mojo:956-1062) in both orders: every x . V_j within
    1e-13 of the exactly rounded sum, across the 4-row groups and the
    16-element blocks."""
    L = O.lib()
    rng = np.random.default_rng(100 * n + nv)
    x = rng.normal(size=n)
    V = rng.normal(size=(nv, n))
    res = np.zeros(nv)
    L.or_red_mdot(a_(x), a_(np.ascontiguousarray(V)), nv, n, a_(res))
    for j in range(nv):
        exact = math.fsum(float(Fr(p) * Fr(q)) for p, q in zip(x, V[j])) if n else 0.0
        scale = max(1.0, math.fsum(abs(p * q) for p, q in zip(x, V[j])))
        assert abs(res[j] - exact) <= 1e-13 * scale, (j, res[j], exact)


def _cancel_vector():
    """1e16 at element 0, ones at 1..15, -1e16 at 16, ones at 17..20.  With
    x = ones a sum of these n = 21 products shows its order: 1e16 + 1 rounds
    back to 1e16 (a tie to even), so which ones meet the 1e16 before the
    -1e16 cancels it decides the result."""
    y = np.ones(21)
    y[0], y[16] = 1e16, -1e16
    return np.ones(21), y


def test_reduction_orders_known_answers(oracle):
    """Results worked out by hand from each order's rule; a mistake in the
    chain of an element, the lane sums, the SW-steps, the tail or the tree
    shape gives another number (the exact sum is 19, a serial sum 4).
    REF dot (simdvec.mojo:332-353): chains c_e = e % 16 over e < 16; lane l
    = c_l + c_{4+l} + c_{8+l} + c_{12+l}: 1e16 (three ones absorbed), 4, 4,
    4; the SW-step over e = 16..19 adds -1e16, 1, 1, 1 to the lanes: 0, 5,
    5, 5; the tail e = 20 gives s = 1; s + ((0 + 5) + (5 + 5)) = 16.
    REF MDot (_mdot_basis, no SW-steps): the same lanes, (1e16 + 4) + 8 =
    1e16 + 12, plus the tail -1e16 + 1 + 1 + 1 + 1 = -1e16 (ones absorbed):
    12.  DEV (tile tree) on a 2048-element tile with 1e16 at 0, ones at 1
    and 2, -1e16 at 512: thread 0's pairs are (1e16 + 1) -> 1e16 and -1e16,
    summing to 0; thread 1's pair gives 1; the trees give 1 (REF: 2)."""
    L = oracle.lib()
    x, y = _cancel_vector()
    res = np.zeros(1)
    with oracle.mode("ref"):
        assert L.or_red_dot(a_(x), a_(y), 21) == 16.0
        L.or_red_mdot(a_(x), a_(y), 1, 21, a_(res))
        assert res[0] == 12.0
    with oracle.mode("dev"):
        assert L.or_red_dot(a_(x), a_(y), 21) != 16.0
    z = np.zeros(2048)
    z[0], z[1], z[2], z[512] = 1e16, 1.0, 1.0, -1e16
    ones = np.ones(2048)
    with oracle.mode("dev"):
        assert L.or_red_dot(a_(z), a_(ones), 2048) == 1.0
        assert L.or_red_total(a_(z), 2048) == 1.0
    with oracle.mode("ref"):
        assert L.or_red_dot(a_(z), a_(ones), 2048) == 2.0
        assert L.or_red_total(a_(z), 2048) == 2.0


def test_normalize_is_a_divide(oracle):
    """VecNormalize (ported.mojo:200-205, simdvec.mojo:296-304): x / ||x||_2,
    a correctly rounded divide per element, not a multiply by the
    reciprocal (which differs in some elements here); returns the norm."""
    L = oracle.lib()
    rng = np.random.default_rng(3)
    x0 = rng.normal(size=257)
    x = x0.copy()
    v = np.zeros(1)
    L.or_vec_normalize(a_(x), a_(v), len(x))
    z = math.sqrt(L.or_red_sumsq(a_(x0), len(x0)))
    assert v[0] == z
    want = np.array([float(Fr(e) / Fr(z)) for e in x0])
    assert np.array_equal(x.view(np.int64), want.view(np.int64))
    assert np.any(x0 * (1.0 / z) != want)
    zero = np.zeros(5)
    L.or_vec_normalize(a_(zero), a_(v), 5)
    assert v[0] == 0.0 and np.all(zero == 0.0)


def test_mult_add_seeded_with_y(oracle):
    """MatMultAdd_SeqAIJ (ported.mojo:394-405): y_i = y_i + sum_j a_ij x_j as
    one serial multiply-add chain that starts from y_i.  Starting from 0 and
    adding y_i at the end gives other bits here."""
    L = oracle.lib()
    ip, ix, va = W.diffusion_3d(4)[0]
    A = oracle.CSR(ip, ix, va)
    rng = np.random.default_rng(2)
    x = rng.normal(size=A.n)
    y0 = rng.normal(size=A.n) * 1e3
    y = y0.copy()
    L.or_mat_mult_add(a_(x), a_(y), a_(A.irow), a_(A.icol), a_(A.data), A.n)
    seeded, late = [], []
    for i in range(A.n):
        s = y0[i]
        t = 0.0
        for j in range(ip[i], ip[i + 1]):
            s = fma_exact(va[j], x[ix[j]], s)
            t = fma_exact(va[j], x[ix[j]], t)
        seeded.append(s)
        late.append(t + y0[i])
    assert np.array_equal(y, np.array(seeded))
    assert not np.array_equal(np.array(seeded), np.array(late))


def _rcm(oracle, n, edges):
    """RCM of the symmetric pattern with the given edges plus the diagonal."""
    rows = [[i] for i in range(n)]
    for a, b in edges:
        rows[a].append(b)
        rows[b].append(a)
    ip = np.zeros(n + 1, dtype=np.int64)
    cols = []
    for i in range(n):
        c = sorted(set(rows[i]))
        cols += c
        ip[i + 1] = len(cols)
    A = oracle.CSR(ip, np.array(cols, dtype=np.int64), np.ones(len(cols)))
    r, _ = oracle.permutations(A, "rcm")
    return r.tolist()


def test_rcm_hand_derived(oracle):
    """MatGetOrdering_RCM (factor.mojo:279-332): breadth first from each
    unvisited node in index order, neighbours in increasing index, each
    component reversed, components in the order found.  The permutations
    below were derived by hand from that rule."""
    # a path 0-1-2-3-4: visit order 0 1 2 3 4, reversed
    assert _rcm(oracle, 5, [(0, 1), (1, 2), (2, 3), (3, 4)]) == [4, 3, 2, 1, 0]
    # the 3 x 3 grid, row-major: visit order 0 1 3 2 4 6 5 7 8, reversed
    A = oracle.CSR(*W.laplacian_2d(3))
    assert oracle.permutations(A, "rcm")[0].tolist() == [8, 7, 5, 6, 4, 2, 3, 1, 0]
    # two components {0, 4, 2} and {1, 5, 3}: 0 4 2 -> 2 4 0, then 1 5 3 -> 3 5 1
    assert _rcm(oracle, 6, [(0, 4), (4, 2), (1, 5), (5, 3)]) == [2, 4, 0, 3, 5, 1]
    # a star whose centre is the last node: 0 3, then 1 and 2 from 3 -> reversed 2 1 3 0
    assert _rcm(oracle, 4, [(0, 3), (1, 3), (2, 3)]) == [2, 1, 3, 0]
