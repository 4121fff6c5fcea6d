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
