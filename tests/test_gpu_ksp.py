"""Whole solves on the GPU against the oracle in the device reduction order:
the same iteration count, converged reason, residual history and iterate,
bit for bit.  Driven through the PETSc surface, which calls mp_ksp_solve_*."""

import numpy as np
import pytest

from _helpers import compare as _compare
from _helpers import same_bits, wide_band
from mpk import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def O(oracle):
    oracle.set_mode("dev")
    yield oracle
    oracle.set_mode("ref")


LAP64 = W.laplacian_2d(64)
ONES = np.ones(64 * 64)
BENCH_B = np.random.default_rng(5).normal(size=64 * 64)


def test_config1_cg_ilu0_ones(native, O):
    """64x64 Poisson, b = ones, cg + ilu(0), rtol 1e-10, max_it 1000."""
    g = _compare(O, LAP64, ONES, "cg", "ilu", rtol=1e-10, max_it=1000)
    assert g[2] == 2


def test_config1_published_count(native, O):
    """The reference's bench RHS gives 77 iterations (README.md:315)."""
    g = _compare(O, LAP64, BENCH_B, "cg", "ilu", rtol=1e-10, atol=1e-50, divtol=1e30, max_it=2000)
    assert g[1] == 77


@pytest.mark.parametrize("ksp_type,pc_type,its", [("cg", "none", 225), ("cg", "icc", 77), ("gmres", "none", 626),
                                                  ("fgmres", "none", 626), ("gmres", "jacobi", 626)])
def test_published_counts(native, O, ksp_type, pc_type, its):
    g = _compare(O, LAP64, BENCH_B, ksp_type, pc_type, rtol=1e-10, atol=1e-50, divtol=1e30, max_it=2000)
    assert g[1] == its


def test_bcgs_band(native, O):
    g = _compare(O, LAP64, BENCH_B, "bcgs", "none", rtol=1e-10, atol=1e-50, divtol=1e30, max_it=2000)
    assert abs(g[1] - 162) <= 20


DIFF16, DIFF16B = W.diffusion_3d(16)
CD12, CD12B = W.convection_diffusion_3d(12)


@pytest.mark.parametrize("ksp_type,pc_type,levels", [
    ("cg", "jacobi", 0), ("cg", "icc", 0), ("cg", "icc", 1), ("cg", "icc", 2), ("cg", "ilu", 0),
    ("cg", "ilu", 2), ("cg", "sor", 0),
    ("bcgs", "jacobi", 0), ("bcgs", "sor", 0), ("gmres", "ilu", 0), ("fgmres", "ilu", 0), ("gmres", "jacobi", 0),
])
def test_diffusion_16(native, O, ksp_type, pc_type, levels):
    _compare(O, DIFF16, DIFF16B, ksp_type, pc_type, levels, rtol=1e-10, max_it=3000)


@pytest.mark.parametrize("opts", [
    dict(), dict(cgstype=1), dict(cgstype=2), dict(orthog=1), dict(restart=10), dict(side=1),
])
def test_gmres_variants_convdiff(native, O, opts):
    _compare(O, CD12, CD12B, "gmres", "ilu", rtol=1e-10, max_it=2000, **opts)


@pytest.mark.parametrize("ksp_type", ["bcgs", "fgmres"])
def test_convdiff_other(native, O, ksp_type):
    _compare(O, CD12, CD12B, ksp_type, "ilu", rtol=1e-10, max_it=2000)


def test_sor_omega(native, O):
    _compare(O, DIFF16, DIFF16B, "bcgs", "sor", omega=1.2, rtol=1e-10, max_it=2000)


def test_nonzero_guess(native, O):
    x0 = np.random.default_rng(11).normal(size=len(DIFF16B))
    for kt, pt in (("cg", "jacobi"), ("gmres", "ilu"), ("bcgs", "jacobi"), ("fgmres", "jacobi")):
        _compare(O, DIFF16, DIFF16B, kt, pt, x0=x0, rtol=1e-8, max_it=2000)


@pytest.mark.parametrize("normtype", [1, 3, 4])
def test_cg_norm_types(native, O, normtype):
    _compare(O, DIFF16, DIFF16B, "cg", "jacobi", normtype=normtype, rtol=1e-10, max_it=300)


def test_max_it_reports_converged_its(native, O):
    g = _compare(O, DIFF16, DIFF16B, "cg", "jacobi", rtol=1e-14, max_it=7)
    assert g[1] == 7 and g[2] == 4


def test_nan_rhs(native, O):
    b = DIFF16B.copy()
    b[3] = np.nan
    g = _compare(O, DIFF16, b, "cg", "jacobi", rtol=1e-10, max_it=50)
    assert g[2] == -9 and g[1] == 0


def test_indefinite_cg(native, O):
    csr = (np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32), np.array([1.0, -1.0]))
    g = _compare(O, csr, np.array([1.0, 1.0]), "cg", "none", rtol=1e-10, max_it=50)
    assert g[2] in (-8, -10)


def test_tiny_divtol(native, O):
    g = _compare(O, DIFF16, DIFF16B, "gmres", "none", rtol=1e-10, divtol=1e-30, max_it=50)
    assert g[2] == -4 and g[1] == 0


@pytest.mark.parametrize("normtype", [2, 3])
def test_cg_jacobi_finalize_steps(native, O, monkeypatch, normtype):
    """CG + Jacobi with its scalar steps in finalize launches (the form large
    systems take, forced here by a zero tile threshold) gives the same bits as
    the oracle, as the folded form does; also a max_it stop and a NaN."""
    monkeypatch.setenv("MPK_CG_FOLD_TILES", "0")
    _compare(O, DIFF16, DIFF16B, "cg", "jacobi", normtype=normtype, rtol=1e-10, max_it=300)
    g = _compare(O, DIFF16, DIFF16B, "cg", "jacobi", rtol=1e-14, max_it=7)
    assert g[1] == 7 and g[2] == 4
    b = DIFF16B.copy()
    b[3] = np.nan
    g = _compare(O, DIFF16, b, "cg", "jacobi", rtol=1e-10, max_it=50)
    assert g[2] == -9


WIDE = wide_band()
WIDE_B = np.random.default_rng(9).normal(size=40000)


@pytest.mark.parametrize("fold", ["default", "0"])
def test_cg_jacobi_int32_columns(native, O, monkeypatch, fold):
    """CG + Jacobi on a matrix whose entries lie 35000 columns from their row,
    so its sliced copy keeps int32 columns (the other systems here use 16-bit
    offsets): the oracle's bits, in the folded and the finalize form."""
    if fold != "default":
        monkeypatch.setenv("MPK_CG_FOLD_TILES", fold)
    g = _compare(O, WIDE, WIDE_B, "cg", "jacobi", rtol=1e-12, max_it=500)
    assert g[2] == 2


def test_cg_jacobi_nontemporal_product(native, O, monkeypatch):
    """CG + Jacobi with the non-temporal product kernels (the default from
    3,000,000 rows, forced by MPK_SELL_NT=1), in both the folded and the
    finalize form of the scalar steps: the oracle's bits."""
    monkeypatch.setenv("MPK_SELL_NT", "1")
    _compare(O, DIFF16, DIFF16B, "cg", "jacobi", rtol=1e-10, max_it=300)
    monkeypatch.setenv("MPK_CG_FOLD_TILES", "0")
    _compare(O, DIFF16, DIFF16B, "cg", "jacobi", rtol=1e-10, max_it=300)
    _compare(O, DIFF16, DIFF16B, "gmres", "jacobi", rtol=1e-10, max_it=300)


def _wide_spectrum(n=400):
    """Diagonal with eigenvalues log-uniform over [1, 1e6]: GMRES needs about
    n iterations, so a restart of 300 fills its basis past 256 vectors."""
    d = np.logspace(0, 6, n)
    ip = np.arange(n + 1, dtype=np.int32)
    return ip, np.arange(n, dtype=np.int32), d


@pytest.mark.parametrize("ksp_type", ["gmres", "fgmres"])
def test_gmres_restart_past_256(native, O, ksp_type):
    """More Gram-Schmidt sums than the finalize keeps in LDS (the basis grows
    to 301 vectors): the oracle's bits.  The reference accepts any restart."""
    csr = _wide_spectrum()
    b = np.random.default_rng(13).normal(size=400)
    g = _compare(O, csr, b, ksp_type, "none", rtol=1e-13, max_it=320, restart=300)
    assert g[1] > 257
