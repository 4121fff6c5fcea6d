"""The oracle against the reference's published answers (no GPU).

* README.md:50-94: tridiagonal n = 64, ILU(0), CG, rtol 1e-12, max_it 500
  prints `1 1.4432899320127035e-15` -- reproduced bit for bit, which pins
  the fused multiply-add placement of the restatement.
* README.md:302-321 (bench/bench.py:64-93,210-239): iteration counts on the
  5-point Laplacian with b = default_rng(5).normal(n), rtol 1e-10, atol 1e-50,
  divtol 1e30, max_it 2000.  BCGS is held to +-20 as test_ksp_parity.py:264-283
  holds it.
* The device reduction order gives the same counts and an iterate within
  1e-12 relative of the reference order (the FP64 tolerance these tests
  hold the device path to).
"""

import numpy as np
import pytest

from mpk import workloads as W


def _bench_b(k):
    return np.random.default_rng(5).normal(size=k * k)


def _solve(O, k, ksp_type, pc_type, mode="ref"):
    O.set_mode(mode)
    try:
        A = O.CSR(*W.laplacian_2d(k))
        return O.ksp_solve(A, _bench_b(k), ksp_type, O.PCSetup(A, pc_type, 0),
                           rtol=1e-10, atol=1e-50, divtol=1e30, max_it=2000)
    finally:
        O.set_mode("ref")


def test_readme_example_bit_exact(oracle):
    O = oracle
    A = O.CSR(*W.tridiagonal(64))
    b = np.ones(64)
    r = O.ksp_solve(A, b, "cg", O.PCSetup(A, "ilu", 0), rtol=1e-12, max_it=500)
    res = A.mult(r.x)
    res = b + (-1.0) * res  # r.aypx(-1.0, b)
    z = np.zeros(2)
    O.lib().or_vec_norm(res.ctypes.data, 1, z.ctypes.data, 64)
    assert r.its == 1
    assert z[0] == 1.4432899320127035e-15


PUBLISHED = [
    (32, "cg", "none", 113), (64, "cg", "none", 225), (128, "cg", "none", 434),
    (32, "cg", "ilu", 40), (64, "cg", "ilu", 77), (128, "cg", "ilu", 146),
    (32, "cg", "icc", 40), (64, "cg", "icc", 77), (128, "cg", "icc", 146),
    (32, "gmres", "none", 232), (64, "gmres", "none", 626),
    (32, "fgmres", "none", 232), (64, "fgmres", "none", 626),
    (32, "gmres", "jacobi", 232), (64, "gmres", "jacobi", 626),
]


@pytest.mark.parametrize("k,ksp_type,pc_type,its", PUBLISHED)
def test_published_iteration_counts(oracle, k, ksp_type, pc_type, its):
    r = _solve(oracle, k, ksp_type, pc_type)
    assert r.reason == 2
    assert r.its == its


@pytest.mark.parametrize("k,its", [(32, 85), (64, 162), (128, 316)])
def test_bcgs_counts_within_band(oracle, k, its):
    r = _solve(oracle, k, "bcgs", "none")
    assert r.reason == 2 and abs(r.its - its) <= 20


def test_gmres_128_hits_max_it_as_converged_its(oracle):
    """setTolerances(max_it=2000) sets convmaxits, so 2000 iterations report
    CONVERGED_ITS (4) in the reference (KSP.py:223-225, ksp.mojo:113-114)."""
    r = _solve(oracle, 128, "gmres", "none")
    assert r.its == 2000 and r.reason == 4


@pytest.mark.parametrize("k,ksp_type,pc_type", [(64, "cg", "ilu"), (64, "cg", "none"), (64, "gmres", "none"),
                                                (32, "fgmres", "none"), (64, "cg", "icc")])
def test_device_order_matches_reference_order(oracle, k, ksp_type, pc_type):
    ref = _solve(oracle, k, ksp_type, pc_type, "ref")
    dev = _solve(oracle, k, ksp_type, pc_type, "dev")
    assert dev.its == ref.its and dev.reason == ref.reason
    rel = np.max(np.abs(dev.x - ref.x)) / np.max(np.abs(ref.x))
    assert rel <= 1e-12, rel


def test_config1_ones_both_orders(oracle):
    """Config 1: 64x64 Poisson, b = ones, cg + ilu(0), rtol 1e-10, max_it 1000."""
    O = oracle
    out = {}
    for mode in ("ref", "dev"):
        O.set_mode(mode)
        A = O.CSR(*W.laplacian_2d(64))
        out[mode] = O.ksp_solve(A, np.ones(4096), "cg", O.PCSetup(A, "ilu", 0), rtol=1e-10, max_it=1000)
    O.set_mode("ref")
    assert out["ref"].its == out["dev"].its and out["ref"].reason == out["dev"].reason == 2
    rel = np.max(np.abs(out["ref"].x - out["dev"].x)) / np.max(np.abs(out["ref"].x))
    assert rel <= 1e-12
    # the true residual meets the tolerance
    res = np.ones(4096) - A.mult(out["ref"].x)
    assert np.linalg.norm(res) / 64.0 < 1e-9
