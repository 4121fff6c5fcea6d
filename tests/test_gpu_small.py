"""The one-workgroup drivers for small systems.  CG: a solve with wave-plan
sweeps and n <= 16384, or without sweeps and n <= 4096, runs as one launch;
GMRES, FGMRES and BiCGStab: n <= 4096.  Each gives the oracle's bits in the
tile-tree order (iterations, reason, residual norm, history and iterate),
the same bits as the general path (MPK_SMALL_ROWS=0); the counter
mpk_small_solves tells the paths apart."""

import numpy as np
import pytest

from _helpers import compare, petsc_solve, rng_vec, same_bits
from mpk import workloads as W

pytestmark = pytest.mark.gpu

MATS = {
    "lap64": lambda: W.laplacian_2d(64),
    "lap20": lambda: W.laplacian_2d(20),
    "diff12": lambda: W.diffusion_3d(12)[0],
    "tri50": lambda: W.tridiagonal(50),
    "lap127": lambda: W.laplacian_2d(127),  # 16129 rows: tiles 0..7, the last one partial
}
# (matrix, pc, rhs, whether the one-workgroup driver takes it): the 3-D
# and 127-wide sweeps run on 2 or 4 waves; 16129 rows without sweeps stay on
# the general path
CASES = [("lap64", "ilu", "ones", 1), ("lap64", "ilu", "normal5", 1), ("lap64", "icc", "normal5", 1),
         ("lap64", "jacobi", "ones", 1), ("lap64", "none", "normal5", 1), ("lap20", "ilu", "r3", 1),
         ("diff12", "icc", "r4", 1), ("diff12", "ilu", "r5", 1), ("diff12", "jacobi", "r6", 1),
         ("tri50", "icc", "ones", 1), ("lap127", "ilu", "r7", 1), ("lap127", "icc", "r9", 1),
         ("lap127", "none", "r8", 0)]


def _rhs(kind, n):
    if kind == "ones":
        return np.ones(n)
    if kind == "normal5":
        return np.random.default_rng(5).normal(size=n)
    return rng_vec(n, int(kind[1:]))


@pytest.mark.parametrize("name,pc,rhs,small", CASES, ids=[f"{a}-{b}-{c}" for a, b, c, _ in CASES])
def test_small_cg_bitwise(native, oracle, monkeypatch, name, pc, rhs, small):
    csr = MATS[name]()
    n = len(csr[0]) - 1
    b = _rhs(rhs, n)
    oracle.set_mode("dev")
    try:
        before = native.mpk_small_solves()
        g = compare(oracle, csr, b, "cg", pc, rtol=1e-10, max_it=1000)
        assert native.mpk_small_solves() == before + small, "the path the solve took"
        monkeypatch.setenv("MPK_SMALL_ROWS", "0")
        h = petsc_solve(csr, b, "cg", pc, rtol=1e-10, max_it=1000)
        assert native.mpk_small_solves() == before + small, "MPK_SMALL_ROWS=0: the general path"
    finally:
        oracle.set_mode("ref")
    assert g[1:3] == h[1:3] and same_bits(g[0], h[0]) and same_bits(g[4], h[4])


def test_small_cg_stops(native, oracle):
    """Stops the general path's way: at max_it (the reference's test reports
    CONVERGED_ITS there, as the oracle does), at once when the initial
    residual meets the tolerance, and with a nonzero guess."""
    csr = MATS["lap64"]()
    n = len(csr[0]) - 1
    oracle.set_mode("dev")
    try:
        g = compare(oracle, csr, np.ones(n), "cg", "ilu", rtol=1e-10, max_it=7)
        assert g[1] == 7
        g = compare(oracle, csr, np.zeros(n), "cg", "jacobi", rtol=1e-10, max_it=100)
        assert g[1] == 0
        x0 = rng_vec(n, 11)
        compare(oracle, csr, rng_vec(n, 12), "cg", "icc", x0=x0, rtol=1e-9, max_it=500)
    finally:
        oracle.set_mode("ref")


# GMRES / FGMRES / BiCGStab: the one-workgroup drivers take n <= 4096 (left or
# right side, a GMRES restart up to 31); the same bits as the general path.
# Options as the oracle takes them: orthog 1 = MGS, cgstype 1 = refine if
# needed, 2 = always, side 1 = right, 2 = symmetric.
KCASES = [
    ("lap20", "gmres", "none", {}, 1), ("lap20", "gmres", "jacobi", {}, 1), ("lap64", "gmres", "ilu", {}, 1),
    ("lap64", "gmres", "none", {"restart": 20}, 1), ("diff12", "gmres", "icc", {}, 1),
    ("lap20", "gmres", "none", {"orthog": 1}, 1), ("lap20", "gmres", "jacobi", {"cgstype": 1}, 1),
    ("lap20", "gmres", "ilu", {"cgstype": 2}, 1), ("lap20", "gmres", "ilu", {"side": 1}, 1),
    ("lap20", "gmres", "jacobi", {"side": 1, "orthog": 1}, 1), ("tri50", "gmres", "icc", {"restart": 31}, 1),
    ("lap20", "gmres", "none", {"restart": 32}, 0), ("lap20", "gmres", "ilu", {"side": 2}, 0),
    ("lap20", "fgmres", "none", {}, 1), ("lap64", "fgmres", "ilu", {}, 1), ("diff12", "fgmres", "jacobi", {}, 1),
    ("lap20", "fgmres", "icc", {"cgstype": 2}, 1), ("lap20", "fgmres", "ilu", {"orthog": 1}, 1),
    ("lap20", "bcgs", "none", {}, 1), ("lap64", "bcgs", "ilu", {}, 1), ("diff12", "bcgs", "jacobi", {}, 1),
    ("lap20", "bcgs", "icc", {"side": 1}, 1), ("lap64", "bcgs", "none", {}, 1),
    ("lap127", "gmres", "none", {}, 0), ("lap127", "bcgs", "ilu", {}, 0),
]


@pytest.mark.parametrize("name,ksp,pc,opts,small", KCASES,
                         ids=[f"{a}-{b}-{c}-{'-'.join(f'{k}{v}' for k, v in d.items())}" for a, b, c, d, _ in KCASES])
def test_small_krylov_bitwise(native, oracle, monkeypatch, name, ksp, pc, opts, small):
    csr = MATS[name]()
    n = len(csr[0]) - 1
    b = rng_vec(n, 21)
    kw = dict(rtol=1e-10, max_it=1500, **opts)
    oracle.set_mode("dev")
    try:
        before = native.mpk_small_solves()
        g = compare(oracle, csr, b, ksp, pc, **kw)
        assert native.mpk_small_solves() == before + small, "the path the solve took"
        monkeypatch.setenv("MPK_SMALL_ROWS", "0")
        h = petsc_solve(csr, b, ksp, pc, **kw)
        assert native.mpk_small_solves() == before + small, "MPK_SMALL_ROWS=0: the general path"
    finally:
        oracle.set_mode("ref")
    assert g[1:3] == h[1:3] and same_bits(g[0], h[0]) and same_bits(g[4], h[4])


@pytest.mark.parametrize("ksp", ["gmres", "fgmres", "bcgs"])
def test_small_krylov_stops(native, oracle, ksp):
    """max_it inside a cycle, a zero right-hand side, and a nonzero guess."""
    csr = MATS["lap20"]()
    n = len(csr[0]) - 1
    oracle.set_mode("dev")
    try:
        before = native.mpk_small_solves()
        g = compare(oracle, csr, rng_vec(n, 3), ksp, "jacobi", rtol=1e-10, max_it=7)
        assert g[1] == 7
        compare(oracle, csr, np.zeros(n), ksp, "none", rtol=1e-10, max_it=100)
        compare(oracle, csr, rng_vec(n, 12), ksp, "ilu", x0=rng_vec(n, 11), rtol=1e-9, max_it=500)
        assert native.mpk_small_solves() == before + 3
    finally:
        oracle.set_mode("ref")
