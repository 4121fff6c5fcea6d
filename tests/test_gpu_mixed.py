"""Mixed-precision solves (FP32 / FP16 inner iterations under FP64 iterative
refinement, and the same solver with FP64 inner iterations as their control):
they reach the FP64 relative residual asked for, checked here with an FP64
residual computed on the host, and their solution agrees with the oracle's
FP64 solve of the same system to within what that residual allows."""

import numpy as np
import pytest

from mpk import workloads as W

pytestmark = pytest.mark.gpu


def _mat(csr):
    from mojopetsc4py import PETSc
    ip, ix, va = csr
    n = len(ip) - 1
    A = PETSc.Mat().createSeqAIJ((n, n), nnz=8)
    A.setValuesCSR(ip, ix, va)
    A.assemble()
    return A


def _pc(A, kind, omega=1.0):
    from mojopetsc4py import PETSc
    pc = PETSc.PC().create()
    pc.setType(kind)
    if kind == "sor":
        pc.setSOROmega(omega)
    pc.setOperators(A)
    pc.setUp()
    return pc


def _host_relres(csr, b, x):
    ip, ix, va = (np.asarray(a) for a in csr)
    n = len(ip) - 1
    rows = np.repeat(np.arange(n), np.diff(ip))
    ax = np.bincount(rows, weights=va * x[ix], minlength=n)
    return np.linalg.norm(b - ax) / np.linalg.norm(b)


def _oracle_x(oracle, csr, b, ksp, pc_kind, **pc_kw):
    oracle.set_mode("ref")
    A = oracle.CSR(*csr)
    pc = oracle.PCSetup(A, pc_kind, **pc_kw)
    r = oracle.ksp_solve(A, b, ksp, pc, rtol=1e-13, atol=0.0, max_it=20000)
    return r.x


CASES = [
    # name, matrix, pc, omega, inner, precision, sweeps
    ("cd12-ilu-gmres-fp32", lambda: W.convection_diffusion_3d(12), "ilu", 1.0, "gmres", "fp32", "exact"),
    ("cd12-ilu-gmres-fp16", lambda: W.convection_diffusion_3d(12), "ilu", 1.0, "gmres", "fp16", "exact"),
    ("diff24-ilu-gmres-fp32", lambda: W.diffusion_3d(24), "ilu", 1.0, "gmres", "fp32", "exact"),
    ("spd3000-sor-bcgs-fp16", lambda: W.random_spd(3000, band=400), "sor", 1.2, "bcgs", "fp16", "exact"),
    ("spd3000-sor-bcgs-fp32", lambda: W.random_spd(3000, band=400), "sor", 1.2, "bcgs", "fp32", "exact"),
    ("lap40-none-gmres-fp32", lambda: (W.laplacian_2d(40), None), "none", 1.0, "gmres", "fp32", "exact"),
    ("cd12-ilu-bcgs-fp32", lambda: W.convection_diffusion_3d(12), "ilu", 1.0, "bcgs", "fp32", "exact"),
    ("cd12-ilu-gmres-fp32-jac3", lambda: W.convection_diffusion_3d(12), "ilu", 1.0, "gmres", "fp32", 3),
    ("diff24-ilu-gmres-fp16-jac2", lambda: W.diffusion_3d(24), "ilu", 1.0, "gmres", "fp16", 2),
    ("spd3000-sor-bcgs-fp16-jac3", lambda: W.random_spd(3000, band=400), "sor", 1.2, "bcgs", "fp16", 3),
    # over 8192 rows with a deep, narrow DAG: ring-planned sweeps
    ("spd20k-sor-bcgs-fp32", lambda: W.random_spd(20000, band=3000), "sor", 1.2, "bcgs", "fp32", "exact"),
    ("lap120-ilu-gmres-fp16", lambda: (W.laplacian_2d(120), None), "ilu", 1.0, "gmres", "fp16", "exact"),
    # the FP64 control: the same inner solver and preconditioner forms
    ("cd12-ilu-gmres-fp64", lambda: W.convection_diffusion_3d(12), "ilu", 1.0, "gmres", "fp64", "exact"),
    ("cd12-ilu-gmres-fp64-jac3", lambda: W.convection_diffusion_3d(12), "ilu", 1.0, "gmres", "fp64", 3),
    ("spd3000-sor-bcgs-fp64", lambda: W.random_spd(3000, band=400), "sor", 1.2, "bcgs", "fp64", "exact"),
    ("spd3000-sor-bcgs-fp64-jac3", lambda: W.random_spd(3000, band=400), "sor", 1.2, "bcgs", "fp64", 3),
    ("spd20k-sor-bcgs-fp64", lambda: W.random_spd(20000, band=3000), "sor", 1.2, "bcgs", "fp64", "exact"),
    ("lap40-none-gmres-fp64", lambda: (W.laplacian_2d(40), None), "none", 1.0, "gmres", "fp64", "exact"),
]


@pytest.mark.parametrize("name,make,pck,omega,inner,prec,sweeps", CASES, ids=[c[0] for c in CASES])
def test_mixed_reaches_fp64_tolerance(native, oracle, name, make, pck, omega, inner, prec, sweeps):
    from mojopetsc4py import PETSc
    from mpk.mixed import MixedSolver
    csr, b = make()
    n = len(csr[0]) - 1
    if b is None:
        b = np.random.default_rng(11).uniform(-1, 1, n)
    A = _mat(csr)
    pc = _pc(A, pck, omega) if pck != "none" else None
    ms = MixedSolver(A, pc, precision=prec, inner=inner, restart=30, sweeps=sweeps)
    x = PETSc.Vec().createSeq(n)
    st = ms.solve(PETSc.Vec().createWithArray(b), x, rtol=1e-10, inner_rtol=1e-4)
    xh = x.getArray().copy()
    assert st["converged"], st
    assert st["relres"] <= 1e-10
    # FP64 residual recomputed on the host
    assert _host_relres(csr, b, xh) <= 1.5e-10
    # error against the oracle's FP64 solve of the same system
    ksp = "bcgs" if inner == "bcgs" else "gmres"
    kw = {"omega": omega} if pck == "sor" else {}
    xo = _oracle_x(oracle, csr, b, ksp, pck, **kw)
    err = np.linalg.norm(xh - xo) / np.linalg.norm(xo)
    assert err <= 1e-7, (st, err)
    ms.destroy()


def test_mixed_rejects_bad_arguments(native):
    from mpk.mixed import MixedSolver
    csr, _ = W.convection_diffusion_3d(6)
    A = _mat(csr)
    with pytest.raises(ValueError):
        MixedSolver(A, None, precision="fp8")
    with pytest.raises(ValueError):
        MixedSolver(A, None, inner="cg")
    with pytest.raises(ValueError):
        MixedSolver(A, None, sweeps=0)
    from mpk._native import NativeError
    with pytest.raises((RuntimeError, NativeError)):
        MixedSolver(A, None, restart=100)
