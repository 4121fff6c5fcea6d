"""Full-size solves on the GPU, at the sizes the benchmark workloads name:

* the FP64 drop-in on diffusion 128^3 (cg + jacobi, cg + icc(0)),
  convection-diffusion 128^3 (gmres(30) + ilu(0)) and the random SPD 1e6
  system (bcgs + sor(1.2)): iteration count, reason, residual norm, history
  and iterate bit for bit with the oracle in the tile-tree order, and within
  1e-12 relative of the oracle in the reference's order -- except cg + icc(0),
  whose two orders drift apart (2.8e-10 in the max norm) and which therefore
  also runs in the reference's order, bit for bit with that oracle;
* the mixed-precision solves of the two mixed workloads, and their FP64
  controls: FP64 relative residual <= 1e-10 recomputed on the host, and the
  solution within 1e-8 of the FP64 oracle's;
* the 4096 batched systems: every one at FP64 relative residual <= 1e-8,
  a sample against the oracle's FP64 solves.

The oracle solves (about a minute each on one core) all start in threads when
the module starts and run while the GPU works."""

import numpy as np
import pytest

import _fullsize as FS
from _helpers import petsc_solve, same_bits

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


@pytest.fixture(scope="module", autouse=True)
def _start_oracle(oracle):
    FS.start(FS.CASES)


def _gpu(case, order):
    import mpk
    sysname, kt, pt, pcopt, opts = FS.CASES[case]
    csr, b = FS.system(sysname)
    with mpk.reduction_order_scope(order):
        return petsc_solve(csr, b, kt, pt, 0, None, **dict(pcopt), **dict(opts))


def _same(g, o):
    assert g[1] == o.its, ("its", g[1], o.its)
    assert g[2] == o.reason, ("reason", g[2], o.reason)
    assert same_bits(np.array([g[3]]), np.array([o.rnorm])), ("rnorm", g[3], o.rnorm)
    assert same_bits(g[4], o.history[: len(g[4])]), "history"
    assert same_bits(g[0], o.x), ("x", FS.rel_gap(g[0], o.x))


@pytest.mark.parametrize("case", list(FS.CASES))
def test_fp64_dropin_tile_order(native, case):
    print(f"[fullsize] {case} tile-tree order", flush=True)
    g = _gpu(case, "dev")
    _same(g, FS.oracle(case, "dev").result())
    gap, gmax = FS.rel_gap(g[0], FS.oracle(case, "ref").result().x)
    print(f"[fullsize] {case}: its {g[1]}, gap to the reference order {gap:.2e} (max norm {gmax:.2e})")
    if case != "cg_icc0":
        assert gmax <= 1e-12


@pytest.mark.parametrize("case", ["cg_icc0", "cg_jacobi"])
def test_fp64_dropin_reference_order(native, case):
    print(f"[fullsize] {case} reference order", flush=True)
    g = _gpu(case, "ref")
    _same(g, FS.oracle(case, "ref").result())


MIXED = [
    # name, system, pc, omega, inner, precision, sweeps, inner rtol (the benchmark's settings)
    ("convdiff128-fp32-jac8", "convdiff_128", "ilu", 1.0, "gmres", "fp32", 8, 1e-2, "gmres_ilu0"),
    ("convdiff128-fp32-exact", "convdiff_128", "ilu", 1.0, "gmres", "fp32", "exact", 1e-3, "gmres_ilu0"),
    ("spd1e6-fp16-jac2", "random_spd_1e6", "sor", 1.2, "bcgs", "fp16", 2, 1e-2, "bcgs_sor"),
    ("spd1e6-fp32-jac2", "random_spd_1e6", "sor", 1.2, "bcgs", "fp32", 2, 1e-2, "bcgs_sor"),
    # the FP64 controls (the same inner solver and preconditioner forms)
    ("convdiff128-fp64-jac8", "convdiff_128", "ilu", 1.0, "gmres", "fp64", 8, 1e-2, "gmres_ilu0"),
    ("spd1e6-fp64-jac2", "random_spd_1e6", "sor", 1.2, "bcgs", "fp64", 2, 1e-2, "bcgs_sor"),
]


def _host_relres(csr, b, x):
    ip, ix, va = (np.asarray(a) for a in csr)
    n = len(ip) - 1
    rows = np.repeat(np.arange(n), np.diff(ip))
    ax = np.bincount(rows, weights=va * x[ix], minlength=n)
    return np.linalg.norm(b - ax) / np.linalg.norm(b)


@pytest.mark.parametrize("name,sysname,pck,omega,inner,prec,sweeps,irtol,fp64case", MIXED, ids=[m[0] for m in MIXED])
def test_mixed_fullsize(native, name, sysname, pck, omega, inner, prec, sweeps, irtol, fp64case):
    from mojopetsc4py import PETSc
    from mpk.mixed import MixedSolver
    print(f"[fullsize] mixed {name}", flush=True)
    csr, b = FS.system(sysname)
    ip, ix, va = csr
    n = len(ip) - 1
    A = PETSc.Mat().createSeqAIJ((n, n), nnz=26)
    A.setValuesCSR(ip, ix, va)
    A.assemble()
    pc = PETSc.PC().create()
    pc.setType(pck)
    if pck == "sor":
        pc.setSOROmega(omega)
    pc.setOperators(A)
    pc.setUp()
    ms = MixedSolver(A, pc, precision=prec, inner=inner, restart=30, sweeps=sweeps)
    x = PETSc.Vec().createSeq(n)
    st = ms.solve(PETSc.Vec().createWithArray(b), x, rtol=1e-10, inner_rtol=irtol)
    xh = x.getArray().copy()
    ms.destroy()
    assert st["converged"] and st["relres"] <= 1e-10, st
    assert _host_relres(csr, b, xh) <= 1.5e-10
    xo = FS.oracle(fp64case, "dev").result().x
    err = np.linalg.norm(xh - xo) / np.linalg.norm(xo)
    print(f"[fullsize] mixed {name}: outer {st.get('outer')}, inner {st.get('inner')}, error vs FP64 {err:.2e}")
    assert err <= 1e-8


def test_batched_fullsize(native, oracle):
    """The 4096 systems of the batched workload in one persistent launch."""
    from mpk import workloads as W
    from mpk.batch import BatchSolver
    print("[fullsize] batched 4096", flush=True)
    systems = W.batch_systems()
    assert len(systems) == 4096
    bs = BatchSolver(systems)
    try:
        x, st = bs.solve(rtol=1e-8, inner_rtol=1e-3, graph=True)
        assert np.all(st[:, 3] == 1.0), np.flatnonzero(st[:, 3] != 1.0)[:10]
        assert np.max(st[:, 0]) <= 1e-8
        for s, (ip, ix, va, b, kind) in enumerate(systems):
            assert _host_relres((ip, ix, va), b, bs.solution(x, s)) <= 1.5e-8, (s, kind)
        oracle.set_mode("ref")
        for s in range(0, 4096, 256):
            ip, ix, va, b, kind = systems[s]
            A = oracle.CSR(ip, ix, va)
            if kind == "cg+jacobi":
                r = oracle.ksp_solve(A, b, "cg", oracle.PCSetup(A, "jacobi"), rtol=1e-13, atol=0.0, max_it=50000)
            else:
                r = oracle.ksp_solve(A, b, "gmres", oracle.PCSetup(A, "ilu", 0), rtol=1e-13, atol=0.0,
                                     max_it=50000, restart=20)
            err = np.linalg.norm(bs.solution(x, s) - r.x) / np.linalg.norm(r.x)
            assert err <= 1e-5, (s, kind, err)
    finally:
        bs.destroy()
