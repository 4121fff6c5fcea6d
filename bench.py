"""Benchmark: time to relative residual 1e-10 and SpMV HBM bandwidth on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--no-cpu]

Workload: 3-D 7-point variable-coefficient diffusion on 128^3 cells
(n = 2,097,152, 14,581,760 nonzeros), face coefficients lognormal(0, 1) from
seed 0, right-hand side uniform[-1, 1] from seed 1, x0 = 0, FP64 values with
int32 indices.  A step is one complete FP64 solve, CG + Jacobi to rtol 1e-10
(PETSc's default convergence test, preconditioned norm), with matrix, vectors
and preconditioner already resident in HBM.  `value` is the mean time to
solution over the K timed steps (lower is better).

Also reported: the SpMV kernel's average launch time from HIP events on the
library's stream and its achieved bandwidth from algorithmic bytes
((8 + 4) nnz + 4 (n + 1) + 2 * 8 n), CG + IC(0) time to solution on the same
system, the mixed-precision solves against plain FP64 (convection-diffusion
128^3 with gmres(30) + ilu(0); a random SPD system of 1e6 rows with bcgs +
sor), the batched solve of 4096 small systems against one-at-a-time solves,
and CPU baselines: the oracle (the C restatement of the reference's
single-threaded path) timed on one core over the full solves of the headline
system and of the two mixed-precision systems, in the reference's reduction
order, beside the GPU's iterations and the iterate gap.

One GPU only: --gpus above 1 exits with a message.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "mojo-petsc4py_amd"), ROOT]

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
_T0 = time.perf_counter()


def _progress(msg: str) -> None:
    """One line on stderr as each leg of the benchmark ends."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)

# HBM bytes per launch of the roofline kernel, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs scaled by a copy kernel of known size
# (profiles/spmv_sell_diffusion128_traffic.md), keyed by (kernel, n, nnz).
MEASURED_TRAFFIC = {("k_spmv_sell", 2097152, 14581760): 190142898, ("k_cg_wdot", 2097152, 14581760): 198584162}


def algo_bytes(n: int, nnz: int) -> int:
    """Algorithmic bytes of one FP64 product over int32 CSR: (8 + 4) nnz +
    4 (n + 1) + 16 n (values and columns, row pointers, x read once, y
    written once); 216,924,164 B for diffusion_3d 128^3."""
    return 12 * nnz + 4 * (n + 1) + 16 * n


def spmv_bytes(n: int, nnz: int, so=None) -> int:
    """The bytes the sliced product actually moves: values, columns (2 bytes
    an entry when the sliced copy `so` holds 16-bit offsets, marked by a
    negative last slice offset; else 4), row pointers, x and y."""
    cb = 4
    if so is not None and int(so.to_numpy(np.int32, (n + 63) // 64 + 1)[-1]) < 0:
        cb = 2
    return (8 + cb) * nnz + 4 * (n + 1) + 16 * n


def build_system(N: int):
    from mojopetsc4py import PETSc
    from mpk import workloads as W
    (ip, ix, va), b = W.diffusion_3d(N)
    n = len(ip) - 1
    A = PETSc.Mat().createSeqAIJ((n, n), nnz=7)
    A.setValuesCSR(ip, ix, va)
    A.assemble()
    bv = PETSc.Vec().createWithArray(b)
    return A, bv, (ip, ix, va), b


def make_ksp(A, pc_type: str):
    from mojopetsc4py import PETSc
    pc = PETSc.PC().create()
    pc.setType(pc_type)
    pc.setOperators(A)
    ksp = PETSc.KSP().create()
    ksp.setPC(pc)
    ksp.setType("cg")
    ksp.setOperators(A)
    ksp.setTolerances(rtol=1e-10, max_it=20000)
    ksp.setUp()
    return ksp


def timed_solves(ksp, bv, x, steps: int, warmup: int):
    from mpk import _native
    lib = _native.lib()
    for _ in range(warmup):
        ksp.solve(bv, x)
    lib.mpk_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ksp.solve(bv, x)
    lib.mpk_synchronize()
    return (time.perf_counter() - t0) / max(steps, 1)


def true_relres(A, bv, x):
    r = A.createVecLeft()
    A.mult(x, r)
    r.aypx(-1.0, bv)
    return r.norm() / bv.norm()


def cpu_baseline(csr, b, x_gpu, max_it: int = 20000):
    """The oracle's full CG + Jacobi solve to rtol 1e-10 on one core, in the
    reference's reduction order: its time, iterations and reason, and the
    gap between its iterate and the GPU's (which computes the tile-tree
    order)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.set_mode("ref")
    A = O.CSR(*csr)
    pc = O.PCSetup(A, "jacobi")
    t0 = time.perf_counter()
    r = O.ksp_solve(A, b, "cg", pc, rtol=1e-10, max_it=max_it)
    dt = time.perf_counter() - t0
    gap2 = float(np.linalg.norm(r.x - x_gpu) / np.linalg.norm(r.x))
    gapm = float(np.max(np.abs(r.x - x_gpu)) / np.max(np.abs(r.x)))
    return dt, r.its, r.reason, r.rnorm, gap2, gapm


def spmv_only(lib, A, bv, n, nnz, reps):
    """The SpMV timing loop and the CG product's (k_cg_wdot), then 10 copies
    of 2^24 doubles (a kernel whose HBM traffic is known: 128 MiB read, 128 MiB
    written per launch), so a counter run can scale its readings."""
    from mpk import _native
    from mpk.device import DeviceBuffer
    d = A._dev()
    xs = bv.duplicate()
    xs.set(1.0)
    ys = bv.duplicate()
    so, sv, sc = d.sell()
    # the first launches of each kernel (code object loading) stay untimed
    lib.mpk_time_spmv(xs._d(), ys._dw(), d.rp.ptr, d.ci.ptr, d.v.ptr, n, so.ptr, sv.ptr, sc.ptr, 5)
    ms = lib.mpk_time_spmv(xs._d(), ys._dw(), d.rp.ptr, d.ci.ptr, d.v.ptr, n, so.ptr, sv.ptr, sc.ptr, reps)
    lib.mpk_time_cg_wdot(xs._d(), ys._dw(), d.rp.ptr, n, so.ptr, sv.ptr, sc.ptr, 5)
    ms_w = lib.mpk_time_cg_wdot(xs._d(), ys._dw(), d.rp.ptr, n, so.ptr, sv.ptr, sc.ptr, reps)
    m = 1 << 24
    src, dst = DeviceBuffer(8 * m), DeviceBuffer(8 * m)
    lib.mp_vec_set(src.ptr, 1.0, m)
    for _ in range(10):
        lib.mp_vec_copy(src.ptr, dst.ptr, m)
    lib.mpk_synchronize()
    _native.check()
    print(json.dumps({"spmv_ms": ms, "cg_wdot_ms": ms_w, "spmv_bytes": spmv_bytes(n, nnz, so), "launches": reps,
                      "copy_bytes_each_way": 8 * m, "copy_launches": 10}))
    return 0


def spmv_cols32(lib, A, bv, n, nnz, reps):
    """The roofline kernel over a sliced copy of the same matrix with int32
    columns only (MPK_SELL_COLS=32 at build), so the report shows the
    product of 12-byte entries beside the 10-byte one it runs."""
    from mpk.device import DeviceBuffer
    d = A._dev()
    padded = lib.mpk_sell_size(d.rp.ptr, n)
    so = DeviceBuffer(4 * ((n + 63) // 64 + 1))
    sv = DeviceBuffer(8 * max(padded, 1))
    sc = DeviceBuffer(6 * max(padded, 1))
    os.environ["MPK_SELL_COLS"] = "32"
    try:
        rc = lib.mpk_sell_build(d.rp.ptr, d.ci.ptr, d.v.ptr, n, so.ptr, sv.ptr, sc.ptr)
    finally:
        del os.environ["MPK_SELL_COLS"]
    if rc != 0:
        return None
    xs = bv.duplicate()
    xs.set(1.0)
    ys = bv.duplicate()
    lib.mpk_time_spmv(xs._d(), ys._dw(), d.rp.ptr, d.ci.ptr, d.v.ptr, n, so.ptr, sv.ptr, sc.ptr, 5)
    ms = lib.mpk_time_spmv(xs._d(), ys._dw(), d.rp.ptr, d.ci.ptr, d.v.ptr, n, so.ptr, sv.ptr, sc.ptr, reps)
    b = spmv_bytes(n, nnz, so)
    gbs = b / (ms * 1e-3) / 1e9
    return {"columns": "int32", "bytes_per_launch": b, "avg_launch_ms": ms, "achieved": gbs,
            "frac": gbs / HBM_PEAK_GBS, "unit": "GB/s"}


def spmv_past_cache(lib, N: int = 256, reps: int = 20):
    """The roofline kernel on diffusion_3d N^3 (1.74 GB at 256^3, past the
    256 MB last-level cache): HIP-event average of `reps` launches, so the
    report shows what it streams from HBM beside the cache-resident 128^3
    figure."""
    A, bv, csr, _ = build_system(N)
    n = A.getSize()[0]
    nnz = int(csr[0][-1])
    d = A._dev()
    so, sv, sc = d.sell()
    xs = bv.duplicate()
    xs.set(1.0)
    ys = bv.duplicate()
    lib.mpk_time_spmv(xs._d(), ys._dw(), d.rp.ptr, d.ci.ptr, d.v.ptr, n, so.ptr, sv.ptr, sc.ptr, 3)
    ms = lib.mpk_time_spmv(xs._d(), ys._dw(), d.rp.ptr, d.ci.ptr, d.v.ptr, n, so.ptr, sv.ptr, sc.ptr, reps)
    b, moved = algo_bytes(n, nnz), spmv_bytes(n, nnz, so)
    gbs = b / (ms * 1e-3) / 1e9
    # the CG iteration's dominant kernel: the same product with its p.w
    # partials (the same algorithmic bytes: the partials are 8 bytes a tile)
    lib.mpk_time_cg_wdot(xs._d(), ys._dw(), d.rp.ptr, n, so.ptr, sv.ptr, sc.ptr, 3)
    ms_w = lib.mpk_time_cg_wdot(xs._d(), ys._dw(), d.rp.ptr, n, so.ptr, sv.ptr, sc.ptr, reps)
    gbs_w = b / (ms_w * 1e-3) / 1e9
    return {"workload": f"diffusion_3d {N}^3", "n": n, "nnz": nnz, "bytes_per_launch": b, "bytes_moved": moved,
            "avg_launch_ms": ms, "achieved": gbs, "frac": gbs / HBM_PEAK_GBS, "unit": "GB/s",
            "moved_gbs": moved / (ms * 1e-3) / 1e9,
            "k_cg_wdot": {"avg_launch_ms": ms_w, "achieved": gbs_w, "frac": gbs_w / HBM_PEAK_GBS,
                          "moved_gbs": moved / (ms_w * 1e-3) / 1e9}}


def _mat(csr):
    from mojopetsc4py import PETSc
    ip, ix, va = csr
    n = len(ip) - 1
    A = PETSc.Mat().createSeqAIJ((n, n), nnz=8)
    A.setValuesCSR(ip, ix, va)
    A.assemble()
    return A


def cpu_sample(csr, b, ksp: str, pc_kind: str, gpu_its: int, gpu_reason: int, gpu_x, **kw):
    """The oracle's (the reference's single-threaded path, restated in C) FP64
    solve on one core, in the reference's reduction order, to the GPU solve's
    tolerances (rtol 1e-10, max_it 20000); its preconditioner set-up is timed
    apart.  Reports its iterations and reason beside the GPU's and the
    iterate gap max|x_ref - x_gpu| / max|x_ref|."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.set_mode("ref")
    A = O.CSR(*csr)
    pc_kw = {k: kw.pop(k) for k in ("omega", "levels") if k in kw}
    t0 = time.perf_counter()
    pc = O.PCSetup(A, pc_kind, *([pc_kw.pop("levels")] if "levels" in pc_kw else []), **pc_kw)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = O.ksp_solve(A, b, ksp, pc, rtol=1e-10, max_it=20000, **kw)
    t_solve = time.perf_counter() - t0
    gap = float(np.max(np.abs(r.x - gpu_x)) / np.max(np.abs(r.x)))
    return {"value": t_solve, "unit": "s", "cores": 1, "kind": "port",
            "sample": f"the oracle's full {ksp}+{pc_kind} solve in the reference's order on one core"
                      f" ({r.its} iterations, {t_solve / max(r.its, 1) * 1e3:.1f} ms/iteration);"
                      f" preconditioner set-up {t_setup:.2f} s apart",
            "its": r.its, "reason": r.reason, "gpu_its": gpu_its, "gpu_reason": gpu_reason,
            "iterate_gap_vs_gpu": gap}


def bench_mixed(lib, which: str, cpu: bool = False):
    """Mixed precision against plain FP64 on the GPU: the FP64 drop-in solve
    (time to rtol 1e-10 and its true residual), then the mixed solve to an
    FP64 relative residual of 1e-10 with exact sweeps and with Jacobi-pass
    sweeps, each with its error against the FP64 solution and beside the
    same solver run in FP64 (its control)."""
    from mojopetsc4py import PETSc
    from mpk import workloads as W
    from mpk.mixed import MixedSolver
    if which == "c3":
        csr, b = W.convection_diffusion_3d(128)
        desc = "convection_diffusion_3d 128^3 upwind, Peclet (10, 5, 2.5); gmres(30) + ilu(0)"
        ksp_t, pck, omega, inner, runs = "gmres", "ilu", 1.0, "gmres", (("fp32", "exact", 1e-3), ("fp32", 8, 1e-2),
                                                                        ("fp16", 8, 1e-2))
    else:
        csr, b = W.random_spd(1000000)
        desc = "random SPD n=1e6, 12 upper offsets in 1..5000 mirrored, diag 1.05 |row|; bcgs + sor(1.2, symmetric)"
        ksp_t, pck, omega, inner, runs = "bcgs", "sor", 1.2, "bcgs", (("fp16", "exact", 1e-2), ("fp16", 2, 1e-2),
                                                                       ("fp32", 2, 1e-2))
    A = _mat(csr)
    bv = PETSc.Vec().createWithArray(b)
    pc = PETSc.PC().create()
    pc.setType(pck)
    if pck == "sor":
        pc.setSOROmega(omega)
    pc.setOperators(A)
    ksp = PETSc.KSP().create()
    ksp.setPC(pc)
    ksp.setType(ksp_t)
    ksp.setOperators(A)
    ksp.setTolerances(rtol=1e-10, max_it=20000)
    ksp.setUp()
    x = bv.duplicate()
    t64 = timed_solves(ksp, bv, x, 1, 1)
    x64 = x.getArray().copy()
    its64 = ksp.getIterationNumber()
    out = {"workload": desc, "fp64": {"seconds": t64, "its": its64,
                                      "reason": ksp.getConvergedReason(), "true_relres": true_relres(A, bv, x)}}
    if cpu:
        if which == "c3":
            out["cpu_baseline"] = cpu_sample(csr, b, "gmres", "ilu", its64, out["fp64"]["reason"], x64, levels=0,
                                             restart=30)
        else:
            out["cpu_baseline"] = cpu_sample(csr, b, "bcgs", "sor", its64, out["fp64"]["reason"], x64, omega=omega)
    def run(prec, sweeps, irt):
        ms = MixedSolver(A, pc, precision=prec, inner=inner, restart=30, sweeps=sweeps)
        xm = bv.duplicate()
        ms.solve(bv, xm, rtol=1e-10, inner_rtol=irt)
        lib.mpk_synchronize()
        t0 = time.perf_counter()
        st = ms.solve(bv, xm, rtol=1e-10, inner_rtol=irt)
        lib.mpk_synchronize()
        t = time.perf_counter() - t0
        err = float(np.linalg.norm(xm.getArray() - x64) / np.linalg.norm(x64))
        ms.destroy()
        return {"seconds": t, "speedup_vs_fp64": t64 / t, "outer": st["outer"], "inner": st["inner"],
                "converged": st["converged"], "fp64_relres": st["relres"], "true_relres": true_relres(A, bv, xm),
                "err_vs_fp64_solution": err, "inner_rtol": irt}

    # the FP64 control of each preconditioner form: the same solver (inner
    # Krylov method, sweeps, refinement, inner tolerance) on FP64 vectors and
    # values, so that its time over a low-precision run's is the precision's
    # own gain ("speedup_vs_fp64" compares with the plain FP64 drop-in solve)
    controls = {}
    for prec, sweeps, irt in runs:
        if (sweeps, irt) not in controls:
            controls[(sweeps, irt)] = run("fp64", sweeps, irt)
            out[f"fp64_{inner}_sweeps_{sweeps}"] = controls[(sweeps, irt)]
        r = run(prec, sweeps, irt)
        r["fp64_same_solver_seconds"] = controls[(sweeps, irt)]["seconds"]
        r["speedup_vs_fp64_same_solver"] = controls[(sweeps, irt)]["seconds"] / r["seconds"]
        out[f"{prec}_{inner}_sweeps_{sweeps}"] = r
    return out


def bench_batched(lib, sample: int = 16, cpu: bool = False):
    """4096 independent 2-D diffusion systems (n log-uniform in [200, 20000],
    lognormal faces; even: cg + jacobi, odd: gmres(20) + ilu(0)), FP32 inner +
    FP64 refinement to 1e-8, one persistent launch replayed from a HIP graph
    (workgroup size and grid autotuned), against the same systems solved one
    at a time through the FP64 drop-in path (a sample, extrapolated)."""
    from mojopetsc4py import PETSc
    from mpk import workloads as W
    from mpk.batch import BatchSolver
    systems = W.batch_systems(count=4096)
    bs = BatchSolver(systems)
    # workgroup and grid, then the tiles (rows per wavefront, product variant,
    # reduction tiling) at the fastest workgroup size, timed on every 4th system
    tune = bs.autotune(candidates=((256, 512), (256, 1024), (512, 256), (512, 512)), sample_every=4)
    bs.solve(graph=True)
    t0 = time.perf_counter()
    _, st = bs.solve(graph=True)
    t = time.perf_counter() - t0
    rng = np.random.default_rng(0)
    tt = 0.0
    picks = rng.choice(len(systems), size=sample, replace=False)
    for s in picks:
        ip, ix, va, b, kind = systems[s]
        A = _mat((ip, ix, va))
        pc = PETSc.PC().create()
        pc.setType("jacobi" if kind == "cg+jacobi" else "ilu")
        pc.setOperators(A)
        ksp = PETSc.KSP().create()
        ksp.setPC(pc)
        ksp.setType("cg" if kind == "cg+jacobi" else "gmres")
        if kind != "cg+jacobi":
            ksp.setGMRESRestart(20)
        ksp.setOperators(A)
        ksp.setTolerances(rtol=1e-9, max_it=20000)
        ksp.setUp()
        bv = PETSc.Vec().createWithArray(b)
        xv = bv.duplicate()
        lib.mpk_synchronize()
        t1 = time.perf_counter()
        ksp.solve(bv, xv)
        lib.mpk_synchronize()
        tt += time.perf_counter() - t1
    per = tt / sample
    bs.destroy()
    cpu_line = None
    if cpu:
        # the oracle's FP64 solves of the same sampled systems, one core
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        O.set_mode("ref")
        tc = 0.0
        for s in picks:
            ip, ix, va, b, kind = systems[s]
            A = O.CSR(ip, ix, va)
            t0 = time.perf_counter()
            if kind == "cg+jacobi":
                O.ksp_solve(A, b, "cg", O.PCSetup(A, "jacobi"), rtol=1e-9, max_it=20000)
            else:
                O.ksp_solve(A, b, "gmres", O.PCSetup(A, "ilu", 0), rtol=1e-9, max_it=20000, restart=20)
            tc += time.perf_counter() - t0
        pc_ = tc / len(picks)
        cpu_line = {"value": 1.0 / pc_, "unit": "solves/s", "cores": 1, "kind": "port",
                    "sample": f"oracle FP64 solves (with set-up) of the {len(picks)} systems the one-at-a-time"
                              f" line solves, on one core, {pc_ * 1e3:.1f} ms each",
                    "vs_one_at_a_time": pc_ / per}
    return {"workload": "4096 x 2-D 5-point diffusion, n log-uniform [200, 20000], cg+jacobi / gmres(20)+ilu(0),"
                        " FP32 inner + FP64 refinement to 1e-8",
            "sum_n": int(sum(len(s_[0]) - 1 for s_ in systems)), "seconds": t, "solves_per_s": len(systems) / t,
            "converged": int(st[:, 3].sum()), "max_relres": float(st[:, 0].max()),
            "autotuned": {k: tune[k] for k in ("workgroup", "grid", "rows_per_wavefront", "spmv_variant",
                                               "reduction_tiling", "sampled_systems", "timings")},
            "one_at_a_time_fp64": {"sample": sample, "ms_per_solve": per * 1e3, "solves_per_s": 1.0 / per},
            "speedup": per * len(systems) / t, "cpu_baseline": cpu_line}


# The reference's published small-system solve times (mojo-petsc4py column,
# README.md:302-322; b = normal(seed 5), x0 = 0, rtol 1e-10, atol 1e-50,
# divtol 1e30, max_it 2000, bench/bench.py:210-228), in ms, with its
# iteration counts.
PUBLISHED_SMALL = {
    ("cg", "none", 32): (0.895, 113), ("cg", "none", 64): (7.597, 225), ("cg", "none", 128): (65.590, 434),
    ("gmres", "none", 32): (3.580, 232), ("gmres", "none", 64): (57.955, 626),
    ("gmres", "none", 128): (720.996, 2000),
    ("fgmres", "none", 32): (3.233, 232), ("fgmres", "none", 64): (47.752, 626),
    ("fgmres", "none", 128): (693.121, 2000),
    ("bcgs", "none", 32): (1.249, 85), ("bcgs", "none", 64): (10.906, 162), ("bcgs", "none", 128): (91.673, 316),
    ("cg", "ilu", 32): (0.784, 40), ("cg", "ilu", 64): (6.067, 77), ("cg", "ilu", 128): (48.259, 146),
    ("cg", "icc", 32): (0.828, 40), ("cg", "icc", 64): (6.008, 77), ("cg", "icc", 128): (48.678, 146),
    ("gmres", "jacobi", 32): (4.485, 232), ("gmres", "jacobi", 64): (51.261, 626),
}


def bench_small(lib, cpu: bool = False, reps: int = 3):
    """Small systems: config 1 (64^2 5-point Laplacian, CG + ILU(0) natural,
    rtol 1e-10, max_it 1000) with b = ones and b = normal(seed 5) -- GPU time
    (the mean of `reps` solves after a first one, which is timed with the
    set-up apart), its, reason, the iterate against the oracle in both
    reduction orders -- then the reference's published 32^2 / 64^2 / 128^2
    solves timed the same way."""
    from mojopetsc4py import PETSc
    from mpk import workloads as W
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    def solve(k, ksp_t, pc_t, b, tol):
        csr = W.laplacian_2d(k)
        A = _mat(csr)
        pc = PETSc.PC().create()
        pc.setType(pc_t)
        pc.setOperators(A)
        ksp = PETSc.KSP().create()
        ksp.setPC(pc)
        ksp.setType(ksp_t)
        ksp.setOperators(A)
        ksp.setTolerances(**tol)
        t0 = time.perf_counter()
        ksp.setUp()
        bv = PETSc.Vec().createWithArray(b)
        x = bv.duplicate()
        ksp.solve(bv, x)
        lib.mpk_synchronize()
        first = time.perf_counter() - t0
        best = timed_solves(ksp, bv, x, reps, 0)
        return csr, x.getArray().copy(), ksp.getIterationNumber(), ksp.getConvergedReason(), best, first

    small0 = lib.mpk_small_solves()
    out = {"config1": {}, "published": []}
    for rhs in ("ones", "normal5"):
        n = 64 * 64
        b = np.ones(n) if rhs == "ones" else np.random.default_rng(5).normal(size=n)
        csr, x, its, reason, best, first = solve(64, "cg", "ilu", b, dict(rtol=1e-10, max_it=1000))
        line = {"rhs": rhs, "gpu_ms": best * 1e3, "first_solve_with_setup_ms": first * 1e3, "its": its,
                "reason": reason}
        if rhs == "normal5":
            line["published_ms"] = PUBLISHED_SMALL[("cg", "ilu", 64)][0]
            line["vs_published"] = line["published_ms"] / line["gpu_ms"]
        A_o = O.CSR(*csr)
        O.set_mode("dev")
        r = O.ksp_solve(A_o, b, "cg", O.PCSetup(A_o, "ilu", 0), rtol=1e-10, max_it=1000)
        line["bitwise_vs_oracle_dev_order"] = bool(np.array_equal(r.x.view(np.int64), x.view(np.int64))
                                                   and r.its == its and r.reason == reason)
        O.set_mode("ref")
        t0 = time.perf_counter()
        r = O.ksp_solve(A_o, b, "cg", O.PCSetup(A_o, "ilu", 0), rtol=1e-10, max_it=1000)
        t_ref = time.perf_counter() - t0
        line["oracle_ref_order"] = {"its": r.its, "reason": r.reason,
                                    "iterate_gap": float(np.max(np.abs(r.x - x)) / np.max(np.abs(r.x)))}
        if cpu:
            line["cpu_baseline"] = {"value": t_ref * 1e3, "unit": "ms", "cores": 1, "kind": "port",
                                    "sample": "the oracle's full CG + ILU(0) solve (reference order, set-up"
                                              " included) on one core"}
        out["config1"][rhs] = line
    tol = dict(rtol=1e-10, atol=1e-50, divtol=1e30, max_it=2000)
    for (ksp_t, pc_t, k), (pub_ms, pub_its) in PUBLISHED_SMALL.items():
        b = np.random.default_rng(5).normal(size=k * k)
        _, _, its, reason, best, _ = solve(k, ksp_t, pc_t, b, tol)
        out["published"].append({"case": f"{ksp_t}+{pc_t} {k}x{k}", "gpu_ms": best * 1e3, "its": its,
                                 "published_ms": pub_ms, "published_its": pub_its, "speedup": pub_ms / (best * 1e3)})
    out["one_launch_solves"] = lib.mpk_small_solves() - small0
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def kernels_only(lib, A, bv, csr, n, nnz, reps: int = 30):
    """Each Vec / Mat / PC operation of the path, `reps` back-to-back calls on
    the 128^3 operator and its vectors (for a kernel-trace run: the per-kernel
    durations are the measurement), host time per call printed beside it."""
    from mojopetsc4py import PETSc
    x, y, z, w = (bv.duplicate() for _ in range(4))
    for v, c in ((x, 1.0), (y, 0.5), (z, 0.25), (w, 0.125)):
        v.set(c)
    vecs = [bv.duplicate() for _ in range(8)]
    for k, v in enumerate(vecs):
        v.set(1.0 / (k + 1))
    alphas = [0.5 ** k for k in range(8)]
    ops = {
        "mult": lambda: A.mult(x, y),
        "multTranspose": lambda: A.multTranspose(x, y),
        "dot": lambda: x.dot(y),
        "norm": lambda: x.norm(),
        "axpy": lambda: y.axpy(0.5, x),
        "aypx": lambda: y.aypx(0.5, x),
        "waxpy": lambda: w.waxpy(0.5, x, y),
        "axpbypcz": lambda: z.axpbypcz(0.5, 0.25, 0.125, x, y),
        "pointwiseMult": lambda: w.pointwiseMult(x, y),
        "maxpy8": lambda: y.maxpy(alphas, vecs),
    }
    out = {}
    for kind in ("jacobi", "sor", "icc", "ilu"):
        pc = PETSc.PC().create()
        pc.setType(kind)
        pc.setOperators(A)
        pc.setUp()
        ops["pc_" + kind] = (lambda pc=pc: pc.apply(x, y))
    for name, op in ops.items():
        op()
        lib.mpk_synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            op()
        lib.mpk_synchronize()
        out[name] = (time.perf_counter() - t0) / reps * 1e6
    # the fused CG + Jacobi kernels: 30 iterations of the solve
    ksp = make_ksp(A, "jacobi")
    ksp.setTolerances(rtol=1e-10, max_it=30)
    t0 = time.perf_counter()
    ksp.solve(bv, x)
    lib.mpk_synchronize()
    out["cg_jacobi_30_its"] = (time.perf_counter() - t0) * 1e6
    # 10 copies of 2^24 doubles (128 MiB read, 128 MiB written each), the
    # kernel of known traffic a counter run scales its readings by
    from mpk import _native
    from mpk.device import DeviceBuffer
    m = 1 << 24
    src, dst = DeviceBuffer(8 * m), DeviceBuffer(8 * m)
    lib.mp_vec_set(src.ptr, 1.0, m)
    for _ in range(10):
        lib.mp_vec_copy(src.ptr, dst.ptr, m)
    lib.mpk_synchronize()
    _native.check()
    print(json.dumps({"n": n, "nnz": nnz, "reps": reps, "host_us_per_call": out, "copy_bytes_each_way": 8 * m,
                      "copy_launches": 10}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="cells per dimension")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-icc", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision comparisons")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched solve")
    ap.add_argument("--no-small", action="store_true", help="skip the small-system solves (config 1 and the"
                    " reference's published 32^2-128^2 table)")
    ap.add_argument("--no-large", action="store_true", help="skip the SpMV timing at 256^3 (past the cache)")
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the SpMV timing loop and a copy of known size (for hardware counter runs)")
    ap.add_argument("--kernels", action="store_true",
                    help="only each Vec / Mat / PC operation, back to back (for kernel-trace runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (its sharding is a design"
              " only); run with --gpus 1", file=sys.stderr)
        return 2

    from mpk import _native
    lib = _native.lib()
    dev_name = _native.device_name()
    N = args.n
    t_build = time.perf_counter()
    A, bv, csr, b = build_system(N)
    n = A.getSize()[0]
    nnz = int(csr[0][-1])
    x = bv.duplicate()
    if args.spmv_only:
        return spmv_only(lib, A, bv, n, nnz, args.spmv_reps)
    if args.kernels:
        return kernels_only(lib, A, bv, csr, n, nnz)

    ksp = make_ksp(A, "jacobi")
    t_solve = timed_solves(ksp, bv, x, args.steps, args.warmup)
    its, reason = ksp.getIterationNumber(), ksp.getConvergedReason()
    relres = true_relres(A, bv, x)
    _progress(f"cg+jacobi {N}^3: {t_solve * 1e3:.2f} ms, {its} its")

    # SpMV: average launch time, HIP events on the library stream
    d = A._dev()
    xs = bv.duplicate()
    xs.set(1.0)
    ys = bv.duplicate()
    xa, ya = xs._d(), ys._dw()
    so, sv, sc = d.sell()
    lib.mpk_time_spmv(xa, ya, d.rp.ptr, d.ci.ptr, d.v.ptr, n, so.ptr, sv.ptr, sc.ptr, 5)
    spmv_ms = lib.mpk_time_spmv(xa, ya, d.rp.ptr, d.ci.ptr, d.v.ptr, n, so.ptr, sv.ptr, sc.ptr, args.spmv_reps)
    row_ms = lib.mpk_time_spmv(xa, ya, d.rp.ptr, d.ci.ptr, d.v.ptr, n, 0, 0, 0, args.spmv_reps)
    # the solve's dominant kernel: the CG product with its p.w partials, 681
    # launches a solve; the same algorithmic bytes as the product
    lib.mpk_time_cg_wdot(xa, ya, d.rp.ptr, n, so.ptr, sv.ptr, sc.ptr, 5)
    wdot_ms = lib.mpk_time_cg_wdot(xa, ya, d.rp.ptr, n, so.ptr, sv.ptr, sc.ptr, args.spmv_reps)
    _native.check()
    sbytes, moved = algo_bytes(n, nnz), spmv_bytes(n, nnz, so)
    spmv_gbs = sbytes / (spmv_ms * 1e-3) / 1e9
    wdot_gbs = sbytes / (wdot_ms * 1e-3) / 1e9
    cols32 = spmv_cols32(lib, A, bv, n, nnz, args.spmv_reps)
    _native.check()
    _progress(f"spmv {spmv_ms * 1e3:.1f} us, cg product {wdot_ms * 1e3:.1f} us")

    extra = {}
    if not args.no_icc:
        ksp2 = make_ksp(A, "icc")
        x2 = bv.duplicate()
        t_icc = timed_solves(ksp2, bv, x2, 1, 1)
        extra["cg_icc0"] = {"seconds": t_icc, "its": ksp2.getIterationNumber(),
                            "reason": ksp2.getConvergedReason(), "true_relres": true_relres(A, bv, x2)}
        _progress(f"cg+icc(0): {t_icc * 1e3:.2f} ms")

    mixed = {}
    if not args.no_mixed:
        mixed["convection_diffusion_128"] = bench_mixed(lib, "c3", cpu=not args.no_cpu)
        _progress("mixed precision: convection-diffusion done")
        mixed["random_spd_1e6"] = bench_mixed(lib, "c4", cpu=not args.no_cpu)
        _progress("mixed precision: random SPD done")
    batched = None if args.no_batched else bench_batched(lib, cpu=not args.no_cpu)
    _progress("batched done")
    small = None if args.no_small else bench_small(lib, cpu=not args.no_cpu)
    _progress("small systems done")

    past = None
    if not args.no_large and N == 128:
        del d, xs, ys
        past = spmv_past_cache(lib)
        _progress("256^3 product done")

    cpu = None
    if not args.no_cpu:
        dt, c_its, c_reason, c_rnorm, gap2, gapm = cpu_baseline(csr, b, x.getArray().copy())
        cpu = {"value": dt, "unit": "s", "cores": 1, "kind": "port",
               "sample": f"the oracle's full CG+Jacobi solve to rtol 1e-10 at {N}^3 in the reference's reduction"
                         f" order, on one core of {cpu_model()} ({dt / max(c_its, 1) * 1e3:.1f} ms/iteration)",
               "its": c_its, "reason": c_reason, "rnorm": c_rnorm, "gpu_its": its, "gpu_reason": reason,
               "iterate_gap_vs_gpu": gap2, "iterate_gap_vs_gpu_max_norm": gapm}

    out = {
        "metric": "time to rel. residual 1e-10 + SpMV HBM GB/s; mixed/batched vs plain FP64",
        "value": t_solve,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_solve * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"diffusion_3d {N}^3 7-point lognormal(0,1), CG+Jacobi rtol 1e-10, FP64 int32 CSR",
                   "n": n, "nnz": nnz, "its": its, "reason": reason, "true_relres": relres,
                   "device": dev_name},
        "roofline": {"bound": "hbm", "kernel": "k_cg_wdot", "achieved": wdot_gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": wdot_gbs / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC.get(("k_cg_wdot", n, nnz)),
                     "bytes_per_launch": sbytes, "bytes_moved": moved, "avg_launch_ms": wdot_ms,
                     "moved_gbs": moved / (wdot_ms * 1e-3) / 1e9,
                     "note": f"the CG product (w = A p with the tile partials of p.w), the dominant kernel of the"
                             f" solve, timed over back-to-back launches; achieved = the FP64 int32-CSR product's"
                             f" algorithmic {sbytes / 1e6:.1f} MB per launch over its time; the sliced copy's 16-bit"
                             f" column offsets move {moved / 1e6:.1f} MB (moved_gbs); the working set stays in the"
                             " 256 MB last-level cache across launches (past_cache: the same kernels on a 1.74 GB"
                             " matrix; cols32: the product over int32 columns)",
                     "spmv": {"kernel": "k_spmv_sell", "avg_launch_ms": spmv_ms, "achieved": spmv_gbs,
                              "frac": spmv_gbs / HBM_PEAK_GBS,
                              "traffic": MEASURED_TRAFFIC.get(("k_spmv_sell", n, nnz)),
                              "row_csr_kernel_ms": row_ms},
                     "cols32": cols32,
                     "past_cache": past},
        "cpu_baseline": cpu,
        "solves": extra,
        "mixed_precision": mixed,
        "batched": batched,
        "small_systems": small,
        "setup_s": time.perf_counter() - t_build,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
