"""Shared helpers for the parity tests: device upload/download through
libmpk and oracle-side equivalents of the C-ABI calls."""

from __future__ import annotations

import numpy as np


def dev(arr):
    """Upload to a new device buffer.  Keep the returned object alive for as
    long as its address is in use: the buffer is freed with it."""
    from mpk.device import DeviceBuffer
    return DeviceBuffer.from_array(np.ascontiguousarray(arr))


def back(buf, dtype, n):
    return buf.to_numpy(dtype, n)


def bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.int64)


def same_bits(a, b):
    return np.array_equal(bits(a), bits(b))


def rng_vec(n, seed):
    return np.random.default_rng(seed).normal(size=n)


class DevMat:
    """A matrix uploaded the way the PETSc layer uploads it (int32 CSR)."""

    def __init__(self, indptr, indices, data):
        self.n = len(indptr) - 1
        nnz = int(indptr[-1])
        self.rp = dev(np.asarray(indptr, dtype=np.int32))
        self.ci = dev(np.asarray(indices[:max(nnz, 1)], dtype=np.int32) if nnz else np.zeros(1, np.int32))
        self.v = dev(np.asarray(data[:max(nnz, 1)], dtype=np.float64) if nnz else np.zeros(1))


def wide_band(n=40000, far=35000):
    """SPD, diagonally dominant: 4 on the diagonal, -1 at distance 1 and -0.5
    at distance `far` (so entries lie up to `far` columns from their row:
    past 32767, where sliced copies keep int32 columns)."""
    from mpk import workloads as W
    rows, cols, vals = [], [], []
    for d, v in ((0, 4.0), (1, -1.0), (-1, -1.0), (far, -0.5), (-far, -0.5)):
        i = np.arange(max(0, -d), min(n, n - d))
        rows.append(i)
        cols.append(i + d)
        vals.append(np.full(len(i), v))
    return W._csr_from_coo(n, np.concatenate(rows).astype(np.int64), np.concatenate(cols).astype(np.int64),
                           np.concatenate(vals))


# ---------------------------------------------------------------- whole solves

def petsc_solve(csr, b, ksp_type, pc_type, levels=0, x0=None, **opt):
    from mojopetsc4py import PETSc
    ip, ix, va = csr
    n = len(ip) - 1
    A = PETSc.Mat().createSeqAIJ((n, n), nnz=7)
    A.setValuesCSR(ip, ix, va)
    A.assemble()
    pc = PETSc.PC().create()
    pc.setType(pc_type)
    if levels:
        pc.setFactorLevels(levels)
    if "omega" in opt:
        pc.setSOROmega(opt.pop("omega"))
    if "ordering" in opt:
        pc.setFactorOrdering(opt.pop("ordering"))
    pc.setOperators(A)
    ksp = PETSc.KSP().create()
    ksp.setPC(pc)
    ksp.setType(ksp_type)
    ksp.setOperators(A)
    tol = {k: opt.pop(k) for k in ("rtol", "atol", "divtol", "max_it") if k in opt}
    if tol:
        ksp.setTolerances(**tol)
    if "restart" in opt:
        ksp.setGMRESRestart(opt.pop("restart"))
    if "orthog" in opt:
        ksp.setGMRESOrthogonalization(opt.pop("orthog"))
    if "cgstype" in opt:
        ksp.setGMRESCGSRefinementType(opt.pop("cgstype"))
    if "side" in opt:
        ksp.setPCSide(opt.pop("side"))
    if "normtype" in opt:
        ksp.setNormType(opt.pop("normtype"))
    assert not opt, opt
    ksp.setConvergenceHistory()
    bv = PETSc.Vec().createWithArray(b)
    xv = PETSc.Vec().createSeq(n)
    if x0 is not None:
        xv.setArray(x0)
        ksp.setInitialGuessNonzero(True)
    ksp.solve(bv, xv)
    its = ksp.getIterationNumber()
    return (xv.getArray().copy(), its, ksp.getConvergedReason(), ksp.getResidualNorm(),
            ksp.getConvergenceHistory()[: its + 1].copy())


def oracle_solve(O, csr, b, ksp_type, pc_type, levels=0, x0=None, **opt):
    A = O.CSR(*csr)
    pcopt = {}
    if "omega" in opt:
        pcopt["omega"] = opt.pop("omega")
    if "ordering" in opt:
        pcopt["ordering"] = opt.pop("ordering")
    pc = O.PCSetup(A, pc_type, levels, **pcopt)
    kw = dict(opt)
    if x0 is not None:
        kw["guess_zero"] = 0
    r = O.ksp_solve(A, b, ksp_type, pc, x0=x0, **kw)
    return r.x, r.its, r.reason, r.rnorm, r.history


def compare(O, csr, b, ksp_type, pc_type, levels=0, x0=None, **opt):
    g = petsc_solve(csr, b, ksp_type, pc_type, levels, x0, **dict(opt))
    o = oracle_solve(O, csr, b, ksp_type, pc_type, levels, x0, **dict(opt))
    assert g[1] == o[1], ("its", g[1], o[1])
    assert g[2] == o[2], ("reason", g[2], o[2])
    assert same_bits(np.array([g[3]]), np.array([o[3]])), ("rnorm", g[3], o[3])
    assert same_bits(g[4], o[4][: len(g[4])]), "history"
    assert same_bits(g[0], o[0]), ("x", np.max(np.abs(g[0] - o[0])))
    return g
