"""Mixed-precision solves beside the drop-in FP64 path.

The inner Krylov iteration (GMRES or BiCGStab), its SpMV and its
preconditioner (ILU or SOR, from a set-up `PETSc.PC`) run on FP32 vectors with
matrix and preconditioner values in FP32, or in FP16 with one FP32 scale per
row; dot products and norms sum in FP64, and the GMRES least squares is FP64.
An FP64 iterative refinement around it computes the true residual b - A x and
updates x in FP64 until ||b - A x|| / ||b|| <= rtol.  precision="fp64" runs the
same inner solver, preconditioner form and refinement on FP64 vectors and
values: the control against which the lower precisions' gain is measured.

    from mpk.mixed import MixedSolver
    ms = MixedSolver(A, pc, precision="fp32", inner="gmres", restart=30)
    stats = ms.solve(b, x, rtol=1e-10)

The library side is mpk_mixed_create / mpk_mixed_solve / mpk_mixed_free.
"""

from __future__ import annotations

import numpy as np

from ._native import check, lib

_PREC = {"fp64": 64, "fp32": 32, "fp16": 16}
_INNER = {"gmres": 0, "bcgs": 1}
_REFINE = {"never": 0, "ifneeded": 1, "always": 2}


class MixedSolver:
    """FP32 / FP16 (or FP64) inner solves under FP64 iterative refinement for
    one matrix and preconditioner.  The matrix and PC must stay unchanged while the
    solver is used (it keeps low-precision copies of their values)."""

    def __init__(self, A, pc=None, precision: str = "fp32", inner: str = "gmres", restart: int = 30,
                 sweeps="exact"):
        """sweeps: "exact" runs the preconditioner's triangular / SOR sweeps in
        dependency order (the block plans); an int k approximates each sweep
        by k Jacobi passes over its entries (bandwidth-bound, unordered)."""
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}")
        if inner not in _INNER:
            raise ValueError(f"inner must be one of {sorted(_INNER)}")
        n, m = A.getSize()
        if n != m:
            raise ValueError("the matrix must be square")
        self._A = A
        self._pc = pc
        self.n = n
        self.precision = precision
        self.inner = inner
        self.restart = int(restart)
        if sweeps == "exact":
            tri = 0
        elif isinstance(sweeps, (int, np.integer)) and sweeps >= 1:
            tri = int(sweeps)
        else:
            raise ValueError('sweeps must be "exact" or a positive number of Jacobi passes')
        self.sweeps = sweeps
        d = A._dev()
        s = d.sell()
        if s is None:
            raise ValueError("the matrix is too large for the sliced copy the mixed solver reads")
        so, sv, sc = s
        if pc is not None:
            pc.setUp()
            pci = pc._pci_block()
            pcf = pc._pcf_block()
        else:
            pci = np.zeros(12, dtype=np.int64)
            pcf = np.zeros(2, dtype=np.float64)
        self._keep = (d, so, sv, sc)
        self._h = lib().mpk_mixed_create(n, d.rp.ptr, d.ci.ptr, d.v.ptr, so.ptr, sv.ptr, sc.ptr,
                                         pci.ctypes.data, pcf.ctypes.data, _PREC[precision], _INNER[inner],
                                         self.restart, tri)
        check()
        if not self._h:
            raise RuntimeError("mpk_mixed_create failed")

    def solve(self, b, x, rtol: float = 1e-10, inner_rtol: float = 1e-4, max_outer: int = 100,
              max_inner: int = 10000, refine: str = "ifneeded", zero_guess: bool = True) -> dict:
        """x ~ A^-1 b to ||b - A x|| / ||b|| <= rtol (FP64 residual).  Returns
        {"relres", "outer", "inner", "converged", "bnorm"}."""
        if self._h is None:
            raise RuntimeError("the solver was destroyed")
        if b.getSize() != self.n or x.getSize() != self.n:
            raise ValueError(f"incompatible vector size: {self.n} is needed")
        cfg = np.zeros(16, dtype=np.float64)
        cfg[:6] = [rtol, inner_rtol, max_outer, max_inner, _REFINE[refine], 1.0 if zero_guess else 0.0]
        rc = lib().mpk_mixed_solve(self._h, b._d(), x._dw(), cfg.ctypes.data)
        check()
        if rc != 0:
            raise RuntimeError("mpk_mixed_solve failed")
        return {"relres": float(cfg[8]), "outer": int(cfg[9]), "inner": int(cfg[10]),
                "converged": bool(cfg[12]), "bnorm": float(cfg[11])}

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            lib().mpk_mixed_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
