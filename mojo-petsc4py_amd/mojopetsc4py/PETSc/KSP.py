"""Krylov solvers, mirroring `petsc4py/PETSc/KSP.pyx` (the reference's
python/mojopetsc4py/PETSc/KSP.py).

Covered types: cg, gmres, fgmres, bcgs.  Defaults, setters, the scalar
blocks handed to the driver (cfg / cfgi, ksp.mojo:9-42), the convergence
test and the result read-back are the reference's; the iteration runs on the
GPU in one native call (mp_ksp_solve_*), with the work arenas in device
memory.
"""

from __future__ import annotations

import math

import numpy as np

from mpk._native import check, lib, reduction_order
from mpk.device import DeviceBuffer

from .Mat import Mat
from .PC import PC
from .Vec import Vec

DEFAULT_RTOL = 1e-5
DEFAULT_ATOL = 1e-50
DEFAULT_DTOL = 1.0e4
DEFAULT_MAX_IT = 1000
GMRES_DEFAULT_MAXK = 30


class KSPType:
    RICHARDSON = "richardson"
    CG = "cg"
    GMRES = "gmres"
    FGMRES = "fgmres"
    BCGS = "bcgs"
    MINRES = "minres"
    PREONLY = "preonly"


SUPPORTED = {KSPType.CG, KSPType.GMRES, KSPType.FGMRES, KSPType.BCGS}


class NormType_:
    NORM_DEFAULT = 0
    NORM_NONE = 1
    NORM_PRECONDITIONED = 2
    NORM_UNPRECONDITIONED = 3
    NORM_NATURAL = 4
    DEFAULT = 0
    NONE = 1
    NO = 1
    PRECONDITIONED = 2
    UNPRECONDITIONED = 3
    NATURAL = 4


class ConvergedReason:
    CONVERGED_ITERATING = 0
    ITERATING = 0
    CONVERGED_RTOL_NORMAL_EQUATIONS = 1
    CONVERGED_RTOL = 2
    CONVERGED_ATOL = 3
    CONVERGED_ITS = 4
    CONVERGED_NEG_CURVE = 5
    CONVERGED_STEP_LENGTH = 6
    CONVERGED_HAPPY_BREAKDOWN = 7
    CONVERGED_ATOL_NORMAL_EQUATIONS = 9
    DIVERGED_NULL = -2
    DIVERGED_ITS = -3
    DIVERGED_DTOL = -4
    DIVERGED_BREAKDOWN = -5
    DIVERGED_BREAKDOWN_BICG = -6
    DIVERGED_NONSYMMETRIC = -7
    DIVERGED_INDEFINITE_PC = -8
    DIVERGED_NANORINF = -9
    DIVERGED_INDEFINITE_MAT = -10
    DIVERGED_PC_FAILED = -11
    DIVERGED_INDEFINITE_PRECONDITIONER = -8
    DIVERGED_PCSETUP_FAILED = -11
    DIVERGED_MAX_IT = -3


class GMRESCGSRefinementType:
    REFINE_NEVER = 0
    REFINE_IFNEEDED = 1
    REFINE_ALWAYS = 2
    CGS_REFINE_NEVER = 0
    CGS_REFINE_IFNEEDED = 1
    CGS_REFINE_ALWAYS = 2


class GMRESOrthogonalizationType:
    CLASSICAL = 0
    MODIFIED = 1
    CGS = 0
    MGS = 1


class KSP:
    """A Krylov solver for a sequential AIJ matrix, iterating on the GPU."""

    Type = KSPType
    ConvergedReason = ConvergedReason
    NormType = NormType_
    GMRESCGSRefinementType = GMRESCGSRefinementType
    GMRESOrthogonalizationType = GMRESOrthogonalizationType

    def __init__(self):
        self._type = KSPType.GMRES
        self._amat = None
        self._pmat = None
        self._pc = PC()
        self._pc_side = 0
        self._normtype = 0
        self._rtol = DEFAULT_RTOL
        self._atol = DEFAULT_ATOL
        self._divtol = DEFAULT_DTOL
        self._max_it = DEFAULT_MAX_IT
        self._min_it = 0
        self._chknorm = -1
        self._convmaxits = 0
        self._guess_zero = 1
        self._restart = GMRES_DEFAULT_MAXK
        self._haptol = 1.0e-30
        self._breakdowntol = 0.1
        self._orthog = 0
        self._cgstype = 0
        self._its = 0
        self._rnorm = 0.0
        self._reason = 0
        self._history = None
        self._histlen = 0
        self._rhs = None
        self._sol = None
        self._setupcalled = False
        self._arena = {}

    # ---------------------------------------------------------- lifecycle
    def create(self, comm=None) -> "KSP":
        return self

    def destroy(self) -> "KSP":
        self._history = None
        self._arena = {}
        return self

    def setType(self, ksp_type) -> None:
        if isinstance(ksp_type, int):
            raise TypeError("expected a Krylov type name")
        if ksp_type not in SUPPORTED:
            raise NotImplementedError(
                f"KSP type {ksp_type!r} is not covered by mojo-petsc4py ({sorted(SUPPORTED)} are)")
        self._type = ksp_type
        self._setupcalled = False

    def getType(self) -> str:
        return self._type

    def setOperators(self, A: Mat = None, P: Mat = None) -> None:
        self._amat = A
        self._pmat = A if P is None else P
        self._pc.setOperators(self._amat, self._pmat)
        self._setupcalled = False

    def getOperators(self):
        return (self._amat, self._pmat)

    def setPC(self, pc: PC) -> None:
        self._pc = pc
        if self._pmat is not None:
            self._pc.setOperators(self._amat, self._pmat)
        self._setupcalled = False

    def getPC(self) -> PC:
        return self._pc

    def setPCSide(self, side) -> None:
        self._pc_side = int(side)

    def getPCSide(self):
        return self._pc_side

    def setNormType(self, normtype) -> None:
        self._normtype = int(normtype)

    def getNormType(self):
        return self._normtype

    def setTolerances(self, rtol=None, atol=None, divtol=None, max_it=None) -> None:
        """A max_it also makes reaching it CONVERGED_ITS, as in the
        reference (KSP.py:212-225)."""
        if rtol is not None:
            self._rtol = float(rtol)
        if atol is not None:
            self._atol = float(atol)
        if divtol is not None:
            self._divtol = float(divtol)
        if max_it is not None:
            self._max_it = int(max_it)
            self._convmaxits = 1

    def getTolerances(self):
        return (self._rtol, self._atol, self._divtol, self._max_it)

    def setMinimumIterations(self, min_it) -> None:
        self._min_it = int(min_it)

    def getMinimumIterations(self) -> int:
        return self._min_it

    def setCheckNorm(self, flag) -> None:
        self._chknorm = int(flag)

    def setInitialGuessNonzero(self, flag) -> None:
        self._guess_zero = 0 if flag else 1

    def getInitialGuessNonzero(self) -> bool:
        return not self._guess_zero

    def setGMRESRestart(self, restart) -> None:
        self._restart = int(restart)
        self._setupcalled = False

    def setGMRESHapTol(self, haptol) -> None:
        self._haptol = float(haptol)

    def setGMRESBreakdownTolerance(self, tol) -> None:
        self._breakdowntol = float(tol)

    def setGMRESOrthogonalization(self, orthog) -> None:
        self._orthog = int(orthog)

    def setGMRESCGSRefinementType(self, cgstype) -> None:
        self._cgstype = int(cgstype)

    def setConvergenceHistory(self, length: int | None = None, reset: bool = False) -> None:
        size = 10000 if length is None or length is True else int(length)
        if size < 0:
            size = 10000
        self._histlen = size
        self._history = np.zeros(size, dtype=np.float64)

    def getConvergenceHistory(self) -> np.ndarray:
        if self._history is None:
            return np.zeros(0, dtype=np.float64)
        return self._history

    def getIterationNumber(self) -> int:
        return self._its

    def getResidualNorm(self) -> float:
        return self._rnorm

    def getConvergedReason(self):
        return self._reason

    def getRhs(self) -> Vec:
        return self._rhs

    def getSolution(self) -> Vec:
        return self._sol

    def setUp(self) -> None:
        if self._amat is None:
            raise RuntimeError("KSP.setOperators() has not been called")
        self._pc.setUp()
        self._setupcalled = True

    # ---------------------------------------------------------- scratch
    def _dev(self, name: str, ndoubles: int) -> int:
        """A device arena of at least `ndoubles`, reused across solves."""
        buf = self._arena.get(name)
        need = 8 * max(int(ndoubles), 1)
        if buf is None or buf.nbytes < need:
            buf = DeviceBuffer(need)
            self._arena[name] = buf
        return buf.ptr

    # ---------------------------------------------------------- solve
    def solve(self, b: Vec, x: Vec) -> None:
        """Solve A x = b on the GPU.  With the default zero initial guess `x`
        is zeroed first, as PETSc's KSPSolve does."""
        if self._amat is None:
            raise RuntimeError("KSP.setOperators() has not been called")
        n = self._amat._nrows
        if self._amat._nrows != self._amat._ncols:
            raise ValueError("the solver matrix must be square")
        if b.getSize() != n or x.getSize() != n:
            raise ValueError(f"incompatible vector size: {n} is needed for both operands")
        if not self._setupcalled:
            self.setUp()
        if self._guess_zero:
            x.zeroEntries()
        self._pc.setUp()
        self._rhs = b
        self._sol = x
        self._resolve_normtype()
        cfg, cfgi = self._state(n)
        d = self._amat._dev()
        s = d.sell()  # None: the solver keeps the CSR row kernel
        if s is not None:
            cfgi[13], cfgi[14], cfgi[15] = s[0].ptr, s[1].ptr, s[2].ptr
        pci = self._pc._pci_block()
        pcf = self._pc._pcf_block()
        L = lib()
        b_addr = b._d()
        x_addr = x._dw()
        if self._type in (KSPType.GMRES, KSPType.FGMRES):
            mk = self._restart
            hs = np.zeros((mk + 2) * (mk + 1) + (mk + 1) * (mk + 1) + 2 * (mk + 1) + 2 * (mk + 2))
            w = self._dev("work", 2 * n)
            vv = self._dev("vv", (mk + 1) * n)
            if self._type == KSPType.GMRES:
                L.mp_ksp_solve_gmres(n, d.rp.ptr, d.ci.ptr, d.v.ptr, pci.ctypes.data, pcf.ctypes.data, 0, w,
                                     vv, hs.ctypes.data, cfg.ctypes.data, cfgi.ctypes.data, b_addr, x_addr)
            else:
                pv = self._dev("prevec", (mk + 2) * n)
                L.mp_ksp_solve_fgmres(n, d.rp.ptr, d.ci.ptr, d.v.ptr, pci.ctypes.data, pcf.ctypes.data, 0, w,
                                      vv, pv, hs.ctypes.data, cfg.ctypes.data, cfgi.ctypes.data, b_addr, x_addr)
        else:
            w = self._dev("work", 7 * n)
            fn = L.mp_ksp_solve_cg if self._type == KSPType.CG else L.mp_ksp_solve_bcgs
            fn(n, d.rp.ptr, d.ci.ptr, d.v.ptr, pci.ctypes.data, pcf.ctypes.data, 0, w, cfg.ctypes.data,
               cfgi.ctypes.data, b_addr, x_addr)
        check()
        self._its = int(cfgi[10])
        self._reason = int(cfgi[11])
        self._rnorm = float(cfg[7])
        full = np.zeros(self._histlen, dtype=np.float64)
        k = min(len(full), len(cfg) - 9)
        full[:k] = cfg[9:9 + k]
        self._history = full

    def _resolve_normtype(self):
        if self._normtype == NormType_.NORM_DEFAULT:
            self._normtype = NormType_.NORM_PRECONDITIONED

    def _state(self, n: int):
        """The scalar blocks of ksp.mojo:9-42 (KSP.py:457-494)."""
        size = 16 + self._max_it + 2
        cfg = np.zeros(size, dtype=np.float64)
        cfg[0] = self._rtol
        cfg[1] = self._atol
        cfg[2] = self._divtol
        cfg[3] = self._haptol
        cfg[4] = self._breakdowntol
        cfgi = np.zeros(16, dtype=np.int64)
        cfgi[0] = self._normtype
        cfgi[1] = self._pc_side
        cfgi[2] = self._guess_zero
        cfgi[3] = self._min_it
        cfgi[4] = self._chknorm
        cfgi[5] = self._convmaxits
        cfgi[6] = self._max_it
        cfgi[7] = self._restart
        cfgi[8] = self._orthog
        cfgi[9] = self._cgstype
        # the right-hand-side norm seeds the test only for a nonzero guess
        cfg[12] = 0.0 if self._guess_zero else self._rhs_norm()
        cfg[14] = 0.0
        cfgi[12] = 0 if self._guess_zero else 1
        return cfg, cfgi

    def _rhs_norm(self) -> float:
        """KSPConvergedDefault's right-hand-side norm (KSP.py:496-507).  The
        reference takes the two 2-norms with numpy on the host; in the
        reference reduction order this does too, so that its bits follow;
        in the tile-tree order they are the device norm."""
        b = self._rhs
        if b is None or self._normtype == NormType_.NORM_NONE:
            return 0.0

        def two_norm(v):
            if reduction_order() == "ref":
                return float(np.linalg.norm(v.getArray()))
            return v.norm()

        if self._normtype == NormType_.UNPRECONDITIONED or self._pc_side == 1:
            return two_norm(b)
        z = b.duplicate()
        self._pc.apply(b, z)
        if self._normtype == NormType_.PRECONDITIONED:
            return two_norm(z)
        return math.sqrt(abs(z.dot(b)))
