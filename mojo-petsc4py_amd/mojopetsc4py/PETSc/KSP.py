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
from dataclasses import dataclass

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


@dataclass
class _KSPOptions:
    """What the solver's setters record; the defaults are PETSc's (KSP.py:
    25-29, 123-158 of the reference)."""

    ksp_type: str = KSPType.GMRES
    side: int = 0               # PC side: left
    normtype: int = NormType_.NORM_DEFAULT
    rtol: float = DEFAULT_RTOL
    atol: float = DEFAULT_ATOL
    divtol: float = DEFAULT_DTOL
    max_it: int = DEFAULT_MAX_IT
    min_it: int = 0
    chknorm: int = -1
    convmaxits: int = 0         # set by setTolerances(max_it=...): max_it then reports CONVERGED_ITS
    guess_zero: int = 1
    restart: int = GMRES_DEFAULT_MAXK
    haptol: float = 1.0e-30
    breakdowntol: float = 0.1
    orthog: int = GMRESOrthogonalizationType.CLASSICAL
    cgstype: int = GMRESCGSRefinementType.REFINE_NEVER


# The scalar blocks the device drivers read (src/ksp.mojo:9-42): slot -> option.
_CFG_SLOTS = ((0, "rtol"), (1, "atol"), (2, "divtol"), (3, "haptol"), (4, "breakdowntol"))
_CFGI_SLOTS = ((0, "normtype"), (1, "side"), (2, "guess_zero"), (3, "min_it"), (4, "chknorm"),
               (5, "convmaxits"), (6, "max_it"), (7, "restart"), (8, "orthog"), (9, "cgstype"))


def _setter(field, convert, resets_setup=False):
    def set_option(self, value):
        setattr(self._opt, field, convert(value))
        if resets_setup:
            self._setupcalled = False

    return set_option


def _getter(field):
    def get_option(self):
        return getattr(self._opt, field)

    return get_option


class _Result:
    """What the last solve reported."""

    __slots__ = ("its", "rnorm", "reason")

    def __init__(self):
        self.its, self.rnorm, self.reason = 0, 0.0, ConvergedReason.CONVERGED_ITERATING


class KSP:
    """A Krylov solver for a sequential AIJ matrix, iterating on the GPU."""

    Type = KSPType
    ConvergedReason = ConvergedReason
    NormType = NormType_
    GMRESCGSRefinementType = GMRESCGSRefinementType
    GMRESOrthogonalizationType = GMRESOrthogonalizationType

    def __init__(self):
        self._opt = _KSPOptions()
        self._amat = self._pmat = None
        self._pc = PC()
        self._last = _Result()
        self._history = None  # the caller's history buffer (setConvergenceHistory)
        self._rhs = self._sol = None
        self._setupcalled = False
        self._arena = {}      # device work blocks, reused across solves

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
        self._opt.ksp_type = ksp_type
        self._setupcalled = False

    getType = _getter("ksp_type")

    @property
    def _type(self):
        return self._opt.ksp_type

    def setOperators(self, A: Mat = None, P: Mat = None) -> None:
        self._amat, self._pmat = A, (A if P is None else P)
        self._pc.setOperators(self._amat, self._pmat)
        self._setupcalled = False

    def getOperators(self):
        return (self._amat, self._pmat)

    def setPC(self, pc: PC) -> None:
        self._pc = pc
        if self._pmat is not None:
            pc.setOperators(self._amat, self._pmat)
        self._setupcalled = False

    def getPC(self) -> PC:
        return self._pc

    def setUp(self) -> None:
        """Set up the preconditioner (the drivers need nothing else)."""
        if self._amat is None:
            raise RuntimeError("KSP.setOperators() has not been called")
        self._pc.setUp()
        self._setupcalled = True

    # ---------------------------------------------------------- options
    setPCSide = _setter("side", int)
    getPCSide = _getter("side")
    setNormType = _setter("normtype", int)
    getNormType = _getter("normtype")
    setMinimumIterations = _setter("min_it", int)
    getMinimumIterations = _getter("min_it")
    setCheckNorm = _setter("chknorm", int)
    setGMRESRestart = _setter("restart", int, resets_setup=True)
    setGMRESHapTol = _setter("haptol", float)
    setGMRESBreakdownTolerance = _setter("breakdowntol", float)
    setGMRESOrthogonalization = _setter("orthog", int)
    setGMRESCGSRefinementType = _setter("cgstype", int)

    def setTolerances(self, rtol=None, atol=None, divtol=None, max_it=None) -> None:
        """Set any of the four; giving max_it also makes reaching it
        CONVERGED_ITS, as in the reference (KSP.py:212-225)."""
        given = {"rtol": rtol, "atol": atol, "divtol": divtol}
        for name, value in given.items():
            if value is not None:
                setattr(self._opt, name, float(value))
        if max_it is not None:
            self._opt.max_it = int(max_it)
            self._opt.convmaxits = 1

    def getTolerances(self):
        o = self._opt
        return (o.rtol, o.atol, o.divtol, o.max_it)

    def setInitialGuessNonzero(self, flag) -> None:
        self._opt.guess_zero = 0 if flag else 1

    def getInitialGuessNonzero(self) -> bool:
        return self._opt.guess_zero == 0

    def setConvergenceHistory(self, length: int | None = None, reset: bool = False) -> None:
        """Record the residual norms of the next solves (10000 entries unless
        a positive length is given)."""
        size = int(length) if (length is not None and length is not True and int(length) >= 0) else 10000
        self._history = np.zeros(size, dtype=np.float64)

    def getConvergenceHistory(self) -> np.ndarray:
        return np.zeros(0, dtype=np.float64) if self._history is None else self._history

    # the last solve's outcome, and its operands
    def getIterationNumber(self) -> int:
        return self._last.its

    def getResidualNorm(self) -> float:
        return self._last.rnorm

    def getConvergedReason(self):
        return self._last.reason

    def getRhs(self) -> Vec:
        return self._rhs

    def getSolution(self) -> Vec:
        return self._sol

    # ---------------------------------------------------------- scratch
    def _dev(self, name: str, ndoubles: int) -> int:
        """A device arena of at least `ndoubles`, reused across solves."""
        need = 8 * max(int(ndoubles), 1)
        buf = self._arena.get(name)
        if buf is None or buf.nbytes < need:
            buf = self._arena[name] = DeviceBuffer(need)
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
        if self._opt.guess_zero:
            x.zeroEntries()
        self._pc.setUp()
        self._rhs, self._sol = b, x
        if self._opt.normtype == NormType_.NORM_DEFAULT:  # resolved per solve (KSP.py:451-455)
            self._opt.normtype = NormType_.NORM_PRECONDITIONED
        cfg, cfgi = self._state(n)
        d = self._amat._dev()
        s = d.sell()  # None: the solver keeps the CSR row kernel
        if s is not None:
            cfgi[13], cfgi[14], cfgi[15] = s[0].ptr, s[1].ptr, s[2].ptr
        pci, pcf = self._pc._pci_block(), self._pc._pcf_block()
        mat = (d.rp.ptr, d.ci.ptr, d.v.ptr, pci.ctypes.data, pcf.ctypes.data, 0)
        ends = (cfg.ctypes.data, cfgi.ctypes.data, b._d(), x._dw())
        L = lib()
        kind = self._opt.ksp_type
        if kind in (KSPType.GMRES, KSPType.FGMRES):
            mk = self._opt.restart
            hs = np.zeros((mk + 2) * (mk + 1) + (mk + 1) * (mk + 1) + 2 * (mk + 1) + 2 * (mk + 2))
            basis = (self._dev("work", 2 * n), self._dev("vv", (mk + 1) * n))
            if kind == KSPType.GMRES:
                L.mp_ksp_solve_gmres(n, *mat, *basis, hs.ctypes.data, *ends)
            else:
                L.mp_ksp_solve_fgmres(n, *mat, *basis, self._dev("prevec", (mk + 2) * n), hs.ctypes.data, *ends)
        else:
            solver = L.mp_ksp_solve_cg if kind == KSPType.CG else L.mp_ksp_solve_bcgs
            solver(n, *mat, self._dev("work", 7 * n), *ends)
        check()
        self._last.its, self._last.reason, self._last.rnorm = int(cfgi[10]), int(cfgi[11]), float(cfg[7])
        if self._history is not None:  # a new array per solve, as the reference hands back
            fresh = np.zeros(len(self._history), dtype=np.float64)
            k = min(len(fresh), len(cfg) - 9)
            fresh[:k] = cfg[9:9 + k]
            self._history = fresh

    def _state(self, n: int):
        """cfg / cfgi as the drivers read them (ksp.mojo:9-42, KSP.py:457-494):
        the options in their slots, the history from cfg[9], and the
        right-hand-side norm and its flag for a nonzero guess."""
        o = self._opt
        cfg = np.zeros(16 + o.max_it + 2, dtype=np.float64)
        cfgi = np.zeros(16, dtype=np.int64)
        for slot, name in _CFG_SLOTS:
            cfg[slot] = getattr(o, name)
        for slot, name in _CFGI_SLOTS:
            cfgi[slot] = getattr(o, name)
        nonzero_guess = not o.guess_zero
        cfg[12] = self._rhs_norm() if nonzero_guess else 0.0  # read only for a nonzero guess
        cfgi[12] = int(nonzero_guess)
        return cfg, cfgi

    def _rhs_norm(self) -> float:
        """KSPConvergedDefault's right-hand-side norm (KSP.py:496-507).  The
        reference takes the two 2-norms with numpy on the host; in the
        reference reduction order this does too, so that its bits follow;
        in the tile-tree order they are the device norm."""
        b = self._rhs
        o = self._opt
        if b is None or o.normtype == NormType_.NORM_NONE:
            return 0.0

        def two_norm(v):
            return float(np.linalg.norm(v.getArray())) if reduction_order() == "ref" else v.norm()

        if o.normtype == NormType_.UNPRECONDITIONED or o.side == 1:
            return two_norm(b)
        z = b.duplicate()
        self._pc.apply(b, z)
        if o.normtype == NormType_.PRECONDITIONED:
            return two_norm(z)
        return math.sqrt(abs(z.dot(b)))
