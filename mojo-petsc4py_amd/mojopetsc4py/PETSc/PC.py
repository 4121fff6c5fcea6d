"""Preconditioners, mirroring `petsc4py/PETSc/PC.pyx` (the reference's
python/mojopetsc4py/PETSc/PC.py).

Covered types: none, jacobi, sor, ilu, icc; every other name raises
NotImplementedError, as in the reference.  Set up follows the reference's
protocol (PC.py:315-492): Jacobi and SOR set up on the GPU; ILU and IC build
their pattern and values on the host with libmpk's set-up routines (the
ILU(0) natural-order fast path, the level-of-fill path with its contiguous
pass and capacity retry, IC always through the level-of-fill path, c = r),
then upload the factor (int32 indices) for the device sweeps.
"""

from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from mpk._native import check, lib
from mpk.device import DeviceBuffer

from .Mat import Mat, MatOrderingType, from_dense
from .Vec import Vec

_NONE = 0
_JACOBI = 2
_SOR = 4
_ILU = 5
_ICC = 6

SOR_ZERO_INITIAL_GUESS = 16
SOR_FORWARD_SWEEP = 1
SOR_BACKWARD_SWEEP = 2
SOR_LOCAL_FORWARD_SWEEP = 4
SOR_LOCAL_BACKWARD_SWEEP = 8
SOR_LOCAL_SYMMETRIC_SWEEP = 12
SOR_EISENSTAT = 32
SOR_APPLY_UPPER = 64
SOR_APPLY_LOWER = 128

PC_JACOBI_DIAGONAL = 0
PC_JACOBI_ROWMAX = 1
PC_JACOBI_ROWL1 = 2
PC_JACOBI_ROWSUM = 3


class PCType:
    NONE = "none"
    JACOBI = "jacobi"
    SOR = "sor"
    LU = "lu"
    SHELL = "shell"
    BJACOBI = "bjacobi"
    ILU = "ilu"
    ICC = "icc"
    ASM = "asm"
    GASM = "gasm"
    KSP = "ksp"
    CHOLESKY = "cholesky"
    PYTHON = "python"
    MAT = "mat"
    FIELDSPLIT = "fieldsplit"
    LSC = "lsc"
    GAMG = "gamg"
    BDDC = "bddc"
    DEFLATION = "deflation"
    HPDDM = "hpddm"


SUPPORTED = {PCType.NONE, PCType.JACOBI, PCType.SOR, PCType.ILU, PCType.ICC}


class Side:
    LEFT = 0
    RIGHT = 1
    SYMMETRIC = 2
    DEFAULT = 0
    L = 0
    R = 1
    S = 2


class FailedReason:
    NOERROR = 0
    FACTOR_NUMERIC_ZEROPIVOT = 1
    FACTOR_NUMERIC_SINGULAR = 2
    FACTOR_INF_OR_NAN = 3
    FACTOR_OTHER = 4
    OUT_OF_MEMORY = 5
    SETUP_ERROR = 6
    SETUP_FAILED = 7
    PROBLEM_TOO_SMALL = 8


def _a(arr: np.ndarray) -> int:
    return arr.ctypes.data


@dataclass
class _PCOptions:
    """What the PC's setters record.  The defaults are PETSc's (PC.py:108-148
    of the reference): natural ordering, level 0, no shift, omega 1, one
    symmetric SOR sweep, the diagonal Jacobi with fix-diagonal on."""

    levels: int = 0
    shift_type: str = "none"
    shift: float = 0.0
    zeropivot: float = 100.0
    ordering: str = MatOrderingType.NATURAL
    diagonal_fill: bool = False
    sor_omega: float = 1.0
    sor_fshift: float = 0.0
    sor_its: int = 1  # recorded, not applied: the apply runs its = lits = 1 (capi.mojo:272-280)
    sor_lits: int = 1
    sor_symmetric: bool = True
    jacobi_type: int = PC_JACOBI_DIAGONAL
    jacobi_abs: bool = False
    jacobi_fix: bool = True
    rowl1_scale: float = 1.0
    use_amat: bool = False
    reuse: bool = False


def _setter(field, convert, resets="setup", doc=None):
    """A PC method storing convert(value) in the options; `resets` is what a
    change invalidates: "setup" (the next apply sets up again), "descriptor"
    (only the packed descriptor) or None."""

    def set_option(self, value):
        setattr(self._opt, field, convert(value))
        if resets == "setup":
            self._setupcalled = False
        elif resets == "descriptor":
            self._pci_cache = None

    set_option.__doc__ = doc
    return set_option


def _getter(field):
    def get_option(self):
        return getattr(self._opt, field)

    return get_option


class PC:
    """A preconditioner for a sequential AIJ matrix, applied on the GPU."""

    Type = PCType
    Side = Side
    FailedReason = FailedReason

    def __init__(self):
        self._type = PCType.NONE
        self._amat = self._pmat = None
        self._opt = _PCOptions()
        self._failed_reason = FailedReason.NOERROR
        self._setupcalled = False
        self._setup_version = None
        self._factor = None      # host factor arrays (int64 / float64)
        self._nlate = 0          # IC entries below the diagonal (see mpk_icc_transpose)
        self._dbufs = {}         # device buffers of the set up
        self._pci_cache = None
        self._sweep_depth = 0    # dependency levels of the deeper sweep

    # ---------------------------------------------------------- lifecycle
    def create(self, comm=None) -> "PC":
        return self

    def destroy(self) -> "PC":
        self._reset_factor()
        return self

    def setType(self, pc_type) -> None:
        if isinstance(pc_type, int):
            raise TypeError("expected a preconditioner type name")
        if pc_type not in SUPPORTED:
            raise NotImplementedError(f"PC type {pc_type!r} is not covered by mojo-petsc4py")
        self._type = pc_type
        self._reset_factor()

    def getType(self) -> str:
        return self._type

    def setOperators(self, A: Mat = None, P: Mat = None) -> None:
        self._amat, self._pmat = A, (A if P is None else P)
        self._reset_factor()

    def getOperators(self):
        return (self._amat, self._pmat)

    setUseAmat = _setter("use_amat", bool, resets=None)
    getUseAmat = _getter("use_amat")
    setReusePreconditioner = _setter("reuse", bool, resets=None)

    def setFailedReason(self, reason) -> None:
        self._failed_reason = reason

    def getFailedReason(self):
        return self._failed_reason

    def reset(self) -> None:
        self._reset_factor()

    def _reset_factor(self):
        self._setupcalled = False
        self._factor = None
        self._dbufs = {}
        self._pci_cache = None

    # ---------------------------------------------------------- options
    setFactorLevels = _setter("levels", int)
    setSOROmega = _setter("sor_omega", float)
    setSORIterations = _setter("sor_its", int, doc="Recorded; the apply runs one sweep pair, as the "
                               "reference's does (capi.mojo:272-280).")
    setSORShift = _setter("sor_fshift", float, resets=None)
    setSORSymmetric = _setter("sor_symmetric", bool, resets="descriptor")
    setJacobiType = _setter("jacobi_type", int)
    setJacobiUseAbs = _setter("jacobi_abs", bool)
    setJacobiFixDiagonal = _setter("jacobi_fix", bool)
    setJacobiRowl1Scale = _setter("rowl1_scale", float)

    def setFactorShift(self, shift_type, amount=None) -> None:
        """Only 'none' (PETSc's default, no pivot retry) is covered."""
        name = str(shift_type)
        if name not in ("none", "positive_definite", "nonzero"):
            raise ValueError(f"unknown shift type: {name}")
        if name != "none":
            raise NotImplementedError(
                f"factor shift {name!r} is not covered by mojo-petsc4py; the"
                " pivot-retry path (MatPivotCheck) is not ported, so only"
                " 'none' reproduces upstream")
        self._opt.shift_type = name
        if amount is not None:
            self._opt.shift = float(amount)
        self._setupcalled = False

    def setFactorPivot(self, flag) -> None:
        tolerance = float(flag) if not isinstance(flag, bool) else (100.0 if flag else 0.0)
        if tolerance != 0.0:
            raise NotImplementedError(
                "a nonzero factor pivot tolerance is not covered by mojo-petsc4py;"
                " MatPivotCheck is not ported, so the tolerance would be ignored")
        self._opt.zeropivot = 0.0
        self._opt.diagonal_fill = bool(flag)
        self._setupcalled = False

    def setFactorOrdering(self, ordering) -> None:
        if ordering not in (MatOrderingType.NATURAL, MatOrderingType.RCM):
            raise NotImplementedError(
                f"matrix ordering {ordering!r} is not covered by mojo-petsc4py (natural and rcm are)")
        self._opt.ordering = ordering
        self._setupcalled = False

    def setFactorReorderForNonzeroDiagonal(self) -> None:
        self._opt.diagonal_fill = True
        self._setupcalled = False

    def setFactorMatrixType(self, mtype) -> None:
        """Only seqaij is covered; accepted and ignored."""

    def setSPAIVerbose(self, level) -> None:
        raise NotImplementedError("spai is not covered by mojo-petsc4py")

    # ---------------------------------------------------------- set up
    def setUp(self) -> "PC":
        pmat = self._pmat
        if pmat is None:
            raise RuntimeError("PCSetOperators() has not been called")
        n = pmat._nrows
        if n != pmat._ncols:
            raise ValueError("the preconditioner matrix must be square")
        pmat._materialize()
        if self._setupcalled and self._setup_version == pmat._version:
            return self
        self._reset_factor()
        if self._type == PCType.JACOBI:
            self._build_jacobi(pmat, n)
        elif self._type == PCType.SOR:
            self._build_sor(pmat, n)
        elif self._type in (PCType.ILU, PCType.ICC):
            self._build_factor(pmat, n)
        check()
        self._setupcalled = True
        self._setup_version = pmat._version
        self._schedule(n)
        return self

    def _schedule(self, n: int):
        """Block plans for the two sweeps of SOR / ILU / IC (see
        mpk_pc_schedule): workgroups run blocks of rows level by level in LDS.
        The descriptor carries them in pci[11]."""
        if self._type not in (PCType.SOR, PCType.ILU, PCType.ICC) or n == 0:
            return
        if self._type != PCType.SOR and self._factor is None:
            return
        L = lib()
        pci = self._pci_block()
        pcf = self._pcf_block()
        size = L.mpk_pc_schedule_size(n, pci.ctypes.data, pcf.ctypes.data)
        check()
        plan = DeviceBuffer(max(int(size), 8))
        depth = L.mpk_pc_schedule(n, pci.ctypes.data, pcf.ctypes.data, plan.ptr, int(size))
        check()
        if depth < 0:
            raise RuntimeError("planning the preconditioner sweeps failed")
        self._dbufs["sched"] = plan
        self._sweep_depth = int(depth)
        self._pci_cache = None

    def _buf(self, name, nbytes):
        b = DeviceBuffer(nbytes)
        self._dbufs[name] = b
        return b

    def _build_jacobi(self, m: Mat, n: int):
        d = m._dev()
        k = max(n, 1)
        diag = self._buf("diag", 8 * k)
        didx = self._buf("didx", 8 * k)
        lib().mp_pc_jacobi_setup(diag.ptr, self._opt.jacobi_type, int(self._opt.jacobi_abs), int(self._opt.jacobi_fix),
                                 self._opt.rowl1_scale, d.rp.ptr, d.ci.ptr, d.v.ptr, n, didx.ptr)
        dsq = self._buf("diagsqrt", 8 * k)
        lib().mp_pc_jacobi_setup_symmetric(dsq.ptr, self._opt.jacobi_type, d.rp.ptr, d.ci.ptr, d.v.ptr, n)

    def _build_sor(self, m: Mat, n: int):
        d = m._dev()
        k = max(n, 1)
        dmark = self._buf("dmark", 4 * k)
        lib().mp_mat_get_diagonal_markers(d.rp.ptr, d.ci.ptr, dmark.ptr, n)
        self._buf("idiag", 8 * k)
        self._buf("mdiag", 8 * k)
        self._buf("sorwork", 8 * 3 * k)

    def _permutations(self, m: Mat, n: int):
        """(row order r, its inverse ic): natural, or the reference's
        index-order RCM (factor.mojo:279-332) computed by the library."""
        if not (self._opt.ordering == MatOrderingType.RCM and n > 0):
            r = np.arange(n, dtype=np.int64)
        else:
            work = np.zeros((4, n), dtype=np.int64)  # permutation, levels, marks, queue
            lib().mp_ordering_rcm(*(_a(row) for row in work), _a(m.irow), _a(m.icol), n)
            r = work[0].copy()
        ic = np.empty_like(r)
        ic[r] = np.arange(n, dtype=np.int64)
        return r, ic

    def _build_factor(self, m: Mat, n: int):
        if n == 0:
            return
        if self._type == PCType.ICC and self._ic0_on_device(m, n):
            return
        if self._type == PCType.ILU and self._ilu0_on_device(m, n):
            return
        r, ic = self._permutations(m, n)
        adiag = _device_diag_markers(m, n)
        scratch = np.zeros(7 * n + 8, dtype=np.int64)
        if self._type == PCType.ILU:
            self._build_ilu(m, n, r, ic, adiag, scratch)
        else:
            self._build_icc(m, n, r, ic, adiag, scratch)
        self._upload_factor(m, n)

    def _build_ilu(self, m, n, r, ic, adiag, scratch):
        """The ILU pattern: the ILU(0) fast path for natural ordering at level
        0, else the level-of-fill pass and its contiguous L/U layout
        (PC.py:392-440).  Every attempt takes fresh output arrays and the
        same scratch, which keeps what the attempt before left in it."""
        L = lib()
        fast = self._opt.levels == 0 and np.array_equal(r, np.arange(n))
        out = {}

        def attempt(cap):
            out.update(bi=np.zeros(n + 1, np.int64), bdiag=np.zeros(n + 1, np.int64), bj=np.zeros(cap, np.int64))
            if fast:
                got = L.mp_ilu_symbolic_ilu0(_a(out["bi"]), _a(out["bj"]), _a(out["bdiag"]), _a(adiag), _a(m.irow),
                                             _a(m.icol), n)
            else:
                out["src"] = np.zeros(cap, np.int64)
                lvl = np.zeros(cap, np.int64)
                got = L.mp_ilu_symbolic(_a(out["bi"]), _a(out["src"]), _a(lvl), _a(out["bdiag"]), _a(m.irow),
                                        _a(m.icol), _a(r), _a(ic), self._opt.levels, cap, _a(scratch), n)
            return got, n - got

        _until_it_fits(attempt, _first_capacity(m, n))
        if not fast:
            L.mp_ilu_contiguous_lu(_a(out["src"]), _a(out["bi"]), _a(out["bdiag"]), _a(out["bj"]), n)
        # the factor is in permuted coordinates and the solution goes back
        # through the permutation itself, c = r (PC.py:436-440); its values
        # come from the device pass (_ilu_numeric_device)
        self._factor = {"bi": out["bi"], "bj": out["bj"], "ba": None, "bdiag": out["bdiag"], "r": r, "c": r,
                        "ic": ic}

    def _build_icc(self, m, n, r, ic, adiag, scratch):
        """The IC pattern through the level-of-fill pass at every level, then
        its contiguous Cholesky layout (PC.py:442-492).  The scratch grows
        only when an attempt needs more than it holds."""
        L = lib()
        rinv = np.empty(n, np.int64)
        rinv[r] = np.arange(n, dtype=np.int64)
        out = {"scratch": scratch}

        def attempt(cap):
            need = 6 * (n + 1) + 2 * cap
            if out["scratch"].size < need:
                out["scratch"] = np.zeros(need, np.int64)
            sc = out["scratch"]
            out.update(ui=np.zeros(n + 1, np.int64), uj=np.zeros(cap, np.int64), udiag=np.zeros(n, np.int64))
            lvl = np.zeros(cap, np.int64)
            got = L.mp_icc_symbolic(_a(out["ui"]), _a(out["uj"]), _a(lvl), _a(out["udiag"]), _a(m.irow), _a(m.icol),
                                    _a(r), _a(rinv), self._opt.levels, cap, _a(sc), n, int(sc.size))
            return got, -got

        cap = _until_it_fits(attempt, _first_capacity(m, n))
        ui, uj = out["ui"], out["uj"]
        bi = np.zeros(n + 1, np.int64)
        bdiag = np.zeros(n + 1, np.int64)
        bdiag[:n] = out["udiag"]
        bj = np.zeros(max(cap, int(ui[-1]) + 1), np.int64)
        L.mp_icc_contiguous_cholesky(_a(uj), _a(ui), _a(bdiag), _a(bi), _a(bdiag), _a(bj), n)
        # the values come from the device pass (_icc_numeric_device)
        self._factor = {"bi": bi, "bj": bj, "ba": None, "bdiag": bdiag, "r": r, "rinv": rinv}

    def _ic0_on_device(self, m: Mat, n: int) -> bool:
        """IC(0) with the natural ordering, analysed on the device
        (mpk_ic0_analyse): the factor layout, its transpose, the row chains'
        visits and the level lists the host passes below give, without them;
        the host keeps only the level bounds, and downloads the factor on
        request (_host_factor).  False, for the host passes, when the library
        declines the matrix (a row not strictly ascending, a column of the
        strict part over 256 entries) or MPK_IC0_DEVICE=0."""
        if (self._opt.levels != 0 or self._opt.ordering == MatOrderingType.RCM
                or os.environ.get("MPK_IC0_DEVICE") == "0"):
            return False
        L = lib()
        d = m._dev()
        nnz = int(L.mpk_ic0_size(d.rp.ptr, d.ci.ptr, n))
        check()
        if nnz < 0:
            return False
        ms = max(nnz - n, 1)
        bi, bj, bd = self._buf("bi", 4 * (n + 1)), self._buf("bj", 4 * nnz), self._buf("bdiag", 4 * (n + 1))
        trp, trow, tpos = self._buf("trp", 4 * (n + 1)), self._buf("trow", 4 * ms), self._buf("tpos", 4 * ms)
        tmid = self._buf("tmid", 4 * n)
        vrow, vpos, lrow = DeviceBuffer(4 * ms), DeviceBuffer(4 * ms), DeviceBuffer(4 * n)
        lptr = np.zeros(n + 1, dtype=np.int64)
        nlev = L.mpk_ic0_analyse(d.rp.ptr, d.ci.ptr, n, bi.ptr, bj.ptr, bd.ptr, trp.ptr, trow.ptr, tpos.ptr,
                                 tmid.ptr, vrow.ptr, vpos.ptr, lrow.ptr, _a(lptr))
        check()
        if nlev < 0:
            raise RuntimeError("the device IC(0) analysis failed")
        r = np.arange(n, dtype=np.int64)
        rd = self._dbufs["r"] = DeviceBuffer.from_array(r.astype(np.int32))
        u0 = DeviceBuffer(8 * nnz)
        ua = self._buf("ba", 8 * nnz)
        # the visits of row k start where its column does in the transpose
        rc = L.mpk_icc_numeric(n, d.rp.ptr, d.ci.ptr, d.v.ptr, rd.ptr, rd.ptr, bi.ptr, bj.ptr, trp.ptr, vrow.ptr,
                               vpos.ptr, lrow.ptr, _a(lptr), nlev, u0.ptr, ua.ptr)
        check()
        if rc != 0:
            raise RuntimeError("the device IC factorisation failed")
        L.mpk_synchronize()
        self._nlate = 0
        self._factor = {"bi": None, "bj": None, "ba": None, "bdiag": None, "r": r, "rinv": r, "nnz": nnz}
        return True

    def _ilu0_on_device(self, m: Mat, n: int) -> bool:
        """ILU(0) with the natural ordering, analysed on the device
        (mpk_ilu0_analyse): the layout of mp_ilu_symbolic_ilu0 and the level
        lists of mpk_ilu_levels, without the host passes; the host keeps only
        the level bounds and downloads the factor on request (_host_factor).
        False, for the host passes, when the library declines the matrix (a
        row not strictly ascending or without its diagonal) or
        MPK_ILU0_DEVICE=0."""
        if (self._opt.levels != 0 or self._opt.ordering == MatOrderingType.RCM
                or os.environ.get("MPK_ILU0_DEVICE") == "0"):
            return False
        L = lib()
        d = m._dev()
        nnz = int(L.mpk_ilu0_size(d.rp.ptr, d.ci.ptr, n))
        check()
        if nnz < 0:
            return False
        bi, bj, bd = self._buf("bi", 4 * (n + 1)), self._buf("bj", 4 * nnz), self._buf("bdiag", 4 * (n + 1))
        lrow = DeviceBuffer(4 * n)
        lptr = np.zeros(n + 1, dtype=np.int64)
        nlev = L.mpk_ilu0_analyse(d.rp.ptr, d.ci.ptr, n, bi.ptr, bj.ptr, bd.ptr, lrow.ptr, _a(lptr))
        check()
        if nlev < 0:
            raise RuntimeError("the device ILU(0) analysis failed")
        r = np.arange(n, dtype=np.int64)
        rd = self._dbufs["r"] = self._dbufs["c"] = DeviceBuffer.from_array(r.astype(np.int32))
        ba = self._buf("ba", 8 * nnz)
        rc = L.mpk_ilu_numeric(n, d.rp.ptr, d.ci.ptr, d.v.ptr, rd.ptr, rd.ptr, bi.ptr, bj.ptr, bd.ptr, lrow.ptr,
                               _a(lptr), nlev, ba.ptr)
        check()
        if rc != 0:
            raise RuntimeError("the device ILU factorisation failed")
        L.mpk_synchronize()
        self._factor = {"bi": None, "bj": None, "ba": None, "bdiag": None, "r": r, "c": r, "ic": r, "nnz": nnz}
        return True

    def _upload_factor(self, m: Mat, n: int):
        f = self._factor
        nnz = int(f["bi"][n]) if self._type == PCType.ICC else int(f["bdiag"][0]) + 1
        for key in ("bi", "bdiag", "r"):
            self._dbufs[key] = DeviceBuffer.from_array(f[key].astype(np.int32))
        self._dbufs["bj"] = DeviceBuffer.from_array(f["bj"][: max(nnz, 1)].astype(np.int32))
        if self._type == PCType.ILU:
            self._dbufs["c"] = self._dbufs["r"]
            self._ilu_numeric_device(m, n, nnz)
            return
        self._icc_numeric_device(m, n, nnz)
        L = lib()
        bi, bj, bd = self._dbufs["bi"].ptr, self._dbufs["bj"].ptr, self._dbufs["bdiag"].ptr
        mt = L.mpk_icc_transpose_size(bi, bd, n)
        check()
        trp = self._buf("trp", 4 * (n + 1))
        trow = self._buf("trow", 4 * max(mt, 1))
        tpos = self._buf("tpos", 4 * max(mt, 1))
        tmid = self._buf("tmid", 4 * max(n, 1))
        self._nlate = L.mpk_icc_transpose(bi, bj, bd, n, trp.ptr, trow.ptr, tpos.ptr, tmid.ptr)
        check()

    def _ilu_numeric_device(self, m: Mat, n: int, nnz: int):
        """MatILUFactorNumeric_SeqAIJ (ported.mojo:957-1037) on the device:
        the host lists the rows level by level (levels of the factor's L
        part), the device eliminates one level per launch into ba.  A factor
        layout the level-by-level pass cannot reproduce (a row without its
        own diagonal slot, which level-of-fill leaves on some nonsymmetric
        patterns) is factored by the serial pass on the host instead."""
        L = lib()
        f = self._factor
        lrow = np.zeros(max(n, 1), dtype=np.int32)
        lptr = np.zeros(n + 1, dtype=np.int64)
        nlev = L.mpk_ilu_levels(_a(f["bi"]), _a(f["bj"]), _a(f["bdiag"]), n, _a(lrow), _a(lptr))
        if nlev == -2:
            ba = np.zeros(max(nnz, 1), dtype=np.float64)
            rtmp = np.zeros(n, dtype=np.float64)
            L.mp_pc_ilu_numeric(_a(ba), _a(rtmp), _a(m.irow), _a(m.icol), _a(m.data), _a(f["bi"]), _a(f["bj"]),
                                _a(f["bdiag"]), _a(f["r"]), _a(f["ic"]), n)
            f["ba"] = ba
            self._dbufs["ba"] = DeviceBuffer.from_array(ba)
            return
        if nlev < 0:
            raise RuntimeError("listing the ILU factor's levels failed")
        d = m._dev()
        ic = DeviceBuffer.from_array(f["ic"].astype(np.int32))
        lrow_d = DeviceBuffer.from_array(lrow)
        ba = self._buf("ba", 8 * max(nnz, 1))
        b = self._dbufs
        rc = L.mpk_ilu_numeric(n, d.rp.ptr, d.ci.ptr, d.v.ptr, b["r"].ptr, ic.ptr, b["bi"].ptr, b["bj"].ptr,
                               b["bdiag"].ptr, lrow_d.ptr, _a(lptr), nlev, ba.ptr)
        check()
        if rc != 0:
            raise RuntimeError("the device ILU factorisation failed")
        L.mpk_synchronize()
        f["nnz"] = nnz

    def _icc_numeric_device(self, m: Mat, n: int, nnz: int):
        """MatICCFactorNumeric_SeqAIJ (ported.mojo:1040-1124) on the device:
        the host replays the up-looking pass's row chains (which rows update
        which, in what order, and the level sets that follow), the device
        does the arithmetic one level per launch.  A layout whose result
        depends on values the reference's dense work row carries from row to
        row is factored by the serial pass on the host instead."""
        L = lib()
        f = self._factor
        bi, bj = f["bi"], f["bj"]
        mm = max(nnz, 1)
        vptr = np.zeros(n + 1, dtype=np.int32)
        vrow = np.zeros(mm, dtype=np.int32)
        vpos = np.zeros(mm, dtype=np.int32)
        lrow = np.zeros(max(n, 1), dtype=np.int32)
        lptr = np.zeros(n + 1, dtype=np.int64)
        nlev = L.mpk_icc_visits(_a(bi), _a(bj), _a(m.irow), _a(m.icol), _a(f["r"]), _a(f["rinv"]), n, _a(vptr),
                                _a(vrow), _a(vpos), _a(lrow), _a(lptr))
        if nlev == -2:
            ua = np.zeros(mm, dtype=np.float64)
            rtmp = np.zeros(n, dtype=np.float64)
            il = np.zeros(n, dtype=np.int64)
            c2r = np.zeros(n, dtype=np.int64)
            L.mp_pc_icc_numeric(_a(ua), _a(rtmp), _a(il), _a(c2r), _a(m.irow), _a(m.icol), _a(m.data), _a(bi),
                                _a(bj), _a(f["bdiag"]), _a(f["r"]), _a(f["rinv"]), n)
            f["ba"] = ua
            self._dbufs["ba"] = DeviceBuffer.from_array(ua)
            return
        if nlev < 0:
            raise RuntimeError("replaying the IC factor's row chains failed")
        d = m._dev()
        b = self._dbufs
        up = {k: DeviceBuffer.from_array(v) for k, v in
              (("riip", f["rinv"].astype(np.int32)), ("vptr", vptr), ("vrow", vrow), ("vpos", vpos),
               ("lrow", lrow))}
        u0 = DeviceBuffer(8 * mm)
        ua = self._buf("ba", 8 * mm)
        rc = L.mpk_icc_numeric(n, d.rp.ptr, d.ci.ptr, d.v.ptr, b["r"].ptr, up["riip"].ptr, b["bi"].ptr,
                               b["bj"].ptr, up["vptr"].ptr, up["vrow"].ptr, up["vpos"].ptr, up["lrow"].ptr,
                               _a(lptr), nlev, u0.ptr, ua.ptr)
        check()
        if rc != 0:
            raise RuntimeError("the device IC factorisation failed")
        L.mpk_synchronize()
        f["nnz"] = nnz

    def _host_factor(self) -> dict:
        """The factor with its values on the host (the values are computed on
        the device and downloaded on first request)."""
        f = self._factor
        if f is not None and f["bi"] is None:  # analysed on the device
            n, b = len(f["r"]), self._dbufs
            for key, count in (("bi", n + 1), ("bdiag", n + 1), ("bj", int(f["nnz"]))):
                f[key] = b[key].to_numpy(np.int32, count).astype(np.int64)
        if f is not None and f["ba"] is None:
            f["ba"] = self._dbufs["ba"].to_numpy(np.float64, max(int(f["nnz"]), 1))
        return f

    # ---------------------------------------------------------- descriptor
    def _sor_flag(self) -> int:
        base = SOR_LOCAL_SYMMETRIC_SWEEP if self._opt.sor_symmetric else SOR_FORWARD_SWEEP
        return base | SOR_ZERO_INITIAL_GUESS

    def _pci_block(self) -> np.ndarray:
        """The descriptor the device solvers read (layout of PC.py:495-547,
        with device addresses and int32 index arrays)."""
        if self._pci_cache is not None:
            return self._pci_cache
        self.setUp()
        pci = [0] * 12
        b = self._dbufs
        if self._type == PCType.JACOBI:
            pci[0] = _JACOBI
            pci[1] = b["diag"].ptr
            pci[2] = b["diagsqrt"].ptr
        elif self._type == PCType.SOR:
            d = self._pmat._dev()
            pci[0] = _SOR
            pci[1:8] = [d.rp.ptr, d.ci.ptr, d.v.ptr, b["dmark"].ptr, b["idiag"].ptr, b["mdiag"].ptr,
                        b["sorwork"].ptr]
            pci[10] = self._sor_flag()
        elif self._type == PCType.ILU and self._factor is not None:
            pci[0] = _ILU
            pci[1:7] = [b["bi"].ptr, b["bj"].ptr, b["ba"].ptr, b["bdiag"].ptr, b["r"].ptr, b["c"].ptr]
        elif self._type == PCType.ICC and self._factor is not None:
            pci[0] = _ICC
            pci[1:11] = [b["bi"].ptr, b["bj"].ptr, b["ba"].ptr, b["bdiag"].ptr, b["r"].ptr, b["trp"].ptr,
                         b["trow"].ptr, b["tpos"].ptr, b["tmid"].ptr, self._nlate]
        if "sched" in b:
            pci[11] = b["sched"].ptr
        self._pci_cache = np.array(pci, dtype=np.int64)
        return self._pci_cache

    def _pcf_block(self) -> np.ndarray:
        return np.array([self._opt.sor_omega, self._opt.sor_fshift], dtype=np.float64)

    # ---------------------------------------------------------- apply
    def _direct_apply(self, x: Vec, y: Vec):
        n = self._pmat._nrows
        L = lib()
        b = self._dbufs
        xa = x._d()
        if self._type == PCType.NONE or (self._type in (PCType.ILU, PCType.ICC) and n == 0):
            L.mp_vec_copy(xa, y._dw(), n)
        elif self._type == PCType.JACOBI:
            L.mp_pc_jacobi_apply(y._dw(), xa, b["diag"].ptr, n)
        else:
            # SOR / ILU / IC through the descriptor, with its level orders
            L.mpk_pc_apply(n, self._pci_block().ctypes.data, self._pcf_block().ctypes.data, xa, y._dw())
        check()

    def apply(self, x: Vec, y: Vec) -> None:
        self.setUp()
        n = self._pmat._nrows
        if x.getSize() != n or y.getSize() != n:
            raise ValueError(f"incompatible vector size: {n} is needed for both operands")
        self._direct_apply(x, y)

    def applyTranspose(self, x: Vec, y: Vec) -> None:
        self.apply(x, y)

    def applySymmetricLeft(self, x: Vec, y: Vec) -> None:
        if self._type != PCType.JACOBI:
            self.apply(x, y)
            return
        self.setUp()
        n = self._pmat._nrows
        if x.getSize() != n or y.getSize() != n:
            raise ValueError(f"incompatible vector size: {n} is needed for both operands")
        xa = x._d()
        lib().mp_pc_jacobi_apply(y._dw(), xa, self._dbufs["diagsqrt"].ptr, n)
        check()

    def applySymmetricRight(self, x: Vec, y: Vec) -> None:
        self.applySymmetricLeft(x, y)

    def getFactorMatrix(self) -> Mat:
        self.setUp()
        f = self._host_factor()
        if f is None:
            raise RuntimeError(f"PC type {self._type} has no factor matrix")
        if self._type == PCType.ILU:
            return _ilu_factor_mat(f)
        return _icc_factor_mat(f)


def _device_diag_markers(m: Mat, n: int) -> np.ndarray:
    """MatGetDiagonalMarkers_SeqAIJ (ported.mojo:354-377) on the device copy
    of the matrix (k_diag_markers), as the int64 array the symbolic phases
    take: the position of a_ii in row i, or the row end."""
    d = m._dev()
    dm = DeviceBuffer(4 * max(n, 1))
    lib().mp_mat_get_diagonal_markers(d.rp.ptr, d.ci.ptr, dm.ptr, n)
    check()
    return dm.to_numpy(np.int32, n).astype(np.int64)


def _first_capacity(m: Mat, n: int) -> int:
    """The first capacity a symbolic factorisation is given (PC.py:393)."""
    return max(4 * n, 4 * int(m.irow[-1]) + 4 * n, 16)


def _until_it_fits(attempt, cap: int) -> int:
    """The symbolic phases' capacity protocol: attempt(cap) returns (count,
    capacity to try next); a negative count means the output did not fit.
    Returns the capacity that fitted."""
    while True:
        got, more = attempt(cap)
        if got >= 0:
            return cap
        cap = int(more)


def _segments(starts, counts):
    """(segment index, position) of every element of the segments
    [starts[k], starts[k] + counts[k]) (a negative count is an empty one)."""
    counts = np.maximum(np.asarray(counts, np.int64), 0)
    seg = np.repeat(np.arange(len(counts)), counts)
    first = np.repeat(np.asarray(starts, np.int64) - np.concatenate(([0], np.cumsum(counts)[:-1])), counts)
    return seg, first + np.arange(int(counts.sum()))


def _ilu_factor_mat(f) -> Mat:
    """L U of the ILU factor, in the original numbering (PC.py:660-689): L is
    unit lower with row i's entries bi[i]..bi[i+1]; U's row i holds the
    entries after bdiag[i+1] and ends on its inverted diagonal at bdiag[i]."""
    bi, bj, ba, bdiag, r, c = (np.asarray(f[k]) for k in ("bi", "bj", "ba", "bdiag", "r", "c"))
    n = len(bi) - 1
    lower = np.eye(n)
    i, q = _segments(bi[:n], np.diff(bi))
    lower[r[i], c[bj[q]]] = ba[q]
    upper = np.zeros((n, n))
    first = bdiag[1:] + 1
    count = bdiag[:n] - bdiag[1:] - 1
    i, q = _segments(first, count)
    upper[r[i], c[bj[q]]] = ba[q]
    upper[r, r] = 1.0 / ba[first + count]
    return from_dense(lower @ upper)


def _icc_factor_mat(f) -> Mat:
    """L D L^T of the IC factor, in the original numbering (PC.py:692-745):
    row i of the stored upper part holds -(u_ij / d_i) at bi[i]..bdiag[i] and
    1 / d_i at bdiag[i]."""
    bi, bj, ba, bdiag, r = (np.asarray(f[k]) for k in ("bi", "bj", "ba", "bdiag", "r"))
    n = len(bi) - 1
    unit = np.zeros((n, n))
    unit[r, r] = 1.0
    i, q = _segments(bi[:n], bdiag[:n] - bi[:n])
    unit[r[bj[q]], r[i]] = -ba[q]
    inv = ba[bdiag[:n]]
    d = np.zeros(n)
    nz = inv != 0.0
    d[r[nz]] = 1.0 / inv[nz]
    return from_dense((unit * d) @ unit.T)
