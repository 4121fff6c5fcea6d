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


class PC:
    """A preconditioner for a sequential AIJ matrix, applied on the GPU."""

    Type = PCType
    Side = Side
    FailedReason = FailedReason

    def __init__(self):
        self._type = PCType.NONE
        self._amat = None
        self._pmat = None
        self._setupcalled = False
        self._setup_version = None
        self._use_amat = False
        self._reuse = False
        self._failed_reason = FailedReason.NOERROR
        self._levels = 0
        self._shift = 0.0
        self._zeropivot = 100.0
        self._shift_type = "none"
        self._ordering = MatOrderingType.NATURAL
        self._diagonal_fill = False
        self._sor_omega = 1.0
        self._sor_fshift = 0.0
        self._sor_its = 1
        self._sor_lits = 1
        self._sor_sym = "symmetric"
        self._fix_diag = True
        self._use_abs = False
        self._jacobi_type = PC_JACOBI_DIAGONAL
        self._rowl1_scale = 1.0
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
        self._amat = A
        self._pmat = A if P is None else P
        self._reset_factor()

    def getOperators(self):
        return (self._amat, self._pmat)

    def setUseAmat(self, flag: bool) -> None:
        self._use_amat = bool(flag)

    def getUseAmat(self) -> bool:
        return self._use_amat

    def setReusePreconditioner(self, flag: bool) -> None:
        self._reuse = bool(flag)

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
    def setFactorLevels(self, levels) -> None:
        self._levels = int(levels)
        self._setupcalled = False

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
        self._shift_type = name
        if amount is not None:
            self._shift = float(amount)
        self._setupcalled = False

    def setFactorPivot(self, flag) -> None:
        tolerance = float(flag) if not isinstance(flag, bool) else (100.0 if flag else 0.0)
        if tolerance != 0.0:
            raise NotImplementedError(
                "a nonzero factor pivot tolerance is not covered by mojo-petsc4py;"
                " MatPivotCheck is not ported, so the tolerance would be ignored")
        self._zeropivot = 0.0
        self._diagonal_fill = bool(flag)
        self._setupcalled = False

    def setFactorOrdering(self, ordering) -> None:
        if ordering not in (MatOrderingType.NATURAL, MatOrderingType.RCM):
            raise NotImplementedError(
                f"matrix ordering {ordering!r} is not covered by mojo-petsc4py (natural and rcm are)")
        self._ordering = ordering
        self._setupcalled = False

    def setFactorReorderForNonzeroDiagonal(self) -> None:
        self._diagonal_fill = True
        self._setupcalled = False

    def setSORSymmetric(self, flag) -> None:
        self._sor_sym = "symmetric" if flag else "forward"
        self._pci_cache = None

    def setSOROmega(self, omega) -> None:
        self._sor_omega = float(omega)
        self._setupcalled = False

    def setSORIterations(self, its) -> None:
        # stored but not applied: the apply runs its = lits = 1, as the
        # reference's does (capi.mojo:272-280)
        self._sor_its = int(its)
        self._setupcalled = False

    def setSORShift(self, fshift) -> None:
        self._sor_fshift = float(fshift)

    def setJacobiType(self, jtype) -> None:
        self._jacobi_type = int(jtype)
        self._setupcalled = False

    def setJacobiUseAbs(self, flag) -> None:
        self._use_abs = bool(flag)
        self._setupcalled = False

    def setJacobiFixDiagonal(self, flag) -> None:
        self._fix_diag = bool(flag)
        self._setupcalled = False

    def setJacobiRowl1Scale(self, scale) -> None:
        self._rowl1_scale = float(scale)
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
        lib().mp_pc_jacobi_setup(diag.ptr, self._jacobi_type, int(self._use_abs), int(self._fix_diag),
                                 self._rowl1_scale, d.rp.ptr, d.ci.ptr, d.v.ptr, n, didx.ptr)
        dsq = self._buf("diagsqrt", 8 * k)
        lib().mp_pc_jacobi_setup_symmetric(dsq.ptr, self._jacobi_type, d.rp.ptr, d.ci.ptr, d.v.ptr, n)

    def _build_sor(self, m: Mat, n: int):
        d = m._dev()
        k = max(n, 1)
        dmark = self._buf("dmark", 4 * k)
        lib().mp_mat_get_diagonal_markers(d.rp.ptr, d.ci.ptr, dmark.ptr, n)
        self._buf("idiag", 8 * k)
        self._buf("mdiag", 8 * k)
        self._buf("sorwork", 8 * 3 * k)

    def _permutations(self, m: Mat, n: int):
        if self._ordering == MatOrderingType.RCM and n > 0:
            perm = np.zeros(n, dtype=np.int64)
            level = np.zeros(n, dtype=np.int64)
            visited = np.zeros(n, dtype=np.int64)
            queue = np.zeros(n, dtype=np.int64)
            lib().mp_ordering_rcm(_a(perm), _a(level), _a(visited), _a(queue), _a(m.irow), _a(m.icol), n)
            r = perm
        else:
            r = np.arange(n, dtype=np.int64)
        ic = np.empty(n, dtype=np.int64)
        ic[r] = np.arange(n, dtype=np.int64)
        return r, ic

    def _build_factor(self, m: Mat, n: int):
        if n == 0:
            return
        r, ic = self._permutations(m, n)
        adiag = np.zeros(n, dtype=np.int64)
        _host_diag_markers(m, adiag)
        scratch = np.zeros(7 * n + 8, dtype=np.int64)
        if self._type == PCType.ILU:
            self._build_ilu(m, n, r, ic, adiag, scratch)
        else:
            self._build_icc(m, n, r, ic, adiag, scratch)
        self._upload_factor(m, n)

    def _build_ilu(self, m, n, r, ic, adiag, scratch):
        L = lib()
        cap = max(4 * n, 4 * int(m._irow[-1]) + 4 * n, 16)
        while True:
            bi = np.zeros(n + 1, dtype=np.int64)
            bj = np.zeros(cap, dtype=np.int64)
            bjlvl = np.zeros(cap, dtype=np.int64)
            bdiag = np.zeros(n + 1, dtype=np.int64)
            if self._levels == 0 and np.array_equal(r, np.arange(n)):
                nnz = L.mp_ilu_symbolic_ilu0(_a(bi), _a(bj), _a(bdiag), _a(adiag), _a(m.irow), _a(m.icol), n)
                contiguous = False
            else:
                bjsrc = np.zeros(cap, dtype=np.int64)
                nnz = L.mp_ilu_symbolic(_a(bi), _a(bjsrc), _a(bjlvl), _a(bdiag), _a(m.irow), _a(m.icol),
                                        _a(r), _a(ic), self._levels, cap, _a(scratch), n)
                contiguous = True
            if nnz >= 0:
                if contiguous:
                    L.mp_ilu_contiguous_lu(_a(bjsrc), _a(bi), _a(bdiag), _a(bj), n)
                break
            cap = -nnz + n
        # the factor lives in permuted coordinates; the solution is mapped back
        # with the permutation itself (PC.py:436-440).  Its values are
        # computed on the device (_ilu_numeric_device), level by level.
        self._factor = {"bi": bi, "bj": bj, "ba": None, "bdiag": bdiag, "r": r, "c": r, "ic": ic}

    def _build_icc(self, m, n, r, ic, adiag, scratch):
        L = lib()
        cap = max(4 * n, 4 * int(m._irow[-1]) + 4 * n, 16)
        rinv = np.empty(n, dtype=np.int64)
        rinv[r] = np.arange(n, dtype=np.int64)
        while True:
            ui = np.zeros(n + 1, dtype=np.int64)
            uj = np.zeros(cap, dtype=np.int64)
            ujlvl = np.zeros(cap, dtype=np.int64)
            udiag = np.zeros(n, dtype=np.int64)
            need = 6 * (n + 1) + 2 * cap
            if scratch.size < need:
                scratch = np.zeros(need, dtype=np.int64)
            nnz = L.mp_icc_symbolic(_a(ui), _a(uj), _a(ujlvl), _a(udiag), _a(m.irow), _a(m.icol), _a(r),
                                    _a(rinv), self._levels, cap, _a(scratch), n, int(scratch.size))
            if nnz >= 0:
                break
            cap = -nnz
        bi = np.zeros(n + 1, dtype=np.int64)
        bdiag = np.zeros(n + 1, dtype=np.int64)
        bj = np.zeros(max(cap, int(ui[-1]) + 1), dtype=np.int64)
        bdiag[:n] = udiag[:n]
        L.mp_icc_contiguous_cholesky(_a(uj), _a(ui), _a(bdiag), _a(bi), _a(bdiag), _a(bj), n)
        # the values are computed on the device (_icc_numeric_device)
        self._factor = {"bi": bi, "bj": bj, "ba": None, "bdiag": bdiag, "r": r, "rinv": rinv}

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
        if f is not None and f["ba"] is None:
            f["ba"] = self._dbufs["ba"].to_numpy(np.float64, max(int(f["nnz"]), 1))
        return f

    # ---------------------------------------------------------- descriptor
    def _sor_flag(self) -> int:
        base = SOR_LOCAL_SYMMETRIC_SWEEP if self._sor_sym == "symmetric" else SOR_FORWARD_SWEEP
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
        return np.array([self._sor_omega, self._sor_fshift], dtype=np.float64)

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


def _host_diag_markers(m: Mat, adiag: np.ndarray) -> None:
    """MatGetDiagonalMarkers_SeqAIJ on the host CSR (ported.mojo:354-377):
    the position of a_ii in row i, or the row end."""
    n = m._nrows
    irow, icol = m.irow, m.icol
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(irow[: n + 1]))
    nnz = len(rows)
    hit = np.nonzero(icol[:nnz] == rows)[0]
    adiag[:] = irow[1: n + 1]
    if len(hit):
        # first match per row
        first = np.full(n, -1, dtype=np.int64)
        rr = rows[hit]
        keep = np.r_[True, rr[1:] != rr[:-1]]
        first[rr[keep]] = hit[keep]
        has = first >= 0
        adiag[has] = first[has]


def _ilu_factor_mat(f) -> Mat:
    """L*U of the ILU factor as a CSR matrix (PC.py:660-689)."""
    bi, bj, ba, bdiag, r, c = f["bi"], f["bj"], f["ba"], f["bdiag"], f["r"], f["c"]
    n = len(bi) - 1
    L = np.eye(n)
    U = np.zeros((n, n))
    for i in range(n):
        for j in range(int(bi[i]), int(bi[i + 1])):
            L[int(r[i]), int(c[int(bj[j])])] = ba[j]
    for i in range(n):
        v = int(bdiag[i + 1]) + 1
        nz = int(bdiag[i] - bdiag[i + 1] - 1)
        for j in range(nz):
            U[int(r[i]), int(c[int(bj[v + j])])] = ba[v + j]
        U[int(r[i]), int(r[i])] = 1.0 / ba[v + nz]
    return from_dense(L @ U)


def _icc_factor_mat(f) -> Mat:
    """L*D*L^T of the IC factor as a CSR matrix (PC.py:692-745)."""
    bi, bj, ba, bdiag, r = f["bi"], f["bj"], f["ba"], f["bdiag"], f["r"]
    n = len(bi) - 1
    unit = np.zeros((n, n), dtype=np.float64)
    d = np.zeros(n, dtype=np.float64)
    for i in range(n):
        row = int(r[i])
        slot = int(bdiag[i])
        unit[row, row] = 1.0
        for j in range(int(bi[i]), slot):
            unit[int(r[int(bj[j])]), row] = -ba[j]
        inv = ba[slot]
        d[row] = 1.0 / inv if inv != 0.0 else 0.0
    return from_dense((unit * d) @ unit.T)
