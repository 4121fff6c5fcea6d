"""ctypes driver for the CPU oracle (oracle.c).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or the timed CPU baseline -- the product
(mojo-petsc4py_amd/) never touches it.

Besides the raw kernels, this module restates the reference's Python-side set
up so that a whole solve can be driven the way `KSP.solve` drives it:
  * PC set up: python/mojopetsc4py/PETSc/PC.py:315-492 (Jacobi, SOR markers,
    the ILU(0) fast path, the level-of-fill ILU with its contiguous pass and
    cap-retry, ICC always through the level-of-fill path, c = r).
  * the descriptor block pci: PC.py:495-547.
  * the KSP state blocks cfg/cfgi and the work arenas: KSP.py:367-494.
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

I = ctypes.c_int64
F = ctypes.c_double

_SIG = {
    "or_set_reduction_mode": ([I], None),
    "or_get_reduction_mode": ([], I),
    "or_red_dot": ([I, I, I], F),
    "or_red_sumsq": ([I, I], F),
    "or_red_asum": ([I, I], F),
    "or_red_total": ([I, I], F),
    "or_red_mdot": ([I, I, I, I, I], None),
    "or_mat_mult": ([I] * 6, None),
    "or_mat_mult_add": ([I] * 6, None),
    "or_mat_mult_transpose": ([I] * 7, None),
    "or_mat_get_diagonal": ([I] * 5, None),
    "or_mat_scale": ([I, I, F], None),
    "or_mat_zero_entries": ([I, I], None),
    "or_vec_pointwise_mult": ([I] * 4, None),
    "or_vec_pointwise_divide": ([I] * 4, None),
    "or_vec_pointwise_min": ([I] * 4, None),
    "or_vec_pointwise_max": ([I] * 4, None),
    "or_vec_pointwise_max_abs": ([I] * 4, None),
    "or_vec_norm": ([I] * 4, None),
    "or_vec_dot": ([I] * 4, None),
    "or_vec_tdot": ([I] * 4, None),
    "or_vec_sum": ([I] * 3, None),
    "or_vec_mean": ([I] * 3, None),
    "or_vec_dot_norm2": ([I] * 5, None),
    "or_vec_max": ([I] * 4, None),
    "or_vec_min": ([I] * 4, None),
    "or_vec_normalize": ([I] * 3, None),
    "or_vec_reciprocal": ([I] * 2, None),
    "or_vec_exp": ([I] * 2, None),
    "or_vec_log": ([I] * 2, None),
    "or_vec_sqrtabs": ([I] * 2, None),
    "or_vec_chop": ([I, I, F], None),
    "or_vec_abs": ([I] * 2, None),
    "or_vec_conjugate": ([I] * 2, None),
    "or_vec_copy": ([I] * 3, None),
    "or_vec_swap": ([I] * 3, None),
    "or_vec_set": ([I, F, I], None),
    "or_vec_zero_entries": ([I] * 2, None),
    "or_vec_scale": ([I, F, I], None),
    "or_vec_shift": ([I, F, I], None),
    "or_vec_axpy": ([I, F, I, I], None),
    "or_vec_aypx": ([I, F, I, I], None),
    "or_vec_axpby": ([I, F, F, I, I], None),
    "or_vec_axpbypcz": ([I, F, F, F, I, I, I], None),
    "or_vec_waxpy": ([I, F, I, I, I], None),
    "or_vec_maxpy": ([I] * 5, None),
    "or_vec_max_pointwise_divide": ([I] * 4, None),
    "or_mat_get_diagonal_markers": ([I] * 4, I),
    "or_pc_jacobi_setup": ([I, I, I, I, F, I, I, I, I, I], None),
    "or_pc_jacobi_setup_symmetric": ([I] * 6, None),
    "or_pc_jacobi_apply": ([I] * 4, None),
    "or_pc_sor_apply": ([I] * 9 + [F, F, I, I], None),
    "or_pc_sor_apply_its": ([I] * 9 + [F, F, I, I, I, I], None),
    "or_pc_ilu_apply": ([I] * 10, None),
    "or_pc_icc_apply": ([I] * 9, I),
    "or_pc_ilu_numeric": ([I] * 11, I),
    "or_pc_icc_numeric": ([I] * 13, None),
    "or_ilu_symbolic_ilu0": ([I] * 7, I),
    "or_ilu_symbolic": ([I] * 12, I),
    "or_icc_symbolic": ([I] * 13, I),
    "or_ilu_contiguous_lu": ([I] * 5, None),
    "or_icc_contiguous_cholesky": ([I] * 7, None),
    "or_ordering_rcm": ([I] * 7, None),
    "or_ksp_converged_default": ([I, F, I, I], I),
    "or_ksp_solve_cg": ([I] * 12, None),
    "or_ksp_solve_bcgs": ([I] * 12, None),
    "or_ksp_solve_gmres": ([I] * 14, None),
    "or_ksp_solve_fgmres": ([I] * 15, None),
}

_lib = None


def build() -> str:
    """Compile oracle.c with the Makefile beside it (gcc, no GPU involved)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        for name, (argt, rest) in _SIG.items():
            fn = getattr(_lib, name)
            fn.argtypes = argt
            fn.restype = rest
    return _lib


def a(arr: np.ndarray) -> int:
    return arr.ctypes.data


class mode:
    """Context manager selecting the reduction order: 'ref' or 'dev'."""

    def __init__(self, which: str):
        self.want = {"ref": 0, "dev": 1}[which]

    def __enter__(self):
        self.old = lib().or_get_reduction_mode()
        lib().or_set_reduction_mode(self.want)
        return self

    def __exit__(self, *exc):
        lib().or_set_reduction_mode(self.old)


def set_mode(which: str) -> None:
    lib().or_set_reduction_mode({"ref": 0, "dev": 1}[which])


# ------------------------------------------------------------------ CSR

class CSR:
    """int64 CSR as the reference's Mat holds it."""

    def __init__(self, indptr, indices, data, ncols=None):
        self.irow = np.ascontiguousarray(indptr, dtype=np.int64)
        self.icol = np.ascontiguousarray(indices, dtype=np.int64)
        self.data = np.ascontiguousarray(data, dtype=np.float64)
        self.n = len(self.irow) - 1
        self.ncols = self.n if ncols is None else int(ncols)
        # the reference's buffers are never empty (Mat.py:402-403)
        if len(self.icol) == 0:
            self.icol = np.zeros(1, dtype=np.int64)
            self.data = np.zeros(1, dtype=np.float64)

    def mult(self, x):
        y = np.zeros(self.n)
        lib().or_mat_mult(a(np.ascontiguousarray(x, dtype=np.float64)), a(y), a(self.irow),
                          a(self.icol), a(self.data), self.n)
        return y

    def dense(self):
        out = np.zeros((self.n, self.ncols))
        for i in range(self.n):
            for j in range(self.irow[i], self.irow[i + 1]):
                out[i, self.icol[j]] += self.data[j]
        return out


# ------------------------------------------------------------------ PC set up

class PCSetup:
    """The oracle's PC: PC.py:315-547 restated on top of the oracle kernels."""

    def __init__(self, A: CSR, kind: str = "none", levels: int = 0, ordering: str = "natural",
                 omega: float = 1.0, fshift: float = 0.0, symmetric: bool = True,
                 jacobi_type: int = 0, use_abs: bool = False, fix_diag: bool = True,
                 rowl1_scale: float = 1.0):
        self.A = A
        self.kind = kind
        self.omega = float(omega)
        self.fshift = float(fshift)
        n = A.n
        self.n = n
        L = lib()
        self.keep = []
        pci = [0] * 12
        if kind == "none":
            pci[0] = 0
        elif kind == "jacobi":
            self.diag = np.zeros(max(n, 1))
            didx = np.zeros(max(n, 1))
            L.or_pc_jacobi_setup(a(self.diag), jacobi_type, int(use_abs), int(fix_diag),
                                 float(rowl1_scale), a(A.irow), a(A.icol), a(A.data), n, a(didx))
            self.diagsqrt = np.zeros(max(n, 1))
            L.or_pc_jacobi_setup_symmetric(a(self.diagsqrt), jacobi_type, a(A.irow), a(A.icol),
                                           a(A.data), n)
            pci[0] = 2
            pci[1] = a(self.diag)
            pci[2] = a(self.diagsqrt)
        elif kind == "sor":
            k = max(n, 1)
            self.dmark = np.zeros(k, dtype=np.int64)
            L.or_mat_get_diagonal_markers(a(A.irow), a(A.icol), a(self.dmark), n)
            self.idiag = np.zeros(k)
            self.mdiag = np.zeros(k)
            self.sorwork = np.zeros(3 * k)
            self.flag = (12 if symmetric else 1) | 16
            pci[0] = 4
            pci[1], pci[2], pci[3] = a(A.irow), a(A.icol), a(A.data)
            pci[4], pci[5], pci[6], pci[7] = a(self.dmark), a(self.idiag), a(self.mdiag), a(self.sorwork)
            pci[10] = self.flag
        elif kind in ("ilu", "icc"):
            self.factor = build_factor(A, kind, levels, ordering)
            f = self.factor
            if kind == "ilu":
                pci[0] = 5
                pci[1:7] = [a(f["bi"]), a(f["bj"]), a(f["ba"]), a(f["bdiag"]), a(f["r"]), a(f["c"])]
            else:
                pci[0] = 6
                pci[1:6] = [a(f["bi"]), a(f["bj"]), a(f["ba"]), a(f["bdiag"]), a(f["r"])]
        else:
            raise NotImplementedError(kind)
        self.pci = np.array(pci, dtype=np.int64)
        self.pcf = np.array([self.omega, self.fshift], dtype=np.float64)
        self.pstmp = np.zeros(max(n, 1))

    def apply(self, x):
        """PC.py:590-623 _direct_apply."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        L = lib()
        A = self.A
        if self.kind == "none":
            y[:] = x
        elif self.kind == "jacobi":
            L.or_pc_jacobi_apply(a(y), a(x), a(self.diag), self.n)
        elif self.kind == "sor":
            L.or_pc_sor_apply(a(x), a(y), a(A.irow), a(A.icol), a(A.data), a(self.dmark),
                              a(self.idiag), a(self.mdiag), a(self.sorwork), self.omega,
                              self.fshift, self.flag, self.n)
        elif self.kind == "ilu":
            f = self.factor
            L.or_pc_ilu_apply(a(x), a(y), a(self.pstmp), a(f["bi"]), a(f["bj"]), a(f["ba"]),
                              a(f["bdiag"]), a(f["r"]), a(f["c"]), self.n)
        else:
            f = self.factor
            L.or_pc_icc_apply(a(x), a(y), a(f["bi"]), a(f["bj"]), a(f["ba"]), a(f["bdiag"]),
                              a(f["r"]), a(self.pstmp), self.n)
        return y


def permutations(A: CSR, ordering: str):
    """PC.py:343-361 _permutations."""
    n = A.n
    if ordering == "rcm" and n > 0:
        perm = np.zeros(n, dtype=np.int64)
        lv = np.zeros(n, dtype=np.int64)
        vis = np.zeros(n, dtype=np.int64)
        q = np.zeros(n, dtype=np.int64)
        lib().or_ordering_rcm(a(perm), a(lv), a(vis), a(q), a(A.irow), a(A.icol), n)
        r = perm
    else:
        r = np.arange(n, dtype=np.int64)
    ic = np.empty(n, dtype=np.int64)
    ic[r] = np.arange(n, dtype=np.int64)
    return r, ic


def symbolic(A: CSR, kind: str, levels: int, ordering: str = "natural"):
    """The symbolic half of PC.py:376-492: returns the factor's index arrays."""
    n = A.n
    L = lib()
    r, ic = permutations(A, ordering)
    adiag = np.zeros(n, dtype=np.int64)
    L.or_mat_get_diagonal_markers(a(A.irow), a(A.icol), a(adiag), n)
    scratch = np.zeros(7 * n + 8, dtype=np.int64)
    cap = max(4 * n, 4 * int(A.irow[-1]) + 4 * n, 16)
    if kind == "ilu":
        while True:
            bi = np.zeros(n + 1, dtype=np.int64)
            bj = np.zeros(cap, dtype=np.int64)
            bjlvl = np.zeros(cap, dtype=np.int64)
            bdiag = np.zeros(n + 1, dtype=np.int64)
            if levels == 0 and np.array_equal(r, np.arange(n)):
                nnz = L.or_ilu_symbolic_ilu0(a(bi), a(bj), a(bdiag), a(adiag), a(A.irow),
                                             a(A.icol), n)
                contiguous = False
            else:
                bjsrc = np.zeros(cap, dtype=np.int64)
                nnz = L.or_ilu_symbolic(a(bi), a(bjsrc), a(bjlvl), a(bdiag), a(A.irow),
                                        a(A.icol), a(r), a(ic), levels, cap, a(scratch), n)
                contiguous = True
            if nnz >= 0:
                if contiguous:
                    L.or_ilu_contiguous_lu(a(bjsrc), a(bi), a(bdiag), a(bj), n)
                break
            cap = -nnz + n
        return {"bi": bi, "bj": bj, "bdiag": bdiag, "r": r, "c": r, "ic": ic, "nnz": int(nnz)}
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
        nnz = L.or_icc_symbolic(a(ui), a(uj), a(ujlvl), a(udiag), a(A.irow), a(A.icol), a(r),
                                a(rinv), levels, cap, a(scratch), n, int(scratch.size))
        if nnz >= 0:
            break
        cap = -nnz
    bi = np.zeros(n + 1, dtype=np.int64)
    bdiag = np.zeros(n + 1, dtype=np.int64)
    bj = np.zeros(max(cap, int(ui[-1]) + 1), dtype=np.int64)
    bdiag[:n] = udiag[:n]
    L.or_icc_contiguous_cholesky(a(uj), a(ui), a(bdiag), a(bi), a(bdiag), a(bj), n)
    return {"bi": bi, "bj": bj, "bdiag": bdiag, "r": r, "rinv": rinv, "nnz": int(ui[-1]),
            "ui": ui, "ujlvl": ujlvl}


def build_factor(A: CSR, kind: str, levels: int = 0, ordering: str = "natural"):
    """PC.py:376-492 _build_factor: symbolic then numeric."""
    n = A.n
    L = lib()
    f = symbolic(A, kind, levels, ordering)
    rtmp = np.zeros(max(n, 1))
    if kind == "ilu":
        bdiag = f["bdiag"]
        ba = np.zeros(max(int(bdiag[0]) + 1, 1))
        L.or_pc_ilu_numeric(a(ba), a(rtmp), a(A.irow), a(A.icol), a(A.data), a(f["bi"]),
                            a(f["bj"]), a(bdiag), a(f["r"]), a(f["ic"]), n)
        f["ba"] = ba
    else:
        ua = np.zeros(max(int(f["ui"][-1]), 1))
        il = np.zeros(max(n, 1), dtype=np.int64)
        c2r = np.zeros(max(n, 1), dtype=np.int64)
        L.or_pc_icc_numeric(a(ua), a(rtmp), a(il), a(c2r), a(A.irow), a(A.icol), a(A.data),
                            a(f["bi"]), a(f["bj"]), a(f["bdiag"]), a(f["r"]), a(f["rinv"]), n)
        f["ba"] = ua
    return f


# ------------------------------------------------------------------ KSP

DEFAULTS = dict(rtol=1e-5, atol=1e-50, divtol=1e4, max_it=1000, restart=30, haptol=1e-30,
                breakdowntol=0.1, orthog=0, cgstype=0, side=0, normtype=0, min_it=0,
                chknorm=-1, convmaxits=0, guess_zero=1)


class Result:
    def __init__(self, x, its, reason, rnorm, history):
        self.x, self.its, self.reason, self.rnorm, self.history = x, its, reason, rnorm, history

    def __repr__(self):
        return f"Result(its={self.its}, reason={self.reason}, rnorm={self.rnorm:.6e})"


def ksp_solve(A: CSR, b, ksp_type: str, pc: PCSetup | None = None, x0=None, **opts) -> Result:
    """KSP.py:367-449 solve, with setTolerances(max_it=...) setting convmaxits
    (KSP.py:212-225) when `max_it` is given, exactly as the Python layer does."""
    o = dict(DEFAULTS)
    if "max_it" in opts:
        o["convmaxits"] = 1
    o.update(opts)
    n = A.n
    L = lib()
    if pc is None:
        pc = PCSetup(A, "none")
    b = np.ascontiguousarray(b, dtype=np.float64).copy()
    x = np.zeros(n) if (x0 is None or o["guess_zero"]) else np.ascontiguousarray(x0, dtype=np.float64).copy()
    if o["guess_zero"]:
        x[:] = 0.0
    normtype = o["normtype"] or 2  # NORM_DEFAULT -> PRECONDITIONED (KSP.py:451-455)
    max_it = int(o["max_it"])
    cfg = np.zeros(16 + max_it + 2)
    cfg[0], cfg[1], cfg[2], cfg[3], cfg[4] = o["rtol"], o["atol"], o["divtol"], o["haptol"], o["breakdowntol"]
    cfgi = np.zeros(16, dtype=np.int64)
    cfgi[0], cfgi[1], cfgi[2], cfgi[3] = normtype, o["side"], o["guess_zero"], o["min_it"]
    cfgi[4], cfgi[5], cfgi[6], cfgi[7] = o["chknorm"], o["convmaxits"], max_it, o["restart"]
    cfgi[8], cfgi[9] = o["orthog"], o["cgstype"]
    cfg[12] = rhs_norm(b, pc, normtype, o["side"]) if not o["guess_zero"] else 0.0
    cfg[14] = 0.0
    cfgi[12] = 0 if o["guess_zero"] else 1
    if ksp_type in ("gmres", "fgmres"):
        mk = int(o["restart"])
        hs = np.zeros((mk + 2) * (mk + 1) + (mk + 1) ** 2 + 2 * (mk + 1) + 2 * (mk + 2))
        work = np.zeros(max(2 * n, 1))
        vv = np.zeros(max((mk + 1) * n, 1))
        if ksp_type == "gmres":
            L.or_ksp_solve_gmres(n, a(A.irow), a(A.icol), a(A.data), a(pc.pci), a(pc.pcf),
                                 a(pc.pstmp), a(work), a(vv), a(hs), a(cfg), a(cfgi), a(b), a(x))
        else:
            prevec = np.zeros(max((mk + 2) * n, 1))
            L.or_ksp_solve_fgmres(n, a(A.irow), a(A.icol), a(A.data), a(pc.pci), a(pc.pcf),
                                  a(pc.pstmp), a(work), a(vv), a(prevec), a(hs), a(cfg), a(cfgi),
                                  a(b), a(x))
    else:
        work = np.zeros(max(7 * n, 1))
        fn = L.or_ksp_solve_cg if ksp_type == "cg" else L.or_ksp_solve_bcgs
        fn(n, a(A.irow), a(A.icol), a(A.data), a(pc.pci), a(pc.pcf), a(pc.pstmp), a(work),
           a(cfg), a(cfgi), a(b), a(x))
    its = int(cfgi[10])
    hist = cfg[9:9 + its + 1].copy()
    return Result(x, its, int(cfgi[11]), float(cfg[7]), hist)


def rhs_norm(b, pc: PCSetup, normtype: int, side: int) -> float:
    """KSP.py:496-507 _rhs_norm.  In the reference order the 2-norms are
    numpy's, as there; in the device order they are the tile-tree norm."""
    L = lib()
    if normtype == 1:
        return 0.0

    def nrm(v):
        if L.or_get_reduction_mode() == 0:
            return float(np.linalg.norm(v))  # KSP.py:502,505 use numpy's norm
        z = np.zeros(2)
        L.or_vec_norm(a(np.ascontiguousarray(v)), 1, a(z), len(v))
        return float(z[0])

    if normtype == 3 or side == 1:
        return nrm(b)
    z = pc.apply(b)
    if normtype == 2:
        return nrm(z)
    d = np.zeros(1)
    L.or_vec_dot(a(z), a(b), a(d), len(b))
    return math.sqrt(abs(float(d[0])))
