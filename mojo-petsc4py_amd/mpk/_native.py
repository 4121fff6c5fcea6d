"""Loader for libmpk.so, the HIP library behind the PETSc surface.

The library is built in-tree by mojo-petsc4py_amd/csrc/build.py.  There is no
CPU fallback: if the library is missing, or no HIP device is usable, `lib()`
raises and nothing runs.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.environ.get("MPK_LIB") or os.path.join(HERE, "libmpk.so")

I = ctypes.c_int64
F = ctypes.c_double

# name -> (argtypes, restype); the mp_* rows mirror the reference's table
# (python/mojopetsc4py/_lib.py:23-87) entry for entry.
SIGNATURES = {
    # Mat
    "mp_mat_mult": ([I, I, I, I, I, I], None),
    "mp_mat_mult_add": ([I, I, I, I, I, I], None),
    "mp_mat_mult_transpose": ([I, I, I, I, I, I, I], None),
    "mp_mat_get_diagonal": ([I, I, I, I, I], None),
    "mp_mat_scale": ([I, I, F], None),
    "mp_mat_zero_entries": ([I, I], None),
    # Vec
    "mp_vec_pointwise_mult": ([I, I, I, I], None),
    "mp_vec_pointwise_divide": ([I, I, I, I], None),
    "mp_vec_pointwise_min": ([I, I, I, I], None),
    "mp_vec_pointwise_max": ([I, I, I, I], None),
    "mp_vec_pointwise_max_abs": ([I, I, I, I], None),
    "mp_vec_norm": ([I, I, I, I], None),
    "mp_vec_dot": ([I, I, I, I], None),
    "mpk_vec_dot_noalloc": ([I, I, I, I], I),
    "mp_vec_tdot": ([I, I, I, I], None),
    "mp_vec_sum": ([I, I, I], None),
    "mp_vec_mean": ([I, I, I], None),
    "mp_vec_dot_norm2": ([I, I, I, I, I], None),
    "mp_vec_max": ([I, I, I, I], None),
    "mp_vec_min": ([I, I, I, I], None),
    "mp_vec_normalize": ([I, I, I], None),
    "mp_vec_reciprocal": ([I, I], None),
    "mp_vec_exp": ([I, I], None),
    "mp_vec_log": ([I, I], None),
    "mp_vec_sqrtabs": ([I, I], None),
    "mp_vec_chop": ([I, I, F], None),
    "mp_vec_abs": ([I, I], None),
    "mp_vec_conjugate": ([I, I], None),
    "mp_vec_copy": ([I, I, I], None),
    "mp_vec_swap": ([I, I, I], None),
    "mp_vec_set": ([I, F, I], None),
    "mp_vec_zero_entries": ([I, I], None),
    "mp_vec_scale": ([I, F, I], None),
    "mp_vec_shift": ([I, F, I], None),
    "mp_vec_axpy": ([I, F, I, I], None),
    "mp_vec_aypx": ([I, F, I, I], None),
    "mp_vec_axpby": ([I, F, F, I, I], None),
    "mp_vec_axpbypcz": ([I, F, F, F, I, I, I], None),
    "mp_vec_waxpy": ([I, F, I, I, I], None),
    "mp_vec_maxpy": ([I, I, I, I, I], None),
    "mp_vec_max_pointwise_divide": ([I, I, I, I], None),
    # PC
    "mp_mat_get_diagonal_markers": ([I, I, I, I], I),
    "mp_pc_jacobi_setup": ([I, I, I, I, F, I, I, I, I, I], None),
    "mp_pc_jacobi_setup_symmetric": ([I, I, I, I, I, I], None),
    "mp_pc_jacobi_apply": ([I, I, I, I], None),
    "mp_pc_sor_apply": ([I, I, I, I, I, I, I, I, I, F, F, I, I], None),
    "mp_pc_ilu_apply": ([I] * 10, None),
    "mp_pc_icc_apply": ([I] * 9, I),
    "mp_pc_ilu_numeric": ([I] * 11, I),
    "mp_pc_icc_numeric": ([I] * 13, None),
    "mp_ilu_symbolic_ilu0": ([I] * 7, I),
    "mp_ilu_symbolic": ([I] * 12, I),
    "mp_icc_symbolic": ([I] * 13, I),
    "mp_ilu_contiguous_lu": ([I] * 5, None),
    "mp_ordering_rcm": ([I] * 7, None),
    "mp_icc_contiguous_cholesky": ([I] * 7, None),
    # KSP
    "mp_ksp_solve_cg": ([I] * 12, None),
    "mp_ksp_solve_bcgs": ([I] * 12, None),
    "mp_ksp_solve_gmres": ([I] * 14, None),
    "mp_ksp_solve_fgmres": ([I] * 15, None),
    # runtime and additions without a reference counterpart
    "mpk_init": ([I], I),
    "mpk_status": ([], I),
    "mpk_last_error": ([I, I], I),
    "mpk_clear_error": ([], None),
    "mpk_device_name": ([I, I], I),
    "mpk_malloc": ([I], I),
    "mpk_free": ([I], None),
    "mpk_memcpy_htod": ([I, I, I], None),
    "mpk_memcpy_dtoh": ([I, I, I], None),
    "mpk_memcpy_dtod": ([I, I, I], None),
    "mpk_memset": ([I, I, I], None),
    "mpk_synchronize": ([], None),
    "mpk_set_reduction_order": ([I], I),
    "mpk_get_reduction_order": ([], I),
    "mpk_set_sweep_wait_limit": ([I], I),
    "mpk_sweep_abort": ([I], I),
    "mpk_small_solves": ([], I),
    "mpk_csr_transpose": ([I] * 7, I),
    "mpk_mat_mult_transpose_t": ([I] * 7, None),
    "mpk_pc_icc_apply_t": ([I] * 13, I),
    "mpk_time_spmv": ([I] * 10, F),
    "mpk_time_cg_wdot": ([I] * 8, F),
    "mpk_sell_size": ([I, I], I),
    "mpk_sell_size_host": ([I, I], I),
    "mpk_sell_build": ([I] * 7, I),
    "mpk_sell_mult": ([I] * 8, None),
    "mpk_batch_create": ([I] * 21, I),
    "mpk_batch_solve": ([I] * 5, I),
    "mpk_batch_free": ([I], None),
    "mpk_batch_variants": ([I, I], I),
    "mpk_mixed_create": ([I] * 13, I),
    "mpk_mixed_solve": ([I] * 4, I),
    "mpk_mixed_free": ([I], None),
    "mpk_pc_schedule_size": ([I] * 3, I),
    "mpk_pc_schedule": ([I] * 5, I),
    "mpk_sweep_plan_bound": ([I] * 3, I),
    "mpk_sweep_plan": ([I] * 8, I),
    "mpk_pc_apply": ([I] * 5, None),
    "mpk_icc_transpose_size": ([I, I, I], I),
    "mpk_icc_transpose": ([I] * 8, I),
    "mpk_ilu_levels": ([I] * 6, I),
    "mpk_ilu_numeric": ([I] * 13, I),
    "mpk_icc_visits": ([I] * 12, I),
    "mpk_icc_numeric": ([I] * 16, I),
    "mpk_ic0_size": ([I] * 3, I),
    "mpk_ic0_analyse": ([I] * 14, I),
    "mpk_ilu0_size": ([I] * 3, I),
    "mpk_ilu0_analyse": ([I] * 8, I),
    "mpk_mat_fold": ([I] * 10, I),
    "mpk_dense_to_csr": ([I] * 7, I),
}

# exported names that only a device call can exercise
MP_NAMES = [k for k in SIGNATURES if k.startswith("mp_")]


class NativeError(RuntimeError):
    pass


_lib = None


def load(init: bool = True) -> ctypes.CDLL:
    """Load libmpk.so and set the signatures; with `init`, also open the HIP
    device and raise if there is none."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise NativeError(
                f"{LIB} is not built; run python mojo-petsc4py_amd/csrc/build.py")
        lib_ = ctypes.CDLL(LIB)
        for name, (argt, rest) in SIGNATURES.items():
            fn = getattr(lib_, name)
            fn.argtypes = argt
            fn.restype = rest
        _lib = lib_
    if init:
        rc = _lib.mpk_init(-1)
        if rc != 0:
            raise NativeError(f"libmpk could not open a HIP device: {last_error()}")
    return _lib


def lib() -> ctypes.CDLL:
    return load(True)


_ready = None


def device_ready() -> bool:
    """True when the library is built and opens a device (checked once).  For
    the optional device paths of host-side work (Mat assembly); the solver
    paths call lib(), which raises instead."""
    global _ready
    if _ready is None:
        try:
            load(True)
            _ready = True
        except (NativeError, OSError):
            _ready = False
    return _ready


def last_error() -> str:
    if _lib is None:
        return ""
    buf = ctypes.create_string_buffer(1024)
    _lib.mpk_last_error(ctypes.addressof(buf), 1024)
    return buf.value.decode(errors="replace")


def check() -> None:
    """Raise if the library has recorded a failure."""
    if _lib is not None and _lib.mpk_status() != 0:
        msg = last_error()
        _lib.mpk_clear_error()
        raise NativeError(msg)


_ORDERS = {"dev": 0, "ref": 1}


def set_reduction_order(order: str) -> str:
    """Select the order of every reduction the library computes: "dev" (the
    default, a fixed tile tree) or "ref" (the reference's SIMD order,
    src/simdvec.mojo:332-473 with SW = 4, BLK = 16).  Returns the previous
    order."""
    if order not in _ORDERS:
        raise ValueError(f"reduction order must be 'dev' or 'ref', not {order!r}")
    old = lib().mpk_set_reduction_order(_ORDERS[order])
    check()
    return "ref" if old == 1 else "dev"


def reduction_order() -> str:
    return "ref" if lib().mpk_get_reduction_order() == 1 else "dev"


class reduction_order_scope:
    """`with reduction_order_scope("ref"): ...` -- the order inside the block."""

    def __init__(self, order: str):
        self.order = order

    def __enter__(self):
        self.old = set_reduction_order(self.order)
        return self

    def __exit__(self, *exc):
        set_reduction_order(self.old)


def device_name() -> str:
    buf = ctypes.create_string_buffer(256)
    lib().mpk_device_name(ctypes.addressof(buf), 256)
    return buf.value.decode(errors="replace")
