"""The C ABI: libmpk.so loads without a GPU and exports every entry point
include/mpk.h declares, with the arity and argument kinds the ctypes table
gives them; the reference's 59 mp_* names (python/mojopetsc4py/_lib.py:23-87)
are all there."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mpk.h")

REFERENCE_NAMES = """
mp_mat_mult mp_mat_mult_add mp_mat_mult_transpose mp_mat_get_diagonal mp_mat_scale mp_mat_zero_entries
mp_vec_pointwise_mult mp_vec_pointwise_divide mp_vec_pointwise_min mp_vec_pointwise_max mp_vec_pointwise_max_abs
mp_vec_norm mp_vec_dot mp_vec_tdot mp_vec_sum mp_vec_mean mp_vec_dot_norm2 mp_vec_max mp_vec_min mp_vec_normalize
mp_vec_reciprocal mp_vec_exp mp_vec_log mp_vec_sqrtabs mp_vec_chop mp_vec_abs mp_vec_conjugate mp_vec_copy
mp_vec_swap mp_vec_set mp_vec_zero_entries mp_vec_scale mp_vec_shift mp_vec_axpy mp_vec_aypx mp_vec_axpby
mp_vec_axpbypcz mp_vec_waxpy mp_vec_maxpy mp_vec_max_pointwise_divide mp_mat_get_diagonal_markers
mp_pc_jacobi_setup mp_pc_jacobi_setup_symmetric mp_pc_jacobi_apply mp_pc_sor_apply mp_pc_ilu_apply mp_pc_icc_apply
mp_pc_ilu_numeric mp_pc_icc_numeric mp_ilu_symbolic_ilu0 mp_ilu_symbolic mp_icc_symbolic mp_ilu_contiguous_lu
mp_ordering_rcm mp_icc_contiguous_cholesky mp_ksp_solve_cg mp_ksp_solve_bcgs mp_ksp_solve_gmres mp_ksp_solve_fgmres
""".split()


def header_decls():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"(void|int64_t|double)\s+(mpk?_\w+)\s*\(([^)]*)\)\s*;", text):
        ret, name, args = m.groups()
        args = [a.strip() for a in args.split(",") if a.strip() and a.strip() != "void"]
        kinds = ["F" if a.startswith("double") else "I" for a in args]
        decls[name] = (ret, kinds)
    return decls


@pytest.fixture(scope="module")
def lib():
    from mpk import _native
    return _native.load(init=False)


def test_reference_names_all_declared():
    d = header_decls()
    assert len(REFERENCE_NAMES) == 59
    assert set(REFERENCE_NAMES) <= set(d)


def test_every_declared_symbol_is_exported(lib):
    for name in header_decls():
        assert hasattr(lib, name), name


def test_ctypes_table_matches_header():
    from mpk import _native
    d = header_decls()
    assert set(_native.SIGNATURES) == set(d)
    for name, (argt, rest) in _native.SIGNATURES.items():
        ret, kinds = d[name]
        got = ["F" if t is ctypes.c_double else "I" for t in argt]
        assert got == kinds, name
        want_ret = {"void": None, "int64_t": ctypes.c_int64, "double": ctypes.c_double}[ret]
        assert rest is want_ret, name


def test_no_device_fails_loudly():
    """Without a GPU the library refuses to run: there is no CPU fallback."""
    from mpk import _native
    lib_ = _native.load(init=False)
    if lib_.mpk_init(-1) == 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(_native.NativeError):
        _native.lib()


def test_sell_slot_count_guard(lib):
    """The sliced copy's padded slot count, host side: 64 slots per slice row
    at the slice's longest row, and -1 once it passes the int32 slot
    indices (2^31 - 1), where the products keep the CSR row kernel."""
    import numpy as np
    from mpk import workloads as W
    ip, _, _ = W.laplacian_2d(20)
    rp = np.asarray(ip, dtype=np.int64)
    n = len(rp) - 1
    lens = np.diff(rp)
    want = sum(64 * int(lens[s:s + 64].max()) for s in range(0, n, 64))
    assert lib.mpk_sell_size_host(rp.ctypes.data, n) == want
    # one dense row: nnz fits int32, but its slice pads to 64 x the row
    fits = np.array([0, (2**31 - 1) // 64], dtype=np.int64)
    assert lib.mpk_sell_size_host(fits.ctypes.data, 1) == 64 * ((2**31 - 1) // 64)
    past = np.array([0, 2**25], dtype=np.int64)
    assert lib.mpk_sell_size_host(past.ctypes.data, 1) == -1
    two = np.array([0, 2**24, 2**25 + 2**24], dtype=np.int64)  # the second row sets the slice's width
    assert lib.mpk_sell_size_host(two.ctypes.data, 2) == -1


def test_sweep_wait_limit_arguments(lib):
    """The sweep waits' poll limit: a value outside 1..2^32-1 is error 4 and
    leaves the limit as it was; a valid one returns the previous limit (the
    default 2^24).  No device is needed."""
    lib.mpk_clear_error()
    old = lib.mpk_set_sweep_wait_limit(0)
    assert lib.mpk_status() == 4
    lib.mpk_clear_error()
    assert lib.mpk_set_sweep_wait_limit(2**32) == old
    assert lib.mpk_status() == 4
    lib.mpk_clear_error()
    assert old == 2**24
    assert lib.mpk_set_sweep_wait_limit(5) == old
    assert lib.mpk_set_sweep_wait_limit(old) == 5
    assert lib.mpk_status() == 0
