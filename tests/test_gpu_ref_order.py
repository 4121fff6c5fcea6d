"""The reference's reduction order on the GPU (mpk_set_reduction_order(1)):
every reduction and whole solve against the oracle in its reference order
(src/simdvec.mojo:332-473 with SW = 4, BLK = 16; _mdot_basis,
src/ksp.mojo:956-1062), bit for bit.  The lengths cross every boundary of
the kernel's loops: the 16-element blocks, the SW-steps and scalar tail, and
the chunks it stages through LDS (2048, 4096 and 8192 elements of each
operand for four, two and one operands)."""

import numpy as np
import pytest

from _helpers import back, compare, dev, rng_vec, same_bits
from mpk import workloads as W

pytestmark = pytest.mark.gpu

LENGTHS = [0, 1, 3, 4, 5, 15, 16, 17, 19, 20, 31, 32, 33, 2047, 2048, 2049, 4095, 4096, 4111, 8191, 8192, 8209,
           16387, 100003, 1048583]


@pytest.fixture(scope="module")
def O(oracle):
    oracle.set_mode("ref")
    return oracle


@pytest.fixture()
def ref_order(native):
    import mpk
    old = mpk.set_reduction_order("ref")
    yield
    mpk.set_reduction_order(old)


def test_order_switch(native):
    import mpk
    old = mpk.set_reduction_order("ref")
    try:
        assert mpk.reduction_order() == "ref"
        assert native.mpk_set_reduction_order(7) == 1  # rejected, order kept
        assert native.mpk_status() == 4
        native.mpk_clear_error()
        assert mpk.reduction_order() == "ref"
    finally:
        mpk.set_reduction_order(old)
    with pytest.raises(ValueError):
        mpk.set_reduction_order("fast")


@pytest.mark.parametrize("n", LENGTHS)
def test_reductions_bitwise(native, O, ref_order, n):
    x, y = rng_vec(n, 21), rng_vec(n, 22)
    dx, dy = dev(x if n else np.zeros(1)), dev(y if n else np.zeros(1))
    for ntype in (0, 1, 2, 3, 4):
        zg, zo = np.zeros(2), np.zeros(2)
        native.mp_vec_norm(dx.ptr, ntype, zg.ctypes.data, n)
        O.lib().or_vec_norm(x.ctypes.data, ntype, zo.ctypes.data, n)
        assert same_bits(zg, zo), (ntype, zg, zo)
    for name in ("dot", "tdot"):
        g, o = np.zeros(1), np.zeros(1)
        getattr(native, "mp_vec_" + name)(dx.ptr, dy.ptr, g.ctypes.data, n)
        getattr(O.lib(), "or_vec_" + name)(x.ctypes.data, y.ctypes.data, o.ctypes.data, n)
        assert same_bits(g, o), name
    g, o = np.zeros(2), np.zeros(2)
    native.mp_vec_dot_norm2(dx.ptr, dy.ptr, g.ctypes.data, g.ctypes.data + 8, n)
    O.lib().or_vec_dot_norm2(x.ctypes.data, y.ctypes.data, o.ctypes.data, o.ctypes.data + 8, n)
    assert same_bits(g, o)
    for name in ("sum", "mean"):
        if name == "mean" and n == 0:
            continue
        g, o = np.zeros(1), np.zeros(1)
        getattr(native, "mp_vec_" + name)(dx.ptr, g.ctypes.data, n)
        getattr(O.lib(), "or_vec_" + name)(x.ctypes.data, o.ctypes.data, n)
        assert same_bits(g, o), name
    if n:
        g, o = np.zeros(1), np.zeros(1)
        xo = x.copy()
        native.mp_vec_normalize(dx.ptr, g.ctypes.data, n)
        O.lib().or_vec_normalize(xo.ctypes.data, o.ctypes.data, n)
        assert same_bits(g, o) and same_bits(back(dx, np.float64, n), xo)


@pytest.mark.parametrize("n", [17, 4111, 100003])
def test_unaligned_operands(native, O, ref_order, n):
    """Operands at odd element offsets (8-byte, not 16-byte, aligned), as the
    GMRES basis vectors of odd-length systems are."""
    x, y = rng_vec(n + 1, 23), rng_vec(n + 1, 24)
    dx, dy = dev(x), dev(y)
    g, o = np.zeros(2), np.zeros(2)
    native.mp_vec_dot_norm2(dx.ptr + 8, dy.ptr + 8, g.ctypes.data, g.ctypes.data + 8, n)
    xs, ys = np.ascontiguousarray(x[1:]), np.ascontiguousarray(y[1:])
    O.lib().or_vec_dot_norm2(xs.ctypes.data, ys.ctypes.data, o.ctypes.data, o.ctypes.data + 8, n)
    assert same_bits(g, o)


def test_special_values(native, O, ref_order):
    """NaN and infinity travel through the chains as through the oracle's
    loops; sums of signed zeros end at +0.0 in both."""
    for arr in (np.array([1.0, np.nan] + [2.0] * 30), np.array([np.inf] + [1.0] * 40),
                np.array([-0.0] * 37), np.array([1e308, 1e308] + [-1e308] * 20)):
        n = len(arr)
        d = dev(arr)
        for ntype in (0, 1, 4):
            zg, zo = np.zeros(2), np.zeros(2)
            native.mp_vec_norm(d.ptr, ntype, zg.ctypes.data, n)
            O.lib().or_vec_norm(arr.ctypes.data, ntype, zo.ctypes.data, n)
            assert np.array_equal(np.isnan(zg), np.isnan(zo))
            assert same_bits(np.nan_to_num(zg), np.nan_to_num(zo)), (arr[:3], ntype, zg, zo)
        g, o = np.zeros(1), np.zeros(1)
        native.mp_vec_sum(d.ptr, g.ctypes.data, n)
        O.lib().or_vec_sum(arr.ctypes.data, o.ctypes.data, n)
        assert np.isnan(g[0]) == np.isnan(o[0]) and same_bits(np.nan_to_num(g), np.nan_to_num(o))


LAP64 = W.laplacian_2d(64)
BENCH_B = np.random.default_rng(5).normal(size=64 * 64)
DIFF16, DIFF16B = W.diffusion_3d(16)
CD12, CD12B = W.convection_diffusion_3d(12)


@pytest.mark.parametrize("ksp_type,pc_type,its", [("cg", "ilu", 77), ("cg", "none", 225), ("cg", "icc", 77),
                                                  ("gmres", "none", 626), ("fgmres", "none", 626),
                                                  ("gmres", "jacobi", 626)])
def test_published_counts(native, O, ref_order, ksp_type, pc_type, its):
    """The reference's bench cases (README.md:302-321) in its own order."""
    g = compare(O, LAP64, BENCH_B, ksp_type, pc_type, rtol=1e-10, atol=1e-50, divtol=1e30, max_it=2000)
    assert g[1] == its


@pytest.mark.parametrize("ksp_type,pc_type,levels", [
    ("cg", "jacobi", 0), ("cg", "icc", 0), ("cg", "icc", 1), ("cg", "ilu", 0), ("cg", "sor", 0),
    ("bcgs", "jacobi", 0), ("bcgs", "sor", 0), ("gmres", "ilu", 0), ("fgmres", "ilu", 0),
])
def test_diffusion_16(native, O, ref_order, ksp_type, pc_type, levels):
    compare(O, DIFF16, DIFF16B, ksp_type, pc_type, levels, rtol=1e-10, max_it=3000)


@pytest.mark.parametrize("opts", [dict(), dict(cgstype=1), dict(cgstype=2), dict(orthog=1), dict(restart=7),
                                  dict(side=1)])
def test_gmres_variants(native, O, ref_order, opts):
    compare(O, CD12, CD12B, "gmres", "ilu", rtol=1e-10, max_it=2000, **opts)


def test_nonzero_guess(native, O, ref_order):
    """A nonzero guess seeds the test with the right-hand side's norm, which
    the reference takes with numpy (KSP.py:496-507)."""
    x0 = np.random.default_rng(11).normal(size=len(DIFF16B))
    for kt, pt in (("cg", "jacobi"), ("gmres", "ilu"), ("bcgs", "jacobi")):
        compare(O, DIFF16, DIFF16B, kt, pt, x0=x0, rtol=1e-8, max_it=2000)


@pytest.mark.parametrize("normtype", [1, 3, 4])
def test_cg_norm_types(native, O, ref_order, normtype):
    compare(O, DIFF16, DIFF16B, "cg", "jacobi", normtype=normtype, rtol=1e-10, max_it=300)


def test_gmres_restart_past_256(native, O, ref_order):
    """Reference-order Gram-Schmidt sums over a basis past 256 vectors."""
    n = 400
    csr = (np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.logspace(0, 6, n))
    b = np.random.default_rng(13).normal(size=n)
    g = compare(O, csr, b, "gmres", "none", rtol=1e-13, max_it=320, restart=300)
    assert g[1] > 257
