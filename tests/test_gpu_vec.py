"""Vec kernels on the GPU against the oracle (device reduction order):
elementwise results bit for bit, reductions bit for bit."""

import numpy as np
import pytest

from _helpers import back, dev, rng_vec, same_bits

pytestmark = pytest.mark.gpu

LENGTHS = [0, 1, 2, 7, 511, 512, 2047, 2048, 2049, 4097, 100003]


@pytest.fixture(scope="module")
def O(oracle):
    oracle.set_mode("dev")
    yield oracle
    oracle.set_mode("ref")


def _oracle(O, name, *args):
    getattr(O.lib(), name)(*args)


@pytest.mark.parametrize("n", LENGTHS)
def test_reductions_bitwise(native, O, n):
    x, y = rng_vec(n, 1), rng_vec(n, 2)
    dx, dy = dev(x if n else np.zeros(1)), dev(y if n else np.zeros(1))
    for ntype in (0, 1, 2, 3, 4):
        zg = np.zeros(2)
        zo = np.zeros(2)
        native.mp_vec_norm(dx.ptr, ntype, zg.ctypes.data, n)
        O.lib().or_vec_norm(x.ctypes.data, ntype, zo.ctypes.data, n)
        assert same_bits(zg, zo), (ntype, zg, zo)
    for name in ("mp_vec_dot", "mp_vec_tdot"):
        g = np.zeros(1)
        o = np.zeros(1)
        getattr(native, name)(dx.ptr, dy.ptr, g.ctypes.data, n)
        _oracle(O, name.replace("mp_", "or_"), x.ctypes.data, y.ctypes.data, o.ctypes.data, n)
        assert same_bits(g, o), name
    g, o = np.zeros(2), np.zeros(2)
    native.mp_vec_dot_norm2(dx.ptr, dy.ptr, g.ctypes.data, g.ctypes.data + 8, n)
    O.lib().or_vec_dot_norm2(x.ctypes.data, y.ctypes.data, o.ctypes.data, o.ctypes.data + 8, n)
    assert same_bits(g, o)
    g, o = np.zeros(1), np.zeros(1)
    native.mp_vec_sum(dx.ptr, g.ctypes.data, n)
    O.lib().or_vec_sum(x.ctypes.data, o.ctypes.data, n)
    assert same_bits(g, o)
    if n:
        for name in ("mp_vec_max", "mp_vec_min"):
            gi, gv = np.zeros(1, np.int64), np.zeros(1)
            oi, ov = np.zeros(1, np.int64), np.zeros(1)
            getattr(native, name)(dx.ptr, gi.ctypes.data, gv.ctypes.data, n)
            _oracle(O, name.replace("mp_", "or_"), x.ctypes.data, oi.ctypes.data, ov.ctypes.data, n)
            assert gi[0] == oi[0] and same_bits(gv, ov), name
        g, o = np.zeros(1), np.zeros(1)
        native.mp_vec_max_pointwise_divide(dx.ptr, dy.ptr, g.ctypes.data, n)
        O.lib().or_vec_max_pointwise_divide(x.ctypes.data, y.ctypes.data, o.ctypes.data, n)
        assert same_bits(g, o)


def test_minmax_ties_and_nan(native, O):
    x = np.array([1.0, 3.0, 3.0, -2.0, np.nan, 3.0, -2.0])
    for name in ("mp_vec_max", "mp_vec_min"):
        gi, gv = np.zeros(1, np.int64), np.zeros(1)
        oi, ov = np.zeros(1, np.int64), np.zeros(1)
        dx = dev(x)
        getattr(native, name)(dx.ptr, gi.ctypes.data, gv.ctypes.data, len(x))
        _oracle(O, name.replace("mp_", "or_"), x.ctypes.data, oi.ctypes.data, ov.ctypes.data, len(x))
        assert gi[0] == oi[0] and same_bits(gv, ov)
    x = np.array([np.nan, 5.0, -1.0])
    gi, gv = np.zeros(1, np.int64), np.zeros(1)
    dx = dev(x)
    native.mp_vec_max(dx.ptr, gi.ctypes.data, gv.ctypes.data, 3)
    assert gi[0] == 0 and np.isnan(gv[0])
    z = np.zeros(2)
    dn = dev(np.array([1.0, np.nan, 4.0]))
    native.mp_vec_norm(dn.ptr, 3, z.ctypes.data, 3)
    assert np.isnan(z[0])


ELEMENTWISE = [
    ("mp_vec_pointwise_mult", "w3"), ("mp_vec_pointwise_divide", "w3"), ("mp_vec_pointwise_min", "w3"),
    ("mp_vec_pointwise_max", "w3"), ("mp_vec_pointwise_max_abs", "w3"),
    ("mp_vec_reciprocal", "x1"), ("mp_vec_sqrtabs", "x1"), ("mp_vec_abs", "x1"), ("mp_vec_conjugate", "x1"),
    ("mp_vec_zero_entries", "x1"),
    ("mp_vec_scale", "xa"), ("mp_vec_shift", "xa"), ("mp_vec_set", "xa"), ("mp_vec_chop", "xt"),
    ("mp_vec_axpy", "yax"), ("mp_vec_aypx", "yax"), ("mp_vec_axpby", "yabx"), ("mp_vec_axpbypcz", "zabcxy"),
    ("mp_vec_waxpy", "waxy"), ("mp_vec_copy", "xy"), ("mp_vec_swap", "xy"),
]


@pytest.mark.parametrize("name,form", ELEMENTWISE)
@pytest.mark.parametrize("n", [1, 7, 2049, 100003])
def test_elementwise_bitwise(native, O, name, form, n):
    x, y, z = rng_vec(n, 3), rng_vec(n, 4) + 3.0, rng_vec(n, 5)
    x[n // 2] = 0.0
    fn, ofn = getattr(native, name), getattr(O.lib(), name.replace("mp_", "or_"))
    hx, hy, hz = x.copy(), y.copy(), z.copy()
    dx, dy, dz = dev(x), dev(y), dev(z)
    A = lambda b: b.ptr  # noqa: E731
    H = lambda a: a.ctypes.data  # noqa: E731
    if form == "w3":
        fn(A(dz), A(dx), A(dy), n)
        ofn(H(hz), H(hx), H(hy), n)
    elif form == "x1":
        fn(A(dx), n)
        ofn(H(hx), n)
    elif form == "xa":
        for alpha in (0.0, 1.0, -1.5):
            fn(A(dx), alpha, n)
            ofn(H(hx), alpha, n)
    elif form == "xt":
        fn(A(dx), n, 0.5)
        ofn(H(hx), n, 0.5)
    elif form == "yax":
        for alpha in (0.0, 1.0, -1.0, 0.37):
            fn(A(dy), alpha, A(dx), n)
            ofn(H(hy), alpha, H(hx), n)
    elif form == "yabx":
        for al, be in ((0.3, 0.0), (0.0, 0.7), (1.25, -0.5)):
            fn(A(dy), al, be, A(dx), n)
            ofn(H(hy), al, be, H(hx), n)
    elif form == "zabcxy":
        fn(A(dz), 0.3, -1.7, 0.9, A(dx), A(dy), n)
        ofn(H(hz), 0.3, -1.7, 0.9, H(hx), H(hy), n)
    elif form == "waxy":
        fn(A(dz), -0.61, A(dx), A(dy), n)
        ofn(H(hz), -0.61, H(hx), H(hy), n)
    elif form == "xy":
        fn(A(dx), A(dy), n)
        ofn(H(hx), H(hy), n)
    for d, h in ((dx, hx), (dy, hy), (dz, hz)):
        assert same_bits(back(d, np.float64, n), h), name


@pytest.mark.parametrize("n", [5, 2049, 65537])
def test_maxpy_and_normalize(native, O, n):
    y = rng_vec(n, 6)
    xs = [rng_vec(n, 10 + k) for k in range(19)]
    alpha = np.random.default_rng(7).normal(size=19)
    dy = dev(y)
    dxs = [dev(v) for v in xs]
    addrs = np.array([d.ptr for d in dxs], dtype=np.int64)
    native.mp_vec_maxpy(dy.ptr, 19, alpha.ctypes.data, addrs.ctypes.data, n)
    haddrs = np.array([v.ctypes.data for v in xs], dtype=np.int64)
    hy = y.copy()
    O.lib().or_vec_maxpy(hy.ctypes.data, 19, alpha.ctypes.data, haddrs.ctypes.data, n)
    assert same_bits(back(dy, np.float64, n), hy)
    g, o = np.zeros(1), np.zeros(1)
    native.mp_vec_normalize(dy.ptr, g.ctypes.data, n)
    O.lib().or_vec_normalize(hy.ctypes.data, o.ctypes.data, n)
    assert same_bits(g, o) and same_bits(back(dy, np.float64, n), hy)


def test_exp_log_close(native, O):
    n = 4097
    x = rng_vec(n, 8)
    for name, arg in (("mp_vec_exp", x), ("mp_vec_log", np.abs(x) + 0.5)):
        d = dev(arg)
        getattr(native, name)(d.ptr, n)
        h = arg.copy()
        getattr(O.lib(), name.replace("mp_", "or_"))(h.ctypes.data, n)
        # device libm against glibc: a few ulp, as the reference allows
        # between its libm and upstream's (1e-9 there)
        np.testing.assert_allclose(back(d, np.float64, n), h, rtol=1e-14, atol=0)
