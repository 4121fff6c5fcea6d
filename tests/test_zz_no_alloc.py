"""mpk_vec_dot_noalloc, the entry point that allocates nothing: its result
against mp_vec_dot's (the entry point it varies) bit for bit, on two calls
in a row with the second smaller, in both reduction orders; sizes past its
fixed scratch are refused with an error and queue nothing."""

import numpy as np
import pytest

from _helpers import dev, rng_vec, same_bits

pytestmark = pytest.mark.gpu

MAX_N = 65536 * 2048


def _both(native, x, y):
    n = len(x)
    dx, dy = dev(x if n else np.zeros(1)), dev(y if n else np.zeros(1))
    out = dev(np.full(1, np.nan))
    assert native.mpk_vec_dot_noalloc(dx.ptr, dy.ptr, out.ptr, n) == 0
    native.mpk_synchronize()
    got = out.to_numpy(np.float64, 1)
    want = np.zeros(1)
    native.mp_vec_dot(dx.ptr, dy.ptr, want.ctypes.data, n)
    assert native.mpk_status() == 0
    return got, want


@pytest.mark.parametrize("order", ["dev", "ref"])
def test_noalloc_dot_matches_mp_vec_dot(native, order):
    import mpk
    old = mpk.set_reduction_order(order)
    try:
        for n in (3_000_017, 40_961, 2048, 17, 0):  # each call smaller than the one before
            x, y = rng_vec(n, 31), rng_vec(n, 32)
            got, want = _both(native, x, y)
            assert same_bits(got, want), (order, n, got, want)
    finally:
        mpk.set_reduction_order(old)


def test_noalloc_dot_rejects_past_its_scratch(native):
    out = dev(np.full(1, 7.0))
    x = dev(np.zeros(4))
    assert native.mpk_vec_dot_noalloc(x.ptr, x.ptr, out.ptr, MAX_N + 1) == -1
    assert native.mpk_status() == 4
    native.mpk_clear_error()
    native.mpk_synchronize()
    assert out.to_numpy(np.float64, 1)[0] == 7.0  # nothing was queued
