"""The two reduction orders of the oracle on the full-size systems: the
reference's SIMD order (REF, src/simdvec.mojo:332-473) and the tile-tree
order the GPU computes by default (DEV).  Both orders must give the same
iteration count and converged reason.  The iterates differ by reduction-order
rounding amplified over the iterations; the north star's 1e-12 relative
holds for every solve except cg + icc(0) on the 128^3 diffusion system, whose
gap is recorded here (the GPU meets 1e-12 there in the reference order, which
the GPU tests check bit for bit).  All eight oracle solves run side by side
(about a minute on eight cores)."""

import numpy as np
import pytest

import _fullsize as FS

GAP_BOUND = {  # relative 2-norm gap allowed between the orders
    "cg_jacobi": 1e-12,
    "cg_icc0": 1e-8,   # measured 2.8e-10: the orders drift apart after ~100 iterations
    "gmres_ilu0": 1e-12,
    "bcgs_sor": 1e-12,
}


@pytest.fixture(scope="module", autouse=True)
def _start_all():
    FS.start(FS.CASES)


@pytest.mark.parametrize("case", list(FS.CASES))
def test_orders_agree(case):
    ref = FS.oracle(case, "ref").result()
    dev = FS.oracle(case, "dev").result()
    assert ref.its == dev.its, (ref.its, dev.its)
    assert ref.reason == dev.reason == 2, (ref.reason, dev.reason)
    gap, gmax = FS.rel_gap(dev.x, ref.x)
    print(f"\n{case}: its {ref.its}, reason {ref.reason}, iterate gap {gap:.3e} (max-norm {gmax:.3e}), "
          f"rnorm ref {ref.rnorm:.6e} dev {dev.rnorm:.6e}")
    assert gap <= GAP_BOUND[case], gap
    # the residual histories agree closely early on and drift apart later
    h = np.abs(dev.history - ref.history) / np.abs(ref.history)
    for k in (10, 50, 100, 150, 200):
        if k < len(h):
            print(f"  history rel. difference at {k}: {h[k]:.2e}")
    assert h[1] < 1e-12
