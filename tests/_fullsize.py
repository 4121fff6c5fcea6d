"""The systems of the full-size parity tests, at the sizes the benchmark
workloads name, and oracle solves of them run side by side in threads (the
oracle's reduction order is per thread, and ctypes releases the GIL for the
length of each C call).

  diffusion 128^3   3-D 7-point lognormal(0, 1) diffusion, n = 2,097,152:
                    cg + jacobi and cg + icc(0), rtol 1e-10
  convdiff 128^3    3-D upwind convection-diffusion (Peclet 10, 5, 2.5):
                    gmres(30) + ilu(0), rtol 1e-10
  random SPD 1e6    n = 1,000,000, 12 upper offsets within 5000 mirrored:
                    bcgs + sor(omega 1.2, symmetric), rtol 1e-10
"""

from __future__ import annotations

import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_lock = threading.Lock()
_systems: dict = {}
_pool = None
_jobs: dict = {}

# name -> (system, ksp type, pc type, pc options, solve options)
CASES = {
    "cg_jacobi": ("diffusion_128", "cg", "jacobi", {}, dict(rtol=1e-10, max_it=10000)),
    "cg_icc0": ("diffusion_128", "cg", "icc", {}, dict(rtol=1e-10, max_it=10000)),
    "gmres_ilu0": ("convdiff_128", "gmres", "ilu", {}, dict(rtol=1e-10, max_it=2000, restart=30)),
    "bcgs_sor": ("random_spd_1e6", "bcgs", "sor", dict(omega=1.2), dict(rtol=1e-10, max_it=2000)),
}


def system(name):
    from mpk import workloads as W
    with _lock:
        if name not in _systems:
            if name == "diffusion_128":
                _systems[name] = W.diffusion_3d(128)
            elif name == "convdiff_128":
                _systems[name] = W.convection_diffusion_3d(128)
            elif name == "random_spd_1e6":
                _systems[name] = W.random_spd(1_000_000)
            else:
                raise KeyError(name)
        return _systems[name]


def _oracle_run(case, order):
    import oracle as O
    O.set_mode(order)  # this thread's order
    sysname, kt, pt, pcopt, opts = CASES[case]
    csr, b = system(sysname)
    A = O.CSR(*csr)
    pc = O.PCSetup(A, pt, 0, **pcopt)
    r = O.ksp_solve(A, b, kt, pc, **opts)
    del pc, A
    return r


def oracle(case, order):
    """A future of the oracle's solve of `case` in `order` ('ref' or 'dev')."""
    global _pool
    import oracle as O
    O.build()
    with _lock:
        if _pool is None:
            _pool = ThreadPoolExecutor(max_workers=8)
        key = (case, order)
        if key not in _jobs:
            _jobs[key] = _pool.submit(_oracle_run, case, order)
        return _jobs[key]


def start(cases, orders=("ref", "dev")):
    for c in cases:
        for o in orders:
            oracle(c, o)


def rel_gap(x, y):
    """||x - y||_2 / ||y||_2 and max |x - y| / max |y|."""
    d = np.asarray(x) - np.asarray(y)
    return (float(np.linalg.norm(d) / np.linalg.norm(y)), float(np.max(np.abs(d)) / np.max(np.abs(y))))
