"""Batched solves (one persistent launch over many independent systems): every
system reaches the FP64 relative residual asked for (checked with an FP64
residual computed on the host), agrees with the oracle's FP64 solve of the
same system, and a HIP-graph replay gives the eager launch's result on
every replay."""

import numpy as np
import pytest

from mpk import workloads as W

pytestmark = pytest.mark.gpu


def _relres(ip, ix, va, b, x):
    n = len(ip) - 1
    rows = np.repeat(np.arange(n), np.diff(ip))
    ax = np.bincount(rows, weights=np.asarray(va) * x[np.asarray(ix)], minlength=n)
    return np.linalg.norm(b - ax) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def batch(native):
    from mpk.batch import BatchSolver
    systems = W.batch_systems(count=48, seed=9, nmin=40, nmax=4000)
    bs = BatchSolver(systems)
    yield systems, bs
    bs.destroy()


@pytest.mark.parametrize("wg,grid", [(128, 16), (256, 8), (512, 64)])
def test_every_system_converges(batch, wg, grid):
    systems, bs = batch
    x, st = bs.solve(rtol=1e-8, inner_rtol=1e-3, workgroup=wg, grid=grid, graph=False)
    for s, (ip, ix, va, b, kind) in enumerate(systems):
        xs = bs.solution(x, s)
        assert st[s, 3] == 1.0, (s, kind, st[s])
        assert st[s, 0] <= 1e-8
        assert _relres(ip, ix, va, b, xs) <= 1.5e-8, (s, kind)


def test_agrees_with_oracle(batch, oracle):
    systems, bs = batch
    x, st = bs.solve(rtol=1e-8, inner_rtol=1e-3, workgroup=256, grid=32, graph=False)
    oracle.set_mode("ref")
    for s in range(0, len(systems), 7):
        ip, ix, va, b, kind = systems[s]
        A = oracle.CSR(ip, ix, va)
        if kind == "cg+jacobi":
            r = oracle.ksp_solve(A, b, "cg", oracle.PCSetup(A, "jacobi"), rtol=1e-13, atol=0.0, max_it=20000)
        else:
            r = oracle.ksp_solve(A, b, "gmres", oracle.PCSetup(A, "ilu", 0), rtol=1e-13, atol=0.0,
                                 max_it=20000, restart=20)
        err = np.linalg.norm(bs.solution(x, s) - r.x) / np.linalg.norm(r.x)
        assert err <= 1e-5, (s, kind, err)


def test_graph_replays_match_eager(batch):
    systems, bs = batch
    x0, st0 = bs.solve(rtol=1e-8, workgroup=256, grid=32, graph=False)
    for rep in range(3):
        x1, st1 = bs.solve(rtol=1e-8, workgroup=256, grid=32, graph=True)
        assert np.array_equal(x1.view(np.int64), x0.view(np.int64)), rep
        assert np.array_equal(st1, st0)
        # eager work between replays
        x2, _ = bs.solve(rtol=1e-8, workgroup=256, grid=32, graph=False)
        assert np.array_equal(x2.view(np.int64), x0.view(np.int64))


def test_autotune_picks_a_timed_candidate(batch):
    systems, bs = batch
    cfg = bs.autotune(candidates=((128, 32), (256, 32)))
    assert (cfg["workgroup"], cfg["grid"]) in ((128, 32), (256, 32))
    assert len(cfg["timings"]) == 2
    x, st = bs.solve(rtol=1e-8)
    assert np.all(st[:, 3] == 1.0)


def test_bad_config_is_rejected(batch):
    from mpk._native import NativeError
    systems, bs = batch
    with pytest.raises((RuntimeError, NativeError)):
        bs.solve(workgroup=100, grid=4)
