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
    cfg = bs.autotune(candidates=((128, 32), (256, 32)), tiles=False)
    assert (cfg["workgroup"], cfg["grid"]) in ((128, 32), (256, 32))
    assert len(cfg["timings"]) == 2
    x, st = bs.solve(rtol=1e-8)
    assert np.all(st[:, 3] == 1.0)


def test_tile_variants_give_the_same_bits(batch):
    """The tile variants change only how loads are grouped (rows per pass,
    entries per round, elements per Gram-Schmidt step), never the order of
    any sum: at one workgroup size and grid every variant gives the default
    tiles' bits, and each system meets the tolerance."""
    from mpk.batch import BatchSolver
    systems, bs = batch
    var = BatchSolver.variants()
    assert len(var) >= 20
    assert {(d["rows_per_thread"], d["spmv_block"], d["gs_tiling"]) for d in var if d["workgroup"] == 512} == {
        (ur, kb, ug) for ur in (2, 4, 8) for kb in (4, 8) for ug in (2, 4, 8)}
    ref, st0 = bs.solve(rtol=1e-8, workgroup=512, grid=32, graph=False, variant=-1)
    assert np.all(st0[:, 3] == 1.0)
    for v, d in enumerate(var):
        if d["workgroup"] != 512:
            continue
        print(f"[batch] variant {v}: {d}", flush=True)
        x, st = bs.solve(rtol=1e-8, grid=32, graph=False, variant=v)
        assert np.array_equal(x.view(np.int64), ref.view(np.int64)), d
        assert np.array_equal(st, st0), d
    for v, d in enumerate(var):
        if d["workgroup"] != 512:
            x, st = bs.solve(rtol=1e-8, grid=16, graph=True, variant=v)
            assert np.all(st[:, 3] == 1.0) and np.max(st[:, 0]) <= 1e-8, d


def test_autotune_chooses_tiles(batch):
    """The autotuner times every tile variant of the fastest workgroup size
    (on a sample of the systems) and reports the three choices."""
    systems, bs = batch
    cfg = bs.autotune(candidates=((512, 32),), sample_every=2)
    assert cfg["workgroup"] == 512 and cfg["grid"] == 32
    assert len(cfg["timings"]) == 18 and cfg["sampled_systems"] == 24
    assert cfg["rows_per_wavefront"] in (128, 256, 512)
    assert cfg["reduction_tiling"] in (2, 4, 8)
    assert cfg["spmv_variant"] in ("4 entries per round", "8 entries per round")
    x, st = bs.solve(rtol=1e-8)  # with the chosen variant
    assert np.all(st[:, 3] == 1.0)


def test_bad_config_is_rejected(batch):
    from mpk._native import NativeError
    systems, bs = batch
    with pytest.raises((RuntimeError, NativeError)):
        bs.solve(workgroup=100, grid=4)
    with pytest.raises((RuntimeError, NativeError, IndexError)):
        bs.solve(grid=4, variant=1000)
