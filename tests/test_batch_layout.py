"""The batched solver's host-built layout (mpk.batch.build_arrays; no GPU
needed): the concatenated CSR and Jacobi diagonal reproduce each system, and
the ILU(0) sweep plans, run level by level here as the batched kernel runs
them (in FP64 on the plans' FP32 coefficients; through the LDS ring of
position mod 2048 for the systems flagged for it), give the oracle's ILU(0)
apply (ILUApplyFactor_SeqAIJ restated) for every GMRES system to within the
FP32 rounding of the coefficients."""

import numpy as np
import pytest

from mpk import workloads as W


@pytest.fixture(scope="module")
def L():
    from mpk import _native
    return _native.load(init=False)


def _systems(count=12, seed=7, nmin=30, nmax=700):
    return W.batch_systems(count=count, seed=seed, nmin=nmin, nmax=nmax)


def _sweeps(H, s, x, coef_f, coef_b):
    """The kernel's ILU(0) apply for system s (FP64 here)."""
    from mpk.batch import _RING
    r0, n, kind, lf0, nlf, lb0, nlb, ring = (int(v) for v in H["sys"][s])
    slot = (lambda p: p % _RING) if ring else (lambda p: p)
    val = np.full(_RING if ring else n, np.nan)
    F = np.zeros(n)
    for lev in range(nlf):
        for p in range(H["flv"][lf0 + lev], H["flv"][lf0 + lev + 1]):
            gp = r0 + p
            i = H["fprow"][gp]
            acc = x[i]
            for e in range(H["fcnt"][gp]):
                acc = acc + coef_f[gp * 4 + e] * val[H["fdep"][gp * 4 + e]]
            val[slot(p)] = acc
            F[i] = acc
    y = np.zeros(n)
    # F by backward position, as the kernel's forward sweep writes it (f2b)
    Fb = np.zeros(n)
    Fb[H["f2b"][r0:r0 + n]] = F[H["fprow"][r0:r0 + n]]
    for lev in range(nlb):
        for p in range(H["blv"][lb0 + lev], H["blv"][lb0 + lev + 1]):
            gp = r0 + p
            i = H["bprow"][gp]
            acc = Fb[p]
            for e in range(H["bcnt"][gp]):
                acc = acc + coef_b[gp * 4 + e] * val[H["bdep"][gp * 4 + e]]
            v = acc * H["dinv64"][gp]
            val[slot(p)] = v
            y[i] = v
    return y


def test_concatenated_csr_and_jacobi(L):
    from mpk.batch import build_arrays
    systems = _systems()
    H = build_arrays(systems, L)
    for s, (ip, ix, va, b, kind) in enumerate(systems):
        r0, n = int(H["row0"][s]), len(ip) - 1
        assert H["sys"][s][0] == r0 and H["sys"][s][1] == n
        q0 = H["rp"][r0]
        assert np.array_equal(H["rp"][r0:r0 + n + 1] - q0, ip)
        nnz = ip[-1]
        assert np.array_equal(H["ci"][q0:q0 + nnz] - r0, ix)
        assert np.array_equal(H["v64"][q0:q0 + nnz], va)
        assert np.array_equal(H["b"][r0:r0 + n], b)
        d = np.array([va[ip[i] + np.nonzero(ix[ip[i]:ip[i + 1]] == i)[0][0]] for i in range(n)])
        assert np.allclose(H["jd"][r0:r0 + n], 1.0 / d, rtol=1e-7)
    assert list(H["order"]) == list(np.argsort(-H["sizes"], kind="stable"))


def test_ilu_plans_reproduce_the_oracle_apply(L, oracle):
    from mpk.batch import build_arrays, _ilu0
    systems = _systems()
    H = build_arrays(systems, L)
    oracle.set_mode("ref")
    checked = 0
    for s, (ip, ix, va, b, kind) in enumerate(systems):
        if kind != "gmres+ilu":
            continue
        n = len(ip) - 1
        A = oracle.CSR(ip, ix, va)
        pc = oracle.PCSetup(A, "ilu", 0)
        # the inverted diagonal in FP64 (the kernel holds it in FP32)
        bi, bj, ba, bd = _ilu0(L, ip, ix, va, n)
        r0 = int(H["row0"][s])
        H["dinv64"] = np.zeros(H["N"])
        H["dinv64"][r0:r0 + n] = ba[bd[H["bprow"][r0:r0 + n]]]  # backward position order
        assert np.array_equal(H["dinv"][r0:r0 + n], H["dinv64"][r0:r0 + n].astype(np.float32))
        x = np.random.default_rng(s).standard_normal(n)
        want = pc.apply(x)
        got32 = _sweeps(H, s, x, H["fcoef"].astype(np.float64), H["bcoef"].astype(np.float64))
        assert np.allclose(got32, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
        checked += 1
    assert checked >= 4


def test_ring_plans_for_grid_systems(L, oracle):
    """Systems of 2-D grids up to 20000 rows read at most two levels back, so
    all take the LDS ring; their ring-slot sweeps still give the oracle's
    ILU(0) apply."""
    from mpk.batch import build_arrays, _ilu0
    systems = _systems(count=6, seed=3, nmin=9000, nmax=20000)
    H = build_arrays(systems, L)
    oracle.set_mode("ref")
    for s, (ip, ix, va, b, kind) in enumerate(systems):
        if kind != "gmres+ilu":
            continue
        assert H["sys"][s][7] == 1
        n = len(ip) - 1
        r0 = int(H["row0"][s])
        bi, bj, ba, bd = _ilu0(L, ip, ix, va, n)
        H["dinv64"] = np.zeros(H["N"])
        H["dinv64"][r0:r0 + n] = ba[bd[H["bprow"][r0:r0 + n]]]
        x = np.random.default_rng(s).standard_normal(n)
        want = oracle.PCSetup(oracle.CSR(ip, ix, va), "ilu", 0).apply(x)
        got = _sweeps(H, s, x, H["fcoef"].astype(np.float64), H["bcoef"].astype(np.float64))
        assert np.allclose(got, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())


def test_ring_check_refuses_a_far_read():
    """Plant the mistake on the CPU: a plan entry that reads a position more
    than the ring behind the end of its level is refused."""
    from mpk.batch import _ring_slots, _RING
    n = _RING + 10
    lvl = np.array([0, _RING + 5, n])
    pcnt = np.zeros(n, np.int32)
    edep = np.zeros(n * 4, np.int32)
    assert _ring_slots(lvl, pcnt, edep, n) is not None
    pcnt[n - 1] = 1  # the last row reads position 0, n - 0 > _RING behind its level's end
    assert _ring_slots(lvl, pcnt, edep, n) is None


def test_bad_kind_is_rejected(L):
    from mpk.batch import build_arrays
    ip, ix, va = W.laplacian_2d(4)
    with pytest.raises(ValueError):
        build_arrays([(ip, ix, va, np.ones(16), "bicg")], L)
