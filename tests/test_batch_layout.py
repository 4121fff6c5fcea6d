"""The batched solver's host-built layout (mpk.batch.build_arrays; no GPU
needed): the concatenated CSR and Jacobi diagonal reproduce each system, and
the ILU(0) sweep plans, run level by level here as the batched kernel runs
them (in FP64 on the plans' FP32 coefficients), give the oracle's ILU(0)
apply (ILUApplyFactor_SeqAIJ restated) for every GMRES system to within the
FP32 rounding of the coefficients."""

import numpy as np
import pytest

from mpk import workloads as W


@pytest.fixture(scope="module")
def L():
    from mpk import _native
    return _native.load(init=False)


def _systems(count=12, seed=7):
    return W.batch_systems(count=count, seed=seed, nmin=30, nmax=700)


def _sweeps(H, s, x, coef_f, coef_b):
    """The kernel's ILU(0) apply for system s (FP64 here)."""
    r0, n, kind, lf0, nlf, lb0, nlb, _ = (int(v) for v in H["sys"][s])
    val = np.zeros(n)
    F = np.zeros(n)
    for lev in range(nlf):
        for p in range(H["flv"][lf0 + lev], H["flv"][lf0 + lev + 1]):
            gp = r0 + p
            i = H["fprow"][gp]
            acc = x[i]
            for e in range(H["fcnt"][gp]):
                acc = acc + coef_f[gp * 4 + e] * val[H["fdep"][gp * 4 + e]]
            val[p] = acc
            F[i] = acc
    y = np.zeros(n)
    for lev in range(nlb):
        for p in range(H["blv"][lb0 + lev], H["blv"][lb0 + lev + 1]):
            gp = r0 + p
            i = H["bprow"][gp]
            acc = F[i]
            for e in range(H["bcnt"][gp]):
                acc = acc + coef_b[gp * 4 + e] * val[H["bdep"][gp * 4 + e]]
            v = acc * H["dinv64"][r0 + i]
            val[p] = v
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
        H["dinv64"][r0:r0 + n] = ba[bd[:n]]
        assert np.array_equal(H["dinv"][r0:r0 + n], ba[bd[:n]].astype(np.float32))
        x = np.random.default_rng(s).standard_normal(n)
        want = pc.apply(x)
        got32 = _sweeps(H, s, x, H["fcoef"].astype(np.float64), H["bcoef"].astype(np.float64))
        assert np.allclose(got32, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
        checked += 1
    assert checked >= 4


def test_bad_kind_is_rejected(L):
    from mpk.batch import build_arrays
    ip, ix, va = W.laplacian_2d(4)
    with pytest.raises(ValueError):
        build_arrays([(ip, ix, va, np.ones(16), "bicg")], L)
