"""Planned sweeps (block plans and ring plans): the plans mpk_pc_schedule
builds for SOR / ILU / IC equal the host planner's output on entries restated
from the factor, every read in them resolves when its level runs (so a sweep
cannot stall), and applies through them give the oracle's result bit for
bit."""

import numpy as np
import pytest

from _helpers import back, rng_vec, same_bits
from mpk import workloads as W

pytestmark = pytest.mark.gpu


def _sweep_entries(kind, pc, A):
    """Per sweep, every entry of every row in the row's own order, as
    (off, col, coef) -- restated from the host factor / matrix."""
    n = A.getSize()[0]

    def pack(rows):
        off = np.zeros(n + 1, np.int64)
        off[1:] = np.cumsum([len(r) for r in rows])
        col = np.array([k for r in rows for k, _ in r] or [0], np.int32)
        coef = np.array([c for r in rows for _, c in r] or [0.0], np.float64)
        return off, col, coef

    if kind == "sor":
        ip, ix, va = A.irow, A.icol, A.data
        dm = np.zeros(n, np.int64)
        for i in range(n):
            hit = np.nonzero(ix[ip[i]:ip[i + 1]] == i)[0]
            dm[i] = ip[i] + hit[0] if len(hit) else ip[i + 1]
        fwd = [[(ix[j], -va[j]) for j in range(ip[i], dm[i])] for i in range(n)]
        bwd = [[(ix[j], -va[j]) for j in range(dm[i] + 1, ip[i + 1])] for i in range(n)]
        return pack(fwd), pack(bwd)
    f = pc._host_factor()
    bi, bj, ba, bd = f["bi"], f["bj"], f["ba"], f["bdiag"]
    if kind == "ilu":
        fwd = [[(bj[j], -ba[j]) for j in range(bi[i], bi[i + 1])] for i in range(n)]
        bwd = [[(bj[j], -ba[j]) for j in range(bd[i + 1] + 1, bd[i])] for i in range(n)]
        return pack(fwd), pack(bwd)
    # IC: backward walks row i's strict part; forward row k gathers, in
    # ascending i, the rows i < k whose strict part holds k
    # the last row's strict part is not applied (the loop starts at n - 2)
    bwd = [[(bj[j], ba[j]) for j in range(bi[i], bd[i])] if i < n - 1 else [] for i in range(n)]
    fwd = [[] for _ in range(n)]
    for i in range(n):
        for j in range(bi[i], bd[i]):
            if bj[j] > i:
                fwd[bj[j]].append((i, ba[j]))
    return pack(fwd), pack(bwd)


MATS = {
    "lap32": lambda: W.laplacian_2d(32),
    "diff8": lambda: W.diffusion_3d(8)[0],
    "cd8": lambda: W.convection_diffusion_3d(8)[0],
    "tri50": lambda: W.tridiagonal(50),
    "diff24": lambda: W.diffusion_3d(24)[0],
    "lap100": lambda: W.laplacian_2d(100),
    "spd20k": lambda: W.random_spd(20000, band=3000)[0],
}


def _setup(name, kind, levels=0, ordering="natural"):
    from mojopetsc4py import PETSc
    ip, ix, va = MATS[name]()
    n = len(ip) - 1
    A = PETSc.Mat().createSeqAIJ((n, n), nnz=7)
    A.setValuesCSR(ip, ix, va)
    A.assemble()
    pc = PETSc.PC().create()
    pc.setType(kind)
    if kind == "sor":
        pc.setSOROmega(1.2)
    else:
        pc.setFactorLevels(levels)
        pc.setFactorOrdering(ordering)
    pc.setOperators(A)
    pc.setUp()
    return (ip, ix, va), A, pc


CASES = [("lap32", "ilu", 0, "natural"), ("diff8", "ilu", 1, "natural"), ("cd8", "ilu", 2, "rcm"),
         ("lap32", "icc", 0, "natural"), ("diff8", "icc", 2, "natural"), ("diff8", "icc", 0, "rcm"),
         ("tri50", "icc", 1, "natural"), ("lap32", "sor", 0, ""), ("cd8", "sor", 0, ""),
         ("tri50", "ilu", 0, "natural"), ("diff24", "ilu", 0, "natural"), ("diff24", "icc", 0, "natural"),
         ("diff24", "icc", 1, "rcm"), ("diff24", "sor", 0, ""), ("lap100", "ilu", 0, "natural"),
         ("lap100", "icc", 1, "natural"), ("spd20k", "sor", 0, ""), ("spd20k", "ilu", 0, "natural")]
# the sweeps of over 8192 rows again, forced into the other kind of plan
MODES = [("auto", c) for c in CASES] + [
    (m, c) for m in ("blocks", "ring") for c in CASES if c[0] in ("diff24", "lap100", "spd20k")]
IDS = [f"{m}-{c[0]}-{c[1]}{c[2]}-{c[3]}" for m, c in MODES]


@pytest.fixture
def plan_mode(monkeypatch, request):
    mode = request.param
    if mode != "auto":
        monkeypatch.setenv("MPK_SWEEP_PLAN", mode)
    else:
        monkeypatch.delenv("MPK_SWEEP_PLAN", raising=False)
    return mode


@pytest.mark.parametrize("plan_mode,case", MODES, ids=IDS, indirect=["plan_mode"])
def test_plan_matches_host_planner(native, plan_mode, case):
    """The plans the library builds from the device factor equal the host
    planner's output on entries restated here from the host factor, byte for
    byte, and every read in them is resolvable when its level runs."""
    from test_sweep_plan import _check, _unpack
    name, kind, levels, ordering = case
    _, A, pc = _setup(name, kind, levels, ordering)
    n = A.getSize()[0]
    raw = back(pc._dbufs["sched"], np.uint8, pc._dbufs["sched"].nbytes)
    at = 0
    depth = 0
    for (off, col, coef), forward in zip(_sweep_entries(kind, pc, A), (True, False)):
        cap = native.mpk_sweep_plan_bound(n, int(off[n]), 8192)
        buf = np.zeros((cap + 7) // 8, np.int64)
        nb = native.mpk_sweep_plan(n, 1 if forward else 0, off.ctypes.data, col.ctypes.data, coef.ctypes.data,
                                   8192, buf.ctypes.data, cap)
        assert nb > 0
        assert np.array_equal(raw[at:at + nb], buf.view(np.uint8)[:nb])
        plan = _unpack(raw[at:at + nb], n)
        assert _check(plan, n, forward, off, col, coef)
        if plan_mode != "auto":
            assert (plan["ring"] > 0) == (plan_mode == "ring")
        depth = max(depth, plan["depth"])
        at += nb
    assert pc._sweep_depth == depth


def _marks(pc, n):
    """Each sweep's hdr[7] in the device plans, with the sweep's plan bytes."""
    raw = back(pc._dbufs["sched"], np.uint8, pc._dbufs["sched"].nbytes)
    out, at = [], 0
    while at < len(raw):
        h = raw[at:at + 64].view(np.int64)
        nb = int(h[4])
        if nb <= 0:
            break
        out.append((int(h[7]), raw[at:at + nb]))
        at += nb
        if len(out) == 2:
            break
    return out


@pytest.mark.parametrize("plan_mode,case", MODES, ids=IDS, indirect=["plan_mode"])
@pytest.mark.parametrize("narrow", ["default", "nowave", "0"])
def test_scheduled_apply_bitwise(native, oracle, monkeypatch, plan_mode, case, narrow):
    """Applies through the plans give the oracle's bits: through the one-wave
    kernel where the planner marks a plan 2, the staged one-workgroup kernel
    where it marks it 1 (every plan it can take with MPK_SWEEP_WAVE=0), the
    block engine elsewhere (everywhere with MPK_SWEEP_NARROW=0)."""
    from mojopetsc4py import PETSc
    from test_sweep_plan import _narrow_rule
    name, kind, levels, ordering = case
    if narrow == "0":
        monkeypatch.setenv("MPK_SWEEP_NARROW", "0")
    if narrow == "nowave":
        monkeypatch.setenv("MPK_SWEEP_WAVE", "0")
    oracle.set_mode("dev")
    try:
        csr, A, pc = _setup(name, kind, levels, ordering)
        n = A.getSize()[0]
        marks = _marks(pc, n)
        assert marks
        for mark, raw in marks:
            want = 0 if narrow == "0" else _narrow_rule(raw, n, wave=narrow == "default")
            assert mark == want
        if plan_mode == "auto" and (name, kind, levels) in (("lap32", "ilu", 0), ("tri50", "ilu", 0)):
            assert [m for m, _ in marks] == [{"default": 2, "nowave": 1, "0": 0}[narrow]] * 2
        Ao = oracle.CSR(*csr)
        if kind == "sor":
            po = oracle.PCSetup(Ao, "sor", omega=1.2)
        else:
            po = oracle.PCSetup(Ao, kind, levels, ordering or "natural")
        for seed in (3, 4):
            b = rng_vec(n, seed)
            y = PETSc.Vec().createSeq(n)
            pc.apply(PETSc.Vec().createWithArray(b), y)
            assert same_bits(y.getArray().copy(), po.apply(b))
    finally:
        oracle.set_mode("ref")
