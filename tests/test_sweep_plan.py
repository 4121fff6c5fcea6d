"""The host half of the planned sweeps (mpk_sweep_plan; no GPU needed): the
plan matches a restatement computed here (block plans and ring plans, and the
choice between them), keeps every row's entries in the row's own order, and
every value a row reads is either computed earlier in its own block (a lower
level; in a ring plan, one still held in the LDS window) or belongs to an
earlier block -- the property the device sweep's progress rests on.  Planted
mistakes in a plan (a row reading a later level of its block; a ring plan
reading a value the window no longer holds) are caught by the same check."""

import numpy as np
import pytest

from mpk import workloads as W


@pytest.fixture(scope="module")
def L():
    from mpk import _native
    return _native.load(init=False)


def _entries(ip, ix, va):
    return np.asarray(ip, np.int64), np.asarray(ix, np.int32), np.asarray(va, np.float64)


def _plan(L, n, forward, off, col, coef, bmax):
    m = int(off[n])
    cap = L.mpk_sweep_plan_bound(n, m, bmax)
    buf = np.zeros((cap + 7) // 8, np.int64)
    nb = L.mpk_sweep_plan(n, 1 if forward else 0, off.ctypes.data, col.ctypes.data, coef.ctypes.data, bmax,
                          buf.ctypes.data, cap)
    assert 0 < nb <= cap
    return _unpack(buf.view(np.uint8)[:nb], n)


def _unpack(raw, n):
    """The plan's arrays, with the entries in CSR form whichever way they are
    stored (fixed slots or CSR)."""
    h = raw[:64].view(np.int64)
    nblk, nent, nlvl, depth, nbytes, w, ring = (int(x) for x in h[:7])
    ints = raw[64:].view(np.int32)
    o = 0

    def take(k):
        nonlocal o
        a = ints[o:o + k].astype(np.int64)
        o += k
        return a
    blk_pos, blk_lev, lvl = take(nblk + 1), take(nblk + 1), take(nlvl)
    prow, pcnt, pent = take(n), take(n), take(n + 1)
    slots = n * w if w else nent
    base = 64 + ((4 * (2 * (nblk + 1) + nlvl + 3 * n + 1) + 7) // 8) * 8
    ecoef = raw[base:base + 8 * slots].view(np.float64)
    edep = raw[base + 8 * slots:base + 12 * slots].view(np.int32).astype(np.int64)
    assert nbytes == len(raw)
    if w:
        assert pcnt.max(initial=0) <= w
        # slot e of position p at e n + p
        take_idx = np.concatenate([np.arange(c) * n + p for p, c in enumerate(pcnt)] or [np.zeros(0, np.int64)])
        ecoef, edep = ecoef[take_idx.astype(np.int64)], edep[take_idx.astype(np.int64)]
        pent = np.concatenate([[0], np.cumsum(pcnt)])
    else:
        assert np.array_equal(np.diff(pent), pcnt)
    return dict(nblk=nblk, depth=depth, w=w, ring=ring, blk_pos=blk_pos, blk_lev=blk_lev, lvl=lvl, prow=prow,
                pent=pent, ecoef=ecoef, edep=edep)


def _slots(off, n):
    """Fixed slots per position: a power of two up to 32, else a multiple of
    8 up to 64 padding the entries at most 4 times; 0 (CSR form) beyond."""
    widest = max((off[i + 1] - off[i] for i in range(n)), default=0)
    w = 4
    while w < widest and w < 32:
        w *= 2
    if w >= widest:
        return w
    w = -(-widest // 8) * 8
    return w if w <= 64 and n * w <= 4 * off[n] else 0


def _restated_ring(n, forward, off, col, coef, bmax, mode):
    """The ring plan when the planner takes it (None otherwise): rows at
    their latest level, values W or more positions behind the end of the
    reading row's level read as published rows, others from slot position
    mod W."""
    if n <= bmax or mode == "blocks":
        return None
    W = 1
    while 2 * W <= bmax:
        W *= 2
    spos = (lambda i: i) if forward else (lambda i: n - 1 - i)
    earlier = (lambda k, i: k < i) if forward else (lambda k, i: k > i)
    deps = [[col[e] for e in range(off[i], off[i + 1]) if earlier(col[e], i)] for i in range(n)]
    asap = np.zeros(n, np.int64)
    for s in range(n):
        i = spos(s)
        asap[i] = max((asap[k] + 1 for k in deps[i]), default=0)
    nl = int(asap.max()) + 1
    # latest level: one before the earliest reader, the last level for none
    readers = [[] for _ in range(n)]
    for i in range(n):
        for k in deps[i]:
            readers[k].append(i)
    lev = np.full(n, nl - 1, np.int64)
    for s in reversed(range(n)):
        i = spos(s)
        lev[i] = min((lev[j] - 1 for j in readers[i]), default=nl - 1)
    assert np.all(lev >= asap)
    prow = sorted((spos(s) for s in range(n)), key=lambda i: lev[i])  # stable: sweep order in a level
    pos = np.empty(n, np.int64)
    pos[prow] = np.arange(n)
    starts = [int(np.sum(lev < l)) for l in range(nl + 1)]
    old = sum(1 for i in range(n) for k in deps[i] if starts[lev[i] + 1] - pos[k] > W)
    if mode != "ring" and (n > 1024 * nl or 64 * old > off[n]):
        return None
    pent, ecoef, edep = [0], [], []
    for i in prow:
        for e in range(off[i], off[i + 1]):
            k = col[e]
            ecoef.append(coef[e])
            if not earlier(k, i):
                edep.append(-k - n - 1)
            elif starts[lev[i] + 1] - pos[k] <= W:
                edep.append(pos[k] % W)
            else:
                edep.append(-k - 1)
        pent.append(len(edep))
    return dict(nblk=1, w=_slots(off, n), ring=W, blk_pos=np.array([0, n]), blk_lev=np.array([0, nl + 1]),
                lvl=np.array(starts), prow=np.array(prow), pent=np.array(pent), ecoef=np.array(ecoef),
                edep=np.array(edep))


def _pencils(n, forward, off, col, bmax):
    """The pencil blocks (rows of each block in sweep order) when the sweep is
    a structured grid's: dependency offsets exactly {1, a, b}, a | b, b | n,
    read as an a x (b/a) x (n/b) grid with x fastest; blocks of t x t columns
    of the (x, y) plane spanning every z (t the largest with t^2 (n/b) <=
    bmax, at least 4 columns and 2 blocks), in row-major tile order, reversed
    for a backward sweep, kept when every dependency comes earlier.  None
    otherwise."""
    if n <= bmax:
        return None
    dists = set()
    for i in range(n):
        for e in range(off[i], off[i + 1]):
            d = i - col[e] if forward else col[e] - i
            if d > 0:
                dists.add(int(d))
    if len(dists) != 3:
        return None
    one, a, b = sorted(dists)
    if one != 1 or b % a or n % b:
        return None
    nx, ny, nz = a, b // a, n // b
    t = 1
    while (t + 1) ** 2 * nz <= bmax:
        t += 1
    tx, ty = min(t, nx), min(t, ny)
    mx, my = -(-nx // tx), -(-ny // ty)
    if tx * ty < 4 or mx * my < 2:
        return None
    blocks = []
    for by in range(my):
        for bx in range(mx):
            blocks.append([z * b + y * nx + x for z in range(nz) for y in range(by * ty, min(ny, (by + 1) * ty))
                           for x in range(bx * tx, min(nx, (bx + 1) * tx))])
    if not forward:
        blocks = [list(reversed(r)) for r in reversed(blocks)]
    rank = np.empty(n, np.int64)
    rank[[i for r in blocks for i in r]] = np.arange(n)
    earlier = (lambda k, i: k < i) if forward else (lambda k, i: k > i)
    for i in range(n):
        for e in range(off[i], off[i + 1]):
            if earlier(col[e], i) and rank[col[e]] >= rank[i]:
                return None
    return blocks


def _restated(n, forward, off, col, coef, bmax, mode="auto"):
    ring = _restated_ring(n, forward, off, col, coef, bmax, mode)
    if ring is not None:
        return ring
    spos = (lambda i: i) if forward else (lambda i: n - 1 - i)
    earlier = (lambda k, i: k < i) if forward else (lambda k, i: k > i)
    blocks = _pencils(n, forward, off, col, bmax) if mode != "ring" else None
    if blocks is None:
        B = n if n <= bmax else min(bmax, max(1024, -(-n // 256)))
        blocks = [[spos(s) for s in range(b * B, min(n, (b + 1) * B))] for b in range(-(-n // B))]
    nblk = len(blocks)
    blk = np.empty(n, np.int64)
    for b, rows in enumerate(blocks):
        blk[rows] = b
    lev = np.zeros(n, np.int64)
    for rows in blocks:
        for i in rows:
            for e in range(off[i], off[i + 1]):
                k = col[e]
                if earlier(k, i) and blk[k] == blk[i]:
                    lev[i] = max(lev[i], lev[k] + 1)
    prow, pos, blk_pos, blk_lev, lvl = [], np.zeros(n, np.int64), [], [], []
    for rows in blocks:
        rows = sorted(rows, key=lambda i: lev[i])  # stable: sweep order inside a level
        p0 = len(prow)
        blk_pos.append(p0)
        blk_lev.append(len(lvl))
        nl = int(max(lev[r] for r in rows)) + 1
        starts = [p0 + sum(1 for r in rows if lev[r] < lv) for lv in range(nl)]
        lvl += starts + [p0 + len(rows)]
        for r in rows:
            pos[r] = len(prow)
            prow.append(r)
    blk_pos.append(n)
    blk_lev.append(len(lvl))
    pent, ecoef, edep = [0], [], []
    for p, i in enumerate(prow):
        b = blk[i]
        for e in range(off[i], off[i + 1]):
            k = col[e]
            ecoef.append(coef[e])
            if not earlier(k, i):
                edep.append(-k - n - 1)
            elif blk[k] == b:
                edep.append(pos[k] - blk_pos[b])
            else:
                edep.append(-k - 1)
        pent.append(len(edep))
    return dict(nblk=nblk, w=_slots(off, n), ring=0, blk_pos=np.array(blk_pos), blk_lev=np.array(blk_lev), lvl=np.array(lvl),
                prow=np.array(prow), pent=np.array(pent), ecoef=np.array(ecoef), edep=np.array(edep))


def _check(plan, n, forward, off, col, coef):
    """Every row's entries in its own order, and every read resolvable when
    the row's level runs: local reads from lower levels of the same block
    (in a ring plan, from the slot of a position still inside the window of
    W positions that ends with the row's level), external reads from earlier
    blocks.  The blocks partition the rows in any order."""
    prow, pent, edep, ecoef = plan["prow"], plan["pent"], plan["edep"], plan["ecoef"]
    W = plan["ring"]
    assert sorted(prow.tolist()) == list(range(n))
    pos = np.empty(n, np.int64)
    pos[prow] = np.arange(n)
    blk = np.empty(n, np.int64)
    for b in range(plan["nblk"]):
        blk[prow[plan["blk_pos"][b]:plan["blk_pos"][b + 1]]] = b
    if W:
        assert plan["nblk"] == 1 and W & (W - 1) == 0
    for b in range(plan["nblk"]):
        p0, p1 = plan["blk_pos"][b], plan["blk_pos"][b + 1]
        if W:
            assert sorted(prow[p0:p1].tolist()) == list(range(n))
        starts = plan["lvl"][plan["blk_lev"][b]:plan["blk_lev"][b + 1]]
        assert starts[0] == p0 and starts[-1] == p1 and np.all(np.diff(starts) > 0)
        for li in range(len(starts) - 1):
            for p in range(starts[li], starts[li + 1]):
                i = prow[p]
                assert pent[p + 1] - pent[p] == off[i + 1] - off[i]
                for e, src in zip(range(pent[p], pent[p + 1]), range(off[i], off[i + 1])):
                    k, d = col[src], edep[e]
                    assert ecoef[e] == coef[src]
                    if d >= 0 and W:
                        # the one position of slot d inside [end - W, end)
                        end = starts[li + 1]
                        q = end - W + (d - (end - W)) % W
                        if not (d < W and q < starts[li] and q >= 0 and prow[q] == k):
                            return False
                    elif d >= 0:
                        if not (p0 + d < starts[li] and prow[p0 + d] == k):
                            return False
                    elif d >= -n and W:
                        # a value older than the window: published at an earlier level
                        if not (-d - 1 == k and pos[k] < starts[li]):
                            return False
                    elif d >= -n:
                        if not (-d - 1 == k and blk[k] < b):
                            return False
                    else:
                        assert -d - n - 1 == k
    return True


def _random(n, per_row, seed):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n):
        for c in set(rng.integers(0, n, per_row).tolist()) | {i}:
            rows.append(i)
            cols.append(c)
    return W._csr_from_coo(n, np.array(rows), np.array(cols), rng.normal(size=len(rows)))


def _lower(ip, ix, va, forward, stray=True):
    """The part of each row a sweep reads: strictly earlier columns, plus
    (stray) a later column in a few rows to exercise the stale path."""
    n = len(ip) - 1
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in range(ip[i], ip[i + 1]):
            k = int(ix[j])
            if (k < i) if forward else (k > i):
                rows.append(i)
                cols.append(k)
                vals.append(va[j])
        if stray and i % 97 == 5 and 0 < i < n - 1:
            rows.append(i)
            cols.append(i + 1 if forward else i - 1)
            vals.append(0.5)
    return _entries(*W._csr_from_coo(n, np.array(rows, np.int64), np.array(cols, np.int64), np.array(vals)))


SMALL = {
    "lap20": lambda: W.laplacian_2d(20),
    "diff7": lambda: W.diffusion_3d(7)[0],
    "cd6": lambda: W.convection_diffusion_3d(6)[0],
    "tri40": lambda: W.tridiagonal(40),
    "rand300": lambda: _random(300, 5, 7),
    "rand200w": lambda: _random(200, 14, 9),
    "rand200x": lambda: _random(200, 70, 10),
    "rand300z": lambda: _random(300, 160, 11),
}


@pytest.mark.parametrize("name", sorted(SMALL))
@pytest.mark.parametrize("forward", [True, False])
@pytest.mark.parametrize("bmax", [8192, 64])
def test_matches_restatement(L, name, forward, bmax):
    off, col, coef = _lower(*SMALL[name](), forward)
    n = len(off) - 1
    got = _plan(L, n, forward, off, col, coef, bmax)
    want = _restated(n, forward, off, col, coef, bmax)
    for k in ("blk_pos", "blk_lev", "lvl", "prow", "pent", "edep"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["ecoef"].view(np.int64), np.asarray(want["ecoef"]).view(np.int64))
    assert got["nblk"] == want["nblk"] and got["w"] == want["w"] and got["ring"] == want["ring"]
    assert _check(got, n, forward, off, col, coef)


@pytest.mark.parametrize("name", ["lap20", "diff7", "tri40", "rand300"])
@pytest.mark.parametrize("forward", [True, False])
@pytest.mark.parametrize("mode", ["blocks", "ring"])
def test_forced_mode_matches_restatement(L, monkeypatch, name, forward, mode):
    """MPK_SWEEP_PLAN forces block or ring plans (a sweep that fits one
    block keeps it)."""
    monkeypatch.setenv("MPK_SWEEP_PLAN", mode)
    off, col, coef = _lower(*SMALL[name](), forward)
    n = len(off) - 1
    got = _plan(L, n, forward, off, col, coef, 64)
    want = _restated(n, forward, off, col, coef, 64, mode)
    for k in ("blk_pos", "blk_lev", "lvl", "prow", "pent", "edep"):
        assert np.array_equal(got[k], want[k]), k
    assert got["ring"] == want["ring"]
    # tri40 fits one block (n <= 64): no ring either way
    assert got["ring"] == (64 if mode == "ring" and name != "tri40" else 0)
    assert _check(got, n, forward, off, col, coef)


@pytest.mark.parametrize("forward", [True, False])
def test_narrow_grid_takes_a_ring(L, forward):
    """A 2-D operator of 10,000 rows (199 levels of at most 100 rows) over
    the 8192-row LDS: a ring plan with W = 8192 and one level per
    anti-diagonal."""
    off, col, coef = _lower(*W.laplacian_2d(100), forward)
    n = len(off) - 1
    plan = _plan(L, n, forward, off, col, coef, 8192)
    assert plan["ring"] == 8192 and plan["nblk"] == 1 and len(plan["lvl"]) == 200
    assert _check(plan, n, forward, off, col, coef)


@pytest.mark.parametrize("forward", [True, False])
def test_grid_operator_blocks(L, monkeypatch, forward):
    """A 3-D operator in several blocks: 20^3 rows in 8 blocks of 1024 (the
    last shorter) with the pencils off, every read resolvable."""
    monkeypatch.setenv("MPK_SWEEP_PLAN", "blocks")
    monkeypatch.setenv("MPK_SWEEP_PENCILS", "0")
    off, col, coef = _lower(*W.diffusion_3d(20)[0], forward)
    n = len(off) - 1
    plan = _plan(L, n, forward, off, col, coef, 1024)
    assert plan["nblk"] == 8
    assert _check(plan, n, forward, off, col, coef)


@pytest.mark.parametrize("name,grid", [("diff20", (20, 20, 20)), ("cd 12x9x10", (12, 9, 10))])
@pytest.mark.parametrize("forward", [True, False])
def test_grid_operator_pencils(L, monkeypatch, name, grid, forward):
    """A 3-D grid operator's sweep in pencils (block plans forced: these
    grids are shallow enough for a ring): blocks of t x t columns spanning
    every z (t = 7 for 20^3 with 1024-row blocks: 9 blocks, the last row and
    column of tiles narrower), matching the restatement, every read
    resolvable, and each row's z chain inside its block."""
    monkeypatch.setenv("MPK_SWEEP_PLAN", "blocks")
    nx, ny, nz = grid
    csr = W.diffusion_3d(20)[0] if name == "diff20" else _box(nx, ny, nz)
    off, col, coef = _lower(*csr, forward)
    n = len(off) - 1
    plan = _plan(L, n, forward, off, col, coef, 1024)
    want = _restated(n, forward, off, col, coef, 1024, "blocks")
    for key in ("nblk", "blk_pos", "blk_lev", "lvl", "prow", "pent", "ecoef", "edep"):
        assert np.array_equal(np.asarray(plan[key]), np.asarray(want[key])), key
    assert _check(plan, n, forward, off, col, coef)
    t = 1
    while (t + 1) ** 2 * nz <= 1024:
        t += 1
    assert plan["nblk"] == -(-nx // min(t, nx)) * -(-ny // min(t, ny)) > 1
    blk = np.empty(n, np.int64)
    for b in range(plan["nblk"]):
        blk[plan["prow"][plan["blk_pos"][b]:plan["blk_pos"][b + 1]]] = b
    plane = nx * ny
    assert np.all(blk[plane:] == blk[:-plane])  # z chains stay in their block


def _box(nx, ny, nz):
    """A 7-point operator on an nx x ny x nz box, x fastest, with unequal
    couplings in each direction."""
    n = nx * ny * nz
    rows, cols, vals = [], [], []
    for i in range(n):
        z, r = divmod(i, nx * ny)
        y, x = divmod(r, nx)
        for dx, dy, dz, v in ((0, 0, -1, -0.3), (0, -1, 0, -0.7), (-1, 0, 0, -1.1), (0, 0, 0, 6.5), (1, 0, 0, -0.9),
                              (0, 1, 0, -0.5), (0, 0, 1, -0.2)):
            if 0 <= x + dx < nx and 0 <= y + dy < ny and 0 <= z + dz < nz:
                rows.append(i)
                cols.append(i + dx + dy * nx + dz * nx * ny)
                vals.append(v)
    return W._csr_from_coo(n, np.array(rows), np.array(cols), np.array(vals))


def test_check_catches_a_bad_plan(L):
    """Plant the mistake in the plan on the CPU: swap two levels of a block so
    rows read values of a level that has not run yet."""
    off, col, coef = _lower(*W.laplacian_2d(12), True)
    n = len(off) - 1
    plan = _plan(L, n, True, off, col, coef, 8192)
    assert _check(plan, n, True, off, col, coef)
    bad = dict(plan)
    lvl = plan["lvl"].copy()
    # move the boundary between levels 1 and 2 so a level-2 row runs with level 1
    lvl[2] += 1
    bad["lvl"] = lvl
    assert not _check(bad, n, True, off, col, coef)


def test_ring_reads_old_values_as_published_rows(L, monkeypatch):
    """A random pattern reads far behind a 64-position window: those entries
    read published rows, the rest slots, and the plan still checks."""
    monkeypatch.setenv("MPK_SWEEP_PLAN", "ring")
    off, col, coef = _lower(*_random(300, 5, 7), True)
    n = len(off) - 1
    plan = _plan(L, n, True, off, col, coef, 64)
    assert plan["ring"] == 64
    d = plan["edep"]
    assert np.any((d < 0) & (d >= -n)) and np.any(d >= 0)
    assert _check(plan, n, True, off, col, coef)


def test_check_catches_a_ring_read_outside_the_window(L, monkeypatch):
    """Plant the mistake on the CPU: shrink the window a ring plan claims
    until some row reads a slot that has been rewritten."""
    monkeypatch.setenv("MPK_SWEEP_PLAN", "ring")
    off, col, coef = _lower(*W.laplacian_2d(20), True)
    n = len(off) - 1
    plan = _plan(L, n, True, off, col, coef, 64)
    assert plan["ring"] == 64 and _check(plan, n, True, off, col, coef)
    bad = dict(plan)
    bad["ring"] = 16
    bad["edep"] = np.where(plan["edep"] >= 0, plan["edep"] % 16, plan["edep"])
    assert not _check(bad, n, True, off, col, coef)


def _narrow_rule(raw, n, wave=True):
    """The one-workgroup-per-block marks restated: 4 or 8 fixed slots, every
    block's levels of at most 256 rows; then 3 (wave blocks: several blocks,
    or reads of published values) or 2 (wave) with at most 12288 levels a
    block and no stale read, else 1 (narrow: one block, no published read,
    at most 4095 levels)."""
    p = _unpack(raw, n)
    d = p["edep"]
    if p["w"] not in (4, 8):
        return 0
    published, stale = bool(np.any((d < 0) & (d >= -n))), bool(np.any(d < -n))
    widest = levels = 0
    for b in range(p["nblk"]):
        lv = p["lvl"][p["blk_lev"][b]:p["blk_lev"][b + 1]]
        if len(lv) > 1:
            widest = max(widest, int(np.diff(lv).max()))
        levels = max(levels, len(lv) - 1)
    if wave and widest <= 256 and levels <= 12288 and not stale:
        return 3 if p["nblk"] > 1 or published else 2
    return 1 if p["nblk"] == 1 and not published and widest <= 256 and levels < 4096 else 0


def _wave_records(raw, n):
    """A wave plan's records restated from its entries: per position the w
    coefficients (-0.0 past the row's count), the w LDS slots (8192 past the
    count), the row and the count, 16-byte aligned."""
    p = _unpack(raw, n)
    h = raw[:64].view(np.int64)
    nblk, nlvl, w = int(h[0]), int(h[2]), int(h[5])
    ints = 2 * (nblk + 1) + nlvl + 3 * n + 1
    at = (64 + (4 * ints + 7) // 8 * 8 + (12 * w * n + 7) // 8 * 8 + 15) // 16 * 16
    rb = (12 * w + 8 + 15) // 16 * 16
    pcnt = np.diff(p["pent"])
    recs = np.zeros((n, rb), np.uint8)
    for q in range(n):
        k = int(pcnt[q])
        c = np.full(w, -0.0)
        dd = np.full(w + 2, 8192, np.int32)
        c[:k] = p["ecoef"][p["pent"][q]:p["pent"][q + 1]]
        dd[:k] = p["edep"][p["pent"][q]:p["pent"][q + 1]]
        dd[w], dd[w + 1] = p["prow"][q], k
        recs[q, :12 * w + 8] = np.concatenate([c.view(np.uint8), dd.view(np.uint8)])
    return at, recs.ravel()


@pytest.mark.parametrize("name,want", [("lap20", 2), ("diff7", 2), ("cd6", 2), ("lap100", 2), ("lap128", 2),
                                       ("lap128-blocks", 3), ("diff20", 0), ("lap300", 0)])
@pytest.mark.parametrize("forward", [True, False])
@pytest.mark.parametrize("stray", [False, True])
def test_narrow_mark(L, monkeypatch, name, want, forward, stray):
    """hdr[7] marks the plans whose blocks one workgroup each runs (levels of
    at most 256 rows): 2 for the wave kernel (a single block or ring with
    every read in LDS: 20^2 to 128^2 grids, small 3-D operators), 3 for the
    wave-block kernel (the 128^2 grid in two blocks, the second reading the
    first's published values), both with their records after the plan; 0 for
    a 20^3 operator (levels of over 256 rows) and a 300-wide grid.
    A stale read (a stray later column) keeps a plan out of the wave form.
    MPK_SWEEP_WAVE=0 leaves the wave plans narrow, MPK_SWEEP_NARROW=0 every
    plan unmarked."""
    mats = dict(SMALL, lap100=lambda: W.laplacian_2d(100), lap128=lambda: W.laplacian_2d(128),
                diff20=lambda: W.diffusion_3d(20)[0], lap300=lambda: W.laplacian_2d(300))
    mats["lap128-blocks"] = mats["lap128"]
    if name.endswith("-blocks"):
        monkeypatch.setenv("MPK_SWEEP_PLAN", "blocks")
    off, col, coef = _lower(*mats[name](), forward, stray)
    if stray and want == 2:
        want = 1
    if stray and want == 3:
        want = 0
    n = len(off) - 1
    m = int(off[n])
    cap = L.mpk_sweep_plan_bound(n, m, 8192)
    buf = np.zeros((cap + 7) // 8, np.int64)
    nb = L.mpk_sweep_plan(n, 1 if forward else 0, off.ctypes.data, col.ctypes.data, coef.ctypes.data, 8192,
                          buf.ctypes.data, cap)
    raw = buf.view(np.uint8)[:nb]
    assert int(raw[:64].view(np.int64)[7]) == _narrow_rule(raw, n) == want
    if want >= 2:
        at, want_bytes = _wave_records(raw, n)
        assert nb == at + len(want_bytes) and nb % 16 == 0
        assert np.array_equal(raw[at:], want_bytes)
    monkeypatch.setenv("MPK_SWEEP_WAVE", "0")
    nb = L.mpk_sweep_plan(n, 1 if forward else 0, off.ctypes.data, col.ctypes.data, coef.ctypes.data, 8192,
                          buf.ctypes.data, cap)
    raw = buf.view(np.uint8)[:nb]
    assert int(buf[7]) == _narrow_rule(raw, n, wave=False) == (1 if want in (1, 2) else 0)
    monkeypatch.setenv("MPK_SWEEP_NARROW", "0")
    nb = L.mpk_sweep_plan(n, 1 if forward else 0, off.ctypes.data, col.ctypes.data, coef.ctypes.data, 8192,
                          buf.ctypes.data, cap)
    assert int(buf[7]) == 0

