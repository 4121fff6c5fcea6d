"""Batched solves: thousands of independent small systems in one persistent
launch.

    from mpk.batch import BatchSolver
    bs = BatchSolver(systems)            # [(indptr, indices, data, b, "cg+jacobi" | "gmres+ilu"), ...]
    cfg = bs.autotune()                  # workgroup, grid and tiles, timed on (a sample of) this batch
    x, stats = bs.solve(rtol=1e-8)

Each workgroup of the persistent kernel takes systems from a queue (largest
first) and solves each alone: FP64 iterative refinement around an FP32 inner
CG + Jacobi or GMRES(20) + ILU(0) (classical Gram-Schmidt), with FP64 dot
products, norms, scalars and least squares.  The launch is captured once into
a HIP graph and replayed.  The library side is mpk_batch_create /
mpk_batch_solve / mpk_batch_free.
"""

from __future__ import annotations

import time

import numpy as np

from ._native import check, lib
from .device import DeviceBuffer

_KIND = {"cg+jacobi": 0, "gmres+ilu": 1}
_W = 4          # sweep entry slots per position (mpk_batch.hip BS_W)
_FIELDS = 8     # per-system descriptor fields


def _a(x: np.ndarray) -> int:
    return x.ctypes.data


def _diag_markers(ip, ix, n):
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(ip))
    adiag = np.asarray(ip[1:], dtype=np.int64).copy()
    hit = np.nonzero(ix == rows)[0]
    if len(hit):
        first = np.full(n, -1, dtype=np.int64)
        rr = rows[hit]
        keep = np.r_[True, rr[1:] != rr[:-1]]
        first[rr[keep]] = hit[keep]
        adiag[first >= 0] = first[first >= 0]
    return adiag


def _ilu0(L, ip, ix, va, n):
    """ILU(0) with natural ordering through the library's host routines (the
    reference's fast path, PC.py:399-405): the factor in the reference layout."""
    irow = np.ascontiguousarray(ip, dtype=np.int64)
    icol = np.ascontiguousarray(ix, dtype=np.int64)
    data = np.ascontiguousarray(va, dtype=np.float64)
    adiag = _diag_markers(irow, icol, n)
    cap = max(4 * n, 4 * int(irow[-1]) + 4 * n, 16)
    bi = np.zeros(n + 1, dtype=np.int64)
    bj = np.zeros(cap, dtype=np.int64)
    bdiag = np.zeros(n + 1, dtype=np.int64)
    nnz = L.mp_ilu_symbolic_ilu0(_a(bi), _a(bj), _a(bdiag), _a(adiag), _a(irow), _a(icol), n)
    if nnz < 0:
        raise RuntimeError("ILU(0) symbolic factorisation failed")
    r = np.arange(n, dtype=np.int64)
    ba = np.zeros(max(int(bdiag[0]) + 1, 1), dtype=np.float64)
    rtmp = np.zeros(n, dtype=np.float64)
    L.mp_pc_ilu_numeric(_a(ba), _a(rtmp), _a(irow), _a(icol), _a(data), _a(bi), _a(bj), _a(bdiag), _a(r),
                        _a(r), n)
    return bi, bj, ba, bdiag


def _plan(L, n, forward, off, col, coef):
    """A single-block sweep plan (mpk_sweep_plan with bmax = n): positions by
    level, entry slots, level starts."""
    cap = L.mpk_sweep_plan_bound(n, int(off[n]), max(n, 1))
    buf = np.zeros((cap + 7) // 8, dtype=np.int64)
    nb = L.mpk_sweep_plan(n, 1 if forward else 0, _a(off), _a(col), _a(coef), max(n, 1), _a(buf), cap)
    if nb <= 0:
        raise RuntimeError("sweep planning failed")
    raw = buf.view(np.uint8)[:nb]
    h = raw[:64].view(np.int64)
    nblk, nent, nlvl, w = int(h[0]), int(h[1]), int(h[2]), int(h[5])
    if nblk != 1 or w != _W:
        raise RuntimeError(f"batched sweeps need one block and {_W} slots per row (got {nblk}, {w})")
    ints = raw[64:].view(np.int32)
    o = 2 * (nblk + 1)
    lvl = ints[o:o + nlvl].copy()
    o += nlvl
    prow = ints[o:o + n].copy()
    pcnt = ints[o + n:o + 2 * n].copy()
    base = 64 + ((4 * (2 * (nblk + 1) + nlvl + 3 * n + 1) + 7) // 8) * 8
    slots = n * w
    # the plan holds slot e of position p at e n + p; the batched kernel reads p w + e
    ecoef = raw[base:base + 8 * slots].view(np.float64).reshape(w, n).T.astype(np.float32).ravel()
    edep = raw[base + 8 * slots:base + 12 * slots].view(np.int32).reshape(w, n).T.copy().ravel()
    return lvl, prow, pcnt, edep, ecoef


_RING = 2048  # LDS ring of sweep values in the batched kernel (BS_RING)


def _ring_slots(lvl, pcnt, edep, n):
    """A single-block plan's value slots (positions, _W per position) as slots
    of the LDS ring (position mod _RING), or None when some entry reads a
    position _RING or more behind the end of its row's level (its slot could
    have been rewritten by then)."""
    pos = np.arange(n)
    lev = np.searchsorted(lvl, pos, side="right") - 1
    end = np.asarray(lvl, np.int64)[lev + 1]
    deps = edep.reshape(n, _W).astype(np.int64)
    used = np.arange(_W)[None, :] < np.asarray(pcnt)[:, None]
    if np.any((end[:, None] - deps)[used] > _RING):
        return None
    return np.where(used, deps % _RING, 0).astype(np.int32).ravel()


def build_arrays(systems, L=None) -> dict:
    """The batch's device layout as host arrays (mpk_batch_create documents
    it): concatenated CSR, Jacobi inverse diagonal, ILU(0) sweep plans."""
    if L is None:
        L = lib()
    count = len(systems)
    if count == 0:
        raise ValueError("empty batch")
    sizes = np.array([len(s[0]) - 1 for s in systems], dtype=np.int64)
    row0 = np.zeros(count + 1, dtype=np.int64)
    row0[1:] = np.cumsum(sizes)
    N = int(row0[-1])
    nnz = np.array([int(s[0][-1]) for s in systems], dtype=np.int64)
    e0 = np.zeros(count + 1, dtype=np.int64)
    e0[1:] = np.cumsum(nnz)
    rp = np.zeros(N + 1, dtype=np.int32)
    ci = np.zeros(int(e0[-1]), dtype=np.int32)
    v64 = np.zeros(int(e0[-1]), dtype=np.float64)
    b = np.zeros(N, dtype=np.float64)
    jd = np.zeros(N, dtype=np.float32)
    fprow = np.zeros(N, np.int32)
    fcnt = np.zeros(N, np.int32)
    bprow = np.zeros(N, np.int32)
    bcnt = np.zeros(N, np.int32)
    fdep = np.zeros(N * _W, np.int32)
    bdep = np.zeros(N * _W, np.int32)
    fcoef = np.zeros(N * _W, np.float32)
    bcoef = np.zeros(N * _W, np.float32)
    dinv = np.zeros(N, np.float32)
    f2b = np.zeros(N, np.int32)
    flv, blv = [], []
    nflv = nblv = 0
    sysd = np.zeros((count, _FIELDS), dtype=np.int64)
    for s, (ip, ix, va, bb, kind) in enumerate(systems):
        if kind not in _KIND:
            raise ValueError(f"system {s}: kind must be one of {sorted(_KIND)}")
        n = int(sizes[s])
        r0, q0 = int(row0[s]), int(e0[s])
        ip = np.asarray(ip, dtype=np.int64)
        rp[r0:r0 + n + 1] = ip + q0
        ci[q0:q0 + nnz[s]] = np.asarray(ix[:nnz[s]], dtype=np.int64) + r0
        v64[q0:q0 + nnz[s]] = va[:nnz[s]]
        b[r0:r0 + n] = bb
        rows = np.repeat(np.arange(n), np.diff(ip))
        dg = np.zeros(n)
        on = np.asarray(ix[:nnz[s]]) == rows
        dg[rows[on]] = np.asarray(va[:nnz[s]])[on]
        jd[r0:r0 + n] = np.where(dg != 0.0, 1.0 / np.where(dg != 0.0, dg, 1.0), 1.0)
        sysd[s, 0], sysd[s, 1], sysd[s, 2] = r0, n, _KIND[kind]
        if kind == "gmres+ilu":
            bi, bj, ba, bd = _ilu0(L, ip, ix, va, n)
            # forward: the L rows, stored in row order; backward: the U rows
            # (stored last row first, each before its inverted diagonal),
            # both with coefficient -ba as the reference's sweeps use them
            fo = np.asarray(bi[:n + 1], dtype=np.int64) - bi[0]
            fsl = slice(int(bi[0]), int(bi[n]))
            fcol = np.asarray(bj[fsl], dtype=np.int32) if fo[-1] else np.zeros(1, np.int32)
            fcf = -np.asarray(ba[fsl]) if fo[-1] else np.zeros(1)
            starts = np.asarray(bd[1:n + 1], dtype=np.int64) + 1
            lens = np.asarray(bd[:n], dtype=np.int64) - starts
            bo = np.zeros(n + 1, np.int64)
            bo[1:] = np.cumsum(lens)
            if bo[-1]:
                idx = np.repeat(starts - bo[:-1], lens) + np.arange(int(bo[-1]))
                bcol, bcf = np.asarray(bj[idx], dtype=np.int32), -np.asarray(ba[idx])
            else:
                bcol, bcf = np.zeros(1, np.int32), np.zeros(1)
            lf, pr, pc_, ed, ec = _plan(L, n, True, fo, np.ascontiguousarray(fcol),
                                        np.ascontiguousarray(fcf, dtype=np.float64))
            fprow[r0:r0 + n], fcnt[r0:r0 + n] = pr, pc_
            fdep[r0 * _W:(r0 + n) * _W], fcoef[r0 * _W:(r0 + n) * _W] = ed, ec
            lb, pr, pc_, ed, ec = _plan(L, n, False, bo, np.ascontiguousarray(bcol),
                                        np.ascontiguousarray(bcf, dtype=np.float64))
            bprow[r0:r0 + n], bcnt[r0:r0 + n] = pr, pc_
            bdep[r0 * _W:(r0 + n) * _W], bcoef[r0 * _W:(r0 + n) * _W] = ed, ec
            rf = _ring_slots(lf, fcnt[r0:r0 + n], fdep[r0 * _W:(r0 + n) * _W], n)
            rb = _ring_slots(lb, pc_, ed, n)
            if rf is not None and rb is not None:
                fdep[r0 * _W:(r0 + n) * _W], bdep[r0 * _W:(r0 + n) * _W] = rf, rb
                sysd[s, 7] = 1
            # by position: the backward sweep's inverted diagonal, and where
            # each forward position's row sits in the backward order
            bpos = np.empty(n, np.int64)
            bpos[bprow[r0:r0 + n]] = np.arange(n)
            f2b[r0:r0 + n] = bpos[fprow[r0:r0 + n]]
            dinv[r0:r0 + n] = np.asarray(ba)[np.asarray(bd)[bprow[r0:r0 + n]]]
            sysd[s, 3], sysd[s, 4] = nflv, len(lf) - 1
            sysd[s, 5], sysd[s, 6] = nblv, len(lb) - 1
            flv.append(lf)
            blv.append(lb)
            nflv += len(lf)
            nblv += len(lb)
    flv = np.concatenate(flv).astype(np.int32) if flv else np.zeros(1, np.int32)
    blv = np.concatenate(blv).astype(np.int32) if blv else np.zeros(1, np.int32)
    order = np.argsort(-sizes, kind="stable").astype(np.int32)
    return dict(sizes=sizes, row0=row0, N=N, sys=sysd, order=order, rp=rp, ci=ci, v64=v64,
                v32=v64.astype(np.float32), jd=jd, fprow=fprow, fcnt=fcnt, fdep=fdep, fcoef=fcoef, flv=flv,
                bprow=bprow, bcnt=bcnt, bdep=bdep, bcoef=bcoef, blv=blv, dinv=dinv, f2b=f2b, b=b)


class BatchSolver:
    """One batch of independent systems, set up once on the device."""

    def __init__(self, systems):
        L = lib()
        self._systems = systems
        H = build_arrays(systems, L)
        self.count = len(systems)
        self.sizes = H["sizes"]
        self.N = N = H["N"]
        self.row0 = H["row0"]
        self.b_host = H["b"]
        self.kinds = [s[4] for s in systems]
        up = DeviceBuffer.from_array
        keys = ("sys", "order", "rp", "ci", "v64", "v32", "jd", "fprow", "fcnt", "fdep", "fcoef", "flv",
                "bprow", "bcnt", "bdep", "bcoef", "blv", "dinv", "f2b", "b")
        self._d = {k: up(H[k]) for k in keys}
        self._d["x"] = DeviceBuffer(8 * N)
        self._d["stats"] = DeviceBuffer(8 * 4 * self.count)
        d = self._d
        self._h = L.mpk_batch_create(self.count, d["sys"].ptr, d["order"].ptr, d["rp"].ptr, d["ci"].ptr,
                                     d["v64"].ptr, d["v32"].ptr, d["jd"].ptr, d["fprow"].ptr, d["fcnt"].ptr,
                                     d["fdep"].ptr, d["fcoef"].ptr, d["flv"].ptr, d["bprow"].ptr,
                                     d["bcnt"].ptr, d["bdep"].ptr, d["bcoef"].ptr, d["blv"].ptr,
                                     d["dinv"].ptr, d["f2b"].ptr, int(self.sizes.max()))
        check()
        if not self._h:
            raise RuntimeError("mpk_batch_create failed")
        self.config = {"workgroup": 256, "grid": 512, "variant": -1}

    @staticmethod
    def variants() -> list:
        """The compiled launch variants: dicts of workgroup, rows_per_thread
        (the product's rows a thread takes per pass; a wavefront covers 64
        times that), spmv_block (the product's entries loaded per round) and
        gs_tiling (elements a thread takes per step of the Gram-Schmidt
        sums), indexed as solve(variant=...) takes them."""
        L = lib()
        buf = np.zeros(4 * 64, dtype=np.int64)
        k = L.mpk_batch_variants(_a(buf), 64)
        return [{"workgroup": int(buf[4 * v]), "rows_per_thread": int(buf[4 * v + 1]),
                 "spmv_block": int(buf[4 * v + 2]), "gs_tiling": int(buf[4 * v + 3])} for v in range(k)]

    def solve(self, rtol: float = 1e-8, inner_rtol: float = 1e-3, max_outer: int = 30, max_inner: int = 4000,
              workgroup: int | None = None, grid: int | None = None, graph: bool = True,
              variant: int | None = None):
        """Solves every system; returns (x, stats) with x the concatenated
        FP64 solutions and stats[s] = (relres, outer, inner, converged).
        variant: an index into variants() (its workgroup size is used), or
        -1 for the default tiles at the workgroup size."""
        if variant is None:
            variant = self.config["variant"] if workgroup is None else -1
        if variant >= 0:
            wg = self.variants()[variant]["workgroup"]
        else:
            wg = workgroup or self.config["workgroup"]
        g = grid or self.config["grid"]
        cfg = np.array([rtol, inner_rtol, max_outer, max_inner, wg, g, 1.0 if graph else 0.0, float(variant)])
        d = self._d
        rc = lib().mpk_batch_solve(self._h, d["b"].ptr, d["x"].ptr, d["stats"].ptr, _a(cfg))
        check()
        if rc != 0:
            raise RuntimeError("mpk_batch_solve failed")
        return d["x"].to_numpy(np.float64, self.N), d["stats"].to_numpy(np.float64, 4 * self.count).reshape(-1, 4)

    def solution(self, x: np.ndarray, s: int) -> np.ndarray:
        return x[self.row0[s]:self.row0[s + 1]]

    def autotune(self, candidates=((128, 512), (256, 256), (256, 512), (256, 1024), (512, 256), (512, 512)),
                 reps: int = 1, tiles: bool = True, sample_every: int = 1, **solve_kw) -> dict:
        """Chooses the launch for this batch by timing solves: first each
        (workgroup size, grid) of `candidates` with the default tiles, then,
        at the fastest of those, every compiled tile variant for that
        workgroup size -- the rows a thread (so a wavefront) takes per pass
        of the product, the product's entry block (its variant), and the
        Gram-Schmidt reduction tiling.  With sample_every = k the timing runs
        on every k-th system (a smaller batch of the same mix) and the choice
        is applied to this one.  Returns the choice and every timing."""
        tb = self if sample_every <= 1 else BatchSolver(self._systems[::sample_every])
        var = self.variants()

        def timed(**kw):
            tb.solve(**kw, **solve_kw)  # warm-up (and graph capture)
            t0 = time.perf_counter()
            for _ in range(reps):
                tb.solve(**kw, **solve_kw)
            return (time.perf_counter() - t0) / reps

        try:
            timings = {}
            for wg, g in candidates:
                timings[f"{wg}x{g}"] = timed(workgroup=wg, grid=g, variant=-1)
            best_wg, best_g = min(candidates, key=lambda c: timings[f"{c[0]}x{c[1]}"])
            best_t, best_v = timings[f"{best_wg}x{best_g}"], -1
            if tiles:
                for v, d in enumerate(var):
                    if d["workgroup"] != best_wg or (d["rows_per_thread"], d["spmv_block"], d["gs_tiling"]) == (4, 4, 4):
                        continue
                    t = timed(grid=best_g, variant=v)
                    timings[f"{best_wg}x{best_g}/ur{d['rows_per_thread']}-kb{d['spmv_block']}-ug{d['gs_tiling']}"] = t
                    if t < best_t:
                        best_t, best_v = t, v
        finally:
            if tb is not self:
                tb.destroy()
        chosen = var[best_v] if best_v >= 0 else {"workgroup": best_wg, "rows_per_thread": 4, "spmv_block": 4,
                                                   "gs_tiling": 4}
        self.config = {"workgroup": best_wg, "grid": best_g, "variant": best_v}
        return {"workgroup": best_wg, "grid": best_g, "variant": best_v,
                "rows_per_wavefront": 64 * chosen["rows_per_thread"],
                "spmv_variant": f"{chosen['spmv_block']} entries per round",
                "reduction_tiling": chosen["gs_tiling"], "timings": timings,
                "sampled_systems": tb.count if tb is not self else self.count}

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            lib().mpk_batch_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
