// mpk_sweep.h -- sync-free triangular sweeps.
//
// A sweep computes rows in dependency order (forward: row i needs rows < i;
// backward: rows > i).  One wave owns 64 consecutive rows of the sweep order;
// each lane computes one row as a serial sum over the row's entries in stored
// order, so the arithmetic of every row is the CPU's.  Rows publish their
// value in `pub`, which is filled with a signalling-NaN sentinel before the
// sweep: one 64-bit relaxed atomic load both tests readiness and fetches the
// value, so no fence orders a flag against the data.  A dependency inside the
// lane's own wave is read from LDS instead.  Blocks take their position in the
// sweep from a ticket counter, so every block a block waits on has already
// started -- there is no deadlock however the hardware orders dispatch.
#pragma once

#include "mpk_internal.h"

namespace mpk {

#define MPK_SENTINEL 0x7FF4D5A17B0E0000ull  // signalling NaN; arithmetic never yields it
#define MPK_QNAN 0x7FF8000000000000ull

__device__ __forceinline__ unsigned long long load_pub(const double *pub, i64 k) {
  return __hip_atomic_load(reinterpret_cast<const unsigned long long *>(pub + k), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_pub(double *pub, i64 k, double v) {
  unsigned long long b = __double_as_longlong(v);
  if (b == MPK_SENTINEL) b = MPK_QNAN;  // a NaN keeps being a NaN
  __hip_atomic_store(reinterpret_cast<unsigned long long *>(pub + k), b, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// Row programs.  Each provides, for sweep row i:
//   stage(i) -> Staged: the accumulator start and the per-row values the
//               finish needs (every load a row makes that does not depend on
//               other rows' results)
//   begin(i), end(i) -> entry range
//   col(j), coef(j) -> dependency row and coefficient; acc = fma(coef, v, acc)
//   fin(i, acc, staged) -> published value (and any side outputs; no loads)
//   stale(k) -> value of a row that is not earlier in the sweep (malformed input)
struct Staged {
  double init, s1, s2;
  i64 o;
};
// Staged values of every position of a planned sweep (see k_sweep_stage).
struct StagedRows {
  double *init, *s1, *s2;
  i64 *o;
};

// One step of a row's sum: acc + coef * v, fused as the CPU loop's multiply-add.
template <class Prog>
__device__ __forceinline__ double sweep_step(const Prog &p, int j, double v, double acc) {
  return fma(p.coef(j), v, acc);
}

// Generic driver: 64-thread blocks (one wave).
template <class Prog>
__global__ void __launch_bounds__(64) k_sweep(Prog prog, i64 n, int forward, unsigned *ticket,
                                              unsigned ticket_base, double *pub, const int64_t *stop) {
  __shared__ unsigned blk_sh;
  __shared__ double lv[64];
  if (threadIdx.x == 0) blk_sh = atomicAdd(ticket, 1u) - ticket_base;
  __syncthreads();
  if (stop && *stop) return;
  const i64 L0 = (i64)blk_sh * 64;
  const int lane = threadIdx.x;
  const i64 L = L0 + lane;
  const bool valid = L < n;
  const i64 i = valid ? (forward ? L : n - 1 - L) : 0;
  bool done = !valid;
  int j = 0, je = 0;
  double acc = 0.0;
  Staged st{0.0, 0.0, 0.0, 0};
  if (valid) {
    st = prog.stage(i);
    acc = st.init;
    j = prog.begin(i);
    je = prog.end(i);
  }
  unsigned long long mask = 0;  // lanes of this wave that have published
  for (;;) {
    bool moved = false;
    if (!done) {
      const int j_in = j;
      while (j < je) {
        const i64 k = prog.col(j);
        const i64 Lk = forward ? k : n - 1 - k;
        double v;
        if (Lk >= L) {
          // not earlier in the sweep: only a malformed matrix (an SOR row
          // without its diagonal) reaches this; it reads the pre-sweep value
          v = prog.stale(k);
        } else if (Lk >= L0) {
          if (!((mask >> (Lk - L0)) & 1ull)) break;
          v = lv[Lk - L0];
        } else {
          unsigned long long b = load_pub(pub, k);
          if (b == MPK_SENTINEL) break;
          v = __longlong_as_double((long long)b);
        }
        acc = sweep_step(prog, j, v, acc);
        j++;
      }
      if (j >= je) {
        double r = prog.fin(i, acc, st);
        lv[lane] = r;
        store_pub(pub, i, r);
        done = true;
      }
      moved = j != j_in || done;
    }
    __syncthreads();  // one wave: orders the LDS writes before the next reads
    const unsigned long long newmask = __ballot(done);
    if (newmask == ~0ull) break;
    // nothing new in this wave: the rows it waits on belong to earlier waves
    if (newmask == mask && __ballot(moved) == 0ull) __builtin_amdgcn_s_sleep(2);
    mask = newmask;
  }
}

// Block schedule ("plan") of a sweep.  The rows are cut into blocks of
// consecutive sweep positions (at most MPK_BLK_ROWS rows).  Inside a block,
// rows are sorted by their local level: 0 when they read no earlier row of the
// block, else 1 + the deepest local level they read.  One workgroup runs a
// block: level by level, its threads compute the level's rows (each a serial
// sum in the row's own entry order), keeping every value in LDS for the rows
// of later levels, with a workgroup barrier between levels.  Reads of rows in
// earlier blocks go through `pub` (sentinel-filled before the sweep).  Blocks
// take tickets in sweep order, so the oldest unfinished block reads only rows
// of finished blocks and the sweep always drains.
//
// The plan stores, per position (level order within each block): the row,
// its entry count, and its entries in the row's order: the coefficient and
// where the value comes from -- d >= 0: LDS slot d; -n <= d < 0: row -d-1 of
// an earlier block; d < -n: row -d-n-1, not earlier in the sweep (stale,
// malformed input only).  When no row has more than MPK_PLAN_W entries they
// sit in MPK_PLAN_W fixed slots per position (one round of loads per row),
// else in CSR form over pent.  Before the block kernel, k_sweep_stage gathers
// each position's Staged values into position order, so a level's loads are
// all contiguous and none depends on another.
#define MPK_BLK_ROWS 8192
#define MPK_PLAN_W 4

struct PlanView {
  i64 nblk, w;
  const int *blk_pos, *blk_lev, *lvl, *prow, *pcnt, *pent, *edep;
  const double *ecoef;
};

// Layout of one sweep's plan: i64 hdr[8] = {nblk, nent, nlvl, depth, bytes,
// w, 0, 0}, then int32 blk_pos[nblk+1], blk_lev[nblk+1], lvl[nlvl], prow[n],
// pcnt[n], pent[n+1], padding to 8, double ecoef[slots], int32 edep[slots]
// with slots = w ? n w : nent.  The plans of a PC's two sweeps sit back to
// back.
__host__ __device__ inline i64 plan_align8(i64 x) { return (x + 7) & ~(i64)7; }
__host__ __device__ inline PlanView plan_view(const char *base, i64 n) {
  const i64 *h = reinterpret_cast<const i64 *>(base);
  PlanView v;
  v.nblk = h[0];
  v.w = h[5];
  const i64 nent = h[1], nlvl = h[2], slots = v.w ? n * v.w : nent;
  v.blk_pos = reinterpret_cast<const int *>(base + 64);
  v.blk_lev = v.blk_pos + v.nblk + 1;
  v.lvl = v.blk_lev + v.nblk + 1;
  v.prow = v.lvl + nlvl;
  v.pcnt = v.prow + n;
  v.pent = v.pcnt + n;
  const i64 ints = 2 * (v.nblk + 1) + nlvl + 3 * n + 1;
  v.ecoef = reinterpret_cast<const double *>(base + 64 + plan_align8(4 * ints));
  v.edep = reinterpret_cast<const int *>(v.ecoef + slots);
  return v;
}

template <class Prog>
__global__ void k_sweep_stage(Prog prog, i64 n, const char *plan, int second, StagedRows st,
                              const int64_t *stop) {
  if (stop && *stop) return;
  if (second) plan += reinterpret_cast<const i64 *>(plan)[4];
  const PlanView P = plan_view(plan, n);
  const i64 p = (i64)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const Staged s = prog.stage(P.prow[p]);
  st.init[p] = s.init;
  st.s1[p] = s.s1;
  st.s2[p] = s.s2;
  st.o[p] = s.o;
}

// The value of entry d (see the plan encoding above).
template <class Prog>
__device__ __forceinline__ double plan_value(const Prog &prog, i64 n, const double *val, const double *pub,
                                             int d) {
  if (d >= 0) return val[d];
  if (d >= -n) {
    unsigned long long bits;
    while ((bits = load_pub(pub, -(i64)d - 1)) == MPK_SENTINEL) __builtin_amdgcn_s_sleep(1);
    return __longlong_as_double((long long)bits);
  }
  return prog.stale(-(i64)d - n - 1);
}

// A position's inputs, loaded one level ahead of their use.
struct PlanRow {
  int p, cnt, base;
  i64 i;
  Staged s;
  double c[MPK_PLAN_W];
  int d[MPK_PLAN_W];
};

__device__ __forceinline__ PlanRow plan_row(const PlanView &P, const StagedRows &st, int p, bool valid) {
  PlanRow r;
  r.p = valid ? p : -1;
  if (!valid) return r;
  r.cnt = P.pcnt[p];
  r.i = P.prow[p];
  r.s = Staged{st.init[p], st.s1[p], st.s2[p], st.o[p]};
  if (P.w == MPK_PLAN_W) {
    r.base = p * MPK_PLAN_W;
#pragma unroll
    for (int e = 0; e < MPK_PLAN_W; e++) {
      r.c[e] = P.ecoef[(i64)r.base + e];
      r.d[e] = P.edep[(i64)r.base + e];
    }
  } else {
    r.base = P.pent[p];
  }
  return r;
}

template <class Prog>
__device__ __forceinline__ void plan_compute(const Prog &prog, i64 n, const PlanView &P, const PlanRow &r,
                                             double *val, int p0, double *pub) {
  double acc = r.s.init;
  if (P.w == MPK_PLAN_W) {
#pragma unroll
    for (int e = 0; e < MPK_PLAN_W; e++)
      if (e < r.cnt) acc = fma(r.c[e], plan_value(prog, n, val, pub, r.d[e]), acc);
  } else {
    for (int e = r.base, ee = r.base + r.cnt; e < ee; e++)
      acc = fma(P.ecoef[e], plan_value(prog, n, val, pub, P.edep[e]), acc);
  }
  const double v = prog.fin(r.i, acc, r.s);
  val[r.p - p0] = v;
  store_pub(pub, r.i, v);
}

// One workgroup per block.  Each thread's first row of the next level is
// loaded while the current level computes (the barrier waits on LDS only, so
// those loads stay in flight across it); further rows of a wide level are
// loaded when they run.
template <class Prog>
__global__ void __launch_bounds__(256) k_sweep_blk(Prog prog, i64 n, const char *plan, int second, StagedRows st,
                                                   unsigned *ticket, double *pub, const int64_t *stop) {
  __shared__ double val[MPK_BLK_ROWS];
  __shared__ unsigned tk_sh;
  if (threadIdx.x == 0) tk_sh = atomicAdd(ticket, 1u);
  __syncthreads();
  if (stop && *stop) return;
  if (second) plan += reinterpret_cast<const i64 *>(plan)[4];
  const PlanView P = plan_view(plan, n);
  const i64 b = tk_sh;
  if (b >= P.nblk) return;
  const int t = (int)threadIdx.x;
  const int p0 = P.blk_pos[b];
  const int l0 = P.blk_lev[b], l1 = P.blk_lev[b + 1] - 1;  // levels l0..l1-1; lvl[l1] ends the block
  int pa = P.lvl[l0], pb = P.lvl[l0 + 1];
  PlanRow cur = plan_row(P, st, pa + t, pa + t < pb);
  for (int l = l0; l < l1; l++) {
    const bool more = l + 1 < l1;
    const int pc = more ? P.lvl[l + 2] : pb;
    const PlanRow nxt = plan_row(P, st, pb + t, more && pb + t < pc);
    if (cur.p >= 0) plan_compute(prog, n, P, cur, val, p0, pub);
    for (int p = pa + t + 256; p < pb; p += 256) plan_compute(prog, n, P, plan_row(P, st, p, true), val, p0, pub);
    __syncthreads();
    cur = nxt;
    pa = pb;
    pb = pc;
  }
}

// Entries of a sweep in the rows' own order, for the host-side planner:
// count[i] (pass 0), then col / coef at off[i].. (pass 1).
template <class Prog>
__global__ void k_sweep_describe(Prog prog, i64 n, int pass, int *count, const i64 *off, int *col,
                                 double *coef) {
  i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int j0 = prog.begin(i), j1 = prog.end(i);
  if (!pass) {
    count[i] = j1 > j0 ? j1 - j0 : 0;
    return;
  }
  i64 o = off[i];
  for (int j = j0; j < j1; j++, o++) {
    col[o] = (int)prog.col(j);
    coef[o] = prog.coef(j);
  }
}

}  // namespace mpk
