// mpk_sweep.h -- sync-free triangular sweeps.
//
// A sweep computes rows in dependency order (forward: row i needs rows < i;
// backward: rows > i).  Each row is one thread's serial sum over the row's
// entries in stored order, so the arithmetic of every row is the CPU's.  Rows
// publish their value in `pub`, which is filled with a signalling-NaN sentinel
// before the sweep: one relaxed atomic load both tests readiness and fetches
// the value, so no fence orders a flag against the data.  Two drivers:
// k_sweep (no schedule: one wave per 64 consecutive rows, tickets in sweep
// order) and k_sweep_blk (a schedule built at set-up: block plans or ring
// plans, below).  Both drain whatever order the hardware dispatches in.
#pragma once

#include "mpk_internal.h"

namespace mpk {

#define MPK_SENTINEL 0x7FF4D5A17B0E0000ull  // signalling NaN; arithmetic never yields it
#define MPK_QNAN 0x7FF8000000000000ull
#define MPK_SENTINEL_F 0x7FA5D5A1u  // the same for FP32 values
#define MPK_QNAN_F 0x7FC00000u

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
__device__ __forceinline__ unsigned load_pub_f(const float *pub, i64 k) {
  return __hip_atomic_load(reinterpret_cast<const unsigned *>(pub + k), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_pub_f(float *pub, i64 k, float v) {
  unsigned b = __float_as_uint(v);
  if (b == MPK_SENTINEL_F) b = MPK_QNAN_F;
  __hip_atomic_store(reinterpret_cast<unsigned *>(pub + k), b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded waits.  A wait for a row another workgroup publishes counts its
// polls; it gives up after the watch's limit, or at its next 256th poll once
// the host raises the abort word (mpk_sweep_abort), and then takes a quiet
// NaN for the value.  The kernel so runs to its end (the rows downstream of
// the wait come out NaN) and, as its last act, raises the fault word, which
// the library reports as an error at its next synchronisation.  Both words
// sit in fine-grained host memory.  A well-formed plan never gives up: its
// blocks take tickets in sweep order, so every awaited row is being computed.
// `gave` is kept in a register by the loops and reported once at the end, so
// no store sits in a loop's branches.
struct SweepWatch {
  unsigned *word;  // [0] abort (set by the host), [1] fault (set by kernels)
  unsigned limit;  // polls before a wait gives up
};
static __device__ SweepWatch g_sweep_watch;  // this file's copy (watch_install)

// Counts one poll; true when the wait must give up.
__device__ __forceinline__ bool watch_expired(unsigned &polls) {
  ++polls;
  const SweepWatch w = g_sweep_watch;
  if (!w.word) return false;
  if (polls >= w.limit) return true;
  return (polls & 255u) == 0 && __hip_atomic_load(w.word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u;
}
__device__ __forceinline__ void watch_report(bool gave) {
  if (!gave) return;
  unsigned *w = g_sweep_watch.word;
  if (w) __hip_atomic_store(w + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Host side: copies the process's watch (words and limit) into this file's
// g_sweep_watch when it changed since the last copy; every launch of a
// kernel that polls goes after it.
static unsigned g_sweep_watch_gen = 0;
static inline void watch_install() {
  unsigned *word = nullptr;
  unsigned limit = 0, gen = 0;
  sweep_watch_state(&word, &limit, &gen);
  if (gen == g_sweep_watch_gen) return;
  const SweepWatch w{word, limit};
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(g_sweep_watch), &w, sizeof w, 0, hipMemcpyHostToDevice));
  if (!failed()) g_sweep_watch_gen = gen;
}

// Value types of a planned sweep: FP64 (the drop-in preconditioners) or FP32
// (the inner solves of the mixed-precision solver).
struct SweepF64 {
  using V = double;
  using Bits = unsigned long long;
  static constexpr Bits kSentinel = MPK_SENTINEL;
  static constexpr Bits kQuiet = MPK_QNAN;
  __device__ static Bits load(const V *pub, i64 k) { return load_pub(pub, k); }
  __device__ static void store(V *pub, i64 k, V v) { store_pub(pub, k, v); }
  __device__ static V value(Bits b) { return __longlong_as_double((long long)b); }
  __device__ static V madd(V a, V b, V c) { return fma(a, b, c); }
};
struct SweepF32 {
  using V = float;
  using Bits = unsigned;
  static constexpr Bits kSentinel = MPK_SENTINEL_F;
  static constexpr Bits kQuiet = MPK_QNAN_F;
  __device__ static Bits load(const V *pub, i64 k) { return load_pub_f(pub, k); }
  __device__ static void store(V *pub, i64 k, V v) { store_pub_f(pub, k, v); }
  __device__ static V value(Bits b) { return __uint_as_float(b); }
  __device__ static V madd(V a, V b, V c) { return fmaf(a, b, c); }
};

// Row programs.  Each provides, for sweep row i:
//   stage(i) -> Staged: the accumulator start and the per-row values the
//               finish needs (every load a row makes that does not depend on
//               other rows' results)
//   begin(i), end(i) -> entry range
//   col(j), coef(j) -> dependency row and coefficient; acc = fma(coef, v, acc)
//   fin(i, acc, staged) -> published value (and any side outputs; no loads)
//   stale(k) -> value of a row that is not earlier in the sweep (malformed input)
struct alignas(16) Staged {
  double init, s1, s2;
  i64 o;
};
// The accumulator a row's sum starts from.
__device__ __forceinline__ double row_start(const Staged &s) { return s.init; }

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
  unsigned polls = 0;
  bool gave = false;  // the wave's wait gave up: unready values read as NaN
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
          if (!((mask >> (Lk - L0)) & 1ull)) {
            if (!gave) break;
            v = __longlong_as_double((long long)MPK_QNAN);
          } else {
            v = lv[Lk - L0];
          }
        } else {
          unsigned long long b = load_pub(pub, k);
          if (b == MPK_SENTINEL) {
            if (!gave) break;
            b = MPK_QNAN;
          }
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
    if (newmask == mask && __ballot(moved) == 0ull) {
      __builtin_amdgcn_s_sleep(2);
      if (watch_expired(polls)) gave = true;  // uniform: polls and the watch are
    }
    mask = newmask;
  }
  if (lane == 0) watch_report(gave);
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
// malformed input only).  Entries sit in w fixed slots per position (w = 4,
// 8, 16 or 32, the smallest that holds the longest row; for longer rows a
// multiple of 8 up to 64 when the padding stays within 4 times the entries),
// slot e of position p at e n + p: the loads of a row are independent of each
// other and the rows of a level (consecutive positions, one per lane) read
// each slot from consecutive addresses.  Else in CSR form over pent.  Before the block kernel, k_sweep_stage
// gathers each position's Staged values into position order, so a level's
// loads are all contiguous and none depends on another.
#define MPK_BLK_ROWS 8192
#define MPK_PLAN_PF 16  // slots of a row held in registers (loaded a level ahead)
// A ring plan (one sweeping workgroup) of rows wider than MPK_PLAN_PF slots
// holds this many: a row past the registers loads the rest when it runs, one
// more round trip on its level's path (with the 1e6-row random SPD matrix's
// lower triangle, 10% of rows pass 16 entries, 0.07% pass 24).
#define MPK_RING_WIDE 24

struct PlanView {
  i64 n, nblk, w;
  int mask;  // LDS slot of position p in block b: (p - blk_pos[b]) & mask
  const int *blk_pos, *blk_lev, *lvl, *prow, *pcnt, *pent, *edep;
  const double *ecoef;
  const char *wrec;  // a wave plan's records (hdr[7] = 2)
};

// A ring plan (hdr[6] = W > 0) is one block over the sweep's global levels
// whose LDS keeps the values of the last W positions: position p lives in
// slot p mod W.  A row reads a value from its slot when it is less than W
// positions behind the end of the row's level (so no row of that level can
// have rewritten the slot), else from `pub` as if from an earlier block.
// Deep, narrow sweeps run that way, with no waits on other blocks.
//
// Layout of one sweep's plan: i64 hdr[8] = {nblk, nent, nlvl, depth, bytes,
// w, W (0: blocks), 0}, then int32 blk_pos[nblk+1], blk_lev[nblk+1], lvl[nlvl], prow[n],
// pcnt[n], pent[n+1], padding to 8, double ecoef[slots], int32 edep[slots]
// with slots = w ? n w : nent.  The plans of a PC's two sweeps sit back to
// back.
__host__ __device__ inline i64 plan_align8(i64 x) { return (x + 7) & ~(i64)7; }
__host__ __device__ inline PlanView plan_view(const char *base, i64 n) {
  const i64 *h = reinterpret_cast<const i64 *>(base);
  PlanView v;
  v.n = n;
  v.nblk = h[0];
  v.w = h[5];
  v.mask = h[6] ? (int)(h[6] - 1) : -1;
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
  v.wrec = base + ((64 + plan_align8(4 * ints) + plan_align8(12 * slots) + 15) & ~(i64)15);
  return v;
}

// Coefficient of plan entry q: the plan's own FP64 values (other sources, such
// as the mixed solver's low-precision copies, provide the same two members).
// A coefficient source: raw(q) loads slot q as stored, conv(raw) gives the
// value the sum takes.  Rows prefetched a level ahead hold the raw values,
// converted only at their use, so that no conversion waits on a load still
// in flight.
struct PlanCoef {
  using Raw = double;
  const double *c;
  static constexpr int kBytes = 8;
  __device__ static PlanCoef make(const PlanView &P, const void *) { return PlanCoef{P.ecoef}; }
  __device__ double operator()(i64 q) const { return c[q]; }
  __device__ Raw raw(i64 q) const { return c[q]; }
  __device__ static double conv(Raw r) { return r; }
};

// Writes each position's staged values (Prog::stage of its row) into
// position order, so that a level's loads of them are contiguous.
template <class Prog, class S>
__global__ void k_sweep_stage(Prog prog, i64 n, const char *plan, int second, S *st, const int64_t *stop) {
  if (stop && *stop) return;
  if (second) plan += reinterpret_cast<const i64 *>(plan)[4];
  const PlanView P = plan_view(plan, n);
  const i64 p = (i64)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  st[p] = prog.stage(P.prow[p]);
}

// A position's inputs, loaded one level ahead of their use: with fixed slots
// every load is independent of the others (one round); in CSR form only the
// entry offset is fetched ahead.
template <int PF, class Tr, class S, class R>
struct PlanRow {
  int p, cnt, base;
  i64 i;
  S s;
  R c[PF];  // raw coefficients (Coef::raw)
  int d[PF];
};

template <int PF, class Tr, class Coef, class S>
__device__ __forceinline__ PlanRow<PF, Tr, S, typename Coef::Raw> plan_row(const PlanView &P, const Coef &coef,
                                                                           const S *st, int p, bool valid) {
  PlanRow<PF, Tr, S, typename Coef::Raw> r;
  r.p = valid ? p : -1;
  if (!valid) return r;
  r.cnt = P.pcnt[p];
  r.i = P.prow[p];
  r.s = st[p];
  if (P.w) {
    r.base = p;
#pragma unroll
    for (int e = 0; e < PF; e++)
      if (e < P.w) {
        r.c[e] = coef.raw((i64)e * P.n + p);
        r.d[e] = P.edep[(i64)e * P.n + p];
      }
  } else {
    r.base = P.pent[p];
  }
  return r;
}

// The values of up to G entries, in two phases so that every load is in
// flight before any is waited on.  Issue: the LDS slot is read for every
// entry (index clamped: a different register from the global loads, so the
// compiler need not drain the global loads first) and the rows of earlier
// blocks are loaded; resolve: pick each entry's value, polling rows still
// holding the sentinel.
template <int G, class Tr>
struct Gather {
  typename Tr::V lv[G];
  typename Tr::Bits gb[G];
};

template <int G, class Tr>
__device__ __forceinline__ void gather_issue(i64 n, const typename Tr::V *val, const typename Tr::V *pub,
                                             const int *d, int cnt, Gather<G, Tr> &g) {
#pragma unroll
  for (int e = 0; e < G; e++) {
    const int de = e < cnt ? d[e] : 0;
    g.lv[e] = val[de >= 0 ? de : 0];
    g.gb[e] = 0;
    if (de < 0 && de >= -n) g.gb[e] = Tr::load(pub, -(i64)de - 1);
  }
}

template <int G, class Tr, class Prog>
__device__ __forceinline__ void gather_resolve(const Prog &prog, i64 n, const typename Tr::V *pub, const int *d,
                                               int cnt, const Gather<G, Tr> &g, typename Tr::V *v, bool &gave) {
#pragma unroll
  for (int e = 0; e < G; e++) {
    if (e < cnt) {
      const int de = d[e];
      if (de >= 0) {
        v[e] = g.lv[e];
      } else if (de >= -n) {
        typename Tr::Bits b = g.gb[e];
        unsigned polls = 0;
        while (b == Tr::kSentinel) {
          __builtin_amdgcn_s_sleep(1);
          if (watch_expired(polls)) {
            gave = true;
            b = Tr::kQuiet;
            break;
          }
          b = Tr::load(pub, -(i64)de - 1);
        }
        v[e] = Tr::value(b);
      } else {
        v[e] = prog.stale(-(i64)de - n - 1);
      }
    }
  }
}

// A row's sum and finish once its slots' values are gathered (register part
// in g); entries beyond the registers (or CSR form) are gathered here.
template <int PF, class Tr, class Prog, class Coef, class S>
__device__ __forceinline__ void plan_finish(const Prog &prog, const Coef &coef, i64 n, const PlanView &P,
                                            const PlanRow<PF, Tr, S, typename Coef::Raw> &r,
                                            const Gather<PF, Tr> &g,
                                            typename Tr::V *val, int p0, typename Tr::V *pub, bool &gave) {
  using V = typename Tr::V;
  V acc = row_start(r.s);
  int j_rest = 0;  // entries still to do (beyond the registers, or CSR form)
  if (P.w) {
    V v[PF];
    gather_resolve<PF, Tr>(prog, n, pub, r.d, r.cnt, g, v, gave);
#pragma unroll
    for (int e = 0; e < PF; e++)
      if (e < r.cnt) acc = Tr::madd(Coef::conv(r.c[e]), v[e], acc);
    j_rest = PF;
  }
  const i64 stride = P.w ? P.n : 1;  // entry j of the row at base + j stride
  for (int j0 = j_rest; j0 < r.cnt; j0 += 8) {
    const int k = r.cnt - j0 < 8 ? r.cnt - j0 : 8;
    int d[8];
    V c[8], v[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const i64 q = r.base + (i64)(j0 + e) * stride;
      d[e] = e < k ? P.edep[q] : 0;
      c[e] = e < k ? coef(q) : V(0);
    }
    Gather<8, Tr> g8;
    gather_issue<8, Tr>(n, val, pub, d, k, g8);
    gather_resolve<8, Tr>(prog, n, pub, d, k, g8, v, gave);
#pragma unroll
    for (int e = 0; e < 8; e++)
      if (e < k) acc = Tr::madd(c[e], v[e], acc);
  }
  const V v = prog.fin(r.i, acc, r.s);
  val[(r.p - p0) & P.mask] = v;
  Tr::store(pub, r.i, v);
}

template <int PF, class Tr, class Prog, class Coef, class S>
__device__ __forceinline__ void plan_compute(const Prog &prog, const Coef &coef, i64 n, const PlanView &P,
                                             const PlanRow<PF, Tr, S, typename Coef::Raw> &r, typename Tr::V *val,
                                             int p0,
                                             typename Tr::V *pub, bool &gave) {
  Gather<PF, Tr> g;
  if (P.w) gather_issue<PF, Tr>(n, val, pub, r.d, r.cnt, g);
  plan_finish<PF, Tr>(prog, coef, n, P, r, g, val, p0, pub, gave);
}

// Ring plans run in workgroup 0 while workgroups 8, 16, .., 8 MPK_RING_PF
// (workgroup 0's XCD under round-robin dispatch) load the data of the levels
// just ahead of it -- each level's rows, staged values and entry slots -- so
// that the sweep's own loads find it in the XCD's L2 rather than in HBM.  The
// sweep publishes its level every MPK_RING_STEP levels (MPK_RING_DONE when it
// ends) in progress; a prefetcher stays at most MPK_RING_AHEAD levels ahead
// and returns once the sweep has ended or the solve is stopped.
#define MPK_RING_PF 8
#define MPK_RING_AHEAD 16
#define MPK_RING_STEP 4
#define MPK_RING_DONE 0x7fffffffu

__device__ __forceinline__ unsigned ring_progress(const unsigned *pr) {
  return __hip_atomic_load(pr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ring_publish(unsigned *pr, unsigned v) {
  __hip_atomic_store(pr, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One word of each 128-byte line of [a, b), the lines spread over the
// workgroup's threads.
__device__ __forceinline__ unsigned touch_span(const void *a, const void *b, int t) {
  unsigned x = 0;
  const char *q0 = reinterpret_cast<const char *>(reinterpret_cast<uintptr_t>(a) & ~(uintptr_t)127);
  for (const char *q = q0 + 128 * (i64)t; q < reinterpret_cast<const char *>(b); q += 128 * 256)
    x ^= *reinterpret_cast<const unsigned *>(q);
  return x;
}

// Prefetcher k (1..MPK_RING_PF) of a ring plan: levels k-1, k-1+MPK_RING_PF, ..
// (coefficients of cbytes bytes at coef, slot-major like the plan's).
template <class S>
__device__ __forceinline__ void ring_prefetch(const PlanView &P, const S *st, const char *coef, int cbytes, int k,
                                              unsigned *progress, unsigned *sink, const int64_t *stop) {
  __shared__ int done_sh;
  const int t = (int)threadIdx.x;
  const int nl = P.blk_lev[1] - 1;
  unsigned acc = 0;
  bool gave = false;
  for (int l = k - 1; l < nl; l += MPK_RING_PF) {
    if (t == 0) {
      unsigned pr = ring_progress(progress);
      unsigned polls = 0;
      while (pr != MPK_RING_DONE && l > (int)pr + MPK_RING_AHEAD) {
        if (stop && *stop) {
          pr = MPK_RING_DONE;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
        if (watch_expired(polls)) {
          gave = true;
          pr = MPK_RING_DONE;
          break;
        }
        pr = ring_progress(progress);
      }
      done_sh = pr == MPK_RING_DONE;
    }
    __syncthreads();
    if (done_sh) break;
    const int s = P.lvl[l], e = P.lvl[l + 1];
    acc ^= touch_span(P.prow + s, P.prow + e, t) ^ touch_span(P.pcnt + s, P.pcnt + e, t);
    acc ^= touch_span(st + s, st + e, t);
    if (P.w) {
      // each slot's run of the level's positions: (slot, line) pairs spread
      // over the threads, four loads in flight per thread
      const int wb = cbytes > 4 ? cbytes : 4;  // the wider of a coefficient and an index
      const int per = (int)(((i64)(e - s) * wb + 127) / 128) + 1;  // lines per run, either array
      const int tot = (int)P.w * per;
#pragma unroll 4
      for (int k2 = t; k2 < tot; k2 += 256) {
        const i64 j = k2 / per, li = k2 % per;
        const char *c0 = coef + ((i64)j * P.n + s) * cbytes, *c1 = coef + ((i64)j * P.n + e) * cbytes;
        const char *q = reinterpret_cast<const char *>(reinterpret_cast<uintptr_t>(c0) & ~(uintptr_t)127) + 128 * li;
        const char *d0 = reinterpret_cast<const char *>(P.edep + (i64)j * P.n + s);
        const char *d1 = reinterpret_cast<const char *>(P.edep + (i64)j * P.n + e);
        const char *qd = reinterpret_cast<const char *>(reinterpret_cast<uintptr_t>(d0) & ~(uintptr_t)127) + 128 * li;
        const unsigned a = q < c1 ? *reinterpret_cast<const unsigned *>(q) : 0u;
        const unsigned b = qd < d1 ? *reinterpret_cast<const unsigned *>(qd) : 0u;
        acc ^= a ^ b;
      }
    } else {
      acc ^= touch_span(coef + (i64)P.pent[s] * cbytes, coef + (i64)P.pent[e] * cbytes, t);
      acc ^= touch_span(P.edep + P.pent[s], P.edep + P.pent[e], t);
    }
    __syncthreads();
  }
  if (acc == 0x9E3779B9u) *sink = acc;  // keeps the loads; the value is never used
  if (t == 0) watch_report(gave);
}

// The levels of block b, holding PF slots of each row in registers.  With a
// progress word (ring plans), thread 0 publishes the level every
// MPK_RING_STEP levels.
template <int PF, class Tr, class Prog, class Coef, class S>
__device__ __forceinline__ void plan_levels(const Prog &prog, const Coef &coef, i64 n, const PlanView &P,
                                            const S *st, i64 b, typename Tr::V *val, typename Tr::V *pub,
                                            unsigned *progress, bool &gave) {
  const int t = (int)threadIdx.x;
  const int p0 = P.blk_pos[b];
  const int l0 = P.blk_lev[b], l1 = P.blk_lev[b + 1] - 1;  // levels l0..l1-1; lvl[l1] ends the block
  // level bounds run one level ahead of the rows' prefetch: at level l the
  // thread holds lvl[l], lvl[l+1], lvl[l+2] and fetches lvl[l+3]
  int pa = P.lvl[l0], pb = P.lvl[l0 + 1];
  int pc = l0 + 1 < l1 ? P.lvl[l0 + 2] : pb;
  PlanRow<PF, Tr, S, typename Coef::Raw> cur = plan_row<PF, Tr>(P, coef, st, pa + t, pa + t < pb);
  for (int l = l0; l < l1; l++) {
    if (progress && t == 0 && ((l - l0) & (MPK_RING_STEP - 1)) == 0) ring_publish(progress, (unsigned)(l - l0));
    const bool more = l + 1 < l1;
    const int pd = l + 2 < l1 ? P.lvl[l + 3] : pc;
    // this level's value loads, then the next level's prefetch, then the sums
    Gather<PF, Tr> g;
    if (cur.p >= 0 && P.w) gather_issue<PF, Tr>(n, val, pub, cur.d, cur.cnt, g);
    const PlanRow<PF, Tr, S, typename Coef::Raw> nxt = plan_row<PF, Tr>(P, coef, st, pb + t, more && pb + t < pc);
    if (cur.p >= 0) plan_finish<PF, Tr>(prog, coef, n, P, cur, g, val, p0, pub, gave);
    for (int p = pa + t + 256; p < pb; p += 256)
      plan_compute<PF, Tr>(prog, coef, n, P, plan_row<PF, Tr>(P, coef, st, p, true), val, p0, pub, gave);
    __syncthreads();
    cur = nxt;
    pa = pb;
    pb = pc;
    pc = pd;
  }
}

// One workgroup per block.  Each thread's first row of the next level is
// loaded while the current level computes (the barrier waits on LDS only, so
// those loads stay in flight across it); further rows of a wide level are
// loaded when they run.  Plans whose rows fit 4 slots keep 4 in registers,
// others 16 (the rest of a longer row is loaded when it runs).  ticket[0]
// orders the blocks of a block plan; ticket[1] is a ring plan's progress
// word (both zeroed before the sweep).  Tr is the value type, Coef the
// coefficient source (lc: its array, when it is not the plan's own).
template <class Prog, class Tr, class Coef, class S>
__global__ void __launch_bounds__(256) k_sweep_blk(Prog prog, i64 n, const char *plan, int second, const S *st,
                                                   const void *lc, unsigned *ticket, typename Tr::V *pub,
                                                   const int64_t *stop) {
  __shared__ typename Tr::V val[MPK_BLK_ROWS];
  __shared__ unsigned tk_sh;
  if (second) plan += reinterpret_cast<const i64 *>(plan)[4];
  const PlanView P = plan_view(plan, n);
  const Coef coef = Coef::make(P, lc);
  bool gave = false;
  if (P.mask >= 0) {  // ring plan
    const unsigned g = blockIdx.x;
    if (g == 0) {
      if (!(stop && *stop)) {
        if (P.w == 4)
          plan_levels<4, Tr>(prog, coef, n, P, st, 0, val, pub, ticket + 1, gave);
        else if (P.w > MPK_PLAN_PF)
          plan_levels<MPK_RING_WIDE, Tr>(prog, coef, n, P, st, 0, val, pub, ticket + 1, gave);
        else
          plan_levels<MPK_PLAN_PF, Tr>(prog, coef, n, P, st, 0, val, pub, ticket + 1, gave);
      }
      if (threadIdx.x == 0) ring_publish(ticket + 1, MPK_RING_DONE);
      watch_report(gave);
    } else if (g % 8 == 0 && g / 8 <= MPK_RING_PF) {
      ring_prefetch(P, st, reinterpret_cast<const char *>(coef.c), Coef::kBytes, (int)(g / 8), ticket + 1,
                    ticket + 2, stop);
    }
    return;
  }
  if (threadIdx.x == 0) tk_sh = atomicAdd(ticket, 1u);
  __syncthreads();
  if (stop && *stop) return;
  const i64 b = tk_sh;
  if (b >= P.nblk) return;
  if (P.w == 4)
    plan_levels<4, Tr>(prog, coef, n, P, st, b, val, pub, nullptr, gave);
  else
    plan_levels<MPK_PLAN_PF, Tr>(prog, coef, n, P, st, b, val, pub, nullptr, gave);
  watch_report(gave);
}

// Narrow plans (the planner marks them in hdr[7]): a single block, or a ring,
// whose levels hold at most MPK_NARROW_ROWS rows and number at most
// MPK_NARROW_LVLS, rows in w = 4 or 8 fixed slots, and no entry reading
// `pub` (every value a row reads is in LDS).  One workgroup runs every level,
// one barrier per level.  Waves 0-3 compute a level, a row per thread, from
// LDS only: the row's record (written one level ahead into one of two LDS
// buffers) and the values of earlier rows.  The other waves stage, in
// MPK_NARROW_GROUPS groups of four: group g holds the loads of the levels
// l = g (mod MPK_NARROW_GROUPS), issued that many levels before the level
// runs, and writes them to LDS the level before.  A stager wave thus keeps
// one level's loads in flight and waits only for its own; every staging load
// is unconditional (a spare lane loads a valid position it then ignores).
// Nothing is polled, so `pub` needs no sentinel fill.
#define MPK_NARROW_ROWS 256
#define MPK_NARROW_LVLS 4096
#define MPK_NARROW_GROUPS 3
#define MPK_NARROW_THREADS ((1 + MPK_NARROW_GROUPS) * MPK_NARROW_ROWS)

template <int SL>
struct NarrowStage {  // one level's rows, structure of arrays
  int cnt[MPK_NARROW_ROWS], row[MPK_NARROW_ROWS];
  double2 st0[MPK_NARROW_ROWS], st1[MPK_NARROW_ROWS];  // a Staged record as two pairs
  double c[SL][MPK_NARROW_ROWS];
  int d[SL][MPK_NARROW_ROWS];
};

template <int SL>
struct NarrowLoad {  // a stager's loads of one position
  double s[4];  // the Staged record's four words
  double c[SL];
  int d[SL];
  int cnt, row;
};

// (loaded into the fields in place: a returned or assigned record is copied
// through scratch memory)
template <int SL>
__device__ __forceinline__ void narrow_load(NarrowLoad<SL> &r, const PlanView &P, const Staged *st, int p) {
  r.cnt = P.pcnt[p];
  r.row = P.prow[p];
  // the 32-byte record as two 16-byte loads (a struct copy goes through scratch)
  const double2 *q = reinterpret_cast<const double2 *>(st + p);
  const double2 a = q[0], b = q[1];
  r.s[0] = a.x;
  r.s[1] = a.y;
  r.s[2] = b.x;
  r.s[3] = b.y;
#pragma unroll
  for (int e = 0; e < SL; e++) {
    r.c[e] = P.ecoef[(i64)e * P.n + p];
    r.d[e] = P.edep[(i64)e * P.n + p];
  }
}

template <int SL>
__device__ __forceinline__ void narrow_put(NarrowStage<SL> &b, int s, const NarrowLoad<SL> &r) {
  b.cnt[s] = r.cnt;
  b.row[s] = r.row;
  b.st0[s] = make_double2(r.s[0], r.s[1]);
  b.st1[s] = make_double2(r.s[2], r.s[3]);
#pragma unroll
  for (int e = 0; e < SL; e++) {
    b.c[e][s] = r.c[e];
    b.d[e][s] = r.d[e];
  }
}

// Position of stager lane s in level l of a narrow block (lv: the block's
// level bounds, nl levels): a valid position also past the level's end or
// past the last level, whose record is then loaded and never read.
__device__ __forceinline__ int narrow_spos(const int *lv, int nl, int s, int l) {
  const int k = l < nl ? l : nl - 1;
  const int p = lv[k] + s;
  return p < lv[k + 1] ? p : lv[k];
}

template <int SL, class Prog>
__device__ __forceinline__ void narrow_levels(const Prog &prog, i64 n, const PlanView &P, const Staged *st, double *pub,
                                              double *val, const int *lv, NarrowStage<SL> *buf) {
  constexpr int G = MPK_NARROW_GROUPS;
  const int t = (int)threadIdx.x;
  const int l0 = P.blk_lev[0], l1 = P.blk_lev[1] - 1;  // levels l0..l1-1; lvl[l1] ends the block
  const int nl = l1 - l0, p0 = P.blk_pos[0];
  // the roles run separate loops with the same barriers, on a branch that is
  // uniform per wave
  const int role = __builtin_amdgcn_readfirstlane(t) / MPK_NARROW_ROWS;  // 0: compute, 1 + g: group g
  if (role > 0) {
    const int g = role - 1, s = t - role * MPK_NARROW_ROWS;
    NarrowLoad<SL> r;
    narrow_load<SL>(r, P, st, narrow_spos(lv, nl, s, g));
    int next = g;  // the level r holds
    if (g == 0) {
      narrow_put<SL>(buf[0], s, r);
      narrow_load<SL>(r, P, st, narrow_spos(lv, nl, s, G));
      next = G;
    }
    __syncthreads();
    for (int l = 0; l < nl; l++) {
      if (next == l + 1) {  // level l + 1 into the buffer level l is not reading
        narrow_put<SL>(buf[(l + 1) & 1], s, r);
        narrow_load<SL>(r, P, st, narrow_spos(lv, nl, s, l + 1 + G));
        next = l + 1 + G;
      }
      __syncthreads();
    }
    return;
  }
  // computing waves: level l from buffer l & 1, thread t's row.  The gather
  // has no branches: an entry past the row's count reads slot 0 and is
  // dropped by a select, so a row's value reads all issue at once.
  __syncthreads();
  // level l's positions lo..hi-1, uniform; the next bound is read a level
  // ahead and taken into a scalar register only at the level's end
  int lo = __builtin_amdgcn_readfirstlane(lv[0]), hi = __builtin_amdgcn_readfirstlane(lv[1]);
  for (int l = 0; l < nl; l++) {
    const int hi_next = lv[l + 2 <= nl ? l + 2 : nl];
    if (lo + t < hi) {
      const NarrowStage<SL> &b = buf[l & 1];
      const int cnt = b.cnt[t];
      const i64 i = b.row[t];
      const double2 a0 = b.st0[t], a1 = b.st1[t];
      const Staged sr{a0.x, a0.y, a1.x, (i64)__double_as_longlong(a1.y)};
      int de[SL];
      double c[SL], v[SL];
#pragma unroll
      for (int e = 0; e < SL; e++) {  // every read in one round, then the selects
        de[e] = b.d[e][t];
        c[e] = b.c[e][t];
      }
#pragma unroll
      for (int e = 0; e < SL; e++) de[e] = e < cnt ? de[e] : 0;
#pragma unroll
      for (int e = 0; e < SL; e++) v[e] = val[de[e] >= 0 ? de[e] : 0];
#pragma unroll
      for (int e = 0; e < SL; e++)  // no -n <= de < 0 in a narrow plan
        if (de[e] < 0) v[e] = prog.stale(-(i64)de[e] - n - 1);
      double acc = row_start(sr);
#pragma unroll
      for (int e = 0; e < SL; e++) acc = e < cnt ? fma(c[e], v[e], acc) : acc;
      const double out = prog.fin(i, acc, sr);
      val[(lo + t - p0) & P.mask] = out;
      store_pub(pub, i, out);
    }
    lo = hi;
    hi = __builtin_amdgcn_readfirstlane(hi_next);
    __syncthreads();
  }
}

template <class Prog, int SL>
__global__ void __launch_bounds__(MPK_NARROW_THREADS) k_sweep_narrow(Prog prog, i64 n, const char *plan, int second,
                                                                     const Staged *st, double *pub,
                                                                     const int64_t *stop) {
  __shared__ double val[MPK_BLK_ROWS];
  __shared__ int lv[MPK_NARROW_LVLS];
  __shared__ NarrowStage<SL> buf[2];
  if (stop && *stop) return;
  if (second) plan += reinterpret_cast<const i64 *>(plan)[4];
  const PlanView P = plan_view(plan, n);
  const int l0 = P.blk_lev[0], l1 = P.blk_lev[1] - 1;
  for (int l = (int)threadIdx.x; l <= l1 - l0; l += MPK_NARROW_THREADS) lv[l] = P.lvl[l0 + l];
  __syncthreads();
  narrow_levels<SL>(prog, n, P, st, pub, val, lv, buf);
}

// Wave plans (the planner marks them hdr[7] = 2): one block (a single-block
// plan or a ring) whose levels hold at most 256 rows, every read in LDS and
// none stale.  All threads first stage every position (prog.stage into st)
// and copy the level bounds into LDS; then nw = 1, 2 or 4 waves (one per 64
// rows of the widest level) run every level, position lo + lane of a level
// on thread `lane`.  A wave's LDS operations take effect in issue order, so
// with one wave a level reads what the level before it wrote with no
// barrier; more waves meet at an LDS-only barrier after each level.  Each
// wave holds the records of its rows of the next kWaveAhead levels in
// registers, each loaded that many levels before its level runs (unrolled,
// so that no set is copied).  A level's step has no branches:
//  - every row sums E entries (E = 2 when no row of the sweep has more,
//    else the slot count); padded entries (coefficient -0.0, slot
//    MPK_WAVE_ZERO holding 0.0) leave a sum unchanged, so each row's sum is
//    the same chain of fused multiply-adds as in the other engines;
//  - a thread past its level's end takes the level's first position, and
//    the levels past the last repeat the last level: such a thread computes
//    that row again from the same values and stores the same bits to the
//    same places.
// (Loads or stores under a branch cost the compiler's count of the loads in
// flight its precision, and the prefetch with it.)
#define MPK_WAVE_LVLS 12288
#define MPK_WAVE_ZERO MPK_BLK_ROWS
#define MPK_WAVE_THREADS 512

template <int SL>
struct WaveSet {  // this lane's records of one level
  static constexpr int kBytes = (12 * SL + 8 + 15) & ~15;  // the plan's record
  double2 c[SL / 2];  // coefficients
  int4 d[SL / 4];     // LDS slots
  int row, pos;
  double2 s0, s1;  // the Staged record
};

// Loads the records of the level at lo..hi-1 for this lane.  Only the words
// used are loaded: a load whose result is partly unused leaves registers the
// compiler may reuse while the load is in flight, at the cost of a wait.
template <int SL>
__device__ __forceinline__ void wave_fetch(WaveSet<SL> &w, const PlanView &P, const Staged *st, int lo, int hi,
                                           int lane) {
  const int p = lo + lane < hi ? lo + lane : lo;
  w.pos = p;
  const char *r = P.wrec + (i64)p * WaveSet<SL>::kBytes;
  const double2 *cq = reinterpret_cast<const double2 *>(r);
  const int4 *dq = reinterpret_cast<const int4 *>(r + 8 * SL);
#pragma unroll
  for (int k = 0; k < SL / 2; k++) w.c[k] = cq[k];
#pragma unroll
  for (int k = 0; k < SL / 4; k++) w.d[k] = dq[k];
  w.row = *reinterpret_cast<const int *>(r + 12 * SL);
  const double2 *sq = reinterpret_cast<const double2 *>(st + p);
  w.s0 = sq[0];
  w.s1 = sq[1];
}

template <int SL>
__device__ __forceinline__ double wave_coef(const WaveSet<SL> &w, int e) {
  return (e & 1) ? w.c[e >> 1].y : w.c[e >> 1].x;
}
template <int SL>
__device__ __forceinline__ int wave_slot(const WaveSet<SL> &w, int e) {
  const int4 q = w.d[e >> 2];
  return (e & 3) == 0 ? q.x : (e & 3) == 1 ? q.y : (e & 3) == 2 ? q.z : q.w;
}

// Level l's bounds (levels past the last: the last level's).
__device__ __forceinline__ void wave_bounds(const int *lv, int nl, int l, int &lo, int &hi) {
  const int k = l < nl ? l : nl - 1;
  lo = lv[k];
  hi = lv[k + 1];
}

// A barrier that orders LDS only: the wait for this wave's LDS writes and
// the barrier, no wait for its loads in flight (a full __syncthreads may
// also wait for its global stores, and with them for every load).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The published words a set's entries read, loaded ahead (an entry reading
// LDS loads row 0's word and ignores it: every load is unconditional).
template <int E>
struct WavePub {
  unsigned long long b[E];
};
template <int SL, int E>
__device__ __forceinline__ void wave_pub_issue(const WaveSet<SL> &w, const double *pub, WavePub<E> &q) {
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int d = wave_slot(w, e);
    q.b[e] = load_pub(pub, d < 0 ? -(i64)d - 1 : 0);
  }
}

// Block b of a wave plan (a wave-block plan when Pub) swept by a whole
// workgroup of MPK_WAVE_THREADS threads: staging by every thread, the levels
// by waves 0..nw-1.  val (MPK_BLK_ROWS + 1 doubles) and lv (MPK_WAVE_LVLS +
// 1 ints) are the caller's LDS.  Every thread must call it; the caller's
// barrier after it orders the results (st, pub and the finish's outputs,
// all in global memory) before any thread reads them.
// With Pub, an entry may read another block's row (or a ring's old row)
// from pub: its word is loaded kPubAhead levels before its level runs, and
// a word still holding the sentinel then is polled until its row is
// published (a uniform branch the wave takes only when its producer lags);
// each row is published as soon as it is computed, since other blocks wait
// on it (without Pub the store is deferred a step, off the level chain).
template <class Prog, int SL, int E, bool Pub>
__device__ __forceinline__ void wave_sweep(const Prog &prog, i64 n, const char *plan, int second, i64 b, Staged *st,
                                           double *pub, double *val, int *lv, int nw, bool &gave) {
  // loads and stores in flight within the 63 a wave can count
  constexpr int kWaveAhead = Pub ? (SL == 4 ? 6 : 4) : (SL == 4 ? 8 : 5);
  // published words kPubAhead levels ahead: their slots come from a set
  // refetched kWaveAhead - 1 - kPubAhead steps earlier, so a deeper lead
  // shortens the records' own (6 sets ahead, 4 slots: 3 measured best)
  constexpr int kPubAhead = Pub ? (SL == 4 ? 3 : 2) : 0;
  constexpr int kStageUnroll = 8;
  if (second) plan += reinterpret_cast<const i64 *>(plan)[4];
  const PlanView P = plan_view(plan, n);
  const int l0 = P.blk_lev[b], nl = P.blk_lev[b + 1] - 1 - l0;  // levels l0..; lvl[l0 + nl] ends the block
  const int p0 = P.blk_pos[b], p1 = P.blk_pos[b + 1];
  // Staging: every position's Staged record, kStageUnroll of them a thread
  // at a time (their loads all issue before the first store).  With one
  // sweeping wave and Pub, the whole workgroup stages only the first kStage0
  // positions before the sweep starts; the other waves stage the rest in
  // position (level) order while it runs, each publishing in LDS the rounds
  // it has finished (after a release fence), and the sweeping wave waits
  // for a round only if it gets there first.
  constexpr int kHelpers = MPK_WAVE_THREADS / 64 - 1;  // waves
  constexpr int kStage0 = kStageUnroll * MPK_WAVE_THREADS;
  constexpr int kRound = kStageUnroll * 64 * kHelpers;  // positions a helper round stages
  __shared__ int stage_rounds[kHelpers];
  const bool helpers = Pub && nw == 1;
  const int t = (int)threadIdx.x;
  const int s0 = helpers && p1 - p0 > kStage0 ? p0 + kStage0 : p1;  // staged before the sweep
  auto stage_span = [=](int q0, int q1, int th, int nth) {
    for (int q = q0 + th; q < q1; q += kStageUnroll * nth) {
      Staged tmp[kStageUnroll];
#pragma unroll
      for (int u = 0; u < kStageUnroll; u++) {
        const int p = q + u * nth;
        if (p < q1) tmp[u] = prog.stage(P.prow[p]);
      }
#pragma unroll
      for (int u = 0; u < kStageUnroll; u++) {
        const int p = q + u * nth;
        if (p < q1) st[p] = tmp[u];
      }
    }
  };
#pragma unroll 4
  for (int l = t; l <= nl; l += MPK_WAVE_THREADS) lv[l] = P.lvl[l0 + l];
  stage_span(p0, s0, t, MPK_WAVE_THREADS);
  if (t < kHelpers) stage_rounds[t] = 0;
  if (t == 0) val[MPK_WAVE_ZERO] = 0.0;
  __syncthreads();  // st and lv (workgroup scope: the stores to st are seen)
  if (t >= 64 * nw) {
    if (helpers) {
      const int h = (t >> 6) - 1, th = t - 64;
      int r = 0;
      for (int q = s0; q < p1; q += kRound) {
        stage_span(q, q + kRound < p1 ? q + kRound : p1, th, 64 * kHelpers);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // this wave's stores, seen in the workgroup
        r++;
        if ((t & 63) == 0) __hip_atomic_store(&stage_rounds[h], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      return;
    }
    // the other waves meet the level barriers too: a barrier counts every
    // wave of the workgroup still running, and the caller's barrier after
    // the sweep must not pair with them
    if (nw > 1)
      for (int l = 0; l < nl; l += kWaveAhead)
        for (int k = 0; k < kWaveAhead; k++) lds_barrier();
    return;
  }
  // positions below `staged` are staged (uniform); waits until `upto` is
  int staged = s0;
  auto wait_staged = [&](int upto) {
    for (;;) {
      int m = 1 << 30;
#pragma unroll
      for (int h = 0; h < kHelpers; h++) {
        const int v = __hip_atomic_load(&stage_rounds[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        m = v < m ? v : m;
      }
      const i64 s = (i64)s0 + (i64)m * kRound;
      staged = s < p1 ? (int)s : p1;
      if (staged >= upto) break;
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  };
  // the first sets in level order: the compiler may not reorder their loads
  // (its waits in the loop are placed for loads issued in that order)
  WaveSet<SL> w[kWaveAhead];
  WavePub<E> q[kWaveAhead];
#pragma unroll
  for (int k = 0; k < kWaveAhead; k++) {
    int lo, hi;
    wave_bounds(lv, nl, k, lo, hi);
    wave_fetch<SL>(w[k], P, st, lo, hi, t);
    __builtin_amdgcn_sched_barrier(0);
  }
  if (Pub) {
#pragma unroll
    for (int k = 0; k < kPubAhead; k++) wave_pub_issue<SL, E>(w[k], pub, q[k]);
  }
  // A step: the value reads; while they are in flight, the step before's
  // deferred work (its published value, the refetch of its set) and the
  // published words of the set kPubAhead levels on; then the chain, the
  // finish and the LDS write the next step reads.  The first step refetches
  // the last set's own level (the same values).
  double pend_out = 0.0;
  int pend_row = 0;
  int flo, fhi;
  wave_bounds(lv, nl, kWaveAhead - 1, flo, fhi);
  for (int l = 0; l < nl; l += kWaveAhead) {
#pragma unroll
    for (int k = 0; k < kWaveAhead; k++) {
      const WaveSet<SL> &cur = w[k];
      WaveSet<SL> &prev = w[(k + kWaveAhead - 1) % kWaveAhead];
      int lo, hi;  // the level this set is refetched for, read with its values
      wave_bounds(lv, nl, l + k + kWaveAhead, lo, hi);
      double v[E];
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int d = wave_slot(cur, e);
        v[e] = val[Pub && d < 0 ? MPK_WAVE_ZERO : d];
      }
      __builtin_amdgcn_sched_barrier(0);
      if (!Pub && l + k > 0) store_pub(pub, pend_row, pend_out);  // uniform
      if (helpers && fhi > staged) wait_staged(fhi);                // uniform, rare
      wave_fetch<SL>(prev, P, st, flo, fhi, t);
      if (Pub) {
        // the set kPubAhead levels on (a level past the last reloads the
        // last level's words: the same, already published)
        const int j = (k + kPubAhead) % kWaveAhead;
        wave_pub_issue<SL, E>(w[j], pub, q[j]);
      }
      __builtin_amdgcn_sched_barrier(0);
      if (Pub) {
        WavePub<E> &qc = q[k];
        bool lag = false;
#pragma unroll
        for (int e = 0; e < E; e++) lag |= wave_slot(cur, e) < 0 && qc.b[e] == MPK_SENTINEL;
        if (__ballot(lag)) {  // uniform; a producer still behind
#pragma unroll
          for (int e = 0; e < E; e++) {
            const int d = wave_slot(cur, e);
            if (d < 0) {
              unsigned polls = 0;
              while (qc.b[e] == MPK_SENTINEL) {
                if (watch_expired(polls)) {
                  gave = true;
                  qc.b[e] = MPK_QNAN;
                  break;
                }
                qc.b[e] = load_pub(pub, -(i64)d - 1);
              }
            }
          }
          // the next sets' words were loaded before this one was
          // published, so likely before theirs were too: load them again
#pragma unroll
          for (int j = 1; j <= kPubAhead; j++)
            wave_pub_issue<SL, E>(w[(k + j) % kWaveAhead], pub, q[(k + j) % kWaveAhead]);
        }
#pragma unroll
        for (int e = 0; e < E; e++)
          if (wave_slot(cur, e) < 0) v[e] = __longlong_as_double((long long)qc.b[e]);
      }
      double acc = cur.s0.x;  // row_start
#pragma unroll
      for (int e = 0; e < E; e++) acc = fma(wave_coef(cur, e), v[e], acc);
      const Staged sr{cur.s0.x, cur.s0.y, cur.s1.x, (i64)__double_as_longlong(cur.s1.y)};
      const double out = prog.fin(cur.row, acc, sr);
      val[(cur.pos - p0) & P.mask] = out;
      if (Pub) store_pub(pub, cur.row, out);  // at once: other blocks wait on it
      pend_out = out;
      pend_row = cur.row;
      flo = lo;
      fhi = hi;
      if (nw > 1) lds_barrier();  // uniform
    }
  }
  if (!Pub) store_pub(pub, pend_row, pend_out);
}

// A wave plan's sweep as its own launch (one workgroup).
template <class Prog, int SL, int E>
__global__ void __launch_bounds__(MPK_WAVE_THREADS) k_sweep_wave(Prog prog, i64 n, const char *plan, int second,
                                                                 Staged *st, double *pub, int nw,
                                                                 const int64_t *stop) {
  __shared__ double val[MPK_BLK_ROWS + 1];
  __shared__ int lv[MPK_WAVE_LVLS + 1];
  if (stop && *stop) return;
  bool gave = false;  // no waits without published reads
  wave_sweep<Prog, SL, E, false>(prog, n, plan, second, 0, st, pub, val, lv, nw, gave);
}

// A wave-block plan's sweep: one workgroup per block, blocks taken in
// sweep order by ticket (so the oldest unfinished block reads only finished
// ones), pub sentinel-filled and the ticket zeroed before the launch.
template <class Prog, int SL, int E>
__global__ void __launch_bounds__(MPK_WAVE_THREADS) k_sweep_wblk(Prog prog, i64 n, const char *plan, int second,
                                                                 Staged *st, double *pub, unsigned *ticket, int nw,
                                                                 const int64_t *stop) {
  __shared__ double val[MPK_BLK_ROWS + 1];
  __shared__ int lv[MPK_WAVE_LVLS + 1];
  __shared__ unsigned tk_sh;
  if (threadIdx.x == 0) tk_sh = atomicAdd(ticket, 1u);
  __syncthreads();
  if (stop && *stop) return;
  const i64 nblk = reinterpret_cast<const i64 *>(second ? plan + reinterpret_cast<const i64 *>(plan)[4] : plan)[0];
  if ((i64)tk_sh >= nblk) return;
  bool gave = false;
  wave_sweep<Prog, SL, E, true>(prog, n, plan, second, tk_sh, st, pub, val, lv, nw, gave);
  watch_report(gave);
}

// A wave plan's kernel form from the registry's word (slot count in bits
// 0-3, 16: a wave plan, 32: no row over 2 entries, bits 6-7: log2 of the
// waves, 256: a wave-block plan): 0 (4 slots, 2 entries), 1 (4 slots), 2 (8
// slots) plus 4 times log2 of the waves; -1 for another plan, wave blocks
// included (one workgroup cannot run them).
__host__ __device__ inline int wave_form(int byte) {
  if (!(byte & 16) || (byte & 256)) return -1;
  return ((byte & 15) == 4 ? ((byte & 32) ? 0 : 1) : 2) + 4 * ((byte >> 6) & 3);
}
// The sweep in its form, by a whole workgroup (wave_sweep).
template <class Prog>
__device__ __forceinline__ void wave_sweep_form(int form, const Prog &prog, i64 n, const char *plan, int second,
                                                Staged *st, double *pub, double *val, int *lv) {
  const int nw = 1 << (form >> 2);
  bool gave = false;  // no waits without published reads
  if ((form & 3) == 0)
    wave_sweep<Prog, 4, 2, false>(prog, n, plan, second, 0, st, pub, val, lv, nw, gave);
  else if ((form & 3) == 1)
    wave_sweep<Prog, 4, 4, false>(prog, n, plan, second, 0, st, pub, val, lv, nw, gave);
  else
    wave_sweep<Prog, 8, 8, false>(prog, n, plan, second, 0, st, pub, val, lv, nw, gave);
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
