// mpk_plan.hip -- the block plan of a triangular or SOR sweep, built on the
// device.  The plan is byte for byte the host planner's (mpk_sweep_plan,
// mpk_factor.cpp, whose comments and mpk_sweep.h's document the layout):
// the same blocks, the same local levels, the same positions, entries, wave
// records and header.  The host planner walks the rows one by one; here
//   - the sweep's global depth (for the ring test) comes from a sync-free
//     level sweep: one wave per 64 consecutive sweep positions, taken in
//     sweep order by ticket, each row's level the largest of its earlier
//     dependencies' plus one, published for later waves (k_plan_glev);
//   - each block's local levels come from one wave that walks the block 64
//     rows at a time, every level kept in LDS (k_plan_local);
//   - each block's positions (rows by level, sweep order within a level)
//     from a counting sort by the same wave (k_plan_pos);
//   - entries, dependencies and wave records one position per thread
//     (k_plan_entries, k_plan_records).
// Ring plans likewise: the latest levels from the same level sweep run in
// reverse sweep order over the dependencies transposed (each row's readers,
// k_plan_tcount / k_plan_tfill); the positions from a stable radix sort of
// the rows by level (sweep order within a level); the window test and the
// entries per row / position.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "mpk_internal.h"
#include "mpk_pc.h"
#include "mpk_sweep.h"

namespace mpk {

__device__ __forceinline__ i64 plan_spos(i64 i, i64 n, int forward) { return forward ? i : n - 1 - i; }

// Global levels of a sweep's DAG (earliest levels: lev[i] = 1 + the largest
// level of an earlier dependency, 0 without one), lev sentinel-filled with
// -1 and the ticket zeroed before the launch; maxlev[0] gets the deepest
// level.  Waits for other waves' rows are bounded (watch_expired); a wave
// that gave up counts the dependency as level 0 and reports it.
__global__ void __launch_bounds__(64) k_plan_glev(i64 n, int forward, const i64 *off, const int *col, int *lev,
                                                  unsigned *ticket, int *maxlev) {
  __shared__ unsigned blk_sh;
  __shared__ int lv[64];
  if (threadIdx.x == 0) blk_sh = atomicAdd(ticket, 1u);
  __syncthreads();
  const int lane = (int)threadIdx.x;
  const i64 L0 = (i64)blk_sh * 64, s = L0 + lane;
  const bool valid = s < n;
  const i64 i = valid ? plan_spos(s, n, forward) : 0;
  bool done = !valid;
  i64 j = valid ? off[i] : 0;
  const i64 je = valid ? off[i + 1] : 0;
  int l = 0;
  unsigned long long mask = 0;
  unsigned polls = 0;
  bool gave = false;
  for (;;) {
    bool moved = false;
    if (!done) {
      const i64 j_in = j;
      while (j < je) {
        const i64 k = col[j], sk = plan_spos(k, n, forward);
        if (sk < s) {  // an earlier row: a dependency
          int v;
          if (sk >= L0) {
            if (!((mask >> (sk - L0)) & 1ull)) {
              if (!gave) break;
              v = 0;
            } else {
              v = lv[sk - L0];
            }
          } else {
            v = __hip_atomic_load(lev + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (v < 0) {
              if (!gave) break;
              v = 0;
            }
          }
          l = v + 1 > l ? v + 1 : l;
        }
        j++;
      }
      if (j >= je) {
        lv[lane] = l;
        __hip_atomic_store(lev + i, l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        done = true;
      }
      moved = j != j_in || done;
    }
    __syncthreads();  // one wave: orders the LDS writes before the next reads
    const unsigned long long nm = __ballot(done);
    if (nm == ~0ull) break;
    if (nm == mask && __ballot(moved) == 0ull) {
      __builtin_amdgcn_s_sleep(2);
      if (watch_expired(polls)) gave = true;  // uniform
    }
    mask = nm;
  }
  int mx = valid ? l : -1;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const int o = __shfl_down(mx, w, 64);
    mx = o > mx ? o : mx;
  }
  if (lane == 0) {
    atomicMax(maxlev, mx);
    watch_report(gave);
  }
}

// Local levels of block b (sweep positions b B .. min(n, (b + 1) B) - 1): a
// row's level is 1 + the largest level of a dependency earlier in the sweep
// and in the same block, 0 without one.  One wave per block takes its rows
// 64 at a time (a row's dependency in an earlier group is done; in its own
// group the wave repeats until it is); lev gets every row's level, nlev[b]
// the block's level count.
__global__ void __launch_bounds__(64) k_plan_local(i64 n, int forward, i64 B, const i64 *off, const int *col,
                                                   int *lev, int *nlev) {
  __shared__ int llev[MPK_BLK_ROWS];
  const int lane = (int)threadIdx.x;
  const i64 b = blockIdx.x, s0 = b * B, s1 = s0 + B < n ? s0 + B : n;
  int mx = -1;
  for (i64 c0 = s0; c0 < s1; c0 += 64) {
    const i64 s = c0 + lane;
    const bool valid = s < s1;
    const i64 i = valid ? plan_spos(s, n, forward) : 0;
    bool done = !valid;
    i64 j = valid ? off[i] : 0;
    const i64 je = valid ? off[i + 1] : 0;
    int l = 0;
    unsigned long long mask = 0;
    for (;;) {
      if (!done) {
        while (j < je) {
          const i64 sk = plan_spos(col[j], n, forward);
          if (sk >= s0 && sk < s) {  // earlier in the sweep and in this block
            if (sk >= c0 && !((mask >> (sk - c0)) & 1ull)) break;
            const int v = llev[sk - s0];
            l = v + 1 > l ? v + 1 : l;
          }
          j++;
        }
        if (j >= je) {
          llev[s - s0] = l;
          done = true;
        }
      }
      __syncthreads();  // one wave: orders the LDS writes before the next reads
      const unsigned long long nm = __ballot(done);
      if (nm == ~0ull) break;
      mask = nm;
    }
    if (valid) {
      lev[i] = l;
      mx = l > mx ? l : mx;
    }
  }
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const int o = __shfl_down(mx, w, 64);
    mx = o > mx ? o : mx;
  }
  if (lane == 0) nlev[b] = mx + 1;
}

// Block b's level starts and positions: the rows sorted by local level, in
// sweep order within a level (a counting sort, stable, 64 rows at a time by
// one wave).  Writes blk_pos[b], blk_lev[b], the block's lvl entries, prow
// and pos (row -> position), and wlev[b] = its widest level.  li[b] is the
// block's first entry in lvl; block 0 also ends blk_pos and blk_lev.
__global__ void __launch_bounds__(64) k_plan_pos(i64 n, int forward, i64 B, i64 nblk, i64 nlvl, const int *lev,
                                                 const i64 *li, const int *nlev, int *blk_pos, int *blk_lev, int *lvl,
                                                 int *prow, int *pos, int *wlev) {
  __shared__ int cnt[MPK_BLK_ROWS + 1];
  const int lane = (int)threadIdx.x;
  const i64 b = blockIdx.x, s0 = b * B, s1 = s0 + B < n ? s0 + B : n;
  const int L = nlev[b];
  for (int l = lane; l <= L; l += 64) cnt[l] = 0;
  __syncthreads();
  for (i64 s = s0 + lane; s < s1; s += 64) atomicAdd(&cnt[lev[plan_spos(s, n, forward)] + 1], 1);
  __syncthreads();
  // cnt[l] = the rows at levels below l (a running prefix over 64 at a time)
  int carry = 0, widest = 0;
  for (int l0 = 0; l0 <= L; l0 += 64) {
    const int l = l0 + lane;
    const int c = l <= L ? cnt[l] : 0;
    int x = c;
#pragma unroll
    for (int w = 1; w < 64; w <<= 1) {
      const int y = __shfl_up(x, w, 64);
      if (lane >= w) x += y;
    }
    if (l <= L) cnt[l] = carry + x;
    if (l >= 1 && l <= L) widest = c > widest ? c : widest;  // rows at level l - 1
    carry += __shfl(x, 63, 64);
  }
  __syncthreads();
  const i64 base = li[b];
  for (int l = lane; l < L; l += 64) lvl[base + l] = (int)(s0 + cnt[l]);
  if (lane == 0) {
    lvl[base + L] = (int)s1;
    blk_pos[b] = (int)s0;
    blk_lev[b] = (int)base;
    if (b == 0) {
      blk_pos[nblk] = (int)n;
      blk_lev[nblk] = (int)nlvl;
    }
  }
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const int o = __shfl_down(widest, w, 64);
    widest = o > widest ? o : widest;
  }
  if (lane == 0) wlev[b] = widest;
  __syncthreads();
  // positions: cnt[l] is the next free position of level l
  for (i64 c0 = s0; c0 < s1; c0 += 64) {
    const i64 s = c0 + lane;
    const bool valid = s < s1;
    const i64 i = valid ? plan_spos(s, n, forward) : 0;
    const int l = valid ? lev[i] : -1;
    int rank = 0, tot = 0;
    for (int q = 0; q < 64; q++) {
      const int lq = __shfl(l, q, 64);
      if (lq == l) {
        tot++;
        if (q < lane) rank++;
      }
    }
    const int at = valid ? cnt[l] : 0;  // every lane reads before any lane writes (one wave, in order)
    __syncthreads();
    if (valid && rank == 0) cnt[l] = at + tot;
    __syncthreads();
    if (valid) {
      const i64 p = s0 + at + rank;
      prow[p] = (int)i;
      pos[i] = (int)p;
    }
  }
}

// Each position's count, entry offset (CSR form) and entries: coefficient
// and dependency slot (d >= 0: the LDS slot of a row of the same block;
// -n <= d < 0: row -d - 1 of an earlier block; d < -n: row -d - n - 1, not
// earlier in the sweep).  flags: 1 when an entry reads an earlier block's
// row, 2 when one reads a row not earlier in the sweep.
__global__ void k_plan_entries(i64 n, int forward, i64 B, i64 w, const i64 *off, const int *col, const double *coef,
                               const int *prow, const int *pos, const int *blk_pos, const i64 *eout, int *pcnt,
                               int *pent, double *ecoef, int *edep, unsigned *flags) {
  const i64 p = (i64)blockIdx.x * 256 + threadIdx.x;
  unsigned f = 0;
  if (p < n) {
    const i64 i = prow[p], si = plan_spos(i, n, forward), b = si / B, o = off[i], cnt = off[i + 1] - o;
    pcnt[p] = (int)cnt;
    pent[p] = w ? 0 : (int)eout[p];
    const i64 q0 = w ? p : eout[p], qs = w ? n : 1;
    for (i64 e = 0; e < cnt; e++) {
      const i64 k = col[o + e], sk = plan_spos(k, n, forward);
      const i64 q = q0 + e * qs;
      ecoef[q] = coef[o + e];
      int d;
      if (sk >= si) {
        d = (int)(-k - n - 1);
        f |= 2u;
      } else if (sk / B == b) {
        d = pos[k] - blk_pos[b];
      } else {
        d = (int)(-k - 1);
        f |= 1u;
      }
      edep[q] = d;
    }
  }
  const unsigned long long any1 = __ballot(f & 1u), any2 = __ballot(f & 2u);
  if ((threadIdx.x & 63) == 0 && (any1 | any2)) atomicOr(flags, (any1 ? 1u : 0u) | (any2 ? 2u : 0u));
}

// The wave records appended to a plan of w (4 or 8) slots: per position its
// w coefficients (-0.0 past its count), w dependency slots (the zero slot
// MPK_WAVE_ZERO past its count), its row and its count, in rb bytes.
__global__ void k_plan_records(i64 n, i64 w, i64 rb, const double *ecoef, const int *edep, const int *prow,
                               const int *pcnt, char *rec) {
  const i64 p = (i64)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  char *r = rec + p * rb;
  double *c = reinterpret_cast<double *>(r);
  int *d = reinterpret_cast<int *>(c + w);
  const int cnt = pcnt[p];
  for (i64 e = 0; e < w; e++) {
    const bool used = e < cnt;
    c[e] = used ? ecoef[e * n + p] : -0.0;
    d[e] = used ? edep[e * n + p] : MPK_WAVE_ZERO;
  }
  d[w] = prow[p];
  d[w + 1] = cnt;
}

// The positions' entry counts in position order (for the CSR form's offsets).
__global__ void k_plan_counts(i64 n, const i64 *off, const int *prow, int *c) {
  const i64 p = (i64)blockIdx.x * 256 + threadIdx.x;
  if (p < n) c[p] = (int)(off[prow[p] + 1] - off[prow[p]]);
}

// The readers of each row, for the latest levels: tcnt[k] += 1 for every
// dependency k of every row (tcnt zeroed before, n + 1 long so that its
// exclusive scan ends with the total), then tcol lists row k's readers from
// toff[k] on (cursor a copy of toff; the order within a row does not matter,
// only the largest level is taken).
__global__ void k_plan_tcount(i64 n, int forward, const i64 *off, const int *col, unsigned long long *tcnt) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const i64 si = plan_spos(i, n, forward);
  for (i64 j = off[i]; j < off[i + 1]; j++)
    if (plan_spos(col[j], n, forward) < si) atomicAdd(tcnt + col[j], 1ull);
}
__global__ void k_plan_tfill(i64 n, int forward, const i64 *off, const int *col, unsigned long long *cursor,
                             int *tcol) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const i64 si = plan_spos(i, n, forward);
  for (i64 j = off[i]; j < off[i + 1]; j++)
    if (plan_spos(col[j], n, forward) < si) tcol[atomicAdd(cursor + col[j], 1ull)] = (int)i;
}

// Latest levels from the reverse sweep's earliest levels: a row's latest
// level is nl - 1 less the longest chain of readers after it, which is its
// earliest level in the reverse sweep over the readers.
__global__ void k_plan_late_out(i64 n, int nl1, const int *rlev, int *late) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) late[i] = nl1 - rlev[i];
}

// Sort inputs: each sweep position's row and its latest level.
__global__ void k_plan_keys(i64 n, int forward, const int *late, int *key, int *row) {
  const i64 s = (i64)blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const i64 i = plan_spos(s, n, forward);
  key[s] = late[i];
  row[s] = (int)i;
}

// After the sort (rows by level, sweep order within one): pos[row] = p, and
// the level starts lvl[l] (l <= nl; lvl[nl] = n) from the sorted levels.
__global__ void k_plan_ring_pos(i64 n, i64 nl, const int *skey, const int *prow, int *pos, int *lvl) {
  const i64 p = (i64)blockIdx.x * 256 + threadIdx.x;
  if (p > n) return;
  const int lo = p == 0 ? -1 : skey[p - 1], hi = p == n ? (int)nl : skey[p];
  for (int l = lo + 1; l <= hi; l++) lvl[l] = (int)p;  // levels starting here (every level if none is empty)
  if (p < n) pos[prow[p]] = (int)p;
}

// Entries that read a row W or more positions before the end of the
// reader's level ("old": their ring slot may have been rewritten).
__global__ void k_plan_old(i64 n, int forward, i64 W, const i64 *off, const int *col, const int *late,
                           const int *pos, const int *lvl, unsigned long long *old) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  unsigned long long c = 0;
  if (i < n) {
    const i64 si = plan_spos(i, n, forward), end = lvl[late[i] + 1];
    for (i64 j = off[i]; j < off[i + 1]; j++) {
      const int k = col[j];
      if (plan_spos(k, n, forward) < si && !(end - pos[k] <= W)) c++;
    }
  }
  if (c) atomicAdd(old, c);
}

// A ring plan's entries, per position: as k_plan_entries, the dependency
// slot being the LDS ring slot (pos mod W) of a row less than W positions
// behind the end of the reader's level, else the published row.
__global__ void k_plan_ring_entries(i64 n, int forward, i64 W, i64 w, const i64 *off, const int *col,
                                    const double *coef, const int *prow, const int *pos, const int *late,
                                    const int *lvl, const i64 *eout, int *pcnt, int *pent, double *ecoef, int *edep,
                                    unsigned *flags) {
  const i64 p = (i64)blockIdx.x * 256 + threadIdx.x;
  unsigned f = 0;
  if (p < n) {
    const i64 i = prow[p], si = plan_spos(i, n, forward), o = off[i], cnt = off[i + 1] - o, end = lvl[late[i] + 1];
    pcnt[p] = (int)cnt;
    pent[p] = w ? 0 : (int)eout[p];
    const i64 q0 = w ? p : eout[p], qs = w ? n : 1;
    for (i64 e = 0; e < cnt; e++) {
      const i64 k = col[o + e];
      const i64 q = q0 + e * qs;
      ecoef[q] = coef[o + e];
      int d;
      if (plan_spos(k, n, forward) >= si) {
        d = (int)(-k - n - 1);
        f |= 2u;
      } else if (end - pos[k] <= W) {
        d = (int)(pos[k] & (W - 1));
      } else {
        d = (int)(-k - 1);
        f |= 1u;
      }
      edep[q] = d;
    }
  }
  const unsigned long long any1 = __ballot(f & 1u), any2 = __ballot(f & 2u);
  if ((threadIdx.x & 63) == 0 && (any1 | any2)) atomicOr(flags, (any1 ? 1u : 0u) | (any2 ? 2u : 0u));
}

// Device scratch of the planner, grown on demand and kept.
struct PlanScratch {
  int *lev = nullptr, *pos = nullptr, *nlev = nullptr, *wlev = nullptr, *maxlev = nullptr, *cnt = nullptr;
  int *late = nullptr, *key = nullptr, *skey = nullptr, *row = nullptr;
  i64 *li = nullptr, *eout = nullptr;
  unsigned *flags = nullptr;
  unsigned long long *old = nullptr, *lw = nullptr;
  i64 *toff = nullptr;
  int *tcol = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_cap = 0;
  i64 cap_n = 0, cap_b = 0, cap_t = 0;
  bool ensure(i64 n, i64 nblk) {
    if (n > cap_n) {
      for (void *q : {(void *)lev, (void *)pos, (void *)eout, (void *)cnt, (void *)late, (void *)key, (void *)skey,
                      (void *)row, (void *)lw, (void *)toff})
        if (q) HIPCHK(hipFree(q));
      HIPCHK(hipMalloc(&lw, (n + 1) * sizeof(unsigned long long)));
      HIPCHK(hipMalloc(&toff, (n + 1) * sizeof(i64)));
      HIPCHK(hipMalloc(&lev, n * sizeof(int)));
      HIPCHK(hipMalloc(&pos, n * sizeof(int)));
      HIPCHK(hipMalloc(&cnt, n * sizeof(int)));
      HIPCHK(hipMalloc(&late, n * sizeof(int)));
      HIPCHK(hipMalloc(&key, n * sizeof(int)));
      HIPCHK(hipMalloc(&skey, n * sizeof(int)));
      HIPCHK(hipMalloc(&row, n * sizeof(int)));
      HIPCHK(hipMalloc(&eout, (n + 1) * sizeof(i64)));
      cap_n = n;
    }
    if (nblk + 1 > cap_b) {
      for (void *q : {(void *)nlev, (void *)wlev, (void *)li})
        if (q) HIPCHK(hipFree(q));
      HIPCHK(hipMalloc(&nlev, (nblk + 1) * sizeof(int)));
      HIPCHK(hipMalloc(&wlev, (nblk + 1) * sizeof(int)));
      HIPCHK(hipMalloc(&li, (nblk + 1) * sizeof(i64)));
      cap_b = nblk + 1;
    }
    if (!maxlev) {
      HIPCHK(hipMalloc(&maxlev, 4 * sizeof(int)));
      HIPCHK(hipMalloc(&flags, 4 * sizeof(unsigned)));
      HIPCHK(hipMalloc(&old, sizeof(unsigned long long)));
    }
    return !failed();
  }
  bool ensure_t(i64 n, i64 m) {  // after ensure(n, ...)
    (void)n;
    if (m > cap_t) {
      if (tcol) HIPCHK(hipFree(tcol));
      HIPCHK(hipMalloc(&tcol, (m > 0 ? m : 1) * sizeof(int)));
      cap_t = m;
    }
    return !failed();
  }
  bool ensure_sort(size_t bytes) {
    if (bytes <= sort_cap && sort_tmp) return true;
    if (sort_tmp) HIPCHK(hipFree(sort_tmp));
    HIPCHK(hipMalloc(&sort_tmp, bytes > 0 ? bytes : 16));
    sort_cap = bytes;
    return !failed();
  }
};

// The CSR form's entry offsets in position order (w = 0): eout [D] and
// pent[n] (the entry count).
static bool plan_eout(i64 n, const i64 *off, const int *prow, int *pent, PlanScratch &S) {
  MPK_LAUNCH(k_plan_counts, nblocks(n, 256), 256, n, off, prow, S.cnt);
  std::vector<int> pc(n);
  copy_dtoh(pc.data(), S.cnt, n * sizeof(int));
  std::vector<i64> eo(n + 1, 0);
  for (i64 p = 0; p < n; p++) eo[p + 1] = eo[p] + pc[p];
  copy_htod(S.eout, eo.data(), (n + 1) * sizeof(i64));
  const int last = (int)eo[n];
  copy_htod(pent + n, &last, sizeof last);
  return !failed();
}

// mark_plan's choice for a plan of nblk blocks, w slots, its deepest
// block's levels and widest level: the wave records (appended when the
// wave form takes it, written by k_plan_records) and hdr[7], hdr[4].
static i64 plan_mark(i64 n, i64 nblk, i64 w, i64 levels, int widest, bool published, bool stale, i64 bytes,
                     i64 cap, const double *ecoef, const int *edep, const int *prow, const int *pcnt, char *dst,
                     int64_t *hdr) {
  const bool wave_w = w == 4 || w == 8;
  if (plan_env_off("MPK_SWEEP_NARROW") || nblk < 1 || !wave_w) return bytes;
  const i64 rb = wave_rec_bytes_rule(w), wbytes = bytes + n * rb;
  const bool narrow = nblk == 1 && !published && widest <= 256 && levels < 4096;
  const bool wave = widest <= 256 && levels <= 12288 && !stale && wbytes <= cap && !plan_env_off("MPK_SWEEP_WAVE");
  if (!wave) {
    hdr[7] = narrow ? 1 : 0;
    return bytes;
  }
  MPK_LAUNCH(k_plan_records, nblocks(n, 256), 256, n, w, rb, ecoef, edep, prow, pcnt, dst + bytes);
  hdr[4] = wbytes;
  hdr[7] = nblk > 1 || published ? 3 : 2;
  return wbytes;
}

// The ring plan of a sweep of depth nl (earliest levels already in S.lev),
// or 0 when the host planner's test sends the sweep to blocks instead.
static i64 device_ring_plan(const DevEntries &E, bool forward, i64 bmax, i64 nl, int mode, char *dst, i64 cap,
                            i64 *depth_out, int *widest_level, int *mark, PlanScratch &S) {
  Ctx *c = ctx();
  const i64 n = E.n, m = E.m;
  const int fw = forward ? 1 : 0;
  const i64 w = plan_w_rule(E.widest, n, m), slots = w ? w * n : m;
  i64 W = 1;
  while (2 * W <= bmax) W *= 2;
  // latest levels: the readers of each row (the dependencies transposed),
  // then the reverse sweep's earliest levels over them
  if (!S.ensure_t(n, m)) return -1;
  HIPCHK(hipMemsetAsync(S.lw, 0, (n + 1) * sizeof(unsigned long long), stream()));
  MPK_LAUNCH(k_plan_tcount, nblocks(n, 256), 256, n, fw, E.off, E.col, S.lw);
  size_t sb = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, S.lw, S.toff, (int)(n + 1), stream()));
  if (!S.ensure_sort(sb)) return -1;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(S.sort_tmp, sb, S.lw, S.toff, (int)(n + 1), stream()));
  HIPCHK(hipMemcpyAsync(S.lw, S.toff, n * sizeof(i64), hipMemcpyDeviceToDevice, stream()));
  MPK_LAUNCH(k_plan_tfill, nblocks(n, 256), 256, n, fw, E.off, E.col, S.lw, S.tcol);
  HIPCHK(hipMemsetAsync(S.lev, 0xff, n * sizeof(int), stream()));
  HIPCHK(hipMemsetAsync(S.maxlev + 1, 0xff, sizeof(int), stream()));
  HIPCHK(hipMemsetAsync(c->tickets, 0, sizeof(unsigned), stream()));
  MPK_LAUNCH(k_plan_glev, nblocks(n, 64), 64, n, 1 - fw, (const i64 *)S.toff, (const int *)S.tcol, S.lev,
             c->tickets, S.maxlev + 1);
  MPK_LAUNCH(k_plan_late_out, nblocks(n, 256), 256, n, (int)(nl - 1), (const int *)S.lev, S.late);
  // positions: rows by latest level, sweep order within a level
  MPK_LAUNCH(k_plan_keys, nblocks(n, 256), 256, n, fw, (const int *)S.late, S.key, S.row);
  const i64 nlvl = nl + 1, bytes = plan_bytes_rule(n, 1, nlvl, slots);
  if (bytes > cap) {
    set_error(5, "sweep plan buffer too small");
    return -1;
  }
  int bits = 1;
  while (((i64)1 << bits) <= nl) bits++;
  int *blk_pos = reinterpret_cast<int *>(dst + 64), *blk_lev = blk_pos + 2, *lvl = blk_lev + 2;
  int *prow = lvl + nlvl, *pcnt = prow + n, *pent = pcnt + n;
  size_t tb = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, S.key, S.skey, S.row, S.pos, (int)n, 0, bits, stream()));
  if (!S.ensure_sort(tb)) return -1;
  // the sorted rows go to the plan's prow once the buffer is cleared; sort
  // into scratch first (S.pos, rewritten below as row -> position)
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(S.sort_tmp, tb, S.key, S.skey, S.row, S.lev, (int)n, 0, bits, stream()));
  const i64 rb = wave_rec_bytes_rule(w), wbytes = bytes + ((w == 4 || w == 8) ? n * rb : 0);
  HIPCHK(hipMemsetAsync(dst, 0, (size_t)std::min(cap, wbytes), stream()));
  HIPCHK(hipMemcpyAsync(prow, S.lev, n * sizeof(int), hipMemcpyDeviceToDevice, stream()));
  MPK_LAUNCH(k_plan_ring_pos, nblocks(n + 1, 256), 256, n, nl, (const int *)S.skey, (const int *)prow, S.pos, lvl);
  // the host planner's test: few enough old reads
  HIPCHK(hipMemsetAsync(S.old, 0, sizeof(unsigned long long), stream()));
  MPK_LAUNCH(k_plan_old, nblocks(n, 256), 256, n, fw, W, E.off, E.col, (const int *)S.late, (const int *)S.pos,
             (const int *)lvl, S.old);
  unsigned long long old = 0;
  copy_dtoh(&old, S.old, sizeof old);
  if (failed()) return -1;
  if (!(mode == 2 || (n <= plan_ring_width_rule() * nl && 64 * (i64)old <= m))) return 0;
  // the ring plan
  const int hb[4] = {0, (int)n, 0, (int)nlvl};
  copy_htod(blk_pos, hb, sizeof hb);
  const i64 ints = 2 * 2 + nlvl + 3 * n + 1;
  double *ecoef = reinterpret_cast<double *>(dst + 64 + ((4 * ints + 7) & ~(i64)7));
  int *edep = reinterpret_cast<int *>(ecoef + slots);
  if (!w && !plan_eout(n, E.off, prow, pent, S)) return -1;
  HIPCHK(hipMemsetAsync(S.flags, 0, sizeof(unsigned), stream()));
  MPK_LAUNCH(k_plan_ring_entries, nblocks(n, 256), 256, n, fw, W, w, E.off, E.col, E.coef, (const int *)prow,
             (const int *)S.pos, (const int *)S.late, (const int *)lvl, (const i64 *)S.eout, pcnt, pent, ecoef, edep,
             S.flags);
  std::vector<int> lv(nlvl);
  copy_dtoh(lv.data(), lvl, nlvl * sizeof(int));
  unsigned flags = 0;
  copy_dtoh(&flags, S.flags, sizeof flags);
  if (failed()) return -1;
  int widest = 0;
  for (i64 l = 0; l + 1 < nlvl; l++) widest = std::max(widest, lv[l + 1] - lv[l]);
  int64_t hdr[8] = {1, m, nlvl, nl, bytes, w, W, 0};
  const i64 total = plan_mark(n, 1, w, nl, widest, flags & 1u, flags & 2u, bytes, cap, ecoef, edep, prow, pcnt,
                              dst, hdr);
  copy_htod(dst, hdr, sizeof hdr);
  if (failed()) return -1;
  *depth_out = nl;
  *widest_level = widest;
  *mark = (int)hdr[7];
  return total;
}
static PlanScratch &plan_scratch() {
  static PlanScratch s;
  return s;
}

i64 device_sweep_plan(const DevEntries &E, bool forward, i64 bmax, char *dst, i64 cap, i64 *depth_out,
                      int *longest, int *widest_level, int *mark) {
  Ctx *c = ctx();
  const i64 n = E.n, m = E.m;
  if (!c || n <= 0) return -1;
  const int fw = forward ? 1 : 0;
  const i64 w = plan_w_rule(E.widest, n, m), slots = w ? w * n : m;
  const int mode = plan_mode_rule();
  const i64 B = plan_rows_rule(n, bmax), nblk = (n + B - 1) / B;
  PlanScratch &S = plan_scratch();
  if (!S.ensure(n, nblk)) return -1;
  watch_install();
  if (n > bmax && mode != 1) {
    // the ring test needs the sweep's depth: a ring is possible when the
    // average level holds at most the ring width
    HIPCHK(hipMemsetAsync(S.lev, 0xff, n * sizeof(int), stream()));
    HIPCHK(hipMemsetAsync(S.maxlev, 0xff, sizeof(int), stream()));
    HIPCHK(hipMemsetAsync(c->tickets, 0, sizeof(unsigned), stream()));
    MPK_LAUNCH(k_plan_glev, nblocks(n, 64), 64, n, fw, E.off, E.col, S.lev, c->tickets, S.maxlev);
    int ml = 0;
    copy_dtoh(&ml, S.maxlev, sizeof ml);
    if (failed()) return -1;
    const i64 nl = (i64)ml + 1;
    if (mode == 2 || !(n > plan_ring_width_rule() * nl)) {
      const i64 r = device_ring_plan(E, forward, bmax, nl, mode, dst, cap, depth_out, widest_level, mark, S);
      if (r != 0) {
        *longest = (int)E.widest;
        return r;
      }
    }
  }
  // blocks: local levels, then the sizes the layout needs
  MPK_LAUNCH(k_plan_local, nblk, 64, n, fw, B, E.off, E.col, S.lev, S.nlev);
  std::vector<int> nlev(nblk);
  copy_dtoh(nlev.data(), S.nlev, nblk * sizeof(int));
  if (failed()) return -1;
  std::vector<i64> li(nblk + 1, 0);
  i64 depth = 0, levels = 0;
  for (i64 b = 0; b < nblk; b++) {
    li[b + 1] = li[b] + nlev[b] + 1;
    depth = std::max<i64>(depth, nlev[b]);
    levels = std::max<i64>(levels, nlev[b]);
  }
  const i64 nlvl = li[nblk], bytes = plan_bytes_rule(n, nblk, nlvl, slots);
  const bool wave_w = w == 4 || w == 8;
  const i64 rb = wave_rec_bytes_rule(w), wbytes = bytes + (wave_w ? n * rb : 0);
  if (bytes > cap) {
    set_error(5, "sweep plan buffer too small");
    return -1;
  }
  copy_htod(S.li, li.data(), (nblk + 1) * sizeof(i64));
  HIPCHK(hipMemsetAsync(dst, 0, (size_t)std::min(cap, wbytes), stream()));
  // the layout (plan_view's)
  int *blk_pos = reinterpret_cast<int *>(dst + 64), *blk_lev = blk_pos + nblk + 1, *lvl = blk_lev + nblk + 1;
  int *prow = lvl + nlvl, *pcnt = prow + n, *pent = pcnt + n;
  const i64 ints = 2 * (nblk + 1) + nlvl + 3 * n + 1;
  double *ecoef = reinterpret_cast<double *>(dst + 64 + ((4 * ints + 7) & ~(i64)7));
  int *edep = reinterpret_cast<int *>(ecoef + slots);
  MPK_LAUNCH(k_plan_pos, nblk, 64, n, fw, B, nblk, nlvl, (const int *)S.lev, (const i64 *)S.li,
             (const int *)S.nlev, blk_pos, blk_lev, lvl, prow, S.pos, S.wlev);
  if (!w && !plan_eout(n, E.off, prow, pent, S)) return -1;  // CSR form's offsets
  HIPCHK(hipMemsetAsync(S.flags, 0, sizeof(unsigned), stream()));
  MPK_LAUNCH(k_plan_entries, nblocks(n, 256), 256, n, fw, B, w, E.off, E.col, E.coef, (const int *)prow,
             (const int *)S.pos, (const int *)blk_pos, (const i64 *)S.eout, pcnt, pent, ecoef, edep, S.flags);
  std::vector<int> wl(nblk);
  copy_dtoh(wl.data(), S.wlev, nblk * sizeof(int));
  unsigned flags = 0;
  copy_dtoh(&flags, S.flags, sizeof flags);
  if (failed()) return -1;
  int widest = 0;
  for (i64 b = 0; b < nblk; b++) widest = std::max(widest, wl[b]);
  const bool published = flags & 1u, stale = flags & 2u;
  int64_t hdr[8] = {nblk, m, nlvl, depth, bytes, w, 0, 0};
  const i64 total = plan_mark(n, nblk, w, levels, widest, published, stale, bytes, cap, ecoef, edep, prow, pcnt, dst,
                              hdr);
  copy_htod(dst, hdr, sizeof hdr);
  if (failed()) return -1;
  *depth_out = depth;
  *longest = (int)E.widest;
  *widest_level = widest;
  *mark = (int)hdr[7];
  return total;
}

}  // namespace mpk
