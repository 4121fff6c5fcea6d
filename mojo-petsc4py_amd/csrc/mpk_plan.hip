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

// The block order.  ord[s]: the row at sweep position s; rank: its inverse;
// rblk: each row's block; block b holds positions bstart[b] .. bstart[b+1]-1.
// Every dependency (an earlier row in the sweep's direction, plan_earlier)
// sits at an earlier position: the identity order (or its reverse) and the
// pencil order (checked by k_pen_topo) both are.
__device__ __forceinline__ bool plan_earlier(i64 k, i64 i, int forward) { return forward ? k < i : k > i; }

// Local levels of block b: a row's level is 1 + the largest level of a
// dependency in the same block, 0 without one.  One wave per block takes its
// rows 64 positions at a time (a row's dependency in an earlier group is
// done; in its own group the wave repeats until it is); lev gets every row's
// level, nlev[b] the block's level count.
__global__ void __launch_bounds__(64) k_plan_local(i64 n, int forward, const i64 *bstart, const int *ord,
                                                   const int *rank, const int *rblk, const i64 *off, const int *col,
                                                   int *lev, int *nlev) {
  __shared__ int llev[MPK_BLK_ROWS];
  const int lane = (int)threadIdx.x;
  const i64 b = blockIdx.x, s0 = bstart[b], s1 = bstart[b + 1];
  int mx = -1;
  for (i64 c0 = s0; c0 < s1; c0 += 64) {
    const i64 s = c0 + lane;
    const bool valid = s < s1;
    const i64 i = valid ? ord[s] : 0;
    bool done = !valid;
    i64 j = valid ? off[i] : 0;
    const i64 je = valid ? off[i + 1] : 0;
    int l = 0;
    unsigned long long mask = 0;
    for (;;) {
      if (!done) {
        while (j < je) {
          const i64 k = col[j];
          if (plan_earlier(k, i, forward) && rblk[k] == b) {  // a dependency in this block
            const i64 sk = rank[k];
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
__global__ void __launch_bounds__(64) k_plan_pos(i64 n, const i64 *bstart, const int *ord, i64 nblk, i64 nlvl,
                                                 const int *lev, const i64 *li, const int *nlev, int *blk_pos,
                                                 int *blk_lev, int *lvl, int *prow, int *pos, int *wlev) {
  __shared__ int cnt[MPK_BLK_ROWS + 1];
  const int lane = (int)threadIdx.x;
  const i64 b = blockIdx.x, s0 = bstart[b], s1 = bstart[b + 1];
  const int L = nlev[b];
  for (int l = lane; l <= L; l += 64) cnt[l] = 0;
  __syncthreads();
  for (i64 s = s0 + lane; s < s1; s += 64) atomicAdd(&cnt[lev[ord[s]] + 1], 1);
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
    const i64 i = valid ? ord[s] : 0;
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
__global__ void k_plan_entries(i64 n, int forward, const int *rblk, i64 w, const i64 *off, const int *col,
                               const double *coef, const int *prow, const int *pos, const int *blk_pos,
                               const i64 *eout, int *pcnt, int *pent, double *ecoef, int *edep, unsigned *flags) {
  const i64 p = (i64)blockIdx.x * 256 + threadIdx.x;
  unsigned f = 0;
  if (p < n) {
    const i64 i = prow[p], b = rblk[i], o = off[i], cnt = off[i + 1] - o;
    pcnt[p] = (int)cnt;
    pent[p] = w ? 0 : (int)eout[p];
    const i64 q0 = w ? p : eout[p], qs = w ? n : 1;
    for (i64 e = 0; e < cnt; e++) {
      const i64 k = col[o + e];
      const i64 q = q0 + e * qs;
      ecoef[q] = coef[o + e];
      int d;
      if (!plan_earlier(k, i, forward)) {
        d = (int)(-k - n - 1);
        f |= 2u;
      } else if (rblk[k] == b) {
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

// Block orders (the host planner's, mpk_sweep_plan / pencil_order): runs of B
// consecutive sweep positions ...
__global__ void k_plan_ident(i64 n, int forward, i64 B, int *rank, int *ord, int *rblk) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const i64 s = plan_spos(i, n, forward);
  rank[i] = (int)s;
  ord[s] = (int)i;
  rblk[i] = (int)(s / B);
}
// ... or pencils: tiles of tx x ty columns of an nx x ny x nz grid (x
// fastest) spanning every z, in row-major tile order, rows in index order
// within a tile (z, then y, then x); reversed for a backward sweep.  A tile
// (bx, by) starts after the full tile rows below it and the tiles to its
// left in its own row.
__global__ void k_pen_rank(i64 n, int forward, i64 nx, i64 ny, i64 nz, i64 tx, i64 ty, i64 mx, i64 nblk, int *rank,
                           int *ord, int *rblk) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const i64 plane = nx * ny, z = i / plane, rem = i - z * plane, y = rem / nx, x = rem - y * nx;
  const i64 by = y / ty, bx = x / tx;
  const i64 wy = ty < ny - by * ty ? ty : ny - by * ty, wx = tx < nx - bx * tx ? tx : nx - bx * tx;
  i64 r = nz * (by * ty * nx + wy * bx * tx) + z * wy * wx + (y - by * ty) * wx + (x - bx * tx);
  i64 t = by * mx + bx;
  if (!forward) {
    r = n - 1 - r;
    t = nblk - 1 - t;
  }
  rank[i] = (int)r;
  ord[r] = (int)i;
  rblk[i] = (int)t;
}
// Pencil detection: the sweep's dependency offsets (forward: i - k, backward:
// k - i) must be exactly three, 1 < a < b.  pm[0]: the smallest, pm[1]: the
// largest (k_pen_span); pm[2]: the smallest strictly between 1 and the
// largest (k_pen_mid); flags bit 0: an offset outside {1, a, b}
// (k_pen_check), bit 1: a dependency at a later position of the order
// (k_pen_topo).  Each row's values are folded over its wave first.
__device__ __forceinline__ i64 pen_dist(i64 i, i64 k, int forward) { return forward ? i - k : k - i; }
__global__ void k_pen_span(i64 n, int forward, const i64 *off, const int *col, int *pm) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  int lo = 0x7fffffff, hi = 0;
  if (i < n)
    for (i64 j = off[i]; j < off[i + 1]; j++) {
      const i64 dd = pen_dist(i, col[j], forward);
      if (dd > 0) {
        lo = dd < lo ? (int)dd : lo;
        hi = dd > hi ? (int)dd : hi;
      }
    }
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const int a = __shfl_down(lo, w, 64), b = __shfl_down(hi, w, 64);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  __shared__ int sl[4], sh[4];
  if ((threadIdx.x & 63) == 0) {
    sl[threadIdx.x >> 6] = lo;
    sh[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // one pair of atomics per workgroup
    for (int k = 1; k < 4; k++) {
      lo = sl[k] < lo ? sl[k] : lo;
      hi = sh[k] > hi ? sh[k] : hi;
    }
    atomicMin(pm, lo);
    atomicMax(pm + 1, hi);
  }
}
__global__ void k_pen_mid(i64 n, int forward, const i64 *off, const int *col, int *pm) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  const i64 top = pm[1];
  int lo = 0x7fffffff;
  if (i < n)
    for (i64 j = off[i]; j < off[i + 1]; j++) {
      const i64 dd = pen_dist(i, col[j], forward);
      if (dd > 1 && dd < top) lo = dd < lo ? (int)dd : lo;
    }
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const int a = __shfl_down(lo, w, 64);
    lo = a < lo ? a : lo;
  }
  __shared__ int sl[4];
  if ((threadIdx.x & 63) == 0) sl[threadIdx.x >> 6] = lo;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; k++) lo = sl[k] < lo ? sl[k] : lo;
    atomicMin(pm + 2, lo);
  }
}
__global__ void k_pen_check(i64 n, int forward, const i64 *off, const int *col, const int *pm, unsigned *flags) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  bool bad = false;
  if (i < n) {
    const i64 a = pm[2], b = pm[1];
    for (i64 j = off[i]; j < off[i + 1]; j++) {
      const i64 dd = pen_dist(i, col[j], forward);
      bad |= dd > 0 && dd != 1 && dd != a && dd != b;
    }
  }
  if (__ballot(bad) && (threadIdx.x & 63) == 0) atomicOr(flags, 1u);
}
__global__ void k_pen_topo(i64 n, int forward, const i64 *off, const int *col, const int *rank, unsigned *flags) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  bool bad = false;
  if (i < n)
    for (i64 j = off[i]; j < off[i + 1]; j++) {
      const i64 k = col[j];
      bad |= plan_earlier(k, i, forward) && rank[k] >= rank[i];
    }
  if (__ballot(bad) && (threadIdx.x & 63) == 0) atomicOr(flags, 2u);
}

// Device scratch of the planner, grown on demand and kept.
struct PlanScratch {
  int *lev = nullptr, *pos = nullptr, *nlev = nullptr, *wlev = nullptr, *maxlev = nullptr, *cnt = nullptr;
  int *late = nullptr, *key = nullptr, *skey = nullptr, *row = nullptr;
  i64 *li = nullptr, *eout = nullptr;
  unsigned *flags = nullptr;
  unsigned long long *old = nullptr, *lw = nullptr;
  i64 *toff = nullptr, *bstart = nullptr;
  int *tcol = nullptr, *ord = nullptr, *rank = nullptr, *rblk = nullptr, *pen = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_cap = 0;
  i64 cap_n = 0, cap_b = 0, cap_t = 0;
  bool ensure(i64 n, i64 nblk) {
    if (n > cap_n) {
      for (void *q : {(void *)lev, (void *)pos, (void *)eout, (void *)cnt, (void *)late, (void *)key, (void *)skey,
                      (void *)row, (void *)lw, (void *)toff, (void *)ord, (void *)rank, (void *)rblk})
        if (q) HIPCHK(hipFree(q));
      HIPCHK(hipMalloc(&ord, n * sizeof(int)));
      HIPCHK(hipMalloc(&rank, n * sizeof(int)));
      HIPCHK(hipMalloc(&rblk, n * sizeof(int)));
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
      for (void *q : {(void *)nlev, (void *)wlev, (void *)li, (void *)bstart})
        if (q) HIPCHK(hipFree(q));
      HIPCHK(hipMalloc(&bstart, (nblk + 1) * sizeof(i64)));
      HIPCHK(hipMalloc(&nlev, (nblk + 1) * sizeof(int)));
      HIPCHK(hipMalloc(&wlev, (nblk + 1) * sizeof(int)));
      HIPCHK(hipMalloc(&li, (nblk + 1) * sizeof(i64)));
      cap_b = nblk + 1;
    }
    if (!maxlev) {
      HIPCHK(hipMalloc(&pen, 8 * sizeof(int)));
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

// The blocks of a block plan, as the host planner picks them: pencils when
// the sweep is a structured grid's and their order is topological for it,
// else runs of plan_rows consecutive positions.  Fills S.ord, S.rank, S.rblk
// and S.bstart; returns the block count, or -1.
static i64 plan_order(const DevEntries &E, bool forward, i64 bmax, int mode, PlanScratch &S) {
  const i64 n = E.n;
  const int fw = forward ? 1 : 0;
  std::vector<i64> bs;
  bool pen = false;
  if (mode != 2 && n > bmax && !plan_env_off("MPK_SWEEP_PENCILS")) {
    const int init[4] = {0x7fffffff, 0, 0x7fffffff, 0};
    copy_htod(S.pen, init, sizeof init);
    MPK_LAUNCH(k_pen_span, nblocks(n, 256), 256, n, fw, E.off, E.col, S.pen);
    MPK_LAUNCH(k_pen_mid, nblocks(n, 256), 256, n, fw, E.off, E.col, S.pen);
    MPK_LAUNCH(k_pen_check, nblocks(n, 256), 256, n, fw, E.off, E.col, (const int *)S.pen,
               reinterpret_cast<unsigned *>(S.pen + 3));
    int pm[4];
    copy_dtoh(pm, S.pen, sizeof pm);
    if (failed()) return -1;
    const i64 a = pm[2], b = pm[1];
    if (pm[0] == 1 && pm[3] == 0 && a < b && b % a == 0 && n % b == 0) {
      const i64 nx = a, ny = b / a, nz = n / b;
      i64 t = 1;
      while ((t + 1) * (t + 1) * nz <= bmax) t++;
      const i64 tx = std::min(t, nx), ty = std::min(t, ny), mx = (nx + tx - 1) / tx, my = (ny + ty - 1) / ty;
      if (tx * ty >= 4 && mx * my >= 2) {
        const i64 nb = mx * my;
        bs.assign(1, 0);
        for (i64 by = 0; by < my; by++)
          for (i64 bx = 0; bx < mx; bx++)
            bs.push_back(bs.back() + std::min(tx, nx - bx * tx) * std::min(ty, ny - by * ty) * nz);
        if (!forward) {
          std::vector<i64> rb(bs.size());
          for (size_t j = 0; j < bs.size(); j++) rb[j] = n - bs[bs.size() - 1 - j];
          bs.swap(rb);
        }
        MPK_LAUNCH(k_pen_rank, nblocks(n, 256), 256, n, fw, nx, ny, nz, tx, ty, mx, nb, S.rank, S.ord, S.rblk);
        MPK_LAUNCH(k_pen_topo, nblocks(n, 256), 256, n, fw, E.off, E.col, (const int *)S.rank,
                   reinterpret_cast<unsigned *>(S.pen + 3));
        unsigned fl = 0;
        copy_dtoh(&fl, S.pen + 3, sizeof fl);
        if (failed()) return -1;
        pen = fl == 0;
      }
    }
  }
  if (!pen) {
    const i64 B = plan_rows_rule(n, bmax), nb = (n + B - 1) / B;
    bs.resize(nb + 1);
    for (i64 b = 0; b <= nb; b++) bs[b] = std::min(n, b * B);
    MPK_LAUNCH(k_plan_ident, nblocks(n, 256), 256, n, fw, B, S.rank, S.ord, S.rblk);
  }
  const i64 nblk = (i64)bs.size() - 1;
  if (!S.ensure(n, nblk)) return -1;
  copy_htod(S.bstart, bs.data(), bs.size() * sizeof(i64));
  return failed() ? -1 : nblk;
}

i64 device_sweep_plan(const DevEntries &E, bool forward, i64 bmax, char *dst, i64 cap, i64 *depth_out,
                      int *longest, int *widest_level, int *mark) {
  Ctx *c = ctx();
  const i64 n = E.n, m = E.m;
  if (!c || n <= 0) return -1;
  const int fw = forward ? 1 : 0;
  const i64 w = plan_w_rule(E.widest, n, m), slots = w ? w * n : m;
  const int mode = plan_mode_rule();
  PlanScratch &S = plan_scratch();
  if (!S.ensure(n, 1)) return -1;
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
  // blocks: their order, local levels, then the sizes the layout needs
  const i64 nblk = plan_order(E, forward, bmax, mode, S);
  if (nblk < 0) return -1;
  MPK_LAUNCH(k_plan_local, nblk, 64, n, fw, (const i64 *)S.bstart, (const int *)S.ord, (const int *)S.rank,
             (const int *)S.rblk, E.off, E.col, S.lev, S.nlev);
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
  MPK_LAUNCH(k_plan_pos, nblk, 64, n, (const i64 *)S.bstart, (const int *)S.ord, nblk, nlvl, (const int *)S.lev,
             (const i64 *)S.li,
             (const int *)S.nlev, blk_pos, blk_lev, lvl, prow, S.pos, S.wlev);
  if (!w && !plan_eout(n, E.off, prow, pent, S)) return -1;  // CSR form's offsets
  HIPCHK(hipMemsetAsync(S.flags, 0, sizeof(unsigned), stream()));
  MPK_LAUNCH(k_plan_entries, nblocks(n, 256), 256, n, fw, (const int *)S.rblk, w, E.off, E.col, E.coef,
             (const int *)prow,
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

// ------------------------------------------------------- IC(0) analysis
// The IC(0) factor's pattern with the natural ordering, and everything the
// host computes from it (mp_icc_symbolic at level 0, mp_icc_contiguous_
// cholesky, mpk_icc_transpose, mpk_icc_visits), built on the device:
//   - row k of the factor holds A's columns above k, ascending, then k
//     itself (level 0 admits no fill; the symbolic pass sorts the strict
//     columns and puts the diagonal slot last, present in A or not);
//   - column k's entries (i, k), i < k, are the rows that update row k; the
//     transpose lists them by ascending i;
//   - the up-looking pass (ported.mojo:1040-1124) visits them in the order
//     its row chains hold them.  Row i joins column k's chain while the pass
//     is at row q, the column before k in row i (or i itself when k is its
//     first strict column), in the order it left q's chain, itself after
//     every row that was in q's chain; each join goes to the chain's head.
//     So column k's visits come by q descending, ties (same q) in the
//     reverse of their order at q, row q itself first.  Unrolled, row i's
//     key is its columns before k in reverse order, then i, then an end
//     mark; keys compare element by element, descending at even depths and
//     ascending at odd ones, the end mark last at even depths and first at
//     odd ones (k_ic0_order);
//   - a row's level is one past its deepest visitor's (k_plan_glev over the
//     transpose), the level lists by a stable radix sort of the rows.
// Rows of A must hold strictly ascending columns within 0..n-1, and no
// column may gather more than kIc0MaxCol entries (each column is sorted by
// one thread); mpk_ic0_size reports -2 otherwise, and the host passes apply.

static const int kIc0MaxCol = 256;

// Per row: the check, the factor row's length (strict columns + 1), and the
// column counts of the strict part.  flags[0] bit 0: a row out of order or
// a column out of range; flags[1]: the longest column (atomicMax).
__global__ void k_ic0_count(i64 n, const int *rp, const int *ci, unsigned long long *rlen, int *ccnt,
                            unsigned *flags) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int prev = -1;
  unsigned long long s = 1;
  bool bad = false;
  for (int j = rp[i]; j < rp[i + 1]; j++) {
    const int c = ci[j];
    bad |= c <= prev || c >= n;
    prev = c;
    if (!bad && c > i) {
      s++;
      const int q = atomicAdd(ccnt + c, 1) + 1;
      if (q > kIc0MaxCol) atomicMax(flags + 1, (unsigned)q);
    }
  }
  rlen[i] = s;
  if (bad) atomicOr(flags, 1u);
}

// The factor rows (bi, bj, bdiag in the contiguous Cholesky layout) and the
// transpose's entries, unordered within a column (cursor: a copy of trp).
__global__ void k_ic0_fill(i64 n, const int *rp, const int *ci, const i64 *b64, int *bi, int *bj, int *bdiag,
                           int *cursor, int *trow, int *tpos) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int b = (int)b64[i], e = (int)b64[i + 1];
  bi[i] = b;
  bdiag[i] = e - 1;
  int t = b;
  for (int j = rp[i]; j < rp[i + 1]; j++) {
    const int c = ci[j];
    if (c > i) {
      bj[t] = c;
      const int p = atomicAdd(cursor + c, 1);
      trow[p] = (int)i;
      tpos[p] = t;
      t++;
    }
  }
  bj[e - 1] = (int)i;
  if (i == n - 1) {
    bi[n] = e;
    bdiag[n] = e;
  }
}

// Element d of the key of row i's entry at position p (bj[p] = k): the
// columns before k backwards, then i, then the end mark (-1).
__device__ __forceinline__ int ic0_key(const int *bi, const int *bj, int i, int p, int d) {
  const int q = p - 1 - d, b = bi[i];
  return q >= b ? bj[q] : (q == b - 1 ? i : -1);
}
// True when the entry (ia, pa) comes before (ib, pb) in the column's visits.
__device__ __forceinline__ bool ic0_before(const int *bi, const int *bj, int ia, int pa, int ib, int pb) {
  for (int d = 0;; d++) {
    const int ea = ic0_key(bi, bj, ia, pa, d), eb = ic0_key(bi, bj, ib, pb, d);
    if (ea == eb) {
      if (ea < 0) return false;  // the same entry
      continue;
    }
    if ((d & 1) == 0) return eb < 0 || (ea >= 0 && ea > eb);
    return ea < 0 || (eb >= 0 && ea < eb);
  }
}

// One thread per column: the transpose by ascending row, the visits in the
// chains' order, tmid (every entry is above the diagonal: the column's end).
__global__ void k_ic0_order(i64 n, const int *trp, int *trow, int *tpos, const int *bi, const int *bj, int *vrow,
                            int *vpos, int *tmid) {
  const i64 k = (i64)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int s0 = trp[k], s1 = trp[k + 1];
  for (int a = s0 + 1; a < s1; a++) {
    const int r = trow[a], p = tpos[a];
    int b = a - 1;
    while (b >= s0 && trow[b] > r) {
      trow[b + 1] = trow[b];
      tpos[b + 1] = tpos[b];
      b--;
    }
    trow[b + 1] = r;
    tpos[b + 1] = p;
  }
  for (int a = s0; a < s1; a++) {
    const int r = trow[a], p = tpos[a];
    int b = a - 1;
    while (b >= s0 && ic0_before(bi, bj, r, p, vrow[b], vpos[b])) {
      vrow[b + 1] = vrow[b];
      vpos[b + 1] = vpos[b];
      b--;
    }
    vrow[b + 1] = r;
    vpos[b + 1] = p;
  }
  tmid[k] = s1;
}

__global__ void k_ic0_widen(i64 n, const int *a, i64 *b) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i <= n) b[i] = a[i];
}
__global__ void k_ic0_iota(i64 n, int *a) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = (int)i;
}
// Level starts from the rows sorted by level (every level 0..nlev-1 holds a
// row: a row's deepest visitor is one level up).
__global__ void k_ic0_lptr(i64 n, i64 nlev, const int *slev, int *lp) {
  const i64 p = (i64)blockIdx.x * 256 + threadIdx.x;
  if (p < n && (p == 0 || slev[p] != slev[p - 1])) lp[slev[p]] = (int)p;
  if (p == 0) lp[nlev] = (int)n;
}

struct Ic0Scratch {
  unsigned long long *rlen = nullptr, *ucnt = nullptr;
  i64 *b64 = nullptr, *t64 = nullptr;
  int *ccnt = nullptr, *cursor = nullptr, *lev = nullptr, *slev = nullptr, *iota = nullptr, *lp = nullptr;
  unsigned *flags = nullptr;
  int *maxlev = nullptr;
  void *tmp = nullptr;
  size_t tmp_cap = 0;
  i64 cap = 0;
  bool ensure(i64 n) {
    if (n > cap) {
      for (void *q : {(void *)rlen, (void *)ucnt, (void *)b64, (void *)t64, (void *)ccnt, (void *)cursor, (void *)lev,
                      (void *)slev, (void *)iota, (void *)lp})
        if (q) HIPCHK(hipFree(q));
      HIPCHK(hipMalloc(&rlen, (n + 1) * sizeof(unsigned long long)));
      HIPCHK(hipMalloc(&ucnt, (n + 1) * sizeof(unsigned long long)));
      HIPCHK(hipMalloc(&b64, (n + 1) * sizeof(i64)));
      HIPCHK(hipMalloc(&t64, (n + 1) * sizeof(i64)));
      for (int **q : {&ccnt, &cursor, &lev, &slev, &iota, &lp}) HIPCHK(hipMalloc(q, (n + 1) * sizeof(int)));
      cap = n;
    }
    if (!flags) {
      HIPCHK(hipMalloc(&flags, 4 * sizeof(unsigned)));
      HIPCHK(hipMalloc(&maxlev, 4 * sizeof(int)));
    }
    return !failed();
  }
  bool ensure_tmp(size_t bytes) {
    if (bytes <= tmp_cap && tmp) return true;
    if (tmp) HIPCHK(hipFree(tmp));
    HIPCHK(hipMalloc(&tmp, bytes > 0 ? bytes : 16));
    tmp_cap = bytes;
    return !failed();
  }
};
static Ic0Scratch &ic0_scratch() {
  static Ic0Scratch s;
  return s;
}

// The counts and the factor's row starts (b64) in the scratch; returns
// bi[n], or -2 as documented, or -1 on error.
static i64 ic0_counts(i64 n, const int *rp, const int *ci, Ic0Scratch &S) {
  if (!S.ensure(n)) return -1;
  HIPCHK(hipMemsetAsync(S.ccnt, 0, (n + 1) * sizeof(int), stream()));
  HIPCHK(hipMemsetAsync(S.flags, 0, 4 * sizeof(unsigned), stream()));
  HIPCHK(hipMemsetAsync(S.rlen + n, 0, sizeof(unsigned long long), stream()));
  MPK_LAUNCH(k_ic0_count, nblocks(n, 256), 256, n, rp, ci, S.rlen, S.ccnt, S.flags);
  size_t tb = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, S.rlen, S.b64, (int)(n + 1), stream()));
  if (!S.ensure_tmp(tb)) return -1;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(S.tmp, tb, S.rlen, S.b64, (int)(n + 1), stream()));
  unsigned flags[2] = {0, 0};
  i64 total = 0;
  copy_dtoh(flags, S.flags, sizeof flags);
  copy_dtoh(&total, S.b64 + n, sizeof total);
  if (failed()) return -1;
  if ((flags[0] & 1u) || flags[1] > (unsigned)kIc0MaxCol || total >= ((i64)1 << 31)) return -2;
  return total;
}

// Level lists of a forward sweep whose rows read only earlier rows (off [D]
// int64[n + 1], col [D] int32): lrow [D] int32[n] the rows level by level,
// ascending within a level; lptr [H] int64[nlev + 1].  Returns nlev, or -1.
static i64 level_lists(i64 n, const i64 *off, const int *col, int *lrow, int64_t *lptr, Ic0Scratch &S) {
  Ctx *c = ctx();
  watch_install();
  HIPCHK(hipMemsetAsync(S.lev, 0xff, n * sizeof(int), stream()));
  HIPCHK(hipMemsetAsync(S.maxlev, 0xff, sizeof(int), stream()));
  HIPCHK(hipMemsetAsync(c->tickets, 0, sizeof(unsigned), stream()));
  MPK_LAUNCH(k_plan_glev, nblocks(n, 64), 64, n, 1, off, col, S.lev, c->tickets, S.maxlev);
  int ml = 0;
  copy_dtoh(&ml, S.maxlev, sizeof ml);
  if (failed()) return -1;
  const i64 nlev = (i64)ml + 1;
  int bits = 1;
  while (((i64)1 << bits) <= nlev) bits++;
  MPK_LAUNCH(k_ic0_iota, nblocks(n, 256), 256, n, S.iota);
  size_t tb = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, S.lev, S.slev, S.iota, lrow, (int)n, 0, bits, stream()));
  if (!S.ensure_tmp(tb)) return -1;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(S.tmp, tb, S.lev, S.slev, S.iota, lrow, (int)n, 0, bits, stream()));
  MPK_LAUNCH(k_ic0_lptr, nblocks(n, 256), 256, n, nlev, (const int *)S.slev, S.lp);
  std::vector<int> lp(nlev + 1);
  copy_dtoh(lp.data(), S.lp, (nlev + 1) * sizeof(int));
  if (failed()) return -1;
  for (i64 l = 0; l <= nlev; l++) lptr[l] = lp[l];
  return nlev;
}

// ------------------------------------------------------ ILU(0) analysis
// The ILU(0) factor with the natural ordering (mp_ilu_symbolic_ilu0's
// layout, factor.mojo:40-67) and its level lists (mpk_ilu_levels), built on
// the device.  Row i's L part (A's columns below i) sits at bi[i]..bi[i+1]-1,
// bi the prefix sums of the L counts; the U rows run from the end of the L
// parts backwards, row n-1 first: row i's strict columns above i ascending,
// then its diagonal slot bdiag[i] = bi[n] - 1 + the sum of (U count + 1)
// over rows i..n-1.  A row's level is one past its deepest L column's.  Rows
// of A must hold strictly ascending columns within 0..n-1 and their
// diagonal (the layouts of rows without one only the host passes handle).

// flags[0] bit 0: a row out of order, out of range or without its diagonal.
__global__ void k_ilu0_count(i64 n, const int *rp, const int *ci, unsigned long long *lcnt,
                             unsigned long long *ucnt, unsigned *flags) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int prev = -1;
  unsigned long long nl = 0, nu = 0;
  bool bad = false, diag = false;
  for (int j = rp[i]; j < rp[i + 1]; j++) {
    const int c = ci[j];
    bad |= c <= prev || c >= n;
    prev = c;
    nl += c < i;
    nu += c > i;
    diag |= c == i;
  }
  lcnt[i] = nl;
  ucnt[i] = nu + 1;
  if (bad || !diag) atomicOr(flags, 1u);
}

__global__ void k_ilu0_fill(i64 n, const int *rp, const int *ci, const i64 *lp64, const i64 *up64, int *bi, int *bj,
                            int *bdiag) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const i64 ltot = lp64[n], utot = up64[n];
  const int d = (int)(ltot - 1 + utot - up64[i]), d1 = (int)(ltot - 1 + utot - up64[i + 1]);
  int l = (int)lp64[i], u = d1 + 1;
  bi[i] = l;
  bdiag[i] = d;
  for (int j = rp[i]; j < rp[i + 1]; j++) {
    const int c = ci[j];
    if (c < i) bj[l++] = c;
    else if (c > i) bj[u++] = c;
  }
  bj[d] = (int)i;
  if (i == n - 1) {
    bi[n] = (int)ltot;
    bdiag[n] = (int)(ltot - 1);
  }
}

// The L and U counts' prefix sums in S.b64 / S.t64; returns the factor's
// entry count, or -2 as documented, or -1 on error.
static i64 ilu0_counts(i64 n, const int *rp, const int *ci, Ic0Scratch &S) {
  if (!S.ensure(n)) return -1;
  HIPCHK(hipMemsetAsync(S.flags, 0, 4 * sizeof(unsigned), stream()));
  HIPCHK(hipMemsetAsync(S.rlen + n, 0, sizeof(unsigned long long), stream()));
  HIPCHK(hipMemsetAsync(S.ucnt + n, 0, sizeof(unsigned long long), stream()));
  MPK_LAUNCH(k_ilu0_count, nblocks(n, 256), 256, n, rp, ci, S.rlen, S.ucnt, S.flags);
  size_t tb = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, S.rlen, S.b64, (int)(n + 1), stream()));
  if (!S.ensure_tmp(tb)) return -1;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(S.tmp, tb, S.rlen, S.b64, (int)(n + 1), stream()));
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(S.tmp, tb, S.ucnt, S.t64, (int)(n + 1), stream()));
  unsigned flags = 0;
  i64 tot[2] = {0, 0};
  copy_dtoh(&flags, S.flags, sizeof flags);
  copy_dtoh(tot, S.b64 + n, sizeof(i64));
  copy_dtoh(tot + 1, S.t64 + n, sizeof(i64));
  if (failed()) return -1;
  if ((flags & 1u) || tot[0] + tot[1] >= ((i64)1 << 31)) return -2;
  return tot[0] + tot[1];
}

}  // namespace mpk

using namespace mpk;

extern "C" {

int64_t mpk_ilu0_size(int64_t irow, int64_t icol, int64_t nrow) {
  if (nrow <= 0) return 0;
  if (!ctx()) return -1;
  return ilu0_counts(nrow, (const int *)irow, (const int *)icol, ic0_scratch());
}

int64_t mpk_ilu0_analyse(int64_t irow, int64_t icol, int64_t nrow, int64_t bi, int64_t bj, int64_t bdiag,
                         int64_t lrow, int64_t lptr) {
  const i64 n = nrow;
  if (n <= 0) return 0;
  if (!ctx()) return -1;
  Ic0Scratch &S = ic0_scratch();
  const int *rp = (const int *)irow, *ci = (const int *)icol;
  const i64 total = ilu0_counts(n, rp, ci, S);
  if (total < 0) {
    if (total == -2) set_error(2, "mpk_ilu0_analyse: the matrix is not one mpk_ilu0_size accepts");
    return -1;
  }
  MPK_LAUNCH(k_ilu0_fill, nblocks(n, 256), 256, n, rp, ci, (const i64 *)S.b64, (const i64 *)S.t64, (int *)bi,
             (int *)bj, (int *)bdiag);
  // levels over the L parts (bi's prefix sums are S.b64)
  return level_lists(n, S.b64, (const int *)bj, (int *)lrow, (int64_t *)lptr, S);
}

int64_t mpk_ic0_size(int64_t irow, int64_t icol, int64_t nrow) {
  if (nrow <= 0) return 0;
  if (!ctx()) return -1;
  return ic0_counts(nrow, (const int *)irow, (const int *)icol, ic0_scratch());
}

int64_t mpk_ic0_analyse(int64_t irow, int64_t icol, int64_t nrow, int64_t bi, int64_t bj, int64_t bdiag, int64_t trp,
                        int64_t trow, int64_t tpos, int64_t tmid, int64_t vrow, int64_t vpos, int64_t lrow,
                        int64_t lptr) {
  const i64 n = nrow;
  if (n <= 0) return 0;
  if (!ctx()) return -1;
  Ic0Scratch &S = ic0_scratch();
  const int *rp = (const int *)irow, *ci = (const int *)icol;
  const i64 total = ic0_counts(n, rp, ci, S);
  if (total < 0) {
    if (total == -2) set_error(2, "mpk_ic0_analyse: the matrix is not one mpk_ic0_size accepts");
    return -1;
  }
  int *TRP = (int *)trp;
  // column starts of the strict part
  size_t tb = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, S.ccnt, TRP, (int)(n + 1), stream()));
  if (!S.ensure_tmp(tb)) return -1;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(S.tmp, tb, S.ccnt, TRP, (int)(n + 1), stream()));
  HIPCHK(hipMemcpyAsync(S.cursor, TRP, (n + 1) * sizeof(int), hipMemcpyDeviceToDevice, stream()));
  MPK_LAUNCH(k_ic0_fill, nblocks(n, 256), 256, n, rp, ci, (const i64 *)S.b64, (int *)bi, (int *)bj, (int *)bdiag,
             S.cursor, (int *)trow, (int *)tpos);
  MPK_LAUNCH(k_ic0_order, nblocks(n, 256), 256, n, (const int *)TRP, (int *)trow, (int *)tpos, (const int *)bi,
             (const int *)bj, (int *)vrow, (int *)vpos, (int *)tmid);
  // levels over the transpose (a row's visitors come before it)
  MPK_LAUNCH(k_ic0_widen, nblocks(n + 1, 256), 256, n, (const int *)TRP, S.t64);
  return level_lists(n, S.t64, (const int *)trow, (int *)lrow, (int64_t *)lptr, S);
}

}  // extern "C"
