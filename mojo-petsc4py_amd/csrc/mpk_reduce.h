// mpk_reduce.h -- the tile-tree reductions.
//
// The order of every floating-point add is a function of n alone:
//   tile k covers elements [2048 k, 2048 k + 2048); thread t (of 256) owns the
//   pairs starting at 2048 k + 512 j + 2 t, j = 0..3; a pair (e, e+1) is
//   v0 + v1, or fma(x0, y0, x1 * y1) for products; the thread sums its pairs as
//   (p0 + p1) + (p2 + p3); a wave sums its 64 lanes by the halving tree
//   u[i] += u[i + w], w = 32, 16, ..., 1; the block combines its four wave sums
//   as (W0 + W2) + (W1 + W3) -- that is tile partial P_k.
//   The finalize block gives thread t the value P_t + P_{t+256} + ... in
//   ascending k (0.0 when t >= T) and applies the same block tree.
// Entries past n contribute +0.0.  Because the order is fixed, a reduction
// gives the same bits on every run and on the CPU restatement of it.
#pragma once

#include "mpk_internal.h"

namespace mpk {

enum RedKind { RK_SUM = 0, RK_ASUM = 1, RK_SUMSQ = 2, RK_DOT = 3 };

struct RedSpec {
  const double *x;
  const double *y;
  int kind;
};

template <int K>
struct RedSet {
  RedSpec s[K];
};

__device__ __forceinline__ double pair_leaf(const RedSpec &s, i64 e, i64 n, bool al) {
  D2 xv = ld2(s.x, e, n, al);
  switch (s.kind) {
    case RK_SUM:
      return xv.a + xv.b;
    case RK_ASUM:
      return fabs(xv.a) + fabs(xv.b);
    case RK_SUMSQ:
      return fma(xv.a, xv.a, xv.b * xv.b);
    default: {
      D2 yv = ld2(s.y, e, n, al);
      return fma(xv.a, yv.a, xv.b * yv.b);
    }
  }
}

// Thread sum over the thread's four pairs of tile `blk`.
__device__ __forceinline__ double thread_sum(const RedSpec &s, i64 n, i64 blk, bool al) {
  const i64 base = blk * MPK_TILE + 2 * (i64)threadIdx.x;
  double p0 = pair_leaf(s, base, n, al);
  double p1 = pair_leaf(s, base + 512, n, al);
  double p2 = pair_leaf(s, base + 1024, n, al);
  double p3 = pair_leaf(s, base + 1536, n, al);
  return (p0 + p1) + (p2 + p3);
}

// K independent reductions over the same length; partial k of tile b goes to
// P[k * gridDim.x + b].  `stop` (may be null) skips the launch when nonzero.
template <int K>
__global__ void __launch_bounds__(MPK_NT) k_partials(RedSet<K> rs, i64 n, double *P,
                                                     const int64_t *stop) {
  if (stop && *stop) return;
  __shared__ double sh[4];
  bool al = true;
#pragma unroll
  for (int k = 0; k < K; k++) al = al && aligned16(rs.s[k].x) && (rs.s[k].kind != RK_DOT || aligned16(rs.s[k].y));
#pragma unroll
  for (int k = 0; k < K; k++) {
    double u = thread_sum(rs.s[k], n, blockIdx.x, al);
    double r = block_tree(u, sh);
    if (threadIdx.x == 0) P[(i64)k * gridDim.x + blockIdx.x] = r;
  }
}

// x . V_j for j < nv (nv <= G) over a basis with leading dimension ldv.
template <int G>
__global__ void __launch_bounds__(MPK_NT) k_partials_mdot(const double *x, const double *V, i64 ldv,
                                                          int nv, i64 n, double *P,
                                                          const int64_t *stop) {
  if (stop && *stop) return;
  __shared__ double sh[4];
  bool al = aligned16(x) && aligned16(V) && (ldv % 2 == 0);
  const i64 base = (i64)blockIdx.x * MPK_TILE + 2 * (i64)threadIdx.x;
  D2 xs[4];
#pragma unroll
  for (int j = 0; j < 4; j++) xs[j] = ld2(x, base + 512 * j, n, al);
  double u[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    u[g] = 0.0;
    if (g < nv) {
      const double *v = V + (i64)g * ldv;
      double p[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        D2 vv = ld2(v, base + 512 * j, n, al);
        p[j] = fma(xs[j].a, vv.a, xs[j].b * vv.b);
      }
      u[g] = (p[0] + p[1]) + (p[2] + p[3]);
    }
  }
#pragma unroll
  for (int g = 0; g < G; g++) {
    if (g < nv) {  // uniform across the block
      double r = block_tree(u[g], sh);
      if (threadIdx.x == 0) P[(i64)g * gridDim.x + blockIdx.x] = r;
    }
  }
}

// Fold T partials of each of K reductions into res[k] (shared, thread 0).
__device__ __forceinline__ void fold_partials(const double *P, int T, int K, double *res) {
  __shared__ double sh[4];
  for (int k = 0; k < K; k++) {
    const double *Pk = P + (i64)k * T;
    double u = (int)threadIdx.x < T ? Pk[threadIdx.x] : 0.0;
    for (int b = threadIdx.x + MPK_NT; b < T; b += MPK_NT) u = u + Pk[b];
    double r = block_tree(u, sh);
    if (threadIdx.x == 0) res[k] = r;
  }
}

// fold_partials for K reductions at once, in the same order for each: every
// thread issues its loads (all of them, up to 2304 partials) before it adds,
// and the K block trees share one barrier, so a block that needs the results
// pays one round trip.
template <int K>
__device__ __forceinline__ void fold_partials_k(const double *P, int T, double *res) {
  __shared__ double sh[K][4];
  const int t = (int)threadIdx.x;
  // thread t's sum P[t] + P[t + 256] + ... in that order; the partials are
  // loaded kFoldRound at a time, all of a round before any is added, so that
  // up to 256 kFoldRound partials cost one round trip
  constexpr int kFoldRound = 8;
  double u[K];
#pragma unroll
  for (int k = 0; k < K; k++) u[k] = t < T ? P[(i64)k * T + t] : 0.0;
  for (int b0 = t + MPK_NT; b0 < T; b0 += kFoldRound * MPK_NT) {
    double v[kFoldRound][K];
#pragma unroll
    for (int m = 0; m < kFoldRound; m++)
#pragma unroll
      for (int k = 0; k < K; k++) v[m][k] = b0 + m * MPK_NT < T ? P[(i64)k * T + b0 + m * MPK_NT] : 0.0;
#pragma unroll
    for (int m = 0; m < kFoldRound; m++)
#pragma unroll
      for (int k = 0; k < K; k++)
        if (b0 + m * MPK_NT < T) u[k] = u[k] + v[m][k];
  }
  const int lane = t & 63, wave = t >> 6;
#pragma unroll
  for (int k = 0; k < K; k++) {
    u[k] = wave_tree(u[k]);
    if (lane == 0) sh[k][wave] = u[k];
  }
  __syncthreads();
  if (t == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) res[k] = (sh[k][0] + sh[k][2]) + (sh[k][1] + sh[k][3]);
  }
}

struct NoEpilogue {
  __device__ void operator()(const double *) const {}
};

// One block: fold the partials, store the K results to `out` (if given) and
// run the epilogue on them in thread 0.  Up to MPK_FIN_LDS results are kept
// in LDS; more (GMRES's Gram-Schmidt sums past a restart of 255) are folded
// straight into `out`, which the caller must then provide (K doubles).
#define MPK_FIN_LDS 256
template <class Epi>
__global__ void __launch_bounds__(MPK_NT) k_finalize(const double *P, int T, int K, double *out,
                                                     Epi epi, const int64_t *stop) {
  if (stop && *stop) return;
  __shared__ double sres[MPK_FIN_LDS];
  double *const res = K <= MPK_FIN_LDS ? sres : out;
  // one round of loads for the common counts (the same sums, in the same
  // order, as fold_partials)
  if (K == 1)
    fold_partials_k<1>(P, T, res);
  else if (K == 2)
    fold_partials_k<2>(P, T, res);
  else if (K == 3)
    fold_partials_k<3>(P, T, res);
  else
    for (int k0 = 0; k0 < K; k0 += 4) {  // four at a time (GMRES's Gram-Schmidt sums)
      const int kk = K - k0 < 4 ? K - k0 : 4;
      if (kk == 4)
        fold_partials_k<4>(P + (i64)k0 * T, T, res + k0);
      else if (kk == 3)
        fold_partials_k<3>(P + (i64)k0 * T, T, res + k0);
      else if (kk == 2)
        fold_partials_k<2>(P + (i64)k0 * T, T, res + k0);
      else
        fold_partials_k<1>(P + (i64)k0 * T, T, res + k0);
      __syncthreads();  // the shared trees are reused by the next four
    }
  if (threadIdx.x == 0) {
    if (out && res != out)
      for (int k = 0; k < K; k++) out[k] = res[k];
    epi(res);
  }
}

// ------------------------------------------------------------ the reference's order
//
// The order of the reference's SIMD loops (simdvec.mojo:332-473 with SW = 4,
// BLK = 16, simd.mojo:18,32-33; _mdot_basis, ksp.mojo:956-1062), opt-in:
//   the first 16 floor(n / 16) elements go to 16 accumulator chains, element
//   e to chain e % 16, each chain a serial fma(x, y, acc) (an add for sums)
//   in ascending e from +0.0; lane l of the SW-wide accumulator is then
//   ((c_l + c_{4+l}) + c_{8+l}) + c_{12+l}; dot and the unary loops run
//   SW-steps of four over the remainder into those lanes (_mdot_basis does
//   not); the last elements go serially into a scalar s from +0.0; the result
//   is s + ((acc_0 + acc_2) + (acc_1 + acc_3)), Mojo's reduce_add over four
//   lanes.
// A chain is a dependency sequence, so one workgroup runs them all: wave k
// runs the 16 chains of reduction k in its first 16 lanes (up to four
// reductions, on the four SIMDs), from chunks of the operands that all 512
// threads stage through LDS (eight 16-byte loads each), double-buffered: the loads of chunk c + 1 are in
// flight while the chains consume chunk c.  In LDS a chunk of an operand is
// stored transposed, chain c's elements in row c, so that a lane reads two
// consecutive steps of its chain with one 16-byte load; the row length is
// chosen so the 16 lanes' loads fall on distinct banks.
#define MPK_REF_NT 512
#define MPK_REF_IT 8  // 16-byte items per thread per chunk
#define MPK_REF_BUF 8192  // doubles in one chunk buffer (two buffers: 128 KB of LDS)

struct RefJob {
  const double *v[4];            // the distinct operands, staged
  const double *sx[4], *sy[4];   // each reduction's operands (read by the tail)
  unsigned code;                 // reduction k: byte k = xi | yi << 2 | kind << 4 (staged indices)
  int nvec, nspec;               // nvec <= 4, nspec <= 4
  int swsteps;                   // 1: dot / unary loops; 0: _mdot_basis (no SW-steps)
  int al;                        // every operand 16-byte aligned
};

// Operand of index i (i < 4) without dynamic indexing.
__device__ __forceinline__ const double *ref_vec(const RefJob &J, int i) {
  const double *p = J.v[0];
  p = i == 1 ? J.v[1] : p;
  p = i == 2 ? J.v[2] : p;
  p = i == 3 ? J.v[3] : p;
  return p;
}

// One chain step: acc = fma(f, g, acc) with (f, g) = (x, y) for a dot, (x, x)
// for a sum of squares, (x, 1) for a sum and (|x|, 1) for an abs-sum.
template <int KIND>
__device__ __forceinline__ double ref_step(double x, double y, double acc) {
  if constexpr (KIND == RK_DOT) return fma(x, y, acc);
  if constexpr (KIND == RK_SUMSQ) return fma(x, x, acc);
  if constexpr (KIND == RK_ASUM) return fma(fabs(x), 1.0, acc);
  return fma(x, 1.0, acc);
}
__device__ __forceinline__ double ref_step_rt(int kind, double x, double y, double acc) {
  switch (kind) {
    case RK_DOT: return ref_step<RK_DOT>(x, y, acc);
    case RK_SUMSQ: return ref_step<RK_SUMSQ>(x, y, acc);
    case RK_ASUM: return ref_step<RK_ASUM>(x, y, acc);
    default: return ref_step<RK_SUM>(x, y, acc);
  }
}

// nblk steps of one chain from its LDS rows xr, yr (16-byte aligned); the
// next eight steps' loads are issued before the current eight are summed.
template <int KIND>
__device__ __forceinline__ double ref_chain(const double *xr, const double *yr, int nblk, double acc) {
  int b = 0;
  double2 xa[4], ya[4];
  if (nblk >= 8)
#pragma unroll
    for (int u = 0; u < 4; u++) {
      xa[u] = *reinterpret_cast<const double2 *>(xr + 2 * u);
      if constexpr (KIND == RK_DOT) ya[u] = *reinterpret_cast<const double2 *>(yr + 2 * u);
    }
  for (; b + 8 <= nblk; b += 8) {
    double2 xb[4], yb[4];
    const bool more = b + 16 <= nblk;
    if (more)
#pragma unroll
      for (int u = 0; u < 4; u++) {
        xb[u] = *reinterpret_cast<const double2 *>(xr + b + 8 + 2 * u);
        if constexpr (KIND == RK_DOT) yb[u] = *reinterpret_cast<const double2 *>(yr + b + 8 + 2 * u);
      }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      double y0 = 0.0, y1 = 0.0;
      if constexpr (KIND == RK_DOT) y0 = ya[u].x, y1 = ya[u].y;
      acc = ref_step<KIND>(xa[u].x, y0, acc);
      acc = ref_step<KIND>(xa[u].y, y1, acc);
    }
    if (more)
#pragma unroll
      for (int u = 0; u < 4; u++) {
        xa[u] = xb[u];
        if constexpr (KIND == RK_DOT) ya[u] = yb[u];
      }
  }
  for (; b < nblk; b++) acc = ref_step<KIND>(xr[b], KIND == RK_DOT ? yr[b] : 0.0, acc);
  return acc;
}

__device__ __forceinline__ void ref_run(const RefJob J, i64 n, double *out) {
  __shared__ __align__(16) double lds[2 * MPK_REF_BUF];
  __shared__ double chv[4][16];
  const int t = (int)threadIdx.x, lane = t & 63, wave = t >> 6;
  const int nvec = J.nvec;
  // blocks per chunk = the row length, in doubles: nvec * 16 * cb <=
  // MPK_REF_BUF (so also nvec * 8 * cb <= MPK_REF_IT * MPK_REF_NT), and
  // cb = 2 mod 32 so that row c starts 4 c dwords (mod 64) into the banks and
  // the 16 lanes' 16-byte reads hit 64 distinct banks
  const int cb = ((MPK_REF_BUF / 16 / nvec - 2) & ~31) + 2, rw = cb, per = 8 * cb, items = nvec * per;
  const i64 nb = n / 16, lim = 16 * nb;
  const i64 nchunks = (nb + cb - 1) / cb;
  const bool chain = wave < J.nspec && lane < 16;
  const int c = lane & 15;
  const unsigned code = (J.code >> (8 * (wave & 3))) & 0xffu;
  const int xi = (int)(code & 3u), yi = (int)((code >> 2) & 3u), kind = (int)(code >> 4);
  // each thread's MPK_REF_IT 16-byte items of a chunk, fixed across chunks:
  // the operand, the pair's place in the chunk and its LDS offset
  const double *src[MPK_REF_IT];
  int el[MPK_REF_IT], roff[MPK_REF_IT];
#pragma unroll
  for (int j = 0; j < MPK_REF_IT; j++) {
    const int idx = t + MPK_REF_NT * j;
    const int vi = idx < items ? idx / per : 0, off = idx - vi * per;
    el[j] = idx < items ? 2 * off : -1;  // chain el % 16 (even), block el / 16
    roff[j] = vi * 16 * rw + ((2 * off) & 15) * rw + ((2 * off) >> 4);
    src[j] = ref_vec(J, vi);
  }
  double2 reg[MPK_REF_IT];
  auto load = [&](i64 chunk) {
#pragma unroll
    for (int j = 0; j < MPK_REF_IT; j++) {
      double2 r = make_double2(0.0, 0.0);
      const i64 e = chunk * 16 * cb + el[j];
      if (el[j] >= 0 && e < lim) {  // e even and lim a multiple of 16: e + 1 < lim too
        const double *p = src[j] + e;
        if (J.al) {
          r = *reinterpret_cast<const double2 *>(p);
        } else {
          r.x = p[0];
          r.y = p[1];
        }
      }
      reg[j] = r;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < MPK_REF_IT; j++)
      if (el[j] >= 0) {
        lds[buf * MPK_REF_BUF + roff[j]] = reg[j].x;
        lds[buf * MPK_REF_BUF + roff[j] + rw] = reg[j].y;  // chain el % 16 + 1
      }
  };
  double acc = 0.0;
  if (nchunks > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (i64 ch = 0; ch < nchunks; ch++) {
    const int buf = (int)(ch & 1);
    if (ch + 1 < nchunks) load(ch + 1);
    if (chain) {
      const double *xr = lds + buf * MPK_REF_BUF + xi * 16 * rw + c * rw;
      const double *yr = lds + buf * MPK_REF_BUF + yi * 16 * rw + c * rw;
      const i64 left = nb - ch * cb;
      const int nblk = left < cb ? (int)left : cb;
      switch (kind) {  // uniform in the wave
        case RK_DOT: acc = ref_chain<RK_DOT>(xr, yr, nblk, acc); break;
        case RK_SUMSQ: acc = ref_chain<RK_SUMSQ>(xr, yr, nblk, acc); break;
        case RK_ASUM: acc = ref_chain<RK_ASUM>(xr, yr, nblk, acc); break;
        default: acc = ref_chain<RK_SUM>(xr, yr, nblk, acc); break;
      }
    }
    if (ch + 1 < nchunks) store(buf ^ 1);
    __syncthreads();
  }
  if (chain) chv[wave][c] = acc;
  __syncthreads();
  // the tail of each reduction, in thread 0 (at most 15 elements each)
  if (t == 0) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (r >= J.nspec) break;
      const int tkind = (int)((J.code >> (8 * r + 4)) & 0xfu);
      const double *x = J.sx[r], *y = J.sy[r];
      double a4[4];
#pragma unroll
      for (int l = 0; l < 4; l++) a4[l] = ((chv[r][l] + chv[r][4 + l]) + chv[r][8 + l]) + chv[r][12 + l];
      i64 i = lim;
      if (J.swsteps)
        for (; i + 4 <= n; i += 4)
#pragma unroll
          for (int l = 0; l < 4; l++) a4[l] = ref_step_rt(tkind, x[i + l], y[i + l], a4[l]);
      double s = 0.0;
      for (; i < n; i++) s = ref_step_rt(tkind, x[i], y[i], s);
      out[r] = s + ((a4[0] + a4[2]) + (a4[1] + a4[3]));
    }
  }
}

__global__ void __launch_bounds__(MPK_REF_NT) k_ref_red(RefJob J, i64 n, double *P, const int64_t *stop);

// x . V_j, j < nv, in _mdot_basis's order: block b takes rows 3b .. 3b + 2
// (a row's result does not depend on how the rows are grouped).
__global__ void __launch_bounds__(MPK_REF_NT) k_ref_mdot(const double *x, const double *V, i64 ldv, int nv, i64 n,
                                                         double *P, const int64_t *stop);

// The job for a set of K <= 4 reductions: distinct operands staged once.
template <int K>
RefJob ref_job(const RedSet<K> &rs) {
  static_assert(K <= 4, "at most four reductions in one pass");
  RefJob J{};
  J.swsteps = 1;
  J.nspec = K;
  J.al = 1;
  auto slot = [&](const double *p) {
    for (int i = 0; i < J.nvec; i++)
      if (J.v[i] == p) return i;
    if (J.nvec == 4) {  // no caller stages more than two
      set_error(4, "reference-order reduction over more than four distinct vectors");
      return 0;
    }
    J.v[J.nvec] = p;
    J.al = J.al && aligned16(p);
    return J.nvec++;
  };
  for (int k = 0; k < K; k++) {
    const int kind = rs.s[k].kind;
    const int xi = slot(rs.s[k].x);
    const int yi = kind == RK_DOT ? slot(rs.s[k].y) : xi;
    J.code |= (unsigned)(xi | yi << 2 | kind << 4) << (8 * k);
    J.sx[k] = rs.s[k].x;
    J.sy[k] = kind == RK_DOT ? rs.s[k].y : rs.s[k].x;
  }
  return J;
}

void ref_launch(const RefJob &J, i64 n, double *P, const int64_t *stop);

// Host side: launch the partials for the K specs and the finalize with `epi`.
// (A last-block finalize inside the partials launch needs a device-scope fence
// per block, an L2 writeback on this GPU, which costs more than the launch.)
// In the reference's order the one workgroup writes each result as a single
// partial, which the finalize then folds with zeros (x + 0.0 is x: the
// reference order never yields -0.0, every chain starting from +0.0).
template <int K, class Epi>
void reduce_launch(const RedSet<K> &rs, i64 n, double *out, Epi epi, const int64_t *stop) {
  if (red_ref()) {
    double *P = partials((size_t)K);
    if (!P) return;
    const RefJob J = ref_job<K>(rs);
    if (failed()) return;
    ref_launch(J, n, P, stop);
    MPK_LAUNCH(k_finalize<Epi>, 1, MPK_NT, P, 1, K, out, epi, stop);
    return;
  }
  unsigned T = ntiles(n);
  if (T == 0) T = 1;  // n == 0 still yields a (zero) result
  double *P = partials((size_t)K * T);
  if (!P) return;
  MPK_LAUNCH(k_partials<K>, T, MPK_NT, rs, n, P, stop);
  MPK_LAUNCH(k_finalize<Epi>, 1, MPK_NT, P, (int)T, K, out, epi, stop);
}

}  // namespace mpk
