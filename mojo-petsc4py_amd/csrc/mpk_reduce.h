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
// run the epilogue on them in thread 0.
template <class Epi>
__global__ void __launch_bounds__(MPK_NT) k_finalize(const double *P, int T, int K, double *out,
                                                     Epi epi, const int64_t *stop) {
  if (stop && *stop) return;
  __shared__ double res[256];
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
    if (out)
      for (int k = 0; k < K; k++) out[k] = res[k];
    epi(res);
  }
}

// Host side: launch the partials for the K specs and the finalize with `epi`.
// (A last-block finalize inside the partials launch needs a device-scope fence
// per block, an L2 writeback on this GPU, which costs more than the launch.)
template <int K, class Epi>
void reduce_launch(const RedSet<K> &rs, i64 n, double *out, Epi epi, const int64_t *stop) {
  unsigned T = ntiles(n);
  if (T == 0) T = 1;  // n == 0 still yields a (zero) result
  double *P = partials((size_t)K * T);
  if (!P) return;
  MPK_LAUNCH(k_partials<K>, T, MPK_NT, rs, n, P, stop);
  MPK_LAUNCH(k_finalize<Epi>, 1, MPK_NT, P, (int)T, K, out, epi, stop);
}

}  // namespace mpk
