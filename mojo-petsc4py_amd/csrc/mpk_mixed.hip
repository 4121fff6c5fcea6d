// mpk_mixed.hip -- mixed-precision solves: the inner Krylov iteration, its
// SpMV and its preconditioner run on FP32 vectors (matrix values in FP32, or in
// FP16 with one FP32 scale per row, FP32 accumulation), inside FP64 iterative
// refinement.  The same solver also runs with FP64 vectors and values (the
// precision's own control: identical inner solver, preconditioner form and
// refinement, so the time ratio of the two is the gain of the precision).  FP64 stays where the accuracy is decided: the true residual
// b - A x, the solution update, every dot product and norm (FP32 data summed
// in FP64), the scalars they feed and the GMRES Hessenberg / Givens least
// squares.  This is a form beside the drop-in path; it does not reproduce the
// reference's arithmetic (the FP64 solvers do), only its answer to the
// refinement tolerance.
//
// Data: the matrix's sliced copy (mpk_sell_build) gives the layout; the low
// precision values sit in the same slots.  The preconditioner's FP64 block
// plans (mpk_pc_schedule) give the sweeps' structure; a low-precision copy of
// their coefficients sits in the same slots.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "mpk_internal.h"
#include "mpk_ops.h"
#include "mpk_pc.h"
#include "mpk_sweep.h"
#include "../../include/mpk.h"

namespace mpk {

// Stored values (CT: __half, float or double) as the inner solve's type VT
// (float, or double with double values), and back.
template <class VT>
__device__ __forceinline__ VT lp_val(float v) {
  return (VT)v;
}
template <class VT>
__device__ __forceinline__ VT lp_val(__half v) {
  return (VT)__half2float(v);
}
template <class VT>
__device__ __forceinline__ VT lp_val(double v) {
  return (VT)v;
}
template <class CT>
__device__ __forceinline__ CT lp_from(double v);
template <>
__device__ __forceinline__ double lp_from<double>(double v) {
  return v;
}
template <>
__device__ __forceinline__ float lp_from<float>(double v) {
  return (float)v;
}
template <>
__device__ __forceinline__ __half lp_from<__half>(double v) {
  return __float2half((float)v);
}
__device__ __forceinline__ float lp_fma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double lp_fma(double a, double b, double c) { return fma(a, b, c); }

// ------------------------------------------------------------ matrix

// Low-precision values in the sliced layout; with rs, value = rs[i] * h and
// rs[i] = max_j |a_ij| (1 for an all-zero row), so |h| <= 1.
template <class CT>
__global__ void k_lp_sell(i64 n, const int *rp, const int *soff, const double *sv, CT *lv, float *rs) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int len = rp[i + 1] - rp[i], base = soff[i >> 6] + (int)(i & 63);
  double s = 1.0;
  if (rs) {
    double m = 0.0;
    for (int k = 0; k < len; k++) m = fmax(m, fabs(sv[base + 64 * k]));
    s = m > 0.0 ? m : 1.0;
    rs[i] = (float)s;
  }
  for (int k = 0; k < len; k++) lv[base + 64 * k] = lp_from<CT>(sv[base + 64 * k] / s);
}

// Windowed products.  The sliced matrices of the inner solve are read in
// blocks of LP_WROWS rows.  When the columns one block reads span fewer than
// lp_xwin entries of x (matrices whose entries stay near the diagonal), the
// block first copies that window of x into LDS with coalesced loads and
// gathers from there; otherwise it gathers from x directly.  The values
// gathered, and so the results, are the same either way.
#define LP_WROWS 512
#define LP_XWIN_BYTES 49152
// entries of x a block's window holds: 12288 FP32, 6144 FP64
template <class VT>
__host__ __device__ constexpr int lp_xwin() {
  return LP_XWIN_BYTES / (int)sizeof(VT);
}

// Column window of each block: win[2b], win[2b+1] = the smallest and largest
// column read by block b's rows (0, -1 when they read none).
__global__ void __launch_bounds__(LP_WROWS) k_col_window(i64 n, const int *rp, const int *soff, const int *sc,
                                                         int *win) {
  __shared__ int lo_sh, hi_sh;
  if (threadIdx.x == 0) {
    lo_sh = 0x7fffffff;
    hi_sh = -1;
  }
  __syncthreads();
  const i64 i = (i64)blockIdx.x * LP_WROWS + threadIdx.x;
  int lo = 0x7fffffff, hi = -1;
  if (i < n) {
    const int len = rp[i + 1] - rp[i], base = soff[i >> 6] + (int)(i & 63);
    for (int k = 0; k < len; k++) {
      const int c = sc[base + 64 * k];
      lo = c < lo ? c : lo;
      hi = c > hi ? c : hi;
    }
  }
  if (hi >= 0) {
    atomicMin(&lo_sh, lo);
    atomicMax(&hi_sh, hi);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    win[2 * blockIdx.x] = hi_sh >= 0 ? lo_sh : 0;
    win[2 * blockIdx.x + 1] = hi_sh;
  }
}

// Copies block b's window of x into xs when it fits; returns whether it did
// (every thread of the block calls this).
template <class VT>
__device__ __forceinline__ bool stage_window(const VT *__restrict__ x, const int *__restrict__ win, unsigned b,
                                             VT *xs, int &lo) {
  lo = win[2 * b];
  const int hi = win[2 * b + 1];
  const bool fit = hi - lo < lp_xwin<VT>();
  if (fit)
    for (int c = lo + (int)threadIdx.x; c <= hi; c += LP_WROWS) xs[c - lo] = x[c];
  __syncthreads();
  return fit;
}

// Row i's serial sum over its sliced entries, eight entries' loads issued
// before their values are gathered (from xs when the window is staged), and
// all eight gathers before the first product: a row of up to eight entries
// costs two dependent rounds of loads.
template <class CT, class VT>
__device__ __forceinline__ VT sliced_row(const int *__restrict__ sc, const CT *__restrict__ sv, int base, int len,
                                         const VT *__restrict__ x, const VT *xs, bool fit, int lo) {
  constexpr int B = 8;
  VT s = VT(0);
  for (int k0 = 0; k0 < len; k0 += B) {
    int c[B];
    VT v[B], xv[B];
#pragma unroll
    for (int e = 0; e < B; e++) {
      const int p = base + 64 * (k0 + e);
      c[e] = k0 + e < len ? sc[p] : 0;
      v[e] = k0 + e < len ? lp_val<VT>(sv[p]) : VT(0);
    }
#pragma unroll
    for (int e = 0; e < B; e++) xv[e] = k0 + e < len ? (fit ? xs[c[e] - lo] : x[c[e]]) : VT(0);
#pragma unroll
    for (int e = 0; e < B; e++)
      if (k0 + e < len) s = lp_fma(v[e], xv[e], s);
  }
  return s;
}

// WIN: the kernel stages windows (matrices whose blocks' windows fit LDS);
// else it gathers from x with no LDS at all.
template <class CT, class VT, bool WIN>
__global__ void __launch_bounds__(LP_WROWS) k_spmv_lp(i64 n, const int *__restrict__ rp, const int *__restrict__ soff,
                                                      const int *__restrict__ sc, const CT *__restrict__ sv,
                                                      const float *__restrict__ rs, const int *__restrict__ win,
                                                      const VT *__restrict__ x, VT *__restrict__ y) {
  __shared__ VT xs[WIN ? lp_xwin<VT>() : 1];
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  int lo = 0;
  const bool fit = WIN ? stage_window(x, win, b, xs, lo) : false;
  const i64 i = (i64)b * LP_WROWS + threadIdx.x;
  if (i >= n) return;
  const VT s = sliced_row<CT, VT>(sc, sv, soff[i >> 6] + (int)(i & 63), rp[i + 1] - rp[i], x, xs, fit, lo);
  y[i] = rs ? s * (VT)rs[i] : s;
}

// ------------------------------------------------------------ inner vectors (VT)

static unsigned lp_grid(i64 n) {
  const i64 g = (n + 255) / 256;
  return (unsigned)(g < 1 ? 1 : g);
}

// out = (VT)(r * scale[0])
template <class VT>
__global__ void k_lp_from_d(i64 n, const double *r, const double *scale, VT *out) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = (VT)(r[i] * scale[0]);
}
// x += scale[0] * d (FP64)
template <class VT>
__global__ void k_lp_update(i64 n, double *x, const double *scale, const VT *d) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] = fma(scale[0], (double)d[i], x[i]);
}
// y = x * s[0]
template <class VT>
__global__ void k_lp_scale(i64 n, const VT *x, const double *s, VT *y, const int *skip) {
  if (skip && *skip) return;
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = x[i] * (VT)s[0];
}
// y += sum_j coef[j] V_j, j < nv (V with leading dimension ld)
template <class VT>
__global__ void k_lp_maxpy(i64 n, VT *y, const VT *V, i64 ld, const double *coef, int nv, const int *skip) {
  if (skip && *skip) return;
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  VT a = y[i];
  for (int j = 0; j < nv; j++) a = lp_fma((VT)coef[j], V[(i64)j * ld + i], a);
  y[i] = a;
}
template <class VT>
__global__ void k_lp_zero(i64 n, VT *y) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = VT(0);
}
// z = x - z
template <class VT>
__global__ void k_lp_rsub(i64 n, const VT *x, VT *z) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) z[i] = x[i] - z[i];
}
template <class VT>
__global__ void k_lp_copy(i64 n, const VT *x, VT *y) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = x[i];
}

// ------------------------------------------------------------ reductions
// VT data summed in FP64: per-block partials, then one block per result.
#define LP_RBLK 512

// P[k * LP_RBLK + b] = sum over block b's share of x_k . y_k (y_k null: x_k . x_k)
template <int K, class VT>
struct LpDots {
  const VT *x[K];
  const VT *y[K];
};
template <int K, class VT>
__global__ void __launch_bounds__(256) k_lp_dots(LpDots<K, VT> d, i64 n, double *P) {
  __shared__ double sh[4];
  double a[K];
#pragma unroll
  for (int k = 0; k < K; k++) a[k] = 0.0;
  for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double xv = d.x[k][i];
      a[k] = fma(xv, d.y[k] ? (double)d.y[k][i] : xv, a[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    const double r = block_tree(a[k], sh);
    if (threadIdx.x == 0) P[(i64)k * gridDim.x + blockIdx.x] = r;
  }
}
// Four consecutive elements of a VT vector (16-byte aligned) in one or two
// 16-byte loads, as doubles.
struct Q4 {
  double x, y, z, w;
};
__device__ __forceinline__ Q4 load_q4(const float *p) {
  const float4 v = *reinterpret_cast<const float4 *>(p);
  return Q4{(double)v.x, (double)v.y, (double)v.z, (double)v.w};
}
__device__ __forceinline__ Q4 load_q4(const double *p) {
  const double2 a = *reinterpret_cast<const double2 *>(p), b = *reinterpret_cast<const double2 *>(p + 2);
  return Q4{a.x, a.y, b.x, b.y};
}

// P[j * LP_RBLK + b] = block b's share of V_j . w, j < nv.  Block b takes one
// contiguous span of the vectors, in groups of 4 elements (V_j at V + j ld,
// ld and the vectors 16-byte aligned: ld a multiple of 8); the last n % 4
// elements go to block 0.  Eight vectors at a time: a thread loads its
// group of w and of all eight (vectors past nv repeat the last one and are
// not stored) before any product, so every load of a step is in flight at
// once.
template <class VT>
__global__ void __launch_bounds__(256) k_lp_mdot(const VT *V, i64 ld, int nv, const VT *w, i64 n, double *P,
                                                 const int *skip) {
  if (skip && *skip) return;
  __shared__ double sh[4];
  const i64 n4 = n / 4, per = (n4 + gridDim.x - 1) / gridDim.x;
  const i64 g0 = (i64)blockIdx.x * per, g1 = g0 + per < n4 ? g0 + per : n4;
  for (int j0 = 0; j0 < nv; j0 += 8) {
    const VT *Vj[8];
#pragma unroll
    for (int k = 0; k < 8; k++) Vj[k] = V + (i64)(j0 + k < nv ? j0 + k : nv - 1) * ld;
    double a[8];
#pragma unroll
    for (int k = 0; k < 8; k++) a[k] = 0.0;
    for (i64 g = g0 + threadIdx.x; g < g1; g += 256) {
      const Q4 wv = load_q4(w + 4 * g);
      Q4 v[8];
#pragma unroll
      for (int k = 0; k < 8; k++) v[k] = load_q4(Vj[k] + 4 * g);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        a[k] = fma(v[k].x, wv.x, a[k]);
        a[k] = fma(v[k].y, wv.y, a[k]);
        a[k] = fma(v[k].z, wv.z, a[k]);
        a[k] = fma(v[k].w, wv.w, a[k]);
      }
    }
    if (blockIdx.x == 0)
      for (i64 i = 4 * n4 + threadIdx.x; i < n; i += 256) {
        const double wi = (double)w[i];
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] = fma((double)Vj[k][i], wi, a[k]);
      }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const double r = block_tree(a[k], sh);
      if (threadIdx.x == 0 && j0 + k < nv) P[(i64)(j0 + k) * gridDim.x + blockIdx.x] = r;
    }
  }
}
// out[v] = sum_b P[v * nb + b], one block per v
__global__ void __launch_bounds__(256) k_lp_fold(const double *P, int nb, double *out, const int *skip) {
  if (skip && *skip) return;
  __shared__ double sh[4];
  double a = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) a += P[(i64)blockIdx.x * nb + b];
  const double r = block_tree(a, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

// ------------------------------------------------------------ sweeps

// Sentinel-fills a sweep's published values and zeroes its block ticket and
// ring progress word.
__device__ __forceinline__ float lp_sentinel(float) { return __uint_as_float(MPK_SENTINEL_F); }
__device__ __forceinline__ double lp_sentinel(double) { return __longlong_as_double((long long)MPK_SENTINEL); }
template <class VT>
__global__ void k_lp_prepare(i64 n, VT *pub, unsigned *ticket) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i == 0) {
    ticket[0] = 0u;
    ticket[1] = 0u;
  }
  if (i < n) pub[i] = lp_sentinel(VT(0));
}

template <class VT>
struct alignas(8) LpStaged {
  VT init, s, d1, d2;
  i64 o;
};
// The row programs of the inner solve add their sum to the staged start in fin.
template <class VT>
__device__ __forceinline__ VT row_start(const LpStaged<VT> &) {
  return VT(0);
}

// The sweep engine's value type for the inner solve's vectors.
template <class VT>
struct LpSweep;
template <>
struct LpSweep<float> {
  using T = SweepF32;
};
template <>
struct LpSweep<double> {
  using T = SweepF64;
};

// Plan coefficients from the solver's copy (slot-major like the plan's own,
// divided by the row scale where there is one).
template <class CT, class VT>
struct LpCoef {
  using Raw = CT;
  const CT *c;
  static constexpr int kBytes = sizeof(CT);
  __device__ static LpCoef make(const PlanView &, const void *lc) { return LpCoef{(const CT *)lc}; }
  __device__ VT operator()(i64 q) const { return lp_val<VT>(c[q]); }
  __device__ Raw raw(i64 q) const { return c[q]; }
  __device__ static VT conv(Raw r) { return lp_val<VT>(r); }
};

// The row programs of the inner solve's sweeps: acc = sum of coef * value in
// VT (coefficients divided by the row scale where there is one), then fin.
template <class VT>
struct LpIluFwd {
  const int *r;
  const VT *b;
  __device__ LpStaged<VT> stage(i64 i) const { return LpStaged<VT>{b[r[i]], VT(1), VT(0), VT(0), 0}; }
  __device__ VT fin(i64, VT acc, const LpStaged<VT> &s) const { return s.init + acc; }
  __device__ VT stale(i64) const { return VT(0); }
};
template <class VT>
struct LpIluBwd {
  const int *c;
  const VT *F, *dinv;
  VT *x;
  __device__ LpStaged<VT> stage(i64 i) const { return LpStaged<VT>{F[i], VT(1), dinv[i], VT(0), c[i]}; }
  __device__ VT fin(i64, VT acc, const LpStaged<VT> &s) const {
    const VT v = (s.init + acc) * s.d1;
    x[s.o] = v;
    return v;
  }
  __device__ VT stale(i64) const { return VT(0); }
};
template <class VT>
struct LpSorFwd {
  const VT *x;
  const float *rs;
  const VT *idiag;
  VT *t, *yout;
  const VT *yin;
  __device__ LpStaged<VT> stage(i64 i) const {
    return LpStaged<VT>{x[i], rs ? (VT)rs[i] : VT(1), idiag[i], VT(0), 0};
  }
  __device__ VT fin(i64 i, VT acc, const LpStaged<VT> &s) const {
    const VT tv = lp_fma(s.s, acc, s.init);
    t[i] = tv;
    const VT v = tv * s.d1;
    if (yout) yout[i] = v;
    return v;
  }
  __device__ VT stale(i64 k) const { return yin[k]; }
};
template <class VT>
struct LpSorBwd {
  const VT *src;
  const float *rs;
  const VT *idiag, *yf;
  VT *y;
  VT omega;
  int use_t;
  __device__ LpStaged<VT> stage(i64 i) const {
    return LpStaged<VT>{src[i], rs ? (VT)rs[i] : VT(1), idiag[i], use_t ? yf[i] : VT(0), 0};
  }
  __device__ VT fin(i64 i, VT acc, const LpStaged<VT> &s) const {
    const VT tv = lp_fma(s.s, acc, s.init);
    const VT v = use_t ? lp_fma(VT(1) - omega, s.d2, tv * s.d1) : tv * s.d1;
    y[i] = v;
    return v;
  }
  __device__ VT stale(i64 k) const { return y[k]; }
};

// Approximate sweeps: k Jacobi passes over a sweep's entries (a truncated
// Neumann series of the triangular solve) instead of the ordered sweep.  Pass
// 1 (u_in null) keeps only the row's own term; pass j adds the entries' values
// from pass j - 1.  Every pass is one bandwidth-bound sweep over the sliced
// entries, with no ordering between rows.
template <class Prog, class CT, class VT, bool WIN>
__global__ void __launch_bounds__(LP_WROWS) k_tri_jacobi(Prog prog, i64 n, const int *__restrict__ rp,
                                                         const int *__restrict__ soff, const int *__restrict__ sc,
                                                         const CT *__restrict__ sv, const int *__restrict__ win,
                                                         const VT *__restrict__ u_in, VT *__restrict__ u_out) {
  __shared__ VT xs[WIN ? lp_xwin<VT>() : 1];
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  int lo = 0;
  const bool fit = WIN && u_in ? stage_window(u_in, win, b, xs, lo) : false;  // uniform over the block
  const i64 i = (i64)b * LP_WROWS + threadIdx.x;
  if (i >= n) return;
  const LpStaged<VT> st = prog.stage(i);
  const VT acc = u_in ? sliced_row<CT, VT>(sc, sv, soff[i >> 6] + (int)(i & 63), rp[i + 1] - rp[i], u_in, xs, fit,
                                           lo)
                      : VT(0);
  u_out[i] = prog.fin(i, acc, st);
}
__global__ void k_f2h(i64 m, const float *x, __half *y) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < m) y[i] = __float2half(x[i]);
}

// A sweep's entries in the sliced layout (for k_tri_jacobi).
struct TriSell {
  int *rp = nullptr, *soff = nullptr, *sc = nullptr, *win = nullptr;
  void *sv = nullptr;
  bool windowed = false;
};

// Column windows of a sliced matrix into win [D]; returns whether the
// windowed kernels pay: at least 7 in 8 blocks' windows fit LDS (xwin
// entries of x).
static bool col_windows(i64 n, const int *rp, const int *soff, const int *sc, int *win, int xwin) {
  const unsigned nb = nblocks(n, LP_WROWS);
  MPK_LAUNCH(k_col_window, nb, LP_WROWS, n, rp, soff, sc, win);
  std::vector<int> h(2 * (size_t)nb);
  copy_dtoh(h.data(), win, h.size() * sizeof(int));
  if (failed()) return false;
  i64 fit = 0;
  for (unsigned b = 0; b < nb; b++) fit += h[2 * b + 1] - h[2 * b] < xwin;
  return 8 * fit >= 7 * (i64)nb;
}

// The inner solve's coefficients of one sweep's plan: lc[slot] = coef / row scale.
template <class CT>
__global__ void k_lp_plan_coef(i64 n, const char *plan, int second, const float *rs, CT *lc) {
  if (second) plan += reinterpret_cast<const i64 *>(plan)[4];
  const PlanView P = plan_view(plan, n);
  const i64 p = (i64)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const double s = rs ? (double)rs[P.prow[p]] : 1.0;
  const i64 base = P.w ? p : P.pent[p], stride = P.w ? n : 1;
  for (int j = 0; j < P.pcnt[p]; j++) {
    const i64 q = base + (i64)j * stride;
    lc[q] = lp_from<CT>(P.ecoef[q] / s);
  }
}

// Inverted diagonals in VT: ILU's stored 1/u_ii; SOR's as MatInvertDiagonalForSOR.
template <class VT>
__global__ void k_lp_ilu_dinv(i64 n, const double *ba, const int *bdiag, VT *dinv) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dinv[i] = (VT)ba[bdiag[i]];
}
template <class VT>
__global__ void k_lp_sor_idiag(i64 n, const double *a, const int *dmark, double omega, double fshift, VT *idiag) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double d = a[dmark[i]];
  idiag[i] = (VT)((omega == 1.0 && fshift <= 0.0) ? 1.0 / d : omega / (fshift + d));
}

static i64 plan_slots(const char *plan_d, int second, i64 n) {
  i64 h[8];
  copy_dtoh(h, plan_d, sizeof h);
  if (second) copy_dtoh(h, plan_d + h[4], sizeof h);
  return h[5] ? n * h[5] : h[1];
}

// ------------------------------------------------------------ the solver

enum { LP_GMRES = 0, LP_BCGS = 1 };
#define LP_MMAX 64

// Device state of the inner iterations (FP64 scalars).
struct LpState {
  double H[(LP_MMAX + 1) * LP_MMAX];  // H(i, j) at j (LP_MMAX + 1) + i
  double cs[LP_MMAX], sn[LP_MMAX], g[LP_MMAX + 1], y[LP_MMAX];
  double h[LP_MMAX + 1], h2[LP_MMAX + 1], coef[LP_MMAX + 1];
  double red[4];
  double beta, inv, res, wnorm;
  double rho, rho_old, alpha, omega, bcoef;
  int norefine, breakdown;
};

// GMRES: start of a cycle, red[0] = w . w
__global__ void k_gm_begin(LpState *s) {
  s->beta = sqrt(s->red[0]);
  s->g[0] = s->beta;
  s->res = s->beta;
  s->inv = s->beta > 0.0 ? 1.0 / s->beta : 0.0;
  s->breakdown = s->beta > 0.0 ? 0 : 1;
}
// coef = -src (the Gram-Schmidt subtraction), unless *skip
__global__ void k_gm_neg(LpState *s, int nv, const double *src, const int *skip) {
  if (skip && *skip) return;
  for (int j = 0; j < nv; j++) s->coef[j] = -src[j];
}
// after the first Gram-Schmidt pass, red[0] = w . w: the second pass runs
// always (refine 2), never (0), or when ||w|| < ||h|| (1, "if needed")
__global__ void k_gm_decide(LpState *s, int nv, int refine) {
  double hn = 0.0;
  for (int j = 0; j < nv; j++) hn = fma(s->h[j], s->h[j], hn);
  s->wnorm = sqrt(s->red[0]);
  s->norefine = !(refine == 2 || (refine == 1 && s->wnorm < sqrt(hn)));
}
// after the (optional) second pass: h += h2, new norm in red[0]
__global__ void k_gm_column(LpState *s, int k) {
  const int nv = k + 1;
  if (!s->norefine) {
    for (int j = 0; j < nv; j++) s->h[j] += s->h2[j];
    s->wnorm = sqrt(s->red[0]);
  }
  double *col = s->H + (i64)k * (LP_MMAX + 1);
  for (int j = 0; j < nv; j++) col[j] = s->h[j];
  col[nv] = s->wnorm;
  for (int j = 0; j < k; j++) {  // previous rotations
    const double a = col[j], b = col[j + 1];
    col[j] = s->cs[j] * a + s->sn[j] * b;
    col[j + 1] = -s->sn[j] * a + s->cs[j] * b;
  }
  const double a = col[k], b = col[k + 1], r = sqrt(a * a + b * b);
  if (r == 0.0) {
    s->breakdown = 1;
    s->res = 0.0;
    s->inv = 0.0;
    return;
  }
  s->cs[k] = a / r;
  s->sn[k] = b / r;
  col[k] = r;
  col[k + 1] = 0.0;
  s->g[k + 1] = -s->sn[k] * s->g[k];
  s->g[k] = s->cs[k] * s->g[k];
  s->res = fabs(s->g[k + 1]);
  s->inv = s->wnorm > 0.0 ? 1.0 / s->wnorm : 0.0;
  if (s->wnorm == 0.0) s->breakdown = 1;
}
// least squares: back substitution on the rotated H, y = R^-1 g (kk columns)
__global__ void k_gm_solve(LpState *s, int kk) {
  for (int i = kk - 1; i >= 0; i--) {
    double t = s->g[i];
    for (int j = i + 1; j < kk; j++) t -= s->H[(i64)j * (LP_MMAX + 1) + i] * s->y[j];
    const double d = s->H[(i64)i * (LP_MMAX + 1) + i];
    s->y[i] = d != 0.0 ? t / d : 0.0;
  }
}

// BiCGStab (right preconditioned) scalar steps
__global__ void k_bc_init(LpState *s) {
  s->rho_old = s->alpha = s->omega = 1.0;
  s->res = sqrt(s->red[0]);
  s->beta = s->res;
  s->breakdown = s->res > 0.0 ? 0 : 1;
}
__global__ void k_bc_rho(LpState *s) {  // red[0] = rhat . r
  const double rho = s->red[0];
  if (rho == 0.0 || s->omega == 0.0) s->breakdown = 1;
  s->bcoef = (s->rho_old != 0.0 && s->omega != 0.0) ? (rho / s->rho_old) * (s->alpha / s->omega) : 0.0;
  s->rho_old = rho;
  s->coef[0] = s->bcoef;                // p = r + bcoef (p - omega v)
  s->coef[1] = -s->bcoef * s->omega;
}
__global__ void k_bc_alpha(LpState *s) {  // red[0] = rhat . v
  s->alpha = s->red[0] != 0.0 ? s->rho_old / s->red[0] : 0.0;
  if (s->red[0] == 0.0) s->breakdown = 1;
  s->coef[2] = -s->alpha;
}
__global__ void k_bc_omega(LpState *s) {  // red[0] = t . s, red[1] = t . t
  s->omega = s->red[1] != 0.0 ? s->red[0] / s->red[1] : 0.0;
  s->coef[3] = s->alpha;
  s->coef[4] = s->omega;
  s->coef[5] = -s->omega;
}
__global__ void k_bc_res(LpState *s) { s->res = sqrt(s->red[0]); }

// p = r + c0 p + c1 v
template <class VT>
__global__ void k_bc_p(i64 n, VT *p, const VT *r, const VT *v, const double *c) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = lp_fma((VT)c[1], v[i], lp_fma((VT)c[0], p[i], r[i]));
}
// y = x + c v
template <class VT>
__global__ void k_lp_waxpy(i64 n, VT *y, const VT *x, const VT *v, const double *c) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = lp_fma((VT)c[0], v[i], x[i]);
}
// d += c0 a + c1 b
template <class VT>
__global__ void k_lp_axpby_add(i64 n, VT *d, const VT *a, const VT *b, const double *c) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) d[i] = lp_fma((VT)c[1], b[i], lp_fma((VT)c[0], a[i], d[i]));
}

// The solver behind a handle: the refinement loop in FP64 around the inner
// solve in VT (float: FP32 or FP16 values; double: FP64 values).
struct MixedBase {
  virtual ~MixedBase() {}
  virtual int solve(const double *B, double *X, double *cfg) = 0;
};

template <class VT>
struct MixedSolver : MixedBase {
  using Tr = typename LpSweep<VT>::T;
  i64 n = 0;
  int prec = 32, inner = LP_GMRES, m = 30, refine = 1;
  DevCSR A64{};
  const int *soff = nullptr, *sc = nullptr;
  void *lv = nullptr;  // the inner solve's matrix values (__half, float or double)
  float *rs = nullptr;  // FP16: one scale per row
  PCView pc;
  void *lc0 = nullptr, *lc1 = nullptr;  // the inner solve's sweep coefficients
  VT *dinv = nullptr, *idiag = nullptr;
  int tri_iters = 0;  // > 0: approximate sweeps by this many Jacobi passes
  TriSell tri[2];
  VT *work = nullptr;     // the inner solve's vectors
  double *r64 = nullptr;  // FP64 residual
  double *P = nullptr;    // reduction partials
  LpState *st = nullptr;
  double *hst = nullptr;  // pinned: res, breakdown
  unsigned *ticket = nullptr;
  LpStaged<VT> *lst = nullptr;  // staged rows of the planned sweeps
  int *winA = nullptr;          // column windows of the inner solve's matrix
  bool winA_on = false;         // products stage them (col_windows)

  ~MixedSolver() override {
    void *bufs[] = {lv,          rs,          lc0,         lc1,       dinv,      idiag,    work,
                    r64,         P,           st,          ticket,    lst,       winA,     tri[0].rp,
                    tri[0].soff, tri[0].sc,   tri[0].sv,   tri[0].win, tri[1].rp, tri[1].soff, tri[1].sc,
                    tri[1].sv,   tri[1].win};
    hipStreamSynchronize(stream());
    for (void *b : bufs)
      if (b) hipFree(b);
    if (hst) hipHostFree(hst);
  }

  // Calls f with a null pointer of the stored value type (CT) of this
  // precision: __half (16), float (32) or double (64).
  template <class F>
  void with_ct(F &&f) const {
    if constexpr (std::is_same<VT, double>::value) {
      f((const double *)nullptr);
    } else {
      if (prec == 16)
        f((const __half *)nullptr);
      else
        f((const float *)nullptr);
    }
  }
  i64 esz() const { return prec == 16 ? 2 : prec == 32 ? 4 : 8; }

  void spmv(const VT *x, VT *y) {
    with_ct([&](auto tag) {
      using CT = std::remove_cv_t<std::remove_pointer_t<decltype(tag)>>;
      if (winA_on)
        MPK_LAUNCH((k_spmv_lp<CT, VT, true>), nblocks(n, LP_WROWS), LP_WROWS, n, A64.rp, soff, sc, (const CT *)lv,
                   (const float *)rs, (const int *)winA, x, y);
      else
        MPK_LAUNCH((k_spmv_lp<CT, VT, false>), nblocks(n, LP_WROWS), LP_WROWS, n, A64.rp, soff, sc, (const CT *)lv,
                   (const float *)rs, (const int *)winA, x, y);
    });
  }

  template <class Prog>
  void tri_pass(const Prog &prog, const TriSell &T, const VT *in, VT *out) {
    with_ct([&](auto tag) {
      using CT = std::remove_cv_t<std::remove_pointer_t<decltype(tag)>>;
      if (T.windowed)
        MPK_LAUNCH((k_tri_jacobi<Prog, CT, VT, true>), nblocks(n, LP_WROWS), LP_WROWS, prog, n, (const int *)T.rp,
                   (const int *)T.soff, (const int *)T.sc, (const CT *)T.sv, (const int *)T.win, in, out);
      else
        MPK_LAUNCH((k_tri_jacobi<Prog, CT, VT, false>), nblocks(n, LP_WROWS), LP_WROWS, prog, n, (const int *)T.rp,
                   (const int *)T.soff, (const int *)T.sc, (const CT *)T.sv, (const int *)T.win, in, out);
    });
  }

  template <class Prog>
  void sweep(const Prog &prog, int second, VT *pub) {
    if (tri_iters > 0) {
      // Jacobi passes, ping-ponging through the 4th PC vector so that the last
      // pass lands in pub
      VT *tmp = pc_scratch() + 3 * vs();
      const TriSell &T = tri[second];
      for (int j = 0; j < tri_iters; j++) {
        VT *out = ((tri_iters - 1 - j) % 2 == 0) ? pub : tmp;
        const VT *in = j == 0 ? nullptr : (((tri_iters - j) % 2 == 0) ? pub : tmp);
        tri_pass(prog, T, in, out);
      }
      return;
    }
    // the planned sweep on the block engine: stage, then the block or ring plan
    watch_install();
    MPK_LAUNCH(k_lp_prepare<VT>, lp_grid(n), 256, n, pub, ticket);
    const void *lc = second ? lc1 : lc0;
    MPK_LAUNCH((k_sweep_stage<Prog, LpStaged<VT>>), lp_grid(n), 256, prog, n, pc_plan(), second, lst,
               (const int64_t *)nullptr);
    with_ct([&](auto tag) {
      using CT = std::remove_cv_t<std::remove_pointer_t<decltype(tag)>>;
      MPK_LAUNCH((k_sweep_blk<Prog, Tr, LpCoef<CT, VT>, LpStaged<VT>>), nblocks(n, 1024), 256, prog, n, pc_plan(),
                 second, (const LpStaged<VT> *)lst, lc, ticket, pub, (const int64_t *)nullptr);
    });
  }
  const char *pc_plan() const {
    return pc.type == PCT_ILU ? pc.ilu.plan : pc.type == PCT_SOR ? pc.sor.plan : nullptr;
  }
  // The inner solve's vectors: the inner solver's (GMRES: V[m+1], w, z, d;
  // BiCGStab: r, rhat, p, v, s, t, phat, shat, d -- d last), then 4 for the
  // preconditioner's sweeps, then the right-hand side
  i64 nsolver() const { return inner == LP_GMRES ? m + 4 : 9; }
  // the stride of the work buffer's vectors: n rounded up to 8 entries, so
  // every vector starts 16-byte aligned (k_lp_mdot's loads)
  i64 vs() const { return (n + 7) & ~(i64)7; }
  VT *d_vec() { return work + (nsolver() - 1) * vs(); }
  VT *pc_scratch() { return work + nsolver() * vs(); }
  VT *rhs_vec() { return work + (nsolver() + 4) * vs(); }

  // The sweeps' entries as sliced matrices in the inner solve's value type
  // (coefficients divided by the row scale where the matrix has one: SOR at
  // FP16).
  bool build_tri() {
    std::vector<float> hrs;
    if (pc.type == PCT_SOR && rs) {
      hrs.resize(n);
      copy_dtoh(hrs.data(), rs, n * sizeof(float));
    }
    for (int w = 0; w < 2; w++) {
      std::vector<i64> off;
      std::vector<int> col;
      std::vector<double> coef;
      if (!pc_sweep_entries(pc, w, off, col, coef)) return false;
      const i64 ns = (n + 63) / 64;
      std::vector<int> rp(n + 1), so(ns + 1, 0);
      for (i64 i = 0; i <= n; i++) rp[i] = (int)off[i];
      i64 tot = 0;
      for (i64 sl = 0; sl < ns; sl++) {
        i64 wmax = 0;
        for (i64 i = sl * 64; i < n && i < sl * 64 + 64; i++) wmax = std::max<i64>(wmax, off[i + 1] - off[i]);
        so[sl] = (int)tot;
        tot += 64 * wmax;
      }
      so[ns] = (int)tot;
      const i64 slots = tot > 0 ? tot : 1;
      std::vector<int> sc(slots, 0);
      std::vector<double> sv(slots, 0.0);
      for (i64 i = 0; i < n; i++) {
        const double sdiv = hrs.empty() ? 1.0 : (double)hrs[i];
        const i64 base = so[i >> 6] + (i & 63);
        for (i64 e = off[i]; e < off[i + 1]; e++) {
          sc[base + 64 * (e - off[i])] = col[e];
          sv[base + 64 * (e - off[i])] = coef[e] / sdiv;
        }
      }
      TriSell &T = tri[w];
      HIPCHK(hipMalloc(&T.rp, (n + 1) * sizeof(int)));
      HIPCHK(hipMalloc(&T.soff, (ns + 1) * sizeof(int)));
      HIPCHK(hipMalloc(&T.sc, slots * sizeof(int)));
      HIPCHK(hipMalloc(&T.sv, slots * esz()));
      if (failed()) return false;
      copy_htod(T.rp, rp.data(), (n + 1) * sizeof(int));
      copy_htod(T.soff, so.data(), (ns + 1) * sizeof(int));
      copy_htod(T.sc, sc.data(), slots * sizeof(int));
      HIPCHK(hipMalloc(&T.win, 2 * (i64)nblocks(n, LP_WROWS) * sizeof(int)));
      T.windowed = col_windows(n, T.rp, T.soff, T.sc, T.win, lp_xwin<VT>());
      if (prec == 64) {
        copy_htod(T.sv, sv.data(), slots * sizeof(double));
      } else {
        std::vector<float> sf(sv.begin(), sv.end());  // rounded to FP32
        if (prec == 16) {
          float *tmp = nullptr;
          HIPCHK(hipMalloc(&tmp, slots * sizeof(float)));
          copy_htod(tmp, sf.data(), slots * sizeof(float));
          MPK_LAUNCH(k_f2h, lp_grid(slots), 256, slots, (const float *)tmp, (__half *)T.sv);
          HIPCHK(hipStreamSynchronize(stream()));
          HIPCHK(hipFree(tmp));
        } else {
          copy_htod(T.sv, sf.data(), slots * sizeof(float));
        }
      }
    }
    return !failed();
  }

  // y = M^-1 x in VT
  void apply_pc(const VT *x, VT *y) {
    VT *S = pc_scratch();
    const i64 ld = vs();
    if (pc.type == PCT_ILU) {
      VT *F = S, *B = S + ld;
      sweep(LpIluFwd<VT>{pc.ilu.r, x}, 0, F);
      sweep(LpIluBwd<VT>{pc.ilu.c, F, dinv, y}, 1, B);
    } else if (pc.type == PCT_SOR) {
      VT *YF = S, *YB = S + ld, *T = S + 2 * ld;
      const int flag = pc.sor.flag;
      const bool fwd = (flag & (1 | 4)) != 0, bwd = (flag & (2 | 8)) != 0;
      if (fwd) sweep(LpSorFwd<VT>{x, rs, idiag, T, bwd ? nullptr : y, y}, 0, YF);
      if (bwd) sweep(LpSorBwd<VT>{fwd ? T : x, rs, idiag, YF, y, (VT)pc.sor.omega, fwd ? 1 : 0}, 1, YB);
    } else {
      MPK_LAUNCH(k_lp_copy<VT>, lp_grid(n), 256, n, x, y);
    }
  }

  // FP64 results of K dot products into st->red
  template <int K>
  void dots(LpDots<K, VT> d) {
    MPK_LAUNCH((k_lp_dots<K, VT>), LP_RBLK, 256, d, n, P);
    MPK_LAUNCH(k_lp_fold, K, 256, (const double *)P, LP_RBLK, st->red, (const int *)nullptr);
  }
  LpDots<1, VT> dot1(const VT *x, const VT *y) { return LpDots<1, VT>{{x}, {y}}; }

  double read_res(int *breakdown) {
    HIPCHK(hipMemcpyAsync(hst, &st->res, sizeof(double), hipMemcpyDeviceToHost, stream()));
    HIPCHK(hipMemcpyAsync(hst + 1, &st->breakdown, sizeof(int), hipMemcpyDeviceToHost, stream()));
    HIPCHK(hipStreamSynchronize(stream()));
    *breakdown = *reinterpret_cast<int *>(hst + 1);
    return hst[0];
  }

  // Inner GMRES(m), left preconditioned, zero guess: d ~ A^-1 rhs to a
  // relative preconditioned residual rtol (or max_it steps).  Returns steps.
  i64 gmres(const VT *rhs, VT *d, double rtol, i64 max_it) {
    const i64 ld = vs();
    VT *V = work, *w = work + (i64)(m + 1) * ld, *z = w + ld;
    MPK_LAUNCH(k_lp_zero<VT>, lp_grid(n), 256, n, d);
    MPK_LAUNCH(k_lp_copy<VT>, lp_grid(n), 256, n, rhs, z);
    i64 its = 0;
    double target = -1.0;
    while (its < max_it) {
      // z holds the current unpreconditioned residual
      apply_pc(z, w);
      dots(dot1(w, nullptr));
      MPK_LAUNCH(k_gm_begin, 1, 1, st);
      MPK_LAUNCH(k_lp_scale<VT>, lp_grid(n), 256, n, (const VT *)w, (const double *)&st->inv, V,
                 (const int *)nullptr);
      int brk = 0;
      const double beta = read_res(&brk);
      if (target < 0.0) target = rtol * beta;
      if (brk || beta <= target) break;
      int kk = 0;
      for (int k = 0; k < m && its < max_it; k++) {
        VT *vk = V + (i64)k * ld, *vn = V + (i64)(k + 1) * ld;
        spmv(vk, z);
        apply_pc(z, w);
        const int nv = k + 1;
        // classical Gram-Schmidt: h = V^T w, w -= V h, then the second pass
        // (h2 = V^T w, w -= V h2, h += h2) when k_gm_decide asks for it; the
        // device skips the second pass's kernels otherwise
        const int *nore = &st->norefine;
        MPK_LAUNCH(k_lp_mdot<VT>, LP_RBLK, 256, (const VT *)V, ld, nv, (const VT *)w, n, P, (const int *)nullptr);
        MPK_LAUNCH(k_lp_fold, nv, 256, (const double *)P, LP_RBLK, st->h, (const int *)nullptr);
        MPK_LAUNCH(k_gm_neg, 1, 1, st, nv, (const double *)st->h, (const int *)nullptr);
        MPK_LAUNCH(k_lp_maxpy<VT>, lp_grid(n), 256, n, w, (const VT *)V, ld, (const double *)st->coef, nv,
                   (const int *)nullptr);
        dots(dot1(w, nullptr));
        MPK_LAUNCH(k_gm_decide, 1, 1, st, nv, refine);
        MPK_LAUNCH(k_lp_mdot<VT>, LP_RBLK, 256, (const VT *)V, ld, nv, (const VT *)w, n, P, nore);
        MPK_LAUNCH(k_lp_fold, nv, 256, (const double *)P, LP_RBLK, st->h2, nore);
        MPK_LAUNCH(k_gm_neg, 1, 1, st, nv, (const double *)st->h2, nore);
        MPK_LAUNCH(k_lp_maxpy<VT>, lp_grid(n), 256, n, w, (const VT *)V, ld, (const double *)st->coef, nv, nore);
        MPK_LAUNCH((k_lp_dots<1, VT>), LP_RBLK, 256, dot1(w, nullptr), n, P);
        MPK_LAUNCH(k_lp_fold, 1, 256, (const double *)P, LP_RBLK, st->red, nore);
        MPK_LAUNCH(k_gm_column, 1, 1, st, k);
        MPK_LAUNCH(k_lp_scale<VT>, lp_grid(n), 256, n, (const VT *)w, (const double *)&st->inv, vn,
                   (const int *)nullptr);
        its++;
        kk = k + 1;
        const double res = read_res(&brk);
        if (brk || res <= target) break;
      }
      MPK_LAUNCH(k_gm_solve, 1, 1, st, kk);
      MPK_LAUNCH(k_lp_maxpy<VT>, lp_grid(n), 256, n, d, (const VT *)V, ld, (const double *)st->y, kk,
                 (const int *)nullptr);
      const double res = read_res(&brk);
      if (brk || res <= target || its >= max_it) break;
      // restart: z = rhs - A d
      spmv(d, z);
      MPK_LAUNCH(k_lp_rsub<VT>, lp_grid(n), 256, n, rhs, z);
    }
    return its;
  }

  // Inner BiCGStab, right preconditioned, zero guess, to a relative residual
  // rtol (or max_it steps).  Returns steps.
  i64 bcgs(const VT *rhs, VT *d, double rtol, i64 max_it) {
    const i64 ld = vs();
    VT *r = work, *rh = r + ld, *p = rh + ld, *v = p + ld, *sv = v + ld, *t = sv + ld, *ph = t + ld, *sh = ph + ld;
    MPK_LAUNCH(k_lp_zero<VT>, lp_grid(n), 256, n, d);
    MPK_LAUNCH(k_lp_copy<VT>, lp_grid(n), 256, n, rhs, r);
    MPK_LAUNCH(k_lp_copy<VT>, lp_grid(n), 256, n, rhs, rh);
    MPK_LAUNCH(k_lp_zero<VT>, lp_grid(n), 256, n, p);
    MPK_LAUNCH(k_lp_zero<VT>, lp_grid(n), 256, n, v);
    dots(dot1(r, nullptr));
    MPK_LAUNCH(k_bc_init, 1, 1, st);
    int brk = 0;
    const double beta = read_res(&brk);
    const double target = rtol * beta;
    i64 its = 0;
    while (!brk && its < max_it) {
      dots(dot1(rh, r));
      MPK_LAUNCH(k_bc_rho, 1, 1, st);
      MPK_LAUNCH(k_bc_p<VT>, lp_grid(n), 256, n, p, (const VT *)r, (const VT *)v, (const double *)st->coef);
      apply_pc(p, ph);
      spmv(ph, v);
      dots(dot1(rh, v));
      MPK_LAUNCH(k_bc_alpha, 1, 1, st);
      MPK_LAUNCH(k_lp_waxpy<VT>, lp_grid(n), 256, n, sv, (const VT *)r, (const VT *)v,
                 (const double *)(st->coef + 2));
      apply_pc(sv, sh);
      spmv(sh, t);
      dots(LpDots<2, VT>{{t, t}, {sv, nullptr}});
      MPK_LAUNCH(k_bc_omega, 1, 1, st);
      MPK_LAUNCH(k_lp_axpby_add<VT>, lp_grid(n), 256, n, d, (const VT *)ph, (const VT *)sh,
                 (const double *)(st->coef + 3));
      MPK_LAUNCH(k_lp_waxpy<VT>, lp_grid(n), 256, n, r, (const VT *)sv, (const VT *)t,
                 (const double *)(st->coef + 5));
      dots(dot1(r, nullptr));
      MPK_LAUNCH(k_bc_res, 1, 1, st);
      its++;
      const double res = read_res(&brk);
      if (res <= target) break;
    }
    return its;
  }

  // Builds the inner solve's copies of the matrix and preconditioner.
  bool create(int64_t irow, int64_t icol, int64_t data, int64_t soff_, int64_t sv, int64_t sc_, int64_t pci,
              int64_t pcf) {
    A64 = DevCSR{n, n, (const int *)irow, (const int *)icol, (const double *)data, (const int *)soff_,
                 (const double *)sv, (const int *)sc_};
    soff = (const int *)soff_;
    sc = (const int *)sc_;
    if (!pc.parse((const int64_t *)pci, (const double *)pcf, n)) return false;
    if (pc.type != PCT_ILU && pc.type != PCT_SOR && pc.type != PCT_NONE) {
      set_error(6, "mixed solve: the preconditioner must be ilu, sor or none");
      return false;
    }
    if ((pc.type == PCT_ILU || pc.type == PCT_SOR) && !pc_plan()) {
      set_error(6, "mixed solve: the preconditioner has no sweep plans (mpk_pc_schedule)");
      return false;
    }
    // the inner solve's matrix values in the sliced slots
    std::vector<int> so(((size_t)n + 63) / 64 + 1);
    copy_dtoh(so.data(), (const void *)soff, so.size() * sizeof(int));
    const i64 total = sell_slots(so.back()), slots = total > 0 ? total : 1;
    HIPCHK(hipMalloc(&lv, slots * esz()));
    if (prec == 16) {
      HIPCHK(hipMalloc(&rs, n * sizeof(float)));
      MPK_LAUNCH(k_lp_sell<__half>, lp_grid(n), 256, n, (const int *)irow, soff, (const double *)sv, (__half *)lv,
                 rs);
    } else if (prec == 32) {
      MPK_LAUNCH(k_lp_sell<float>, lp_grid(n), 256, n, (const int *)irow, soff, (const double *)sv, (float *)lv,
                 (float *)nullptr);
    } else {
      MPK_LAUNCH(k_lp_sell<double>, lp_grid(n), 256, n, (const int *)irow, soff, (const double *)sv,
                 (double *)lv, (float *)nullptr);
    }
    HIPCHK(hipMalloc(&winA, 2 * (i64)nblocks(n, LP_WROWS) * sizeof(int)));
    winA_on = col_windows(n, (const int *)irow, soff, sc, winA, lp_xwin<VT>());
    // the inner solve's preconditioner
    if (const char *plan = pc_plan()) {
      void **dst[2] = {&lc0, &lc1};
      for (int sw = 0; sw < 2; sw++) {
        const i64 ns = plan_slots(plan, sw, n);
        HIPCHK(hipMalloc(dst[sw], (ns > 0 ? ns : 1) * esz()));
        const float *rsw = pc.type == PCT_SOR ? rs : nullptr;
        if (prec == 16)
          MPK_LAUNCH(k_lp_plan_coef<__half>, lp_grid(n), 256, n, plan, sw, rsw, (__half *)*dst[sw]);
        else if (prec == 32)
          MPK_LAUNCH(k_lp_plan_coef<float>, lp_grid(n), 256, n, plan, sw, rsw, (float *)*dst[sw]);
        else
          MPK_LAUNCH(k_lp_plan_coef<double>, lp_grid(n), 256, n, plan, sw, rsw, (double *)*dst[sw]);
      }
      if (tri_iters > 0 && !build_tri()) return false;
      if (pc.type == PCT_ILU) {
        HIPCHK(hipMalloc(&dinv, n * sizeof(VT)));
        MPK_LAUNCH(k_lp_ilu_dinv<VT>, lp_grid(n), 256, n, pc.ilu.ba, pc.ilu.bdiag, dinv);
      } else {
        HIPCHK(hipMalloc(&idiag, n * sizeof(VT)));
        MPK_LAUNCH(k_lp_sor_idiag<VT>, lp_grid(n), 256, n, pc.sor.a, pc.sor.dmark, pc.sor.omega, pc.sor.fshift,
                   idiag);
      }
    }
    const i64 nvec = nsolver() + 5;
    HIPCHK(hipMalloc(&work, nvec * vs() * sizeof(VT)));
    HIPCHK(hipMalloc(&r64, n * sizeof(double)));
    HIPCHK(hipMalloc(&P, (LP_MMAX + 1) * LP_RBLK * sizeof(double)));
    HIPCHK(hipMalloc(&st, sizeof(LpState)));
    HIPCHK(hipHostMalloc(&hst, 4 * sizeof(double), hipHostMallocDefault));
    HIPCHK(hipMalloc(&ticket, 64 * sizeof(unsigned)));
    HIPCHK(hipMalloc(&lst, (n > 0 ? n : 1) * sizeof(LpStaged<VT>)));
    if (!failed()) HIPCHK(hipMemset(st, 0, sizeof(LpState)));
    sync();
    return !failed();
  }

  // Iterative refinement (mpk_mixed_solve).
  int solve(const double *B, double *X, double *cfg) override {
    const double rtol = cfg[0], irtol = cfg[1];
    const i64 max_outer = (i64)cfg[2], max_inner = (i64)cfg[3];
    refine = (int)cfg[4];
    const bool zero = cfg[5] != 0.0;
    double *R = r64;
    const double bn = norm2_host(B, n);
    cfg[11] = bn;
    VT *rhs = rhs_vec(), *d = d_vec();
    if (zero) vset(X, 0.0, n, nullptr);
    i64 outer = 0, ninner = 0;
    double rel = 0.0;
    for (;;) {
      mpk::spmv(A64, X, R, 0, nullptr);  // R = b - A x in FP64
      vaypx(R, -1.0, B, n, nullptr);
      const double rn = norm2_host(R, n);
      rel = bn > 0.0 ? rn / bn : rn;
      if (failed() || rel <= rtol || outer >= max_outer || !(rn > 0.0)) break;
      // right-hand side scaled to unit norm; x += ||r|| d
      double scl[2] = {1.0 / rn, rn};
      copy_htod(st->red + 2, scl, sizeof scl);
      MPK_LAUNCH(k_lp_from_d<VT>, lp_grid(n), 256, n, (const double *)R, (const double *)(st->red + 2), rhs);
      ninner += inner == LP_GMRES ? gmres(rhs, d, irtol, max_inner) : bcgs(rhs, d, irtol, max_inner);
      MPK_LAUNCH(k_lp_update<VT>, lp_grid(n), 256, n, X, (const double *)(st->red + 3), (const VT *)d);
      outer++;
    }
    sync();
    cfg[8] = rel;
    cfg[9] = (double)outer;
    cfg[10] = (double)ninner;
    cfg[12] = rel <= rtol ? 1.0 : 0.0;
    return failed() ? -1 : 0;
  }
};

}  // namespace mpk

using namespace mpk;

extern "C" {

// See include/mpk.h.
int64_t mpk_mixed_create(int64_t n, int64_t irow, int64_t icol, int64_t data, int64_t soff, int64_t sv,
                         int64_t sc, int64_t pci, int64_t pcf, int64_t prec, int64_t inner, int64_t restart,
                         int64_t tri_iters) {
  if (n <= 0 || !ctx()) return 0;
  if (prec != 64 && prec != 32 && prec != 16) {
    set_error(6, "mixed solve: precision must be 64, 32 or 16, got %lld", (long long)prec);
    return 0;
  }
  if (inner != LP_GMRES && inner != LP_BCGS) {
    set_error(6, "mixed solve: inner solver must be 0 (gmres) or 1 (bcgs)");
    return 0;
  }
  if (inner == LP_GMRES && (restart < 1 || restart > LP_MMAX)) {
    set_error(6, "mixed solve: restart must be in 1..%d", LP_MMAX);
    return 0;
  }
  auto make = [&](auto *S) -> int64_t {
    S->n = n;
    S->prec = (int)prec;
    S->inner = (int)inner;
    S->m = (int)restart;
    S->tri_iters = tri_iters > 0 ? (int)tri_iters : 0;
    if (!S->create(irow, icol, data, soff, sv, sc, pci, pcf)) {
      delete S;
      return 0;
    }
    return (int64_t) static_cast<MixedBase *>(S);
  };
  return prec == 64 ? make(new MixedSolver<double>()) : make(new MixedSolver<float>());
}

// Iterative refinement: x (FP64, [D]) from b (FP64, [D]).  cfg [H] float64:
// in  {rtol, inner_rtol, max_outer, max_inner, refine (gmres: 0 never,
//      1 if needed, 2 always), x_is_zero}
// out {cfg[8] final relative residual, cfg[9] outer steps, cfg[10] inner
//      steps, cfg[11] ||b||, cfg[12] converged (1) or not (0)}.
int64_t mpk_mixed_solve(int64_t h, int64_t b, int64_t x, int64_t cfg_) {
  MixedBase *S = (MixedBase *)h;
  if (!S || !ctx()) return -1;
  return S->solve((const double *)b, (double *)x, (double *)cfg_);
}

void mpk_mixed_free(int64_t h) { delete (MixedBase *)h; }

}  // extern "C"
