// mpk_vec.hip -- Vec kernels and the mp_vec_* exports.
//
// Every elementwise kernel keeps the reference expression shape
// (src/simdvec.mojo:40-319) with fma() where that expression contracts, so
// its result is bitwise identical to the CPU path for the same inputs.
// Reductions use the tile-tree order of mpk_reduce.h.
#include <cmath>
#include <cstring>

#include "mpk_internal.h"
#include "mpk_ops.h"
#include "mpk_reduce.h"
#include "../../include/mpk.h"

namespace mpk {

// ------------------------------------------------------------ elementwise

// Each thread handles the pairs at tile_base + 512 j + 2 t, j = 0..3: four
// 16-byte accesses per operand in flight, 2048 elements per block.
template <class Op>
__global__ void __launch_bounds__(MPK_NT) k_ew(i64 n, Op op, const int64_t *stop) {
  if (stop && *stop) return;
  const i64 base = (i64)blockIdx.x * MPK_TILE + 2 * (i64)threadIdx.x;
#pragma unroll
  for (int j = 0; j < 4; j++) op(base + 512 * j, n);
}

template <class Op>
static void ew(i64 n, const Op &op, const int64_t *stop = nullptr) {
  if (n <= 0 || !ctx()) return;
  MPK_LAUNCH(k_ew<Op>, ntiles(n), MPK_NT, n, op, stop);
}

// w = f(x, y) elementwise
template <class F>
struct Map2 {
  double *w;
  const double *x, *y;
  bool al;
  F f;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    D2 a = ld2(x, e, n, al), b = ld2(y, e, n, al);
    st2(w, e, n, al, D2{f(a.a, b.a), f(a.b, b.b)});
  }
};
template <class F>
struct Map1 {  // x = f(x) in place, or w = f(x)
  double *w;
  const double *x;
  bool al;
  F f;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    D2 a = ld2(x, e, n, al);
    st2(w, e, n, al, D2{f(a.a), f(a.b)});
  }
};

struct FMul { __device__ double operator()(double a, double b) const { return a * b; } };
struct FDiv { __device__ double operator()(double a, double b) const { return a / b; } };
struct FMin { __device__ double operator()(double a, double b) const { return a < b ? a : b; } };
struct FMax { __device__ double operator()(double a, double b) const { return a > b ? a : b; } };
struct FMaxAbs {
  __device__ double operator()(double a, double b) const {
    double p = fabs(a), q = fabs(b);
    return p > q ? p : q;
  }
};
struct FRecip { __device__ double operator()(double a) const { return a == 0.0 ? 0.0 : 1.0 / a; } };
struct FExp { __device__ double operator()(double a) const { return exp(a); } };
struct FLog { __device__ double operator()(double a) const { return log(a); } };
struct FSqrtAbs { __device__ double operator()(double a) const { return sqrt(fabs(a)); } };
struct FAbs { __device__ double operator()(double a) const { return fabs(a); } };
struct FChop {
  double tol;
  __device__ double operator()(double a) const { return fabs(a) < tol ? 0.0 : a; }
};
struct FScale {
  double s;
  __device__ double operator()(double a) const { return a * s; }
};
struct FShift {
  double s;
  __device__ double operator()(double a) const { return a + s; }
};
// the divide, not a multiply by the reciprocal (simdvec.mojo:296-304)
struct FDivBy {
  double z;
  __device__ double operator()(double a) const { return a / z; }
};
struct FSet {
  double s;
  __device__ double operator()(double) const { return s; }
};

// y = y + alpha x  ->  fma(alpha, x, y)
struct OpAxpy {
  double *y;
  const double *x;
  double alpha;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    D2 a = ld2(x, e, n, al), b = ld2(y, e, n, al);
    st2(y, e, n, al, D2{fma(alpha, a.a, b.a), fma(alpha, a.b, b.b)});
  }
};
// y = x + alpha y  ->  fma(alpha, y, x)
struct OpAypx {
  double *y;
  const double *x;
  double alpha;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    D2 a = ld2(x, e, n, al), b = ld2(y, e, n, al);
    st2(y, e, n, al, D2{fma(alpha, b.a, a.a), fma(alpha, b.b, a.b)});
  }
};
// y = alpha x + beta y  ->  fma(alpha, x, beta y)
struct OpAxpby {
  double *y;
  const double *x;
  double alpha, beta;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    D2 a = ld2(x, e, n, al), b = ld2(y, e, n, al);
    st2(y, e, n, al, D2{fma(alpha, a.a, beta * b.a), fma(alpha, a.b, beta * b.b)});
  }
};
// z = (alpha x + beta y) + gamma z  ->  fma(gamma, z, fma(alpha, x, beta y))
struct OpAxpbypcz {
  double *z;
  const double *x, *y;
  double alpha, beta, gamma;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    D2 a = ld2(x, e, n, al), b = ld2(y, e, n, al), c = ld2(z, e, n, al);
    st2(z, e, n, al,
        D2{fma(gamma, c.a, fma(alpha, a.a, beta * b.a)), fma(gamma, c.b, fma(alpha, a.b, beta * b.b))});
  }
};
// w = alpha x + y  ->  fma(alpha, x, y)
struct OpWaxpy {
  double *w;
  const double *x, *y;
  double alpha;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    D2 a = ld2(x, e, n, al), b = ld2(y, e, n, al);
    st2(w, e, n, al, D2{fma(alpha, a.a, b.a), fma(alpha, a.b, b.b)});
  }
};
struct OpSwap {
  double *x, *y;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    D2 a = ld2(x, e, n, al), b = ld2(y, e, n, al);
    st2(x, e, n, al, b);
    st2(y, e, n, al, a);
  }
};

// y = y + sum_k alpha_k x_k, terms added in k order per element -- the same
// bits as k sequential axpy sweeps (ported.mojo:329-335), in one pass.
#define MPK_MAXPY_G 16
struct OpMaxpy {
  double *y;
  const double *xs[MPK_MAXPY_G];
  double a[MPK_MAXPY_G];
  int nv;
  bool zero_first;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    D2 acc = zero_first ? D2{0.0, 0.0} : ld2(y, e, n, al);
#pragma unroll
    for (int k = 0; k < MPK_MAXPY_G; k++) {
      if (k < nv) {
        D2 v = ld2(xs[k], e, n, al);
        acc.a = fma(a[k], v.a, acc.a);
        acc.b = fma(a[k], v.b, acc.b);
      }
    }
    st2(y, e, n, al, acc);
  }
};

// ------------------------------------------------------------ host helpers

void vaxpy(double *y, double alpha, const double *x, i64 n, const int64_t *stop) {
  ew(n, OpAxpy{y, x, alpha, aligned16(y) && aligned16(x)}, stop);
}
void vaypx_raw(double *y, double alpha, const double *x, i64 n, const int64_t *stop) {
  ew(n, OpAypx{y, x, alpha, aligned16(y) && aligned16(x)}, stop);
}
void vaxpby_raw(double *y, double alpha, double beta, const double *x, i64 n, const int64_t *stop) {
  ew(n, OpAxpby{y, x, alpha, beta, aligned16(y) && aligned16(x)}, stop);
}
void vaxpbypcz(double *z, double alpha, double beta, double gamma, const double *x, const double *y,
               i64 n, const int64_t *stop) {
  ew(n, OpAxpbypcz{z, x, y, alpha, beta, gamma, aligned16(z) && aligned16(x) && aligned16(y)}, stop);
}
void vwaxpy(double *w, double alpha, const double *x, const double *y, i64 n, const int64_t *stop) {
  ew(n, OpWaxpy{w, x, y, alpha, aligned16(w) && aligned16(x) && aligned16(y)}, stop);
}
void vscale(double *x, double alpha, i64 n, const int64_t *stop) {
  ew(n, Map1<FScale>{x, x, aligned16(x), FScale{alpha}}, stop);
}
void vdivide(double *x, double z, i64 n, const int64_t *stop) {
  ew(n, Map1<FDivBy>{x, x, aligned16(x), FDivBy{z}}, stop);
}
void vset(double *x, double alpha, i64 n, const int64_t *stop) {
  ew(n, Map1<FSet>{x, x, aligned16(x), FSet{alpha}}, stop);
}
void vcopy(const double *x, double *y, i64 n) {
  if (n <= 0 || x == y || !ctx()) return;
  HIPCHK(hipMemcpyAsync(y, x, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
}
void vpointwise_mult(double *w, const double *x, const double *y, i64 n, const int64_t *stop) {
  ew(n, Map2<FMul>{w, x, y, aligned16(w) && aligned16(x) && aligned16(y), FMul{}}, stop);
}
void vmaxpy(double *y, int nv, const double *alpha_host, const double *const *xs, i64 n,
            bool zero_first, const int64_t *stop) {
  if (n <= 0) return;
  if (nv <= 0) {
    if (zero_first) vset(y, 0.0, n, stop);
    return;
  }
  for (int k0 = 0; k0 < nv; k0 += MPK_MAXPY_G) {
    OpMaxpy op;
    op.y = y;
    op.nv = nv - k0 < MPK_MAXPY_G ? nv - k0 : MPK_MAXPY_G;
    op.zero_first = zero_first && k0 == 0;
    op.al = aligned16(y);
    for (int k = 0; k < MPK_MAXPY_G; k++) {
      op.xs[k] = k < op.nv ? xs[k0 + k] : nullptr;
      op.a[k] = k < op.nv ? alpha_host[k0 + k] : 0.0;
      if (k < op.nv) op.al = op.al && aligned16(op.xs[k]);
    }
    ew(n, op, stop);
  }
}

// Device-side coefficients: y = (zero_first ? 0 : y) + sum_k coef[k] x_k with
// x_k = base + k * ld and coef read from device memory (solver basis updates).
struct OpMaxpyDev {
  double *y;
  const double *base;
  i64 ld;
  const double *coef;
  int k0, nv;
  double sign;
  bool zero_first;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    D2 acc = zero_first ? D2{0.0, 0.0} : ld2(y, e, n, al);
    for (int k = 0; k < nv; k++) {
      double a = sign * coef[k0 + k];
      D2 v = ld2(base + (i64)(k0 + k) * ld, e, n, al);
      acc.a = fma(a, v.a, acc.a);
      acc.b = fma(a, v.b, acc.b);
    }
    st2(y, e, n, al, acc);
  }
};
void vmaxpy_dev(double *y, const double *base, i64 ld, const double *coef_dev, int nv, double sign,
                i64 n, bool zero_first, const int64_t *stop) {
  if (n <= 0) return;
  if (nv <= 0) {
    if (zero_first) vset(y, 0.0, n, stop);
    return;
  }
  bool al = aligned16(y) && aligned16(base) && (ld % 2 == 0);
  for (int k0 = 0; k0 < nv; k0 += MPK_MAXPY_G) {
    int g = nv - k0 < MPK_MAXPY_G ? nv - k0 : MPK_MAXPY_G;
    ew(n, OpMaxpyDev{y, base, ld, coef_dev, k0, g, sign, zero_first && k0 == 0, al}, stop);
  }
}

// ------------------------------------------------------------ the reference's order

__global__ void __launch_bounds__(MPK_REF_NT) k_ref_red(RefJob J, i64 n, double *P, const int64_t *stop) {
  if (stop && *stop) return;
  ref_run(J, n, P);
}

void ref_launch(const RefJob &J, i64 n, double *P, const int64_t *stop) {
  MPK_LAUNCH(k_ref_red, 1, MPK_REF_NT, J, n, P, stop);
}

__global__ void __launch_bounds__(MPK_REF_NT) k_ref_mdot(const double *x, const double *V, i64 ldv, int nv, i64 n,
                                                         double *P, const int64_t *stop) {
  if (stop && *stop) return;
  const int j0 = 3 * (int)blockIdx.x;
  const int ns = nv - j0 < 3 ? nv - j0 : 3;
  RefJob J;
  J.swsteps = 0;
  J.nspec = ns;
  J.nvec = 1 + ns;
  J.al = aligned16(x) && aligned16(V) && ldv % 2 == 0;
  J.v[0] = x;
  J.v[1] = V + (i64)j0 * ldv;
  J.v[2] = V + (i64)(j0 + (ns > 1 ? 1 : 0)) * ldv;
  J.v[3] = V + (i64)(j0 + (ns > 2 ? 2 : 0)) * ldv;
  J.code = 0;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    J.code |= (unsigned)(0 | (r < 3 ? 1 + r : 3) << 2 | RK_DOT << 4) << (8 * r);
    J.sx[r] = x;
    J.sy[r] = J.v[r < 3 ? 1 + r : 3];
  }
  ref_run(J, n, P + j0);
}

// ------------------------------------------------------------ reductions to host

// Launch `K` tile-tree reductions and return the results on the host.
template <int K>
static void reduce_host(const RedSet<K> &rs, i64 n, double *res) {
  Ctx *c = ctx();
  if (!c) {
    for (int k = 0; k < K; k++) res[k] = NAN;
    return;
  }
  reduce_launch<K>(rs, n, c->dres, NoEpilogue{}, nullptr);
  HIPCHK(hipMemcpyAsync(c->hres, c->dres, K * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < K; k++) res[k] = c->hres[k];
}

double dot_host(const double *x, const double *y, i64 n) {
  RedSet<1> rs{{{x, y, RK_DOT}}};
  double r;
  reduce_host<1>(rs, n, &r);
  return r;
}
double reduce1_host(const double *x, i64 n, int kind) {
  RedSet<1> rs{{{x, x, kind}}};
  double r;
  reduce_host<1>(rs, n, &r);
  return r;
}
double norm2_host(const double *x, i64 n) {
  if (!n) return 0.0;
  return sqrt(reduce1_host(x, n, RK_SUMSQ));
}

// ---- ordered scans: the infinity norm, max/min with index, max |x/y|

struct ValIdx {
  double v;
  i64 i;
};

// mode 0: max |x| adopting a NaN (norm_infinity, simdvec.mojo:476-488)
// mode 1: max with first index, x[0] NaN sticks (VecMax_Seq)
// mode 2: min with first index, x[0] NaN sticks (VecMin_Seq)
// mode 3: max |x/y| with x where y == 0, NaN never adopted (max_pointwise_divide)
__device__ __forceinline__ ValIdx vi_leaf(int mode, const double *x, const double *y, i64 i) {
  double v = x[i];
  if (mode == 0) return ValIdx{fabs(v), i};
  if (mode == 3) {
    double yi = y[i];
    return ValIdx{fabs(yi == 0.0 ? v : v / yi), i};
  }
  return ValIdx{v, i};
}
// a beats b?  (associative and commutative)
__device__ __forceinline__ bool vi_better(int mode, const ValIdx &a, const ValIdx &b) {
  bool an = a.v != a.v, bn = b.v != b.v;
  if (mode == 0) {  // NaN dominates; otherwise larger; ties by index
    if (an != bn) return an;
    if (an) return a.i < b.i;
    if (a.v != b.v) return a.v > b.v;
    return a.i < b.i;
  }
  if (mode == 3) {  // NaN loses; larger wins (value only matters)
    if (an != bn) return bn;
    if (a.v != b.v) return a.v > b.v;
    return a.i < b.i;
  }
  // modes 1, 2: NaN at index 0 dominates, other NaNs lose
  bool a0 = an && a.i == 0, b0 = bn && b.i == 0;
  if (a0 || b0) return a0;
  if (an != bn) return bn;
  if (a.v != b.v) return mode == 1 ? a.v > b.v : a.v < b.v;
  return a.i < b.i;
}

__global__ void __launch_bounds__(MPK_NT) k_vi_partials(int mode, const double *x, const double *y,
                                                        i64 n, ValIdx *P) {
  __shared__ ValIdx sh[MPK_NT];
  ValIdx best{NAN, -1};
  bool have = false;
  for (i64 i = (i64)blockIdx.x * MPK_TILE + threadIdx.x; i < n && i < (i64)(blockIdx.x + 1) * MPK_TILE;
       i += MPK_NT) {
    ValIdx c = vi_leaf(mode, x, y, i);
    if (!have || vi_better(mode, c, best)) best = c;
    have = true;
  }
  if (!have) best = ValIdx{NAN, -1};
  sh[threadIdx.x] = best;
  __syncthreads();
  for (int w = MPK_NT / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) {
      ValIdx a = sh[threadIdx.x], b = sh[threadIdx.x + w];
      if (b.i >= 0 && (a.i < 0 || vi_better(mode, b, a))) sh[threadIdx.x] = b;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) P[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(MPK_NT) k_vi_final(int mode, const ValIdx *P, int T, ValIdx *out) {
  __shared__ ValIdx sh[MPK_NT];
  ValIdx best{NAN, -1};
  for (int b = threadIdx.x; b < T; b += MPK_NT) {
    ValIdx c = P[b];
    if (c.i >= 0 && (best.i < 0 || vi_better(mode, c, best))) best = c;
  }
  sh[threadIdx.x] = best;
  __syncthreads();
  for (int w = MPK_NT / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) {
      ValIdx a = sh[threadIdx.x], b = sh[threadIdx.x + w];
      if (b.i >= 0 && (a.i < 0 || vi_better(mode, b, a))) sh[threadIdx.x] = b;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

static ValIdx scan_host(int mode, const double *x, const double *y, i64 n) {
  Ctx *c = ctx();
  ValIdx r{NAN, -1};
  if (!c || n <= 0) return r;
  unsigned T = ntiles(n);
  ValIdx *P = reinterpret_cast<ValIdx *>(partials((size_t)2 * T + 2));
  ValIdx *out = reinterpret_cast<ValIdx *>(c->dres);
  MPK_LAUNCH(k_vi_partials, T, MPK_NT, mode, x, y, n, P);
  MPK_LAUNCH(k_vi_final, 1, MPK_NT, mode, (const ValIdx *)P, (int)T, out);
  HIPCHK(hipMemcpyAsync(c->hres, c->dres, sizeof(ValIdx), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  memcpy(&r, c->hres, sizeof r);
  return r;
}

}  // namespace mpk

using namespace mpk;

#define DP(a) reinterpret_cast<double *>(a)
#define CDP(a) reinterpret_cast<const double *>(a)
#define AL3(a, b, c) (aligned16(DP(a)) && aligned16(DP(b)) && aligned16(DP(c)))

extern "C" {

// ---- pointwise (ported.mojo:75-98)
void mp_vec_pointwise_mult(int64_t w, int64_t x, int64_t y, int64_t n) {
  ew(n, Map2<FMul>{DP(w), CDP(x), CDP(y), AL3(w, x, y), FMul{}});
}
void mp_vec_pointwise_divide(int64_t w, int64_t x, int64_t y, int64_t n) {
  ew(n, Map2<FDiv>{DP(w), CDP(x), CDP(y), AL3(w, x, y), FDiv{}});
}
void mp_vec_pointwise_min(int64_t w, int64_t x, int64_t y, int64_t n) {
  ew(n, Map2<FMin>{DP(w), CDP(x), CDP(y), AL3(w, x, y), FMin{}});
}
void mp_vec_pointwise_max(int64_t w, int64_t x, int64_t y, int64_t n) {
  ew(n, Map2<FMax>{DP(w), CDP(x), CDP(y), AL3(w, x, y), FMax{}});
}
void mp_vec_pointwise_max_abs(int64_t w, int64_t x, int64_t y, int64_t n) {
  ew(n, Map2<FMaxAbs>{DP(w), CDP(x), CDP(y), AL3(w, x, y), FMaxAbs{}});
}

// ---- reductions (ported.mojo:100-168)
void mp_vec_norm(int64_t x, int64_t ntype, int64_t z, int64_t n) {
  double z0 = 0.0, z1 = 0.0;
  if (n) {
    if (ntype == 1 || ntype == 2 || ntype == 4) {
      if (ntype == 4) {
        z0 = reduce1_host(CDP(x), n, RK_ASUM);
        z1 = norm2_host(CDP(x), n);
      } else {
        z0 = norm2_host(CDP(x), n);
      }
    } else if (ntype == 3) {
      z0 = scan_host(0, CDP(x), nullptr, n).v;
    } else if (ntype == 0) {
      z0 = reduce1_host(CDP(x), n, RK_ASUM);
    }
  }
  DP(z)[0] = z0;
  if (ntype == 4) DP(z)[1] = z1;
}
void mp_vec_dot(int64_t x, int64_t y, int64_t val, int64_t n) { *DP(val) = dot_host(CDP(x), CDP(y), n); }

// mp_vec_dot's kernels over the fixed scratch array instead of the grown
// partials, the result left in device memory: nothing is allocated, nothing
// waits.  See include/mpk.h.
int64_t mpk_vec_dot_noalloc(int64_t x, int64_t y, int64_t val, int64_t n) {
  Ctx *c = ctx_ready();  // never initialises: initialising allocates
  if (!c) {
    set_error(4, "mpk_vec_dot_noalloc: the library is not initialised (call mpk_init first)");
    return -1;
  }
  if (!c->fixed) {
    set_error(4, "mpk_vec_dot_noalloc: no fixed scratch on this device");
    return -1;
  }
  if (n < 0 || !val) {
    set_error(4, "mpk_vec_dot_noalloc: n must be >= 0 and val non-null");
    return -1;
  }
  const RedSet<1> rs{{{CDP(x), CDP(y), RK_DOT}}};
  if (red_ref()) {
    const RefJob J = ref_job<1>(rs);
    if (failed()) return -1;
    ref_launch(J, n, c->fixed, nullptr);
    MPK_LAUNCH(k_finalize<NoEpilogue>, 1, MPK_NT, (const double *)c->fixed, 1, 1, DP(val), NoEpilogue{},
               nullptr);
    return failed() ? -1 : 0;
  }
  const unsigned T = n > 0 ? ntiles(n) : 1u;
  if ((size_t)T > c->fixed_cap) {
    set_error(4, "mpk_vec_dot_noalloc: n = %lld needs %u partials, more than the %zu of the fixed scratch "
              "(at most %lld elements)", (long long)n, T, c->fixed_cap, (long long)c->fixed_cap * MPK_TILE);
    return -1;
  }
  MPK_LAUNCH(k_partials<1>, T, MPK_NT, rs, n, c->fixed, nullptr);
  MPK_LAUNCH(k_finalize<NoEpilogue>, 1, MPK_NT, (const double *)c->fixed, (int)T, 1, DP(val), NoEpilogue{},
             nullptr);
  return failed() ? -1 : 0;
}
void mp_vec_tdot(int64_t x, int64_t y, int64_t val, int64_t n) { *DP(val) = dot_host(CDP(y), CDP(x), n); }
void mp_vec_dot_norm2(int64_t x, int64_t y, int64_t sval, int64_t rval, int64_t n) {
  RedSet<2> rs{{{CDP(x), CDP(y), RK_DOT}, {CDP(x), CDP(x), RK_SUMSQ}}};
  double r[2];
  reduce_host<2>(rs, n, r);
  *DP(sval) = r[0];
  *DP(rval) = r[1];
}
void mp_vec_sum(int64_t x, int64_t val, int64_t n) { *DP(val) = reduce1_host(CDP(x), n, RK_SUM); }
void mp_vec_mean(int64_t x, int64_t val, int64_t n) {
  *DP(val) = reduce1_host(CDP(x), n, RK_SUM) / (double)n;
}
void mp_vec_max(int64_t x, int64_t idx, int64_t z, int64_t n) {
  if (!n) {
    *DP(z) = -1.7976931348623157e308;
    *reinterpret_cast<int64_t *>(idx) = -1;
    return;
  }
  ValIdx r = scan_host(1, CDP(x), nullptr, n);
  *DP(z) = r.v;
  *reinterpret_cast<int64_t *>(idx) = r.i;
}
void mp_vec_min(int64_t x, int64_t idx, int64_t z, int64_t n) {
  if (!n) {
    *DP(z) = 1.7976931348623157e308;
    *reinterpret_cast<int64_t *>(idx) = -1;
    return;
  }
  ValIdx r = scan_host(2, CDP(x), nullptr, n);
  *DP(z) = r.v;
  *reinterpret_cast<int64_t *>(idx) = r.i;
}
void mp_vec_normalize(int64_t x, int64_t v, int64_t n) {
  double z = norm2_host(CDP(x), n);
  if (z > 0.0) vdivide(DP(x), z, n, nullptr);
  *DP(v) = z;
}
void mp_vec_max_pointwise_divide(int64_t x, int64_t y, int64_t rval, int64_t n) {
  if (!n) {
    *DP(rval) = 0.0;
    return;
  }
  ValIdx r = scan_host(3, CDP(x), CDP(y), n);
  double m = r.v;
  *DP(rval) = (m > 0.0) ? m : 0.0;
}

// ---- in place (ported.mojo:200-275)
void mp_vec_reciprocal(int64_t x, int64_t n) { ew(n, Map1<FRecip>{DP(x), CDP(x), aligned16(DP(x)), FRecip{}}); }
void mp_vec_exp(int64_t x, int64_t n) { ew(n, Map1<FExp>{DP(x), CDP(x), aligned16(DP(x)), FExp{}}); }
void mp_vec_log(int64_t x, int64_t n) { ew(n, Map1<FLog>{DP(x), CDP(x), aligned16(DP(x)), FLog{}}); }
void mp_vec_sqrtabs(int64_t x, int64_t n) {
  ew(n, Map1<FSqrtAbs>{DP(x), CDP(x), aligned16(DP(x)), FSqrtAbs{}});
}
void mp_vec_chop(int64_t x, int64_t n, double tol) {
  ew(n, Map1<FChop>{DP(x), CDP(x), aligned16(DP(x)), FChop{tol}});
}
void mp_vec_abs(int64_t x, int64_t n) { ew(n, Map1<FAbs>{DP(x), CDP(x), aligned16(DP(x)), FAbs{}}); }
void mp_vec_conjugate(int64_t x, int64_t n) { (void)x, (void)n; }
void mp_vec_copy(int64_t x, int64_t y, int64_t n) { vcopy(CDP(x), DP(y), n); }
void mp_vec_swap(int64_t x, int64_t y, int64_t n) {
  ew(n, OpSwap{DP(x), DP(y), aligned16(DP(x)) && aligned16(DP(y))});
}
void mp_vec_set(int64_t x, double alpha, int64_t n) { vset(DP(x), alpha == 0.0 ? 0.0 : alpha, n, nullptr); }
void mp_vec_zero_entries(int64_t x, int64_t n) { vset(DP(x), 0.0, n, nullptr); }
void mp_vec_scale(int64_t x, double alpha, int64_t n) { vscale(DP(x), alpha, n, nullptr); }
void mp_vec_shift(int64_t x, double alpha, int64_t n) {
  ew(n, Map1<FShift>{DP(x), CDP(x), aligned16(DP(x)), FShift{alpha}});
}

// ---- axpy family (ported.mojo:278-335)
void mp_vec_axpy(int64_t y, double alpha, int64_t x, int64_t n) { vaxpy(DP(y), alpha, CDP(x), n, nullptr); }
void mp_vec_aypx(int64_t y, double alpha, int64_t x, int64_t n) { vaypx(DP(y), alpha, CDP(x), n, nullptr); }
void mp_vec_axpby(int64_t y, double alpha, double beta, int64_t x, int64_t n) {
  if (beta == 0.0)
    ew(n, Map1<FScale>{DP(y), CDP(x), aligned16(DP(y)) && aligned16(CDP(x)), FScale{alpha}});
  else if (alpha == 0.0)
    vscale(DP(y), beta, n, nullptr);
  else
    vaxpby_raw(DP(y), alpha, beta, CDP(x), n, nullptr);
}
void mp_vec_axpbypcz(int64_t z, double alpha, double beta, double gamma, int64_t x, int64_t y, int64_t n) {
  vaxpbypcz(DP(z), alpha, beta, gamma, CDP(x), CDP(y), n, nullptr);
}
void mp_vec_waxpy(int64_t w, double alpha, int64_t x, int64_t y, int64_t n) {
  vwaxpy(DP(w), alpha, CDP(x), CDP(y), n, nullptr);
}
void mp_vec_maxpy(int64_t y, int64_t nv, int64_t alpha, int64_t x, int64_t n) {
  // alpha: host array of nv doubles; x: host array of nv device addresses
  vmaxpy(DP(y), (int)nv, CDP(alpha), reinterpret_cast<const double *const *>(x), n, false, nullptr);
}

}  // extern "C"

namespace mpk {
// y = x + alpha y with the reference's special cases (ported.mojo:285-294)
void vaypx(double *y, double alpha, const double *x, i64 n, const int64_t *stop) {
  if (alpha == 0.0)
    vcopy(x, y, n);
  else if (alpha == 1.0)
    vaxpy(y, alpha, x, n, stop);
  else
    vaypx_raw(y, alpha, x, n, stop);
}
}  // namespace mpk
