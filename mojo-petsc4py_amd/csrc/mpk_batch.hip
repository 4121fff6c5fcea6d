// mpk_batch.hip -- thousands of independent small systems in one persistent
// launch.  Every workgroup takes systems from a queue (largest first) and
// solves each one alone: FP64 iterative refinement (true residual b - A x and
// the update of x in FP64) around an FP32 inner solve -- CG with Jacobi, or
// GMRES(m) with ILU(0) and modified Gram-Schmidt -- whose dot products,
// norms, scalars and Hessenberg least squares are FP64.  Everything a system
// needs is synchronised by the workgroup's own barriers; workgroups never
// wait on each other.
//
// Layout (concatenated over the systems, rows of system s at row0..row0+n):
// rp int32[N+1] (global entry offsets), ci int32 (global columns), v64 float64
// and v32 float32 values; jd32 float32 (Jacobi 1/a_ii); ILU(0) sweeps as
// single-block plans (positions by level): for each of the forward and
// backward sweep, prow int32[N] (local row of position), pcnt int32[N], 4
// fixed entry slots per position (dep int32 = position slot, coef float32),
// level starts int32 (per system, at the system's level offset), and dinv32
// float32[N] (the factor's inverted diagonal).
#include <cmath>
#include <cstring>
#include <vector>

#include "mpk_internal.h"
#include "../../include/mpk.h"

namespace mpk {

enum { BS_CG = 0, BS_GMRES = 1 };
#define BS_M 20      // GMRES restart
#define BS_W 4       // sweep entry slots per position
#define BS_FIELDS 8  // per-system descriptor: row0, n, kind, lf0, nlf, lb0, nlb, 0

struct BatchArgs {
  i64 nsys;
  const i64 *sys;    // [nsys][BS_FIELDS]
  const int *order;  // processing order (largest first)
  const int *rp, *ci;
  const double *v64;
  const float *v32, *jd32;
  const int *fprow, *fcnt, *fdep, *bprow, *bcnt, *bdep, *flv, *blv;
  const float *fcoef, *bcoef, *dinv32;
  const double *b;
  double *x;
  double *stats;  // [nsys][4]: relres, outer, inner, converged
  float *work;    // per workgroup: ws floats
  double *work64; // per workgroup: ws64 doubles
  i64 ws, ws64;
  unsigned *queue;
  double rtol, inner_rtol;
  int max_outer, max_inner;
};

// Sum over the workgroup, result in every thread.  `sh` holds BS/64 doubles
// plus one broadcast slot.
template <int BS>
__device__ __forceinline__ double wg_sum(double v, double *sh) {
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) v += __shfl_down(v, w, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < BS / 64; k++) s += sh[k];
    sh[BS / 64] = s;
  }
  __syncthreads();
  return sh[BS / 64];
}

template <int BS>
struct Sys {
  const BatchArgs &a;
  i64 row0, n;
  double *sh;
  int t;
  __device__ double dot(const float *x, const float *y) const {
    double s = 0.0;
    for (i64 i = t; i < n; i += BS) s = fma((double)x[i], (double)y[i], s);
    return wg_sum<BS>(s, sh);
  }
  // y = A x (FP32 values; x, y local vectors)
  __device__ void spmv32(const float *x, float *y) const {
    for (i64 i = t; i < n; i += BS) {
      const int j0 = a.rp[row0 + i], j1 = a.rp[row0 + i + 1];
      float s = 0.0f;
      for (int j = j0; j < j1; j++) s = fmaf(a.v32[j], x[a.ci[j] - row0], s);
      y[i] = s;
    }
    __syncthreads();
  }
  // r64 = b - A x (FP64), returns ||r64||
  __device__ double residual(double *r64) const {
    double s = 0.0;
    for (i64 i = t; i < n; i += BS) {
      const int j0 = a.rp[row0 + i], j1 = a.rp[row0 + i + 1];
      double ax = 0.0;
      for (int j = j0; j < j1; j++) ax = fma(a.v64[j], a.x[a.ci[j]], ax);
      const double r = a.b[row0 + i] - ax;
      r64[i] = r;
      s = fma(r, r, s);
    }
    return sqrt(wg_sum<BS>(s, sh));
  }
  // y = M^-1 x for ILU(0): forward sweep into val (by position), scatter F by
  // row, backward sweep, y by row
  __device__ void ilu(const float *x, float *y, float *val, float *F, i64 lf0, i64 nlf, i64 lb0, i64 nlb) const {
    for (i64 l = 0; l < nlf; l++) {
      const int pa = a.flv[lf0 + l], pb = a.flv[lf0 + l + 1];
      for (int p = pa + t; p < pb; p += BS) {
        const i64 gp = row0 + p;
        const int i = a.fprow[gp];
        float acc = x[i];
        for (int e = 0; e < a.fcnt[gp]; e++) acc = fmaf(a.fcoef[gp * BS_W + e], val[a.fdep[gp * BS_W + e]], acc);
        val[p] = acc;
        F[i] = acc;
      }
      __syncthreads();
    }
    for (i64 l = 0; l < nlb; l++) {
      const int pa = a.blv[lb0 + l], pb = a.blv[lb0 + l + 1];
      for (int p = pa + t; p < pb; p += BS) {
        const i64 gp = row0 + p;
        const int i = a.bprow[gp];
        float acc = F[i];
        for (int e = 0; e < a.bcnt[gp]; e++) acc = fmaf(a.bcoef[gp * BS_W + e], val[a.bdep[gp * BS_W + e]], acc);
        const float v = acc * a.dinv32[row0 + i];
        val[p] = v;
        y[i] = v;
      }
      __syncthreads();
    }
  }
};

template <int BS>
__global__ void __launch_bounds__(BS) k_batch(BatchArgs a) {
  __shared__ double sh[BS / 64 + 1];
  __shared__ double H[(BS_M + 1) * BS_M], cs[BS_M], sn[BS_M], g[BS_M + 1], yv[BS_M];
  __shared__ int sys_sh;
  const int t = threadIdx.x;
  float *W = a.work + (i64)blockIdx.x * a.ws;
  double *r64 = a.work64 + (i64)blockIdx.x * a.ws64;
  for (;;) {
    if (t == 0) sys_sh = (int)atomicAdd(a.queue, 1u);
    __syncthreads();
    const i64 q = sys_sh;
    __syncthreads();
    if (q >= a.nsys) return;
    const i64 s = a.order[q];
    const i64 *d = a.sys + s * BS_FIELDS;
    Sys<BS> S{a, d[0], d[1], sh, t};
    const i64 n = S.n, row0 = S.row0;
    const int kind = (int)d[2];
    // local FP32 vectors
    float *rhs = W, *dv = W + n, *r = W + 2 * n, *z = W + 3 * n, *p = W + 4 * n, *w = W + 5 * n,
          *F = W + 6 * n, *val = W + 7 * n, *V = W + 8 * n;
    for (i64 i = t; i < n; i += BS) a.x[row0 + i] = 0.0;
    __syncthreads();
    double bn = 0.0;
    for (i64 i = t; i < n; i += BS) bn = fma(a.b[row0 + i], a.b[row0 + i], bn);
    bn = sqrt(wg_sum<BS>(bn, sh));
    int outer = 0, inner = 0;
    double rel = 0.0;
    for (;;) {
      const double rn = S.residual(r64);
      rel = bn > 0.0 ? rn / bn : rn;
      if (rel <= a.rtol || outer >= a.max_outer || !(rn > 0.0)) break;
      const double inv = 1.0 / rn;
      for (i64 i = t; i < n; i += BS) {
        rhs[i] = (float)(r64[i] * inv);
        dv[i] = 0.0f;
      }
      __syncthreads();
      if (kind == BS_CG) {
        // CG + Jacobi on A d = rhs (||rhs|| = 1), zero guess
        for (i64 i = t; i < n; i += BS) {
          r[i] = rhs[i];
          z[i] = a.jd32[row0 + i] * rhs[i];
          p[i] = z[i];
        }
        __syncthreads();
        double rz = S.dot(r, z);
        for (int it = 0; it < a.max_inner; it++) {
          S.spmv32(p, w);
          const double pq = S.dot(p, w);
          if (!(pq != 0.0)) break;
          const float al = (float)(rz / pq);
          for (i64 i = t; i < n; i += BS) {
            dv[i] = fmaf(al, p[i], dv[i]);
            r[i] = fmaf(-al, w[i], r[i]);
            z[i] = a.jd32[row0 + i] * r[i];
          }
          __syncthreads();
          inner++;
          const double rr = S.dot(r, r);
          if (sqrt(rr) <= a.inner_rtol) break;
          const double rz2 = S.dot(r, z);
          const float be = (float)(rz2 / rz);
          rz = rz2;
          for (i64 i = t; i < n; i += BS) p[i] = fmaf(be, p[i], z[i]);
          __syncthreads();
        }
      } else {
        // GMRES(m) + ILU(0), left preconditioned, MGS, zero guess
        double target = -1.0;
        int its = 0;
        for (;;) {
          // z = rhs - A dv (dv = 0 on the first cycle)
          S.spmv32(dv, w);
          for (i64 i = t; i < n; i += BS) z[i] = rhs[i] - w[i];
          __syncthreads();
          S.ilu(z, w, val, F, d[3], d[4], d[5], d[6]);
          const double beta = sqrt(S.dot(w, w));
          if (target < 0.0) target = a.inner_rtol * beta;
          if (!(beta > target)) break;
          for (i64 i = t; i < n; i += BS) V[i] = (float)(w[i] / beta);
          if (t == 0) {
            g[0] = beta;
            for (int k = 1; k <= BS_M; k++) g[k] = 0.0;
          }
          __syncthreads();
          int kk = 0;
          bool done = false;
          for (int k = 0; k < BS_M && its < a.max_inner; k++) {
            const float *vk = V + (i64)k * n;
            float *vn = V + (i64)(k + 1) * n;
            S.spmv32(vk, z);
            S.ilu(z, w, val, F, d[3], d[4], d[5], d[6]);
            for (int j = 0; j <= k; j++) {
              const float *vj = V + (i64)j * n;
              const double h = S.dot(vj, w);
              if (t == 0) H[k * (BS_M + 1) + j] = h;
              const float hf = (float)h;
              for (i64 i = t; i < n; i += BS) w[i] = fmaf(-hf, vj[i], w[i]);
              __syncthreads();
            }
            const double hn = sqrt(S.dot(w, w));
            for (i64 i = t; i < n; i += BS) vn[i] = hn > 0.0 ? (float)(w[i] / hn) : 0.0f;
            if (t == 0) {
              double *col = H + k * (BS_M + 1);
              col[k + 1] = hn;
              for (int j = 0; j < k; j++) {
                const double x0 = col[j], x1 = col[j + 1];
                col[j] = cs[j] * x0 + sn[j] * x1;
                col[j + 1] = -sn[j] * x0 + cs[j] * x1;
              }
              const double x0 = col[k], x1 = col[k + 1], rr = sqrt(x0 * x0 + x1 * x1);
              cs[k] = rr > 0.0 ? x0 / rr : 1.0;
              sn[k] = rr > 0.0 ? x1 / rr : 0.0;
              col[k] = rr;
              col[k + 1] = 0.0;
              g[k + 1] = -sn[k] * g[k];
              g[k] = cs[k] * g[k];
              sh[BS / 64] = fabs(g[k + 1]);
            }
            __syncthreads();
            const double res = sh[BS / 64];
            __syncthreads();
            its++;
            inner++;
            kk = k + 1;
            if (res <= target || !(hn > 0.0)) {
              done = true;
              break;
            }
          }
          if (t == 0) {
            for (int i = kk - 1; i >= 0; i--) {
              double acc = g[i];
              for (int j = i + 1; j < kk; j++) acc -= H[j * (BS_M + 1) + i] * yv[j];
              const double dd = H[i * (BS_M + 1) + i];
              yv[i] = dd != 0.0 ? acc / dd : 0.0;
            }
          }
          __syncthreads();
          for (i64 i = t; i < n; i += BS) {
            float acc = dv[i];
            for (int j = 0; j < kk; j++) acc = fmaf((float)yv[j], V[(i64)j * n + i], acc);
            dv[i] = acc;
          }
          __syncthreads();
          if (done || its >= a.max_inner) break;
        }
      }
      // x += ||r|| d in FP64
      for (i64 i = t; i < n; i += BS) a.x[row0 + i] = fma(rn, (double)dv[i], a.x[row0 + i]);
      __syncthreads();
      outer++;
    }
    if (t == 0) {
      double *st = a.stats + s * 4;
      st[0] = rel;
      st[1] = outer;
      st[2] = inner;
      st[3] = rel <= a.rtol ? 1.0 : 0.0;
    }
    __syncthreads();
  }
}

__global__ void k_batch_reset(unsigned *q) { *q = 0u; }

struct Batch {
  BatchArgs a{};
  i64 nmax = 0;
  int bs = 256, grid = 0;
  unsigned *queue = nullptr;
  float *work = nullptr;
  double *work64 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  BatchArgs captured{};  // the arguments the graph was captured with
  ~Batch() {
    hipStreamSynchronize(stream());
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    if (queue) hipFree(queue);
    if (work) hipFree(work);
    if (work64) hipFree(work64);
  }
  bool size_work(int g) {
    if (g <= grid && work) return true;
    if (work) HIPCHK(hipFree(work));
    if (work64) HIPCHK(hipFree(work64));
    work = nullptr;
    work64 = nullptr;
    a.ws = (8 + BS_M + 1) * nmax;
    a.ws64 = nmax;
    HIPCHK(hipMalloc(&work, (size_t)g * a.ws * sizeof(float)));
    HIPCHK(hipMalloc(&work64, (size_t)g * a.ws64 * sizeof(double)));
    grid = g;
    a.work = work;
    a.work64 = work64;
    return !failed();
  }
  void launch(int bs_, int g) {
    MPK_LAUNCH(k_batch_reset, 1, 1, queue);
    switch (bs_) {
      case 128:
        MPK_LAUNCH(k_batch<128>, g, 128, a);
        break;
      case 512:
        MPK_LAUNCH(k_batch<512>, g, 512, a);
        break;
      default:
        MPK_LAUNCH(k_batch<256>, g, 256, a);
    }
  }
};

}  // namespace mpk

using namespace mpk;

extern "C" {

// See include/mpk.h.
int64_t mpk_batch_create(int64_t nsys, int64_t sys, int64_t order, int64_t rp, int64_t ci, int64_t v64,
                         int64_t v32, int64_t jd32, int64_t fprow, int64_t fcnt, int64_t fdep, int64_t fcoef,
                         int64_t flv, int64_t bprow, int64_t bcnt, int64_t bdep, int64_t bcoef, int64_t blv,
                         int64_t dinv32, int64_t nmax) {
  if (nsys <= 0 || !ctx()) return 0;
  if (nmax <= 0) {
    set_error(7, "batch: nmax must be positive");
    return 0;
  }
  Batch *B = new Batch();
  B->nmax = nmax;
  BatchArgs &a = B->a;
  a.nsys = nsys;
  a.sys = (const i64 *)sys;
  a.order = (const int *)order;
  a.rp = (const int *)rp;
  a.ci = (const int *)ci;
  a.v64 = (const double *)v64;
  a.v32 = (const float *)v32;
  a.jd32 = (const float *)jd32;
  a.fprow = (const int *)fprow;
  a.fcnt = (const int *)fcnt;
  a.fdep = (const int *)fdep;
  a.fcoef = (const float *)fcoef;
  a.flv = (const int *)flv;
  a.bprow = (const int *)bprow;
  a.bcnt = (const int *)bcnt;
  a.bdep = (const int *)bdep;
  a.bcoef = (const float *)bcoef;
  a.blv = (const int *)blv;
  a.dinv32 = (const float *)dinv32;
  HIPCHK(hipMalloc(&B->queue, 64 * sizeof(unsigned)));
  a.queue = B->queue;
  if (failed()) {
    delete B;
    return 0;
  }
  return (int64_t)B;
}

// One batched solve: x [D] (concatenated) from b [D]; stats [D] float64
// [nsys][4] = {relres, outer, inner, converged}.  cfg [H] float64 in {rtol,
// inner_rtol, max_outer, max_inner, workgroup size (128/256/512), grid
// (workgroups), use_graph}.  With use_graph the queue reset and the
// persistent kernel are captured once into a HIP graph and replayed.
// Returns 0, or -1 on error.
int64_t mpk_batch_solve(int64_t h, int64_t b, int64_t x, int64_t stats, int64_t cfg_) {
  Batch *B = (Batch *)h;
  const double *cfg = (const double *)cfg_;
  if (!B || !ctx()) return -1;
  BatchArgs &a = B->a;
  a.b = (const double *)b;
  a.x = (double *)x;
  a.stats = (double *)stats;
  a.rtol = cfg[0];
  a.inner_rtol = cfg[1];
  a.max_outer = (int)cfg[2];
  a.max_inner = (int)cfg[3];
  const int bs = (int)cfg[4], g = (int)cfg[5];
  const bool graph = cfg[6] != 0.0;
  if (bs != 128 && bs != 256 && bs != 512) {
    set_error(7, "batch: workgroup size must be 128, 256 or 512");
    return -1;
  }
  if (g < 1 || g > 65536) {
    set_error(7, "batch: grid must be 1..65536");
    return -1;
  }
  if (!B->size_work(g)) return -1;
  if (graph) {
    // re-capture when the launch shape or any argument changes
    if (B->exec && (B->bs != bs || B->grid != g || std::memcmp(&B->captured, &a, sizeof a) != 0)) {
      HIPCHK(hipGraphExecDestroy(B->exec));
      HIPCHK(hipGraphDestroy(B->graph));
      B->exec = nullptr;
      B->graph = nullptr;
    }
    if (!B->exec) {
      HIPCHK(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
      B->launch(bs, g);
      HIPCHK(hipStreamEndCapture(stream(), &B->graph));
      if (failed()) return -1;
      HIPCHK(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
      B->bs = bs;
      B->captured = a;
    }
    HIPCHK(hipGraphLaunch(B->exec, stream()));
  } else {
    B->launch(bs, g);
  }
  sync();
  return failed() ? -1 : 0;
}

void mpk_batch_free(int64_t h) { delete (Batch *)h; }

}  // extern "C"
