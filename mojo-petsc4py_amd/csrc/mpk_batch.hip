// mpk_batch.hip -- thousands of independent small systems in one persistent
// launch.  Every workgroup takes systems from a queue (largest first) and
// solves each one alone: FP64 iterative refinement (true residual b - A x and
// the update of x in FP64) around an FP32 inner solve -- CG with Jacobi, or
// GMRES(m) with ILU(0) and classical Gram-Schmidt -- whose dot products,
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
// level starts int32 (per system, at the system's level offset), dinv32
// float32[N] (the factor's inverted diagonal, in backward position order) and
// f2b int32[N] (forward position -> backward position of the same row).
#include <cmath>
#include <cstring>
#include <vector>

#include "mpk_internal.h"
#include "../../include/mpk.h"

namespace mpk {

enum { BS_CG = 0, BS_GMRES = 1 };
#define BS_M 20      // GMRES restart
#define BS_W 4       // sweep entry slots per position
#define BS_FIELDS 8  // per-system descriptor: row0, n, kind, lf0, nlf, lb0, nlb, ring
#define BS_RING 2048 // LDS ring of sweep values (systems whose plans fit it)

struct BatchArgs {
  i64 nsys;
  const i64 *sys;    // [nsys][BS_FIELDS]
  const int *order;  // processing order (largest first)
  const int *rp, *ci;
  const double *v64;
  const float *v32, *jd32;
  const int *fprow, *fcnt, *fdep, *bprow, *bcnt, *bdep, *flv, *blv;
  const float *fcoef, *bcoef, *dinv32;
  const int *f2b;  // forward position -> backward position of the same row
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

// Sum over the workgroup, result in every thread: wave sums into one of two
// LDS buffers (alternating per call, so one barrier suffices: a buffer is
// written again only after every thread has passed the next call's
// barrier), then every thread adds the wave sums in wave order.
template <int BS>
__device__ __forceinline__ double wg_sum(double v, double *sh, int &par) {
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) v += __shfl_down(v, w, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double *buf = sh + (par ? BS / 64 : 0);
  par ^= 1;
  if (lane == 0) buf[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < BS / 64; k++) s += buf[k];
  return s;
}

// A system's solve by one workgroup of BS threads.  The tile parameters are
// the autotuner's choices: UR, the rows a thread takes per pass of the FP32
// product (a wave covers 64 UR rows a pass); KB, the product's variant --
// the entries of a row loaded per round before their gathers (a 5-point row
// takes two rounds at 4, one at 8); UG, the reduction tiling of the
// Gram-Schmidt sums and updates -- the elements a thread takes per step for
// four basis vectors at once.
template <int BS, int UR_, int KB_, int UG_>
struct Sys {
  const BatchArgs &a;
  i64 row0, n;
  double *sh;
  int t;
  int par = 0;  // wg_sum's buffer
  __device__ double wsum(double v) { return wg_sum<BS>(v, sh, par); }
  // The thread's elements are i = t, t + BS, ...; every loop below takes
  // them U at a time and issues all of a group's loads before it uses any,
  // so a thread waits about one memory latency per group rather than one
  // per element (the order of the arithmetic is the element order).
  static constexpr int U = 8;
  static constexpr int UG = UG_;  // the same for the Gram-Schmidt passes (four basis vectors each)
  __device__ double dot(const float *__restrict__ x, const float *__restrict__ y) {
    double s = 0.0;
    for (i64 i0 = t; i0 < n; i0 += U * BS) {
      float xv[U], yv[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const i64 i = i0 + (i64)u * BS;
        xv[u] = i < n ? x[i] : 0.0f;
        yv[u] = i < n ? y[i] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if (i0 + (i64)u * BS < n) s = fma((double)xv[u], (double)yv[u], s);
    }
    return wsum(s);
  }
  // y = A x (FP32 values; x, y local vectors): UR rows at a time, their
  // entries KB at a time -- every index and value of a block loaded first,
  // then every gathered x, so a block costs two memory round trips (each
  // row's sum still in entry order)
  static constexpr int UR = UR_, KB = KB_;
  __device__ void spmv32(const float *__restrict__ x, float *__restrict__ y) const {
    for (i64 i0 = t; i0 < n; i0 += UR * BS) {
      int j0[UR], len[UR], lmax = 0;
#pragma unroll
      for (int u = 0; u < UR; u++) {
        const i64 i = i0 + (i64)u * BS;
        j0[u] = i < n ? a.rp[row0 + i] : 0;
        len[u] = i < n ? a.rp[row0 + i + 1] - j0[u] : 0;
        lmax = len[u] > lmax ? len[u] : lmax;
      }
      float sum[UR];
#pragma unroll
      for (int u = 0; u < UR; u++) sum[u] = 0.0f;
      for (int k0 = 0; k0 < lmax; k0 += KB) {
        int c[UR][KB];
        float v[UR][KB], xv[UR][KB];
#pragma unroll
        for (int u = 0; u < UR; u++)
#pragma unroll
          for (int q = 0; q < KB; q++) {
            const bool in = k0 + q < len[u];
            c[u][q] = in ? a.ci[j0[u] + k0 + q] - (int)row0 : 0;
            v[u][q] = in ? a.v32[j0[u] + k0 + q] : 0.0f;
          }
#pragma unroll
        for (int u = 0; u < UR; u++)
#pragma unroll
          for (int q = 0; q < KB; q++) xv[u][q] = k0 + q < len[u] ? x[c[u][q]] : 0.0f;
#pragma unroll
        for (int u = 0; u < UR; u++)
#pragma unroll
          for (int q = 0; q < KB; q++)
            if (k0 + q < len[u]) sum[u] = fmaf(v[u][q], xv[u][q], sum[u]);
      }
#pragma unroll
      for (int u = 0; u < UR; u++)
        if (i0 + (i64)u * BS < n) y[i0 + (i64)u * BS] = sum[u];
    }
    __syncthreads();
  }
  // r64 = b - A x (FP64), returns ||r64||
  __device__ double residual(double *__restrict__ r64) {
    double s = 0.0;
    for (i64 i0 = t; i0 < n; i0 += U * BS) {
      int j0[U], len[U], lmax = 0;
#pragma unroll
      for (int u = 0; u < U; u++) {
        const i64 i = i0 + (i64)u * BS;
        j0[u] = i < n ? a.rp[row0 + i] : 0;
        len[u] = i < n ? a.rp[row0 + i + 1] - j0[u] : 0;
        lmax = len[u] > lmax ? len[u] : lmax;
      }
      double ax[U];
#pragma unroll
      for (int u = 0; u < U; u++) ax[u] = 0.0;
      for (int k = 0; k < lmax; k++) {
        double v[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const bool in = k < len[u];
          v[u] = in ? a.v64[j0[u] + k] : 0.0;
          xv[u] = in ? a.x[a.ci[j0[u] + k]] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
          if (k < len[u]) ax[u] = fma(v[u], xv[u], ax[u]);
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const i64 i = i0 + (i64)u * BS;
        if (i < n) {
          const double r = a.b[row0 + i] - ax[u];
          r64[i] = r;
          s = fma(r, r, s);
        }
      }
    }
    return sqrt(wsum(s));
  }
  // Classical Gram-Schmidt over the basis V_0..V_k (V_j at V + j n):
  // h_j = V_j . w for every j in one pass over w and the basis (four basis
  // vectors per pass, each thread's elements in order), then the workgroup
  // sums into hsh[0..k] (LDS; buf holds (BS_M + 1) BS/64 partial sums).
  __device__ void mdot(const float *__restrict__ V, const float *__restrict__ w, int k, double *hsh,
                       double *buf) const {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j0 = 0; j0 <= k; j0 += 4) {
      const float *vq[4];
#pragma unroll
      for (int q = 0; q < 4; q++) vq[q] = V + (i64)(j0 + q <= k ? j0 + q : k) * n;
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      for (i64 i0 = t; i0 < n; i0 += UG * BS) {
        float wv[UG], vv[4][UG];
#pragma unroll
        for (int u = 0; u < UG; u++) {
          const i64 i = i0 + (i64)u * BS;
          wv[u] = i < n ? w[i] : 0.0f;
#pragma unroll
          for (int q = 0; q < 4; q++) vv[q][u] = i < n ? vq[q][i] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
          for (int u = 0; u < UG; u++)
            if (i0 + (i64)u * BS < n) acc[q] = fma((double)vv[q][u], (double)wv[u], acc[q]);
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        double v = acc[q];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0 && j0 + q <= k) buf[(j0 + q) * (BS / 64) + wave] = v;
      }
    }
    __syncthreads();
    if (t <= k) {
      double sum = 0.0;
      for (int q = 0; q < BS / 64; q++) sum += buf[t * (BS / 64) + q];
      hsh[t] = sum;
    }
    __syncthreads();
  }
  // w -= sum_j h_j V_j (j <= k, in j order for each element), four basis
  // vectors per pass
  __device__ void maxpy(const float *__restrict__ V, const double *hsh, int k, float *__restrict__ w) const {
    for (int j0 = 0; j0 <= k; j0 += 4) {
      const int nq = k - j0 + 1 < 4 ? k - j0 + 1 : 4;
      float hq[4];
      const float *vq[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        hq[q] = q < nq ? (float)hsh[j0 + q] : 0.0f;
        vq[q] = V + (i64)(q < nq ? j0 + q : j0) * n;
      }
      for (i64 i0 = t; i0 < n; i0 += UG * BS) {
        float wv[UG], vv[4][UG];
#pragma unroll
        for (int u = 0; u < UG; u++) {
          const i64 i = i0 + (i64)u * BS;
          wv[u] = i < n ? w[i] : 0.0f;
#pragma unroll
          for (int q = 0; q < 4; q++) vv[q][u] = i < n ? vq[q][i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < UG; u++) {
          const i64 i = i0 + (i64)u * BS;
          if (i < n) {
            float x = wv[u];
#pragma unroll
            for (int q = 0; q < 4; q++)
              if (q < nq) x = fmaf(-hq[q], vv[q][u], x);
            w[i] = x;
          }
        }
      }
    }
  }
  // y = a x + y over the thread's elements (the axpy-shaped loops of the
  // solvers), U at a time
  __device__ void axpy(float al, const float *__restrict__ x, float *__restrict__ y) const {
    for (i64 i0 = t; i0 < n; i0 += U * BS) {
      float xv[U], yv[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const i64 i = i0 + (i64)u * BS;
        xv[u] = i < n ? x[i] : 0.0f;
        yv[u] = i < n ? y[i] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const i64 i = i0 + (i64)u * BS;
        if (i < n) y[i] = fmaf(al, xv[u], yv[u]);
      }
    }
  }
  // A position's inputs for a sweep, all stored in position order (entry
  // count, slots, start value, inverted diagonal, output index), so each
  // level's are loaded two levels ahead with no load waiting on another.
  struct Row {
    int p, o, cnt;
    float x0, dv;
    float c[BS_W];
    int d[BS_W];
  };
  __device__ Row row_pos(const int *omap, const int *pcnt, const float *coef, const int *dep, const float *xpos,
                         const float *dpos, int p, bool valid) const {
    Row r;
    r.p = valid ? p : -1;
    if (!valid) return r;
    const i64 gp = row0 + p;
    r.o = omap[gp];
    r.cnt = pcnt[gp];
#pragma unroll
    for (int e = 0; e < BS_W; e++) {
      r.c[e] = coef[gp * BS_W + e];
      r.d[e] = dep[gp * BS_W + e];
    }
    r.x0 = xpos[p];
    r.dv = dpos ? dpos[gp] : 1.0f;
    return r;
  }
  // One sweep over levels l0..l0+nl of a single-block plan, each thread's
  // row of a level loaded two levels ahead.  Values by position in vb (the
  // LDS ring when mask = BS_RING - 1, the slots then holding position mod
  // BS_RING; else a global vector by position), and in out at each
  // position's output index.
  __device__ void sweep(const int *lv, i64 l0, i64 nl, const int *omap, const int *pcnt, const float *coef,
                        const int *dep, const float *xpos, const float *dpos, float *vb, int mask, float *out) const {
    if (nl <= 0) return;
    int pa = lv[l0], pb = lv[l0 + 1], pc = nl > 1 ? lv[l0 + 2] : pb;
    Row cur = row_pos(omap, pcnt, coef, dep, xpos, dpos, pa + t, pa + t < pb);
    Row nxt = row_pos(omap, pcnt, coef, dep, xpos, dpos, pb + t, nl > 1 && pb + t < pc);
    for (i64 l = 0; l < nl; l++) {
      const int pd = l + 2 < nl ? lv[l0 + l + 3] : pc;
      const Row nx2 = row_pos(omap, pcnt, coef, dep, xpos, dpos, pc + t, l + 2 < nl && pc + t < pd);
      if (cur.p >= 0) finish(cur, vb, mask, out, dpos != nullptr);
      for (int p = pa + t + BS; p < pb; p += BS)
        finish(row_pos(omap, pcnt, coef, dep, xpos, dpos, p, true), vb, mask, out, dpos != nullptr);
      __syncthreads();
      cur = nxt;
      nxt = nx2;
      pa = pb;
      pb = pc;
      pc = pd;
    }
  }
  __device__ void finish(const Row &r, float *vb, int mask, float *out, bool scale) const {
    float acc = r.x0;
#pragma unroll
    for (int e = 0; e < BS_W; e++)
      if (e < r.cnt) acc = fmaf(r.c[e], vb[r.d[e]], acc);
    const float v = scale ? acc * r.dv : acc;
    vb[r.p & mask] = v;
    out[r.o] = v;
  }
  // y = M^-1 x for ILU(0).  The forward sweep's start values x[row] are
  // first copied into position order (xs); the forward sweep writes F in the
  // backward sweep's position order (f2b), which starts from it and writes y
  // by row, times the inverted diagonal (dinv32, backward position order).
  // ring: the system's plans fit the LDS ring sr, else values go through val
  // (global, by position).
  __device__ void ilu(const float *__restrict__ x, float *y, float *val, float *F, float *xs, i64 lf0, i64 nlf,
                      i64 lb0, i64 nlb, bool ring, float *sr) const {
    for (i64 p0 = t; p0 < n; p0 += U * BS) {
      int iv[U];
#pragma unroll
      for (int u = 0; u < U; u++) iv[u] = p0 + (i64)u * BS < n ? a.fprow[row0 + p0 + (i64)u * BS] : 0;
#pragma unroll
      for (int u = 0; u < U; u++)
        if (p0 + (i64)u * BS < n) xs[p0 + (i64)u * BS] = x[iv[u]];
    }
    __syncthreads();
    float *vb = ring ? sr : val;
    const int mask = ring ? BS_RING - 1 : -1;
    sweep(a.flv, lf0, nlf, a.f2b, a.fcnt, a.fcoef, a.fdep, xs, nullptr, vb, mask, F);
    sweep(a.blv, lb0, nlb, a.bprow, a.bcnt, a.bcoef, a.bdep, F, a.dinv32, vb, mask, y);
  }
};

template <int BS, int UR, int KB, int UG>
__global__ void __launch_bounds__(BS) k_batch(BatchArgs a) {
  __shared__ double sh[2 * (BS / 64) + 1];  // wg_sum's two buffers, then a broadcast slot
  __shared__ double H[(BS_M + 1) * BS_M], cs[BS_M], sn[BS_M], g[BS_M + 1], yv[BS_M];
  __shared__ double hcol[BS_M + 1], mbuf[(BS_M + 1) * (BS / 64)];
  __shared__ int sys_sh;
  __shared__ float sring[BS_RING];
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
    Sys<BS, UR, KB, UG> S{a, d[0], d[1], sh, t};
    const i64 n = S.n, row0 = S.row0;
    const int kind = (int)d[2];
    // local FP32 vectors
    float *rhs = W, *dv = W + n, *r = W + 2 * n, *z = W + 3 * n, *p = W + 4 * n, *w = W + 5 * n,
          *F = W + 6 * n, *val = W + 7 * n, *V = W + 8 * n;
    for (i64 i = t; i < n; i += BS) a.x[row0 + i] = 0.0;
    __syncthreads();
    double bn = 0.0;
    for (i64 i = t; i < n; i += BS) bn = fma(a.b[row0 + i], a.b[row0 + i], bn);
    bn = sqrt(S.wsum(bn));
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
        // (element-wise loops touch only the thread's own elements; a barrier
        // is needed only before a product reads other threads' entries, and
        // the reductions' own barrier provides it)
        for (i64 i = t; i < n; i += BS) {
          r[i] = rhs[i];
          z[i] = a.jd32[row0 + i] * rhs[i];
          p[i] = z[i];
        }
        double rz = S.dot(r, z);
        for (int it = 0; it < a.max_inner; it++) {
          S.spmv32(p, w);
          const double pq = S.dot(p, w);
          if (!(pq != 0.0)) break;
          const float al = (float)(rz / pq);
          for (i64 i0 = t; i0 < n; i0 += S.U * BS) {
            float pv[S.U], dvv[S.U], wv[S.U], rv[S.U], jv[S.U];
#pragma unroll
            for (int u = 0; u < S.U; u++) {
              const i64 i = i0 + (i64)u * BS;
              const bool in = i < n;
              pv[u] = in ? p[i] : 0.0f;
              dvv[u] = in ? dv[i] : 0.0f;
              wv[u] = in ? w[i] : 0.0f;
              rv[u] = in ? r[i] : 0.0f;
              jv[u] = in ? a.jd32[row0 + i] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < S.U; u++) {
              const i64 i = i0 + (i64)u * BS;
              if (i < n) {
                dv[i] = fmaf(al, pv[u], dvv[u]);
                const float rn = fmaf(-al, wv[u], rv[u]);
                r[i] = rn;
                z[i] = jv[u] * rn;
              }
            }
          }
          inner++;
          const double rr = S.dot(r, r);
          if (sqrt(rr) <= a.inner_rtol) break;
          const double rz2 = S.dot(r, z);
          const float be = (float)(rz2 / rz);
          rz = rz2;
          for (i64 i0 = t; i0 < n; i0 += S.U * BS) {
            float pv[S.U], zv[S.U];
#pragma unroll
            for (int u = 0; u < S.U; u++) {
              const i64 i = i0 + (i64)u * BS;
              pv[u] = i < n ? p[i] : 0.0f;
              zv[u] = i < n ? z[i] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < S.U; u++)
              if (i0 + (i64)u * BS < n) p[i0 + (i64)u * BS] = fmaf(be, pv[u], zv[u]);
          }
          __syncthreads();
        }
      } else {
        // GMRES(m) + ILU(0), left preconditioned, classical Gram-Schmidt, zero guess
        double target = -1.0;
        int its = 0;
        for (;;) {
          // z = rhs - A dv (dv = 0 on the first cycle)
          S.spmv32(dv, w);
          for (i64 i = t; i < n; i += BS) z[i] = rhs[i] - w[i];
          __syncthreads();
          S.ilu(z, w, val, F, r, d[3], d[4], d[5], d[6], d[7] != 0, sring);
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
            S.ilu(z, w, val, F, r, d[3], d[4], d[5], d[6], d[7] != 0, sring);
            // classical Gram-Schmidt (PETSc's default orthogonalisation)
            S.mdot(V, w, k, hcol, mbuf);
            if (t == 0)
              for (int j = 0; j <= k; j++) H[k * (BS_M + 1) + j] = hcol[j];
            S.maxpy(V, hcol, k, w);
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
              sh[2 * (BS / 64)] = fabs(g[k + 1]);
            }
            __syncthreads();
            const double res = sh[2 * (BS / 64)];
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
          for (int j = 0; j < kk; j++) S.axpy((float)yv[j], V + (i64)j * n, dv);
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

// The compiled launch variants (workgroup, UR, KB, UG), in the order
// mpk_batch_variants lists them: every tile choice at 512 threads (the
// fastest workgroup size on the batched workload), the default tiles
// (4, 4, 4) at 128 and 256.
#define MPK_BATCH_VARIANTS(X)                                                                                   \
  X(512, 4, 4, 4) X(512, 2, 4, 4) X(512, 8, 4, 4) X(512, 2, 8, 4) X(512, 4, 8, 4) X(512, 8, 8, 4)            \
  X(512, 2, 4, 2) X(512, 4, 4, 2) X(512, 8, 4, 2) X(512, 2, 8, 2) X(512, 4, 8, 2) X(512, 8, 8, 2)            \
  X(512, 2, 4, 8) X(512, 4, 4, 8) X(512, 8, 4, 8) X(512, 2, 8, 8) X(512, 4, 8, 8) X(512, 8, 8, 8)            \
  X(128, 4, 4, 4) X(256, 4, 4, 4)
struct BatchVariant {
  int bs, ur, kb, ug;
};
static const BatchVariant kBatchVariants[] = {
#define MPK_BV_ROW(bs, ur, kb, ug) {bs, ur, kb, ug},
    MPK_BATCH_VARIANTS(MPK_BV_ROW)
#undef MPK_BV_ROW
};
static const int kNumBatchVariants = (int)(sizeof kBatchVariants / sizeof kBatchVariants[0]);
// The variant with the default tiles at workgroup size bs, or -1.
static int batch_default_variant(int bs) {
  for (int v = 0; v < kNumBatchVariants; v++)
    if (kBatchVariants[v].bs == bs && kBatchVariants[v].ur == 4 && kBatchVariants[v].kb == 4 &&
        kBatchVariants[v].ug == 4)
      return v;
  return -1;
}

struct Batch {
  BatchArgs a{};
  i64 nmax = 0;
  int grid = 0;  // the workgroups the work buffers are sized for
  int cap_variant = -1, cap_grid = 0;  // the launch the graph was captured with
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
  void launch(int variant, int g) {
    MPK_LAUNCH(k_batch_reset, 1, 1, queue);
    int v = 0;
#define MPK_BV_LAUNCH(bs, ur, kb, ug)                            \
  if (variant == v++) {                                          \
    MPK_LAUNCH((k_batch<bs, ur, kb, ug>), g, bs, a);             \
    return;                                                      \
  }
    MPK_BATCH_VARIANTS(MPK_BV_LAUNCH)
#undef MPK_BV_LAUNCH
  }
};

}  // namespace mpk

using namespace mpk;

extern "C" {

// See include/mpk.h.
int64_t mpk_batch_create(int64_t nsys, int64_t sys, int64_t order, int64_t rp, int64_t ci, int64_t v64,
                         int64_t v32, int64_t jd32, int64_t fprow, int64_t fcnt, int64_t fdep, int64_t fcoef,
                         int64_t flv, int64_t bprow, int64_t bcnt, int64_t bdep, int64_t bcoef, int64_t blv,
                         int64_t dinv32, int64_t f2b, int64_t nmax) {
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
  a.f2b = (const int *)f2b;
  HIPCHK(hipMalloc(&B->queue, 64 * sizeof(unsigned)));
  a.queue = B->queue;
  if (failed()) {
    delete B;
    return 0;
  }
  return (int64_t)B;
}

// See include/mpk.h.
int64_t mpk_batch_variants(int64_t out, int64_t cap) {
  int64_t *o = (int64_t *)out;
  for (int v = 0; v < kNumBatchVariants && v < cap; v++) {
    o[4 * v] = kBatchVariants[v].bs;
    o[4 * v + 1] = kBatchVariants[v].ur;
    o[4 * v + 2] = kBatchVariants[v].kb;
    o[4 * v + 3] = kBatchVariants[v].ug;
  }
  return kNumBatchVariants;
}

// One batched solve: x [D] (concatenated) from b [D]; stats [D] float64
// [nsys][4] = {relres, outer, inner, converged}.  cfg [H] float64 in {rtol,
// inner_rtol, max_outer, max_inner, workgroup size (128/256/512), grid
// (workgroups), use_graph, variant}: variant an index into
// mpk_batch_variants' list (its workgroup size must be cfg[4]), or -1 for
// the default tiles at that size.  With use_graph the queue reset and the
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
  const int variant = cfg[7] < 0.0 ? batch_default_variant(bs) : (int)cfg[7];
  if (variant < 0 || variant >= kNumBatchVariants || kBatchVariants[variant].bs != bs) {
    set_error(7, "batch: variant %d is not a compiled variant at workgroup size %d", variant, bs);
    return -1;
  }
  if (g < 1 || g > 65536) {
    set_error(7, "batch: grid must be 1..65536");
    return -1;
  }
  if (!B->size_work(g)) return -1;
  if (graph) {
    // re-capture when the launch shape or any argument changes
    if (B->exec && (B->cap_variant != variant || B->cap_grid != g || std::memcmp(&B->captured, &a, sizeof a) != 0)) {
      HIPCHK(hipGraphExecDestroy(B->exec));
      HIPCHK(hipGraphDestroy(B->graph));
      B->exec = nullptr;
      B->graph = nullptr;
    }
    if (!B->exec) {
      HIPCHK(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
      B->launch(variant, g);
      HIPCHK(hipStreamEndCapture(stream(), &B->graph));
      if (failed()) return -1;
      HIPCHK(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
      B->cap_variant = variant;
      B->cap_grid = g;
      B->captured = a;
    }
    HIPCHK(hipGraphLaunch(B->exec, stream()));
  } else {
    B->launch(variant, g);
  }
  sync();
  return failed() ? -1 : 0;
}

void mpk_batch_free(int64_t h) { delete (Batch *)h; }

}  // extern "C"
