// mpk_ksp.hip -- device-resident Krylov drivers: CG, BiCGStab, GMRES, FGMRES.
//
// Each driver runs the reference's iteration (src/ksp.mojo:271-903) as a
// stream of kernels.  The solver's scalars live in a device-side KspState; the
// scalar steps between vector operations (alpha and beta, the convergence
// test, breakdown checks, the Givens updates and the back substitution) run in
// the one-block finalize of the reduction that feeds them.  When a step
// decides the solve is over it sets `stop`, and every kernel already queued
// behind it returns at once.  The host queues iterations in chunks and reads
// one flag per chunk (CG, BiCGStab) or per restart cycle (GMRES, FGMRES).
//
// Scalar arithmetic follows the reference expression by expression, with
// fma() where it contracts, so given the same reductions the iterates are
// bitwise those of the CPU path.
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "mpk_internal.h"
#include "mpk_ops.h"
#include "mpk_pc.h"
#include "mpk_progs.h"
#include "mpk_reduce.h"
#include "../../include/mpk.h"

namespace mpk {

enum {
  KSP_CONVERGED_RTOL = 2,
  KSP_CONVERGED_ATOL = 3,
  KSP_CONVERGED_ITS = 4,
  KSP_CONVERGED_HAPPY_BREAKDOWN = 7,
  KSP_DIVERGED_NULL = -2,
  KSP_DIVERGED_ITS = -3,
  KSP_DIVERGED_DTOL = -4,
  KSP_DIVERGED_BREAKDOWN = -5,
  KSP_DIVERGED_INDEFINITE_PC = -8,
  KSP_DIVERGED_NANORINF = -9,
  KSP_DIVERGED_INDEFINITE_MAT = -10,
};
enum { NORM_NONE = 1, NORM_PRECONDITIONED = 2, NORM_UNPRECONDITIONED = 3, NORM_NATURAL = 4 };
enum { PC_LEFT = 0, PC_RIGHT = 1, PC_SYMMETRIC = 2 };

struct KspState {
  // configuration (cfg / cfgi, ksp.mojo:9-42)
  double rtol, atol, divtol, haptol, breakdowntol, rhs_norm;
  i64 normtype, side, guess_zero, min_it, chknorm, convmaxits, max_it, m, orthog, cgstype, seed_rhs;
  // KSPConvergedDefault bookkeeping
  double rnorm0, ttol;
  i64 seeded;
  // results
  i64 its, reason;
  double rnorm;
  i64 stop;        // the solve is over
  i64 inner_stop;  // GMRES: the Arnoldi loop is over (or the solve is)
  i64 skip_ref;    // GMRES CGS: no refinement pass this step
  i64 skip_build;  // GMRES: BuildSoln has nothing (more) to do
  i64 special;     // BiCGStab: the d2 == 0 exit needs x += alpha p
  i64 stopped_at_init;
  // CG
  i64 i;
  double beta, betaold, dpi, dpiold, a, bcoef, dp;
  // BiCGStab
  double rho, rhoold, alpha, omega, omegaold, d1, d2, cz;
  // GMRES
  i64 it, itcount, hapend, breason, nbuild;
  double res, prev_rnorm, rnorm_cycle, tt, hnrm;
  // device arrays
  double *hist;  // max_it + 2
  double *HH, *HES, *CC, *SS, *RS, *LHH;
};

// ------------------------------------------------------------ scalar steps

// The fields of the state the convergence test and CG's norm step read,
// loaded at once so their latency overlaps a kernel's other loads (none of
// them changes once rnorm0 is seeded).
struct ConvView {
  i64 convmaxits, max_it, normtype, chknorm, min_it;
  double ttol, atol, divtol, rnorm0;
  __device__ static ConvView load(const KspState *s) {
    return ConvView{s->convmaxits, s->max_it, s->normtype, s->chknorm, s->min_it,
                    s->ttol,       s->atol,   s->divtol,   s->rnorm0};
  }
};

// KSPConvergedDefault (ksp.mojo:96-152) once rnorm0 is seeded: reads the
// state (a KspState or a ConvView) only.
template <class S>
__device__ i64 conv_check(const S *s, i64 n, double rnorm) {
  if (s->convmaxits && n >= s->max_it) return KSP_CONVERGED_ITS;
  if (s->normtype == NORM_NONE) return 0;
  if (n <= s->chknorm) return 0;
  // MAX_REAL, not infinity, as the reference compares (ksp.mojo:85,138)
  if (rnorm != rnorm || fabs(rnorm) == 1.7976931348623157e308) return KSP_DIVERGED_NANORINF;
  if (n < s->min_it) return 0;
  if (rnorm <= s->ttol) {
    if (rnorm < s->atol) return KSP_CONVERGED_ATOL;
    return KSP_CONVERGED_RTOL;
  }
  if (rnorm >= s->divtol * s->rnorm0) return KSP_DIVERGED_DTOL;
  return 0;
}

// KSPConvergedDefault (ksp.mojo:96-152): seeds rnorm0 and ttol at n = 0.
__device__ i64 conv_default(KspState *s, i64 n, double rnorm) {
  if (s->convmaxits && n >= s->max_it) return KSP_CONVERGED_ITS;
  if (s->normtype == NORM_NONE) return 0;
  if (!n && !s->seeded) {
    double rnorm0 = rnorm;
    if (s->seed_rhs) {
      double snorm = s->rhs_norm;
      if (!snorm) snorm = rnorm;
      rnorm0 = snorm;
    }
    s->rnorm0 = rnorm0;
    double ttol = s->rtol * rnorm0;
    s->ttol = ttol > s->atol ? ttol : s->atol;
    s->seeded = 1;
  }
  return conv_check(s, n, rnorm);
}

__device__ void finish(KspState *s, double rnorm, i64 its, i64 reason) {
  s->rnorm = rnorm;
  s->its = its;
  s->reason = reason;
  s->stop = 1;
  s->inner_stop = 1;
  s->skip_ref = 1;
  s->skip_build = 1;
}

__device__ i64 sign_real(double x) { return x < 0.0 ? -1 : (x == 0.0 ? 0 : 1); }

// CG: the head of iteration i (ksp.mojo:337-350)
__device__ void cg_prep(KspState *s) {
  i64 i = s->i;
  i64 its = i + 1;
  if (s->beta == 0.0) {
    finish(s, s->dp, its, KSP_CONVERGED_ATOL);
    return;
  } else if (i > 0 && s->beta * s->betaold < 0.0) {
    finish(s, s->dp, its, KSP_DIVERGED_INDEFINITE_PC);
    return;
  }
  s->bcoef = i == 0 ? 0.0 : s->beta / s->betaold;
  s->dpiold = s->dpi;
}

enum {
  EPI_CG_INIT,
  EPI_CG_DPI,
  EPI_CG_NORM,
  EPI_BCGS_INIT,
  EPI_BCGS_RHO,
  EPI_BCGS_D1,
  EPI_BCGS_D2,
  EPI_BCGS_DP,
  EPI_GM_START,
  EPI_GM_CGS_A,
  EPI_GM_CGS_NORM,
  EPI_GM_CGS_B,
  EPI_GM_MGS,
  EPI_GM_STEP,
  EPI_FGM_START,
  EPI_FGM_STEP,
};

// Givens update of the new Hessenberg column (ksp.mojo:589-629); returns
// the reason and writes the residual estimate.
__device__ i64 update_hessenberg(KspState *s, i64 it, i64 hapend, double *res) {
  const i64 m = s->m;
  double *hh = s->HH, *cc = s->CC, *ss = s->SS, *rs = s->RS;
  if (it > 0) {
    // rotation j rewrites entries j-1 and j of the column; entry j, its
    // second result, is carried to rotation j+1 in a register, so that no
    // load of the chain waits on the store before it (the arrays are
    // distinct, which lets the next rotations' loads issue early)
    double *__restrict__ col = hh + it * (m + 2);
    const double *__restrict__ c = cc;
    const double *__restrict__ sn = ss;
    double carry = col[0];
#pragma unroll 4
    for (i64 j = 1; j <= it; j++) {
      const double tt = carry, nx = col[j];
      col[j - 1] = fma(c[j - 1], tt, sn[j - 1] * nx);
      carry = fma(c[j - 1], nx, -(sn[j - 1] * tt));
    }
    col[it] = carry;
  }
  if (!hapend) {
    i64 base = it * (m + 2) + it;
    double a = hh[base], b = hh[base + 1];
    double tt = sqrt(fma(a, a, b * b));
    if (tt == 0.0) {
      *res = 0.0;
      return KSP_DIVERGED_NULL;
    }
    cc[it] = a / tt;
    ss[it] = b / tt;
    rs[it + 1] = -ss[it] * rs[it];
    rs[it] = cc[it] * rs[it];
    hh[base] = fma(cc[it], a, ss[it] * b);
    *res = fabs(rs[it + 1]);
    return 0;
  }
  *res = 0.0;
  return 0;
}

// CG's residual norm after iteration i from {z.z, r.r, z.r} (ksp.mojo:364-372)
template <class S>
__device__ double cg_norm(const S *s, i64 i, const double *r) {
  if (s->normtype == NORM_PRECONDITIONED && s->chknorm < i + 2) return sqrt(r[0]);
  if (s->normtype == NORM_UNPRECONDITIONED && s->chknorm < i + 2) return sqrt(r[1]);
  if (s->normtype == NORM_NATURAL) return sqrt(fabs(r[2]));
  return 0.0;
}

__device__ void ksp_epilogue(KspState *s, int which, int arg, const double *r) {
  const i64 m = s->m;
  switch (which) {
    // ---------------------------------------------------------------- CG
    case EPI_CG_INIT: {  // r = {z.z, r.r, z.r} (ksp.mojo:313-335)
      double dp;
      if (s->normtype == NORM_PRECONDITIONED)
        dp = sqrt(r[0]);
      else if (s->normtype == NORM_UNPRECONDITIONED)
        dp = sqrt(r[1]);
      else if (s->normtype == NORM_NATURAL)
        dp = sqrt(fabs(r[2]));
      else
        dp = 0.0;
      s->dp = dp;
      s->hist[0] = dp;
      i64 reason = conv_default(s, 0, dp);
      if (reason) {
        finish(s, dp, 0, reason);
        return;
      }
      s->beta = r[2];
      s->i = 0;
      s->betaold = 1.0;
      s->dpi = 0.0;
      cg_prep(s);
      return;
    }
    case EPI_CG_DPI: {  // r = {p.w} (ksp.mojo:350-362)
      s->dpi = r[0];
      s->betaold = s->beta;
      if (s->dpi == 0.0 || (s->i > 0 && sign_real(s->dpi) * sign_real(s->dpiold) < 0)) {
        finish(s, s->dp, s->i + 1, KSP_DIVERGED_INDEFINITE_MAT);
        return;
      }
      s->a = s->beta / s->dpi;
      return;
    }
    case EPI_CG_NORM: {  // r = {z.z, r.r, z.r} (ksp.mojo:364-386)
      i64 i = s->i, its = i + 1;
      const double dp = cg_norm(s, i, r);
      s->dp = dp;
      s->hist[its] = dp;
      i64 reason = conv_default(s, its, dp);
      if (reason) {
        finish(s, dp, its, reason);
        return;
      }
      s->beta = r[2];
      s->i = i + 1;
      if (s->i >= s->max_it) {
        finish(s, dp, its, KSP_DIVERGED_ITS);
        return;
      }
      cg_prep(s);
      return;
    }
    // ---------------------------------------------------------------- BiCGStab
    case EPI_BCGS_INIT: {  // r = {r.r} (ksp.mojo:426-436)
      double dp = s->normtype != NORM_NONE ? sqrt(r[0]) : 0.0;
      s->dp = dp;
      s->hist[0] = dp;
      i64 reason = conv_default(s, 0, dp);
      if (reason) {
        finish(s, dp, 0, reason);
        s->stopped_at_init = 1;
        return;
      }
      s->rhoold = 1.0;
      s->alpha = 1.0;
      s->omegaold = 1.0;
      s->omega = 1.0;
      s->rho = 1.0;
      s->i = 0;
      s->its = 0;
      return;
    }
    case EPI_BCGS_RHO: {  // r = {r.rp} (ksp.mojo:448-451)
      s->rho = r[0];
      s->beta = (s->rho / s->rhoold) * (s->alpha / s->omegaold);
      s->cz = -s->omegaold * s->beta;
      return;
    }
    case EPI_BCGS_D1: {  // r = {v.rp}
      s->d1 = r[0];
      if (s->d1 == 0.0) {
        finish(s, s->dp, s->its, KSP_DIVERGED_BREAKDOWN);
        return;
      }
      s->alpha = s->rho / s->d1;
      return;
    }
    case EPI_BCGS_D2: {  // r = {s.t, t.t, s.s} (ksp.mojo:466-483)
      s->d1 = r[0];
      s->d2 = r[1];
      if (s->d2 == 0.0) {
        s->d1 = r[2];
        if (s->d1 != 0.0) {
          finish(s, s->dp, s->its, KSP_DIVERGED_BREAKDOWN);
          return;
        }
        s->special = 1;  // x += alpha p, then converged
        finish(s, 0.0, s->its + 1, KSP_CONVERGED_RTOL);
        return;
      }
      s->omega = s->d1 / s->d2;
      return;
    }
    case EPI_BCGS_DP: {  // r = {r.r} (ksp.mojo:486-504)
      i64 i = s->i;
      if (s->normtype != NORM_NONE && s->chknorm < i + 2) s->dp = sqrt(r[0]);
      s->rhoold = s->rho;
      s->omegaold = s->omega;
      s->its += 1;
      s->hist[s->its] = s->dp;
      i64 reason = conv_default(s, s->its, s->dp);
      if (reason) {
        finish(s, s->dp, s->its, reason);
        return;
      }
      if (s->rho == 0.0) {
        finish(s, s->dp, s->its, KSP_DIVERGED_BREAKDOWN);
        return;
      }
      s->i = i + 1;
      if (s->i >= s->max_it) finish(s, s->dp, s->its, KSP_DIVERGED_ITS);
      return;
    }
    // ---------------------------------------------------------------- GMRES
    case EPI_GM_START: {  // r = {v0.v0}: the cycle head (ksp.mojo:700-720)
      double res = sqrt(r[0]);
      s->res = res;
      if (s->prev_rnorm > 0.0 && fabs(res - s->prev_rnorm) > s->breakdowntol * s->rnorm_cycle) {
        finish(s, res, s->its, KSP_DIVERGED_BREAKDOWN);
        return;
      }
      s->RS[0] = res;
      s->rnorm_cycle = res;
      s->hist[s->its] = res;
      s->reason = !res ? KSP_CONVERGED_ATOL : conv_default(s, s->its, res);
      s->it = 0;
      s->hapend = 0;
      s->inner_stop = s->reason != 0 || 0 >= m || s->its >= s->max_it;
      return;
    }
    case EPI_FGM_START: {  // r = {v0.v0} (ksp.mojo:809-817)
      double res = sqrt(r[0]);
      s->res = res;
      s->RS[0] = res;
      s->hist[s->its] = res;
      s->reason = conv_default(s, s->its, res);
      s->it = 0;
      s->hapend = 0;
      s->inner_stop = s->reason != 0 || 0 >= m || s->its >= s->max_it;
      return;
    }
    case EPI_GM_CGS_A: {  // r[j] = v_new . v_j, j <= it (ksp.mojo:533-566)
      i64 it = s->it;
      for (i64 j = 0; j <= it; j++) {
        s->HH[it * (m + 2) + j] = 0.0;
        s->HES[it * (m + 1) + j] = 0.0;
      }
      for (i64 j = 0; j <= it; j++) s->LHH[j] = -r[j];
      for (i64 j = 0; j <= it; j++) {
        s->HH[it * (m + 2) + j] = s->HH[it * (m + 2) + j] - s->LHH[j];
        s->HES[it * (m + 1) + j] = s->HES[it * (m + 1) + j] - s->LHH[j];
      }
      if (s->cgstype == 1) {
        double hnrm = 0.0;
        for (i64 j = 0; j <= it; j++) hnrm = fma(s->LHH[j], s->LHH[j], hnrm);
        s->hnrm = sqrt(hnrm);
        s->skip_ref = 1;  // decided after the norm of the projected vector
      } else {
        s->skip_ref = s->cgstype == 2 ? 0 : 1;
      }
      return;
    }
    case EPI_GM_CGS_NORM: {  // r = {v_new . v_new}: refine if ||v_new|| < hnrm
      s->skip_ref = (sqrt(r[0]) < s->hnrm) ? 0 : 1;
      return;
    }
    case EPI_GM_CGS_B: {  // refinement accumulates (ksp.mojo:575-586)
      i64 it = s->it;
      for (i64 j = 0; j <= it; j++) s->LHH[j] = -r[j];
      for (i64 j = 0; j <= it; j++) {
        s->HH[it * (m + 2) + j] = s->HH[it * (m + 2) + j] - s->LHH[j];
        s->HES[it * (m + 1) + j] = s->HES[it * (m + 1) + j] - s->LHH[j];
      }
      return;
    }
    case EPI_GM_MGS: {  // r = {v_new . v_k}, k = arg (ksp.mojo:511-530)
      i64 it = s->it, k = arg;
      double d = r[0];
      s->HH[it * (m + 2) + k] = d;
      s->HES[it * (m + 1) + k] = d;
      s->LHH[k] = -d;
      return;
    }
    case EPI_GM_STEP:
    case EPI_FGM_STEP: {  // r = {v_new . v_new} (ksp.mojo:728-755, 835-860)
      i64 it = s->it;
      double tt = sqrt(r[0]);
      s->tt = tt;
      s->HH[it * (m + 2) + it + 1] = tt;
      s->HES[it * (m + 1) + it + 1] = tt;
      double hapbnd = fabs(tt / s->RS[it]);
      hapbnd = hapbnd > s->haptol ? s->haptol : hapbnd;
      i64 hapend;
      if (which == EPI_GM_STEP)
        hapend = tt < hapbnd ? 1 : 0;
      else
        hapend = tt > hapbnd ? 0 : 1;
      s->hapend = hapend;
      double res;
      i64 reason = update_hessenberg(s, it, hapend, &res);
      if (which == EPI_GM_STEP) {
        s->res = res;
        s->it = it + 1;
        s->its += 1;
        s->hist[s->its] = res;
        if (reason) {
          s->reason = reason;
        } else {
          reason = conv_default(s, s->its, res);
          s->reason = reason;
          if (hapend) {
            if (s->normtype == NORM_NONE)
              s->reason = KSP_CONVERGED_HAPPY_BREAKDOWN;
            else if (reason == 0)
              s->reason = KSP_DIVERGED_BREAKDOWN;
          }
        }
      } else {
        s->res = res;
        if (reason) {
          s->reason = reason;  // FGMRES breaks before counting the step
        } else {
          s->it = it + 1;
          s->its += 1;
          s->hist[s->its] = res;
          reason = conv_default(s, s->its, res);
          s->reason = reason;
          if (hapend && reason == 0) s->reason = KSP_DIVERGED_BREAKDOWN;
        }
      }
      s->inner_stop = s->reason != 0 || s->it >= m || s->its >= s->max_it;
      return;
    }
  }
}

struct KspEpi {
  KspState *s;
  int which, arg;
  __device__ void operator()(const double *r) const { ksp_epilogue(s, which, arg, r); }
};

// ------------------------------------------------------------ scalar kernels

// p = z + bcoef p with VecAYPX's special cases (ported.mojo:285-294); p = z at i = 0
struct OpCgP {
  double *p;
  const double *z;
  const KspState *s;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    D2 zv = ld2(z, e, n, al);
    if (s->i == 0) {
      st2(p, e, n, al, zv);
      return;
    }
    double b = s->bcoef;
    if (b == 0.0) {
      st2(p, e, n, al, zv);
      return;
    }
    D2 pv = ld2(p, e, n, al);
    if (b == 1.0)
      st2(p, e, n, al, D2{fma(b, zv.a, pv.a), fma(b, zv.b, pv.b)});
    else
      st2(p, e, n, al, D2{fma(b, pv.a, zv.a), fma(b, pv.b, zv.b)});
  }
};
// CG, w = A p (MatMult_SeqAIJ, ported.mojo:380-391, over the sliced copy) and
// the tile partial of p.w (ksp.mojo:338) in one pass.  Block b owns tile b
// (2048 rows), 4 * MPK_NT threads: thread T computes rows T and T + 1024 one
// lane per row, as k_spmv_sell does (each w row is its serial sum).  The even
// lane of each row pair takes its neighbour's p and w by a shuffle and forms
// the pair's leaf, which goes to LDS where the tile reduction keeps it (pair
// 512 j + 2t is thread t's j-th); the leaves are then folded in that
// reduction's order, so the partial is k_partials'.
template <bool NT, class Cols>
__device__ __forceinline__ void wdot_rows(const double *__restrict__ sv, const Cols &cols,
                                          const double *__restrict__ p, int q0, int len0, int e0, int q1, int len1,
                                          int e1, double &w0, double &w1) {
  const int len = len0 > len1 ? len0 : len1;
  for (int k = 0; k < len; k += 4) {
    int c0[4], c1[4];
    double v0[4], v1[4], x0[4], x1[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (k + j < len0) {
        c0[j] = cols.template col<NT>(q0 + 64 * (k + j), e0);
        v0[j] = ld_stream<NT>(sv + q0 + 64 * (k + j));
      }
      if (k + j < len1) {
        c1[j] = cols.template col<NT>(q1 + 64 * (k + j), e1);
        v1[j] = ld_stream<NT>(sv + q1 + 64 * (k + j));
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (k + j < len0) x0[j] = p[c0[j]];
      if (k + j < len1) x1[j] = p[c1[j]];
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (k + j < len0) w0 = fma(v0[j], x0[j], w0);
      if (k + j < len1) w1 = fma(v1[j], x1[j], w1);
    }
  }
}

// At most 64 VGPRs, so that two 1024-thread blocks fit a CU (8 waves a SIMD).
template <bool NT>
__global__ void __launch_bounds__(4 * MPK_NT) __attribute__((amdgpu_waves_per_eu(8))) k_cg_wdot(i64 n, const int *__restrict__ rp,
                                                        const int *__restrict__ soff, const double *__restrict__ sv,
                                                        const int *__restrict__ sc, const double *__restrict__ p,
                                                        double *__restrict__ w, double *P, const int64_t *stop) {
  __shared__ double lf[4 * MPK_NT];
  __shared__ double sh[4];
  // the stop flag and the row bounds in one round of loads (indices clamped
  // to row n - 1, n >= 1, so that no load waits on a branch), before the
  // entries are read
  const bool stopped = stop && *stop != 0;  // the solve's own flag (null: never stopped)
  const unsigned tile = xcd_block(blockIdx.x, gridDim.x);
  const int T = (int)threadIdx.x;
  const i64 e0 = (i64)tile * MPK_TILE + T, e1 = e0 + MPK_TILE / 2;
  const i64 c0 = e0 < n ? e0 : n - 1, c1 = e1 < n ? e1 : n - 1;
  const int r00 = rp[c0], r01 = rp[c0 + 1], r10 = rp[c1], r11 = rp[c1 + 1];
  const int o0 = soff[c0 >> 6], o1 = soff[c1 >> 6];
  const short *sd = sell_offsets16(soff, sc, n);
  const int len0 = e0 < n ? r01 - r00 : 0, len1 = e1 < n ? r11 - r10 : 0;
  const int q0 = e0 < n ? o0 + (int)(e0 & 63) : 0, q1 = e1 < n ? o1 + (int)(e1 & 63) : 0;
  if (stopped) return;
  // both rows' sums in order (sell_row's), four entries of each in flight,
  // over the 16-bit column offsets when the copy has them
  double w0 = 0.0, w1 = 0.0;
  if (sd)
    wdot_rows<NT>(sv, SellCols16{sd}, p, q0, len0, (int)e0, q1, len1, (int)e1, w0, w1);
  else
    wdot_rows<NT>(sv, SellCols32{sc}, p, q0, len0, (int)e0, q1, len1, (int)e1, w0, w1);
  double p0 = 0.0, p1 = 0.0;
  if (e0 < n) {
    st_stream<NT>(w + e0, w0);
    p0 = p[e0];
  }
  if (e1 < n) {
    st_stream<NT>(w + e1, w1);
    p1 = p[e1];
  }
  // pair (r, r + 1), r = T even, in each half: r = 512 j + 2t with j = 2 h + (T >= 512)
  const double pa = __shfl_down(p0, 1, 64), wa = __shfl_down(w0, 1, 64);
  const double pb = __shfl_down(p1, 1, 64), wb = __shfl_down(w1, 1, 64);
  if ((T & 1) == 0) {
    const int t = (T & 511) >> 1, j = T >> 9;
    lf[j * MPK_NT + t] = fma(p0, w0, pa * wa);
    lf[(j + 2) * MPK_NT + t] = fma(p1, w1, pb * wb);
  }
  __syncthreads();
  // thread t's sum over its four pairs, then the tile tree (block_tree's)
  double u = 0.0;
  if (T < MPK_NT) u = wave_tree((lf[T] + lf[T + MPK_NT]) + (lf[T + 2 * MPK_NT] + lf[T + 3 * MPK_NT]));
  if (T < MPK_NT && (T & 63) == 0) sh[T >> 6] = u;
  __syncthreads();
  if (T == 0) P[tile] = (sh[0] + sh[2]) + (sh[1] + sh[3]);
}

// x += a p; r += (-a) w  (two VecAXPY sweeps fused, ksp.mojo:363-364)
struct OpCgXR {
  double *x, *r;
  const double *p, *w;
  const KspState *s;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    double a = s->a;
    D2 xv = ld2(x, e, n, al), pv = ld2(p, e, n, al), rv = ld2(r, e, n, al), wv = ld2(w, e, n, al);
    st2(x, e, n, al, D2{fma(a, pv.a, xv.a), fma(a, pv.b, xv.b)});
    st2(r, e, n, al, D2{fma(-a, wv.a, rv.a), fma(-a, wv.b, rv.b)});
  }
};
// CG with Jacobi, one pass for the second half of an iteration.  With Pw,
// every block first folds the T tile partials of p.w (fold_partials_k, k_finalize's order)
// and takes the p.w step from them (ksp.mojo:350-362): the breakdown test
// and a = beta / p.w, reading only state that step leaves alone; the block
// of tile 0 also runs the step on the state.  Unless the solve ends there:
// x = x + a p, r = r - a w (OpCgXR), z = r .* diag (PCApply_Jacobi), and the
// tile partials of z.z, r.r, z.r (folded by k_cg_pz).  Each thread holds the
// elements the tile reduction gives it, so every value and every partial is
// the one the separate kernels produce.  w may alias z (the reference's CG
// keeps W = Z).
template <bool al>  // all six vectors 16-byte aligned (a constant, so every pair is one 16-byte access)
__global__ void __launch_bounds__(MPK_NT) k_cg_xrz(i64 n, double *x, double *r, double *z, const double *p,
                                                   const double *w, const double *dg, const double *Pw, int T,
                                                   KspState *s, double *P, const int64_t *stop) {
  __shared__ double sh[4];
  __shared__ double res[1];
  __shared__ double sa;
  __shared__ int brk;
  // every load is issued before any is waited on: the stop flag, the state
  // the p.w step reads, the tile's operands, then the partials; a stopped
  // solve returns before anything is written
  const bool stopped = stop && *stop != 0;  // the solve's own flag (null: never stopped)
  const i64 si = s->i;
  const double dpiold = s->dpiold, sbeta = s->beta, s_a = s->a;
  const i64 base = (i64)blockIdx.x * MPK_TILE + 2 * (i64)threadIdx.x;
  D2 xv_[4], pv_[4], rv_[4], wv_[4], dv_[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const i64 e = base + 512 * j;
    if (e < n) {
      xv_[j] = ld2(x, e, n, al);
      pv_[j] = ld2(p, e, n, al);
      rv_[j] = ld2(r, e, n, al);
      wv_[j] = ld2(w, e, n, al);
      dv_[j] = ld2(dg, e, n, al);
    }
  }
  double a = s_a;  // without Pw the p.w step ran in a finalize (the stop flag covers its breakdown)
  if (Pw) {
    fold_partials_k<1>(Pw, T, res);
    if (threadIdx.x == 0) {
      const double dpi = res[0];
      brk = stopped || dpi == 0.0 || (si > 0 && sign_real(dpi) * sign_real(dpiold) < 0);
      sa = sbeta / dpi;
      if (!stopped && blockIdx.x == 0) ksp_epilogue(s, EPI_CG_DPI, 0, res);
    }
    __syncthreads();
    if (brk) return;
    a = sa;
  } else if (stopped) {
    return;
  }
  double lz[4], lr[4], lzr[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const i64 e = base + 512 * j;
    D2 rn{0.0, 0.0}, zn{0.0, 0.0};
    if (e < n) {
      const D2 xv = xv_[j], pv = pv_[j], rv = rv_[j], wv = wv_[j], dv = dv_[j];
      st2(x, e, n, al, D2{fma(a, pv.a, xv.a), fma(a, pv.b, xv.b)});
      rn = D2{fma(-a, wv.a, rv.a), fma(-a, wv.b, rv.b)};
      if (e + 1 >= n) rn.b = 0.0;
      st2(r, e, n, al, rn);
      zn = D2{rn.a * dv.a, rn.b * dv.b};
      st2(z, e, n, al, zn);
    }
    lz[j] = fma(zn.a, zn.a, zn.b * zn.b);
    lr[j] = fma(rn.a, rn.a, rn.b * rn.b);
    lzr[j] = fma(zn.a, rn.a, zn.b * rn.b);
  }
  const double u[3] = {(lz[0] + lz[1]) + (lz[2] + lz[3]), (lr[0] + lr[1]) + (lr[2] + lr[3]),
                       (lzr[0] + lzr[1]) + (lzr[2] + lzr[3])};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double t = block_tree(u[k], sh);
    if (threadIdx.x == 0) P[(i64)k * gridDim.x + blockIdx.x] = t;
  }
}

// CG with Jacobi, the head of iteration i + 1 in one pass.  Every block folds
// the tile partials of z.z, r.r, z.r from iteration i (k_finalize's order)
// and takes the norm step's decisions from them (ksp.mojo:364-386 and the
// checks of 337-350: convergence, max_it, beta = 0, a sign change of beta)
// reading only state that step leaves alone; i comes from the host, which
// queued the iterations.  The block of tile 0 also runs the step on the
// state.  Unless the solve ends there, p = z + bcoef p over the block's tile
// with OpCgP's arithmetic (VecAYPX, ported.mojo:285-294).
template <bool al>  // p and z 16-byte aligned
__global__ void __launch_bounds__(MPK_NT) k_cg_pz(i64 n, double *p, const double *z, const double *Pn, int T,
                                                  KspState *s, i64 i, const int64_t *stop) {
  __shared__ double res[3];
  __shared__ double sb;
  __shared__ int brk;
  // every load is issued before any is waited on: the stop flag, the state
  // the norm step's decisions read (none of it changes in that step), the
  // tile's operands, then the partials; a stopped solve returns before
  // anything is written
  const bool stopped = stop && *stop != 0;  // the solve's own flag (null: never stopped)
  const ConvView cv = ConvView::load(s);
  const double betaold = s->betaold;
  const i64 base = (i64)blockIdx.x * MPK_TILE + 2 * (i64)threadIdx.x;
  D2 zv_[4], pv_[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const i64 e = base + 512 * j;
    if (e < n) {
      zv_[j] = ld2(z, e, n, al);
      pv_[j] = ld2(p, e, n, al);
    }
  }
  fold_partials_k<3>(Pn, T, res);
  if (threadIdx.x == 0) {
    const double beta = res[2];
    bool end = conv_check(&cv, i + 1, cg_norm(&cv, i, res)) != 0 || i + 1 >= cv.max_it;
    end = end || beta == 0.0 || beta * betaold < 0.0;
    brk = stopped || end;
    sb = end ? 0.0 : beta / betaold;
    if (!stopped && blockIdx.x == 0) ksp_epilogue(s, EPI_CG_NORM, 0, res);
  }
  __syncthreads();
  if (brk) return;
  const double b = sb;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const i64 e = base + 512 * j;
    if (e >= n) break;
    const D2 zv = zv_[j];
    if (b == 0.0) {
      st2(p, e, n, al, zv);
      continue;
    }
    const D2 pv = pv_[j];
    if (b == 1.0)
      st2(p, e, n, al, D2{fma(b, zv.a, pv.a), fma(b, zv.b, pv.b)});
    else
      st2(p, e, n, al, D2{fma(b, pv.a, zv.a), fma(b, pv.b, zv.b)});
  }
}

// p = (1 r + cz v) + beta p  (VecAXPBYPCZ(P, 1, -omegaold beta, beta, R, V))
struct OpBcgsP {
  double *p;
  const double *r, *v;
  const KspState *s;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    double cz = s->cz, bt = s->beta;
    D2 rv = ld2(r, e, n, al), vv = ld2(v, e, n, al), pv = ld2(p, e, n, al);
    st2(p, e, n, al, D2{fma(bt, pv.a, fma(1.0, rv.a, cz * vv.a)), fma(bt, pv.b, fma(1.0, rv.b, cz * vv.b))});
  }
};
// w = (-coef) x + y with coef a device scalar (VecWAXPY)
struct OpWaxpyDev {
  double *w;
  const double *x, *y;
  const double *coef;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    double c = -*coef;
    D2 xv = ld2(x, e, n, al), yv = ld2(y, e, n, al);
    st2(w, e, n, al, D2{fma(c, xv.a, yv.a), fma(c, xv.b, yv.b)});
  }
};
// x = (alpha p + omega s) + 1 x  (VecAXPBYPCZ(X, alpha, omega, 1, P, S))
struct OpBcgsX {
  double *x;
  const double *p, *sv;
  const KspState *s;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    double al_ = s->alpha, om = s->omega;
    D2 pv = ld2(p, e, n, al), sv2 = ld2(sv, e, n, al), xv = ld2(x, e, n, al);
    st2(x, e, n, al,
        D2{fma(1.0, xv.a, fma(al_, pv.a, om * sv2.a)), fma(1.0, xv.b, fma(al_, pv.b, om * sv2.b))});
  }
};
// x += alpha p, once, for BiCGStab's d2 == 0 exit
struct OpBcgsSpecial {
  double *x;
  const double *p;
  const KspState *s;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n || !s->special) return;
    double a = s->alpha;
    D2 xv = ld2(x, e, n, al), pv = ld2(p, e, n, al);
    st2(x, e, n, al, D2{fma(a, pv.a, xv.a), fma(a, pv.b, xv.b)});
  }
};
// x = x / d with d a device scalar, when d > 0 (VecNormalize's divide)
struct OpDivDev {
  double *x;
  const double *d;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    double z = *d;
    if (!(z > 0.0)) return;
    D2 xv = ld2(x, e, n, al);
    st2(x, e, n, al, D2{xv.a / z, xv.b / z});
  }
};
// x = x * (1 / d): FGMRES's VecScale by the reciprocal (ksp.mojo:820,844)
struct OpScaleRecipDev {
  double *x;
  const double *d;
  const i64 *skip_if_hapend;  // null: always
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    if (skip_if_hapend && *skip_if_hapend) return;
    double sc = 1.0 / *d;
    D2 xv = ld2(x, e, n, al);
    st2(x, e, n, al, D2{xv.a * sc, xv.b * sc});
  }
};
// y = y + c x with c = sign * coef[k] a device scalar (MGS axpy)
struct OpAxpyDev {
  double *y;
  const double *x;
  const double *coef;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    double c = *coef;
    D2 xv = ld2(x, e, n, al), yv = ld2(y, e, n, al);
    st2(y, e, n, al, D2{fma(c, xv.a, yv.a), fma(c, xv.b, yv.b)});
  }
};

struct OpCopy {
  double *y;
  const double *x;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    st2(y, e, n, al, ld2(x, e, n, al));
  }
};

template <class Op>
__global__ void __launch_bounds__(MPK_NT) k_ew2(i64 n, Op op, const int64_t *stop) {
  if (stop && *stop) return;
  const i64 base = (i64)blockIdx.x * MPK_TILE + 2 * (i64)threadIdx.x;
#pragma unroll
  for (int j = 0; j < 4; j++) op(base + 512 * j, n);
}
template <class Op>
static void ew2(i64 n, const Op &op, const int64_t *stop) {
  if (n <= 0) return;
  MPK_LAUNCH(k_ew2<Op>, ntiles(n), MPK_NT, n, op, stop);
}

// Back substitution of KSPGMRESBuildSoln / _fgmres_buildsoln (ksp.mojo:632-658,
// 886-900) in one thread.  Sets nbuild = it + 1 vectors to combine, or
// skip_build when there is nothing to combine.
__device__ void gmres_backsub(KspState *s, int flexible) {
  const i64 m = s->m;
  i64 it = s->it - 1;
  double *hh = s->HH, *rs = s->RS;
  s->breason = 0;
  s->skip_build = 1;
  s->nbuild = 0;
  if (it < 0) return;
  if (hh[it * (m + 2) + it] != 0.0) {
    rs[it] = rs[it] / hh[it * (m + 2) + it];
  } else {
    if (!flexible) {
      s->breason = KSP_DIVERGED_BREAKDOWN;
      return;
    }
    rs[it] = 0.0;
  }
  for (i64 ii = 1; ii <= it; ii++) {
    i64 k = it - ii;
    double tt = rs[k];
    // (no stores in the sum: unrolled, its loads run ahead of the chain)
    const double *__restrict__ h = hh;
    const double *__restrict__ r = rs;
#pragma unroll 4
    for (i64 j = k + 1; j <= it; j++) tt = fma(-h[j * (m + 2) + k], r[j], tt);
    if (!flexible && hh[k * (m + 2) + k] == 0.0) {
      s->breason = KSP_DIVERGED_BREAKDOWN;
      return;
    }
    rs[k] = tt / hh[k * (m + 2) + k];
  }
  s->nbuild = it + 1;
  s->skip_build = 0;
}
__global__ void k_gmres_backsub(KspState *s, int flexible, const int64_t *stop) {
  if (stop && *stop) return;
  gmres_backsub(s, flexible);
}

// The tail of a GMRES cycle (ksp.mojo:756-768) / FGMRES cycle (861-867).
__device__ void gmres_cycle_end(KspState *s, int flexible) {
  if (s->stop) return;
  if (!flexible) {
    if (s->reason == 0 && s->breason != 0) s->reason = s->breason;
    s->itcount += s->it;
    if (s->itcount >= s->max_it) {
      if (s->reason == 0) s->reason = KSP_DIVERGED_ITS;
      finish(s, s->res, s->its, s->reason);
      return;
    }
    s->prev_rnorm = s->res;
    if (s->reason != 0) finish(s, s->res, s->its, s->reason);
  } else {
    if (s->its >= s->max_it) {
      if (s->reason == 0) s->reason = KSP_DIVERGED_ITS;
      finish(s, s->res, s->its, s->reason);
      return;
    }
    if (s->reason != 0) finish(s, s->res, s->its, s->reason);
  }
}
__global__ void k_gmres_cycle_end(KspState *s, int flexible) { gmres_cycle_end(s, flexible); }

// Combination for BuildSoln: y = sum_k rs[k] v_k, k < nbuild (device count).
struct OpBuild {
  double *y;
  const double *base;
  i64 ld;
  const KspState *s;
  bool al;
  __device__ void operator()(i64 e, i64 n) const {
    if (e >= n) return;
    i64 nb = s->nbuild;
    D2 acc{0.0, 0.0};  // VecMAXPBY with beta = 0: the zeroed destination
#pragma unroll 4
    for (i64 k = 0; k < nb; k++) {
      double c = s->RS[k];
      D2 v = ld2(base + k * ld, e, n, al);
      acc.a = fma(c, v.a, acc.a);
      acc.b = fma(c, v.b, acc.b);
    }
    st2(y, e, n, al, acc);
  }
};

// ------------------------------------------------------------ host driver

// Device state, scalar block and flag read-back of the solver, reused across
// solves (one solve runs at a time on the library's stream).
struct SolveBuffers {
  KspState *d = nullptr;
  double *buf = nullptr;
  size_t cap = 0;
  int64_t *hflag = nullptr;
  hipEvent_t ev[2];
};
static SolveBuffers &solve_buffers() {
  static SolveBuffers b;
  return b;
}

struct Solve {
  i64 n;
  DevCSR A;
  std::unique_ptr<SellTemp> sell;
  PCView pc;
  KspState h;        // host copy of the initial state
  KspState *d = nullptr;  // device state
  double *dbuf = nullptr; // history + GMRES scalars
  int64_t *stop = nullptr, *inner = nullptr, *skipref = nullptr, *skipbuild = nullptr;
  int64_t *hflag = nullptr;  // pinned host copy of `stop`
  hipEvent_t ev[2];

  bool start(i64 n_, const int64_t *irow, const int64_t *icol, const int64_t *data, const int64_t *pci,
             const double *pcf, const double *cfg, const int64_t *cfgi) {
    n = n_;
    A = DevCSR{n, n, (const int *)irow, (const int *)icol, (const double *)data};
    // the sliced copy of A for the products: the caller's (cfgi[13..15]) or
    // one built for this solve
    if (cfgi[13] && cfgi[14] && cfgi[15]) {
      A.sell_off = (const int *)cfgi[13];
      A.sell_v = (const double *)cfgi[14];
      A.sell_c = (const int *)cfgi[15];
    } else {
      sell.reset(new SellTemp(A));
    }
    if (!pc.parse(pci, pcf, n)) return false;
    memset(&h, 0, sizeof h);
    h.rtol = cfg[0];
    h.atol = cfg[1];
    h.divtol = cfg[2];
    h.haptol = cfg[3];
    h.breakdowntol = cfg[4];
    h.rhs_norm = cfg[12];
    h.normtype = cfgi[0];
    h.side = cfgi[1];
    h.guess_zero = cfgi[2];
    h.min_it = cfgi[3];
    h.chknorm = cfgi[4];
    h.convmaxits = cfgi[5];
    h.max_it = cfgi[6];
    h.m = cfgi[7];
    h.orthog = cfgi[8];
    h.cgstype = cfgi[9];
    h.seed_rhs = cfgi[12];
    h.prev_rnorm = -1.0;
    i64 m = h.m > 0 ? h.m : 1;
    size_t nh = (size_t)h.max_it + 2;
    size_t ns = (size_t)((m + 2) * (m + 1) + (m + 1) * (m + 1) + 2 * (m + 1) + 2 * (m + 2));
    SolveBuffers &B = solve_buffers();
    if (!B.d) {
      HIPCHK(hipMalloc(&B.d, sizeof(KspState)));
      HIPCHK(hipHostMalloc(&B.hflag, 2 * sizeof(int64_t), hipHostMallocDefault));
      HIPCHK(hipEventCreateWithFlags(&B.ev[0], hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&B.ev[1], hipEventDisableTiming));
    }
    if (nh + ns > B.cap) {
      if (B.buf) {
        HIPCHK(hipStreamSynchronize(stream()));
        HIPCHK(hipFree(B.buf));
      }
      B.cap = nh + ns;
      HIPCHK(hipMalloc(&B.buf, B.cap * sizeof(double)));
    }
    d = B.d;
    dbuf = B.buf;
    hflag = B.hflag;
    ev[0] = B.ev[0];
    ev[1] = B.ev[1];
    HIPCHK(hipMemsetAsync(dbuf, 0, (nh + ns) * sizeof(double), stream()));
    if (failed()) return false;
    h.hist = dbuf;
    h.HH = dbuf + nh;
    h.HES = h.HH + (m + 2) * (m + 1);
    h.CC = h.HES + (m + 1) * (m + 1);
    h.SS = h.CC + (m + 1);
    h.RS = h.SS + (m + 1);
    h.LHH = h.RS + (m + 2);
    HIPCHK(hipMemcpyAsync(d, &h, sizeof h, hipMemcpyHostToDevice, stream()));
    stop = &d->stop;
    inner = &d->inner_stop;
    skipref = &d->skip_ref;
    skipbuild = &d->skip_build;
    return !failed();
  }

  // Queue a flag read-back behind the work queued so far (`stop`, or
  // another flag of the state).
  void mark(int slot, const int64_t *flag = nullptr) {
    HIPCHK(hipMemcpyAsync(hflag + slot, flag ? flag : stop, sizeof(int64_t), hipMemcpyDeviceToHost, stream()));
    HIPCHK(hipEventRecord(ev[slot], stream()));
  }
  // Within a GMRES / FGMRES cycle: after every chunk of Arnoldi steps queued
  // (step s done), read inner_stop back a chunk behind; true when the loop is
  // over, so the steps left would only return on the flag and need not be
  // queued.  The first call of a cycle has s = 0.
  bool inner_over(i64 s, i64 m, int chunk, int &slot, bool &pending) {
    if (s == 0) pending = false;
    if ((s + 1) % chunk != 0 || s + 1 >= m) return false;
    mark(slot, inner);
    const bool over = pending && done(slot ^ 1);
    pending = true;
    slot ^= 1;
    return over;
  }
  bool done(int slot) {
    HIPCHK(hipEventSynchronize(ev[slot]));
    return failed() || hflag[slot] != 0;
  }

  void finish_out(double *cfg, int64_t *cfgi) {
    sync();
    KspState out;
    copy_dtoh(&out, d, sizeof out);
    i64 its = out.its;
    std::vector<double> hist((size_t)h.max_it + 2, 0.0);
    copy_dtoh(hist.data(), h.hist, hist.size() * sizeof(double));
    cfg[7] = out.rnorm;
    cfgi[10] = its;
    cfgi[11] = out.reason;
    // the history is written into cfg[9 + k] (ksp.mojo:325,375,...), which also
    // holds cfg[12] and cfg[14]; entries the solve wrote replace them
    i64 lim = its < h.max_it + 1 ? its : h.max_it + 1;
    for (i64 k = 0; k <= lim; k++) cfg[9 + k] = hist[(size_t)k];
    cfg[5] = out.rnorm0;
    cfg[6] = out.ttol;
    if (lim < 5) cfg[14] = out.seeded ? 1.0 : cfg[14];
  }

  template <int K>
  void reduce(const RedSet<K> &rs, int which, int arg, const int64_t *stp) {
    reduce_launch<K>(rs, n, nullptr, KspEpi{d, which, arg}, stp);
  }

  // KSP_PCApplyBAorAB (ksp.mojo:246-265)
  void ba_or_ab(const double *x, double *y, double *w, const int64_t *stp) {
    if (h.side == PC_RIGHT) {
      pc.apply(x, w, stp);
      spmv(A, w, y, 0, stp);
    } else if (h.side == PC_LEFT) {
      spmv(A, x, w, 0, stp);
      pc.apply(w, y, stp);
    } else {
      pc.apply_symmetric(x, w, stp);
      spmv(A, w, y, 0, stp);
      copy_if(y, w, stp);
      pc.apply_symmetric(w, y, stp);
    }
  }
  // KSPInitialResidual (ksp.mojo:221-243)
  void initial_residual(i64 guess_zero, double *vsoln, double *vt1, double *vt2, double *vres, const double *vb,
                        const int64_t *stp) {
    if (!guess_zero) {
      spmv(A, vsoln, vt1, 0, stp);
      copy_if(vb, vt2, stp);
      vaxpy(vt2, -1.0, vt1, n, stp);
    } else {
      copy_if(vb, vt2, stp);
    }
    if (h.side == PC_SYMMETRIC)
      pc.apply_symmetric(vt2, vres, stp);
    else if (h.side == PC_LEFT)
      pc.apply(vt2, vres, stp);
    else
      copy_if(vt2, vres, stp);
  }
  // KSPUnwindPreconditioner (ksp.mojo:211-218)
  void unwind(double *vsoln, double *vt1, const int64_t *stp) {
    copy_if(vsoln, vt1, stp);
    if (h.side == PC_RIGHT) {
      pc.apply(vsoln, vt1, stp);
      copy_if(vt1, vsoln, stp);
    }
  }
  // a copy that honours the stop flag
  void copy_if(const double *x, double *y, const int64_t *stp) {
    if (x == y) return;
    if (!stp) {
      vcopy(x, y, n);
      return;
    }
    ew2(n, OpCopy{y, x, aligned16(y) && aligned16(x)}, stp);
  }
};

// ------------------------------------------------------------ small systems
//
// A CG solve of a small system (n <= MPK_SMALL_ROWS) in one launch of one
// workgroup of MPK_WAVE_THREADS threads.  The steps of an iteration run one
// after another, separated by barriers, with the arithmetic of the general
// path's launches: the same pair operations (OpCgP, OpCgXR, the Jacobi
// product), serial row sums, tile-tree reductions (threads 0-255 and 256-511
// are two 256-thread blocks, each taking tiles as k_partials' blocks do,
// then threads 0-255 fold the partials as k_finalize does), the same
// epilogues in thread 0, and the sweeps of wave plans (wave_sweep) -- so its
// iterates are the general path's bit for bit.  The loop ends when an
// epilogue sets the stop flag (at the latest at max_it).

#define MPK_SMALL_ROWS 16384
#define MPK_SMALL_PLAIN 4096  // the limit for a PC without sweeps, and for GMRES / FGMRES / BiCGStab
#define MPK_SMALL_TILES (MPK_SMALL_ROWS / MPK_TILE)

// The reductions' LDS (small_reduce).
struct SmallRedLds {
  double wsum[8][MPK_SMALL_TILES][4];  // per tile, its four wave sums
  double fsum[8][4];                   // the fold's wave sums
  double res[8];
};
template <int LVLS>
struct SmallLdsT : SmallRedLds {
  double val[MPK_BLK_ROWS + 1];  // the sweeps' values
  int lv[LVLS + 1];              // the sweeps' level bounds
};
using SmallLds = SmallLdsT<MPK_WAVE_LVLS>;

struct SmallCG {
  i64 n, max_it;
  const int *rp, *ci;
  const double *v;
  double *X, *R, *Z, *P, *W;  // W may be Z
  KspState *s;
  const double *jac;  // Jacobi: the inverted diagonal
  IluFactor ilu;
  IccFactor icc;
  int form0, form1;  // the sweeps' wave forms
  double *S;         // the sweeps' scratch
  Staged *st;
};

// K reductions (RedSet) of length n <= MPK_SMALL_ROWS into L.res, in the
// tile-tree order (mpk_reduce.h).  Every thread calls it.
template <int K, class LD>
__device__ __forceinline__ void small_reduce(const RedSet<K> &rs, i64 n, LD &L) {
  const int t = (int)threadIdx.x, tl = t & 255, lane = t & 63, wv = (t >> 6) & 3;
  const int T = (int)((n + MPK_TILE - 1) / MPK_TILE);
  bool al = true;
#pragma unroll
  for (int k = 0; k < K; k++) al = al && aligned16(rs.s[k].x) && (rs.s[k].kind != RK_DOT || aligned16(rs.s[k].y));
  for (int j = t >> 8; j < T; j += 2) {  // block t >> 8 takes every other tile
    // every sum's leaves (and so all their loads) before any wave tree
    const i64 base = (i64)j * MPK_TILE + 2 * tl;
    double u[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double p0 = pair_leaf(rs.s[k], base, n, al), p1 = pair_leaf(rs.s[k], base + 512, n, al);
      const double p2 = pair_leaf(rs.s[k], base + 1024, n, al), p3 = pair_leaf(rs.s[k], base + 1536, n, al);
      u[k] = (p0 + p1) + (p2 + p3);
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double w = wave_tree(u[k]);
      if (lane == 0) L.wsum[k][j][wv] = w;
    }
  }
  __syncthreads();
  if (t < 256) {  // the fold: thread t's partial P_t (T <= 256), then the block tree
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double u = t < T ? (L.wsum[k][t][0] + L.wsum[k][t][2]) + (L.wsum[k][t][1] + L.wsum[k][t][3]) : 0.0;
      const double r = wave_tree(u);
      if (lane == 0) L.fsum[k][wv] = r;
    }
  }
  __syncthreads();
  if (t == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) L.res[k] = (L.fsum[k][0] + L.fsum[k][2]) + (L.fsum[k][1] + L.fsum[k][3]);
  }
}

// y = A x, each row's serial sum from 0.0 in CSR order (MatMult_SeqAIJ,
// ported.mojo:380-391).  A thread takes 2 rows at a time and issues the loads
// of up to 8 entries of each before any product, so that the rows' loads
// overlap; longer rows finish in a loop.
__device__ __forceinline__ void small_spmv(const SmallCG &a, const double *x, double *y) {
  constexpr int R = 2, E = 8;
  const int t = (int)threadIdx.x;
  for (i64 i0 = t; i0 < a.n; i0 += R * MPK_WAVE_THREADS) {
    int j0[R], j1[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const i64 i = i0 + (i64)r * MPK_WAVE_THREADS;
      j0[r] = i < a.n ? a.rp[i] : 0;
      j1[r] = i < a.n ? a.rp[i + 1] : 0;
    }
    int c[R][E];
    double v[R][E], xv[R][E], acc[R];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int e = 0; e < E; e++)
        if (j0[r] + e < j1[r]) {
          c[r][e] = a.ci[j0[r] + e];
          v[r][e] = a.v[j0[r] + e];
        }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int e = 0; e < E; e++)
        if (j0[r] + e < j1[r]) xv[r][e] = x[c[r][e]];
#pragma unroll
    for (int r = 0; r < R; r++) {
      acc[r] = 0.0;
#pragma unroll
      for (int e = 0; e < E; e++)
        if (j0[r] + e < j1[r]) acc[r] = fma(v[r][e], xv[r][e], acc[r]);
      for (int j = j0[r] + E; j < j1[r]; j++) acc[r] = fma(a.v[j], x[a.ci[j]], acc[r]);
      const i64 i = i0 + (i64)r * MPK_WAVE_THREADS;
      if (i < a.n) y[i] = acc[r];
    }
  }
}

// z = M^-1 r for the PC type PCT (PCView::apply's arithmetic), ending with a
// barrier.
template <int PCT, class LD>
__device__ __forceinline__ void small_pc(const SmallCG &a, const double *r, double *z, LD &L) {
  const i64 n = a.n;
  const int t = (int)threadIdx.x;
  if constexpr (PCT == PCT_JACOBI) {
    for (i64 i = t; i < n; i += MPK_WAVE_THREADS) z[i] = r[i] * a.jac[i];
  } else if constexpr (PCT == PCT_ILU) {
    double *F = a.S, *B = a.S + n;
    const IluFactor &f = a.ilu;
    wave_sweep_form(a.form0, IluFwd{f.bi, f.bj, f.r, f.ba, r}, n, f.plan, 0, a.st, F, L.val, L.lv);
    __syncthreads();
    wave_sweep_form(a.form1, IluBwd{f.bj, f.bdiag, f.c, f.ba, F, z}, n, f.plan, 1, a.st, B, L.val, L.lv);
  } else if constexpr (PCT == PCT_ICC) {
    double *Y = a.S, *F = a.S + n, *B = a.S + 2 * n;
    const IccFactor &f = a.icc;
    wave_sweep_form(a.form0, IccFwd{f.trp, f.trow, f.tpos, f.tmid, f.udiag, f.c, f.ua, r, F}, n, f.plan, 0, a.st,
                    Y, L.val, L.lv);
    __syncthreads();
    if (f.nlate) {
      for (i64 k = t; k < n; k += MPK_WAVE_THREADS) icc_late_row(k, f.trp, f.trow, f.tpos, f.tmid, f.ua, Y, F);
      __syncthreads();
    }
    wave_sweep_form(a.form1, IccBwd{f.ui, f.uj, f.udiag, f.c, f.ua, F, z, n - 1}, n, f.plan, 1, a.st, B, L.val,
                    L.lv);
  } else if (r != z) {
    for (i64 i = t; i < n; i += MPK_WAVE_THREADS) z[i] = r[i];
  }
  __syncthreads();
}

__device__ __forceinline__ bool small_stopped(const KspState *s) {
  return *reinterpret_cast<const volatile int64_t *>(&s->stop) != 0;
}

// The iterations of KSPSolve_CG (ksp.mojo:337-389) after the initial
// residual, its PC apply and EPI_CG_INIT, as the general path queues them.
// The solve's state lives in LDS for the kernel's length (thread 0's
// epilogues are serial chains of its fields), copied back at the end.
struct SmallCgLds : SmallLds {
  KspState ks;
};

template <int PCT>
__global__ void __launch_bounds__(MPK_WAVE_THREADS) k_cg_small(SmallCG a) {
  __shared__ SmallCgLds L;
  const i64 n = a.n;
  const int t = (int)threadIdx.x;
  if (t == 0) L.ks = *a.s;
  __syncthreads();
  KspState *s = &L.ks;
  const bool al_p = aligned16(a.P) && aligned16(a.Z);
  const bool al_x = aligned16(a.X) && aligned16(a.R) && aligned16(a.P) && aligned16(a.W);
  for (i64 it = 0; it <= a.max_it + 1; it++) {
    if (small_stopped(s)) break;  // the flag is read after a barrier: uniform
    for (i64 e = 2 * (i64)t; e < n; e += 2 * MPK_WAVE_THREADS) OpCgP{a.P, a.Z, s, al_p}(e, n);
    __syncthreads();
    small_spmv(a, a.P, a.W);
    __syncthreads();
    small_reduce<1>(RedSet<1>{{{a.P, a.W, RK_DOT}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_CG_DPI, 0, L.res);
    __syncthreads();
    if (small_stopped(s)) break;
    for (i64 e = 2 * (i64)t; e < n; e += 2 * MPK_WAVE_THREADS) OpCgXR{a.X, a.R, a.P, a.W, s, al_x}(e, n);
    __syncthreads();
    small_pc<PCT>(a, a.R, a.Z, L);
    small_reduce<3>(RedSet<3>{{{a.Z, a.Z, RK_SUMSQ}, {a.R, a.R, RK_SUMSQ}, {a.Z, a.R, RK_DOT}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_CG_NORM, 0, L.res);
    __syncthreads();
  }
  __syncthreads();
  if (t == 0) *a.s = L.ks;
}

// Systems of at most MPK_SMALL_LDS_ROWS rows without sweeps (no PC, or
// Jacobi) keep their vectors in LDS as well: k_cg_small's steps, on LDS
// copies of X, R, Z (= W) and P, every step's loads and stores then local.
// The same operations on the same values, so the same bits; the vectors
// and the state go back to global memory at the end.
#define MPK_SMALL_LDS_ROWS 2048

struct SmallCgVecLds {
  double vec[4][MPK_SMALL_LDS_ROWS] __attribute__((aligned(16)));  // X, R, Z (= W), P
  double wsum[8][MPK_SMALL_TILES][4];
  double fsum[8][4];
  double res[8];
  KspState ks;
};

// This thread's rows (t, t + 512, ..) of the matrix held in registers for
// the solve: at most 8 entries a row (else `fits` is false and the product
// reads the matrix as small_spmv does).
// RR rows a thread: systems of up to RR * 512 rows.
template <int RR = (MPK_SMALL_LDS_ROWS + MPK_WAVE_THREADS - 1) / MPK_WAVE_THREADS>
struct SmallRows {
  int cnt[RR];
  int c[RR][8];
  double v[RR][8];
};
template <int RR>
__device__ __forceinline__ bool small_rows_load(const SmallCG &a, SmallRows<RR> &m) {
  const int t = (int)threadIdx.x;
  bool fits = a.n <= (i64)RR * MPK_WAVE_THREADS;
#pragma unroll
  for (int r = 0; r < RR; r++) {
    const i64 i = t + (i64)r * MPK_WAVE_THREADS;
    const int j0 = i < a.n ? a.rp[i] : 0, j1 = i < a.n ? a.rp[i + 1] : 0;
    m.cnt[r] = j1 - j0;
    fits = fits && j1 - j0 <= 8;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const bool in = j0 + e < j1;
      m.c[r][e] = in ? a.ci[j0 + e] : 0;
      m.v[r][e] = in ? a.v[j0 + e] : 0.0;
    }
  }
  return __syncthreads_and(fits) != 0;
}
// y = A x from the rows in registers: each row's serial sum from 0.0 in CSR
// order, as small_spmv.
template <int RR>
__device__ __forceinline__ void small_spmv_rows(const SmallRows<RR> &m, i64 n, const double *x, double *y) {
  const int t = (int)threadIdx.x;
#pragma unroll
  for (int r = 0; r < RR; r++) {
    const i64 i = t + (i64)r * MPK_WAVE_THREADS;
    double xv[8];
#pragma unroll
    for (int e = 0; e < 8; e++) xv[e] = x[m.c[r][e]];
    double acc = 0.0;
#pragma unroll
    for (int e = 0; e < 8; e++)
      if (e < m.cnt[r]) acc = fma(m.v[r][e], xv[e], acc);
    if (i < n) y[i] = acc;
  }
}

template <int PCT>
__global__ void __launch_bounds__(MPK_WAVE_THREADS) k_cg_small_lds(SmallCG g) {
  __shared__ SmallCgVecLds L;
  const i64 n = g.n;
  const int t = (int)threadIdx.x;
  SmallRows<> rows;
  const bool fits = small_rows_load(g, rows);
  for (i64 i = t; i < n; i += MPK_WAVE_THREADS) {
    L.vec[0][i] = g.X[i];
    L.vec[1][i] = g.R[i];
    L.vec[2][i] = g.Z[i];
    L.vec[3][i] = g.P[i];
  }
  if (t == 0) L.ks = *g.s;
  __syncthreads();
  SmallCG a = g;
  a.X = L.vec[0], a.R = L.vec[1], a.Z = L.vec[2], a.W = L.vec[2], a.P = L.vec[3];
  KspState *s = &L.ks;
  const bool al_p = true, al_x = true;  // LDS rows are 16-byte aligned
  for (i64 it = 0; it <= a.max_it + 1; it++) {
    if (small_stopped(s)) break;  // the flag is read after a barrier: uniform
    for (i64 e = 2 * (i64)t; e < n; e += 2 * MPK_WAVE_THREADS) OpCgP{a.P, a.Z, s, al_p}(e, n);
    __syncthreads();
    if (fits)
      small_spmv_rows(rows, n, a.P, a.W);
    else
      small_spmv(a, a.P, a.W);
    __syncthreads();
    small_reduce<1>(RedSet<1>{{{a.P, a.W, RK_DOT}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_CG_DPI, 0, L.res);
    __syncthreads();
    if (small_stopped(s)) break;
    for (i64 e = 2 * (i64)t; e < n; e += 2 * MPK_WAVE_THREADS) OpCgXR{a.X, a.R, a.P, a.W, s, al_x}(e, n);
    __syncthreads();
    if (PCT == PCT_JACOBI)
      for (i64 i = t; i < n; i += MPK_WAVE_THREADS) a.Z[i] = a.R[i] * a.jac[i];
    else
      for (i64 i = t; i < n; i += MPK_WAVE_THREADS) a.Z[i] = a.R[i];
    __syncthreads();
    small_reduce<3>(RedSet<3>{{{a.Z, a.Z, RK_SUMSQ}, {a.R, a.R, RK_SUMSQ}, {a.Z, a.R, RK_DOT}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_CG_NORM, 0, L.res);
    __syncthreads();
  }
  __syncthreads();
  for (i64 i = t; i < n; i += MPK_WAVE_THREADS) {
    g.X[i] = L.vec[0][i];
    g.R[i] = L.vec[1][i];
    g.Z[i] = L.vec[2][i];
    g.P[i] = L.vec[3][i];
  }
  if (t == 0) *g.s = L.ks;
}

// ------------------------------------------------------------ small GMRES / FGMRES / BiCGStab
//
// The same one-workgroup form for KSPSolve_GMRES (ksp.mojo:667-769),
// KSPSolve_FGMRES (775-868) and the iterations of KSPSolve_BCGS (437-505):
// each step of the general path's launch sequence in order, with its
// arithmetic (the pair operations, serial row sums, tile-tree reductions and
// the epilogues in thread 0), and a step the general path would skip on a
// flag (stop, inner_stop, skip_ref, skip_build) skipped on the same flag,
// read after the barrier that follows the epilogue setting it.

#define MPK_SMALL_M 31  // the largest restart of the one-workgroup GMRES
#define MPK_SMALL_GM_ROWS 1024  // the limit of the LDS-resident GMRES / FGMRES
#define MPK_SMALL_GM_BAS 12     // the basis vectors they keep in LDS (a multiple of 4)

// The LDS of the GMRES / FGMRES / BiCGStab drivers: the sweeps' and
// reductions' space (levels up to the rows, n <= MPK_SMALL_PLAIN), the
// Gram-Schmidt sums of a step, and the solve's state with its Hessenberg,
// Givens and right-hand-side arrays, moved here for the kernel's length so
// that the epilogues' serial steps in thread 0 touch LDS, not global memory.
struct SmallGmFields {
  double mres[MPK_SMALL_M + 1];
  double mw[MPK_SMALL_M + 1][2][4];  // the Gram-Schmidt sums' wave sums per tile (n <= 4096: two tiles)
  KspState ks;
  double HH[(MPK_SMALL_M + 2) * (MPK_SMALL_M + 1)], HES[(MPK_SMALL_M + 1) * (MPK_SMALL_M + 1)];
  double CC[MPK_SMALL_M + 1], SS[MPK_SMALL_M + 1], RS[MPK_SMALL_M + 2], LHH[MPK_SMALL_M + 1];
};
struct SmallKspLds : SmallLdsT<MPK_SMALL_PLAIN>, SmallGmFields {};

// The state into LDS (with the GMRES arrays when `arrays`: m <= MPK_SMALL_M);
// returns the LDS copy, which the kernel then uses throughout.  Every thread
// calls it.
template <class LD>
__device__ __forceinline__ KspState *small_state_in(KspState *g, LD &L, bool arrays) {
  const int t = (int)threadIdx.x;
  if (arrays) {
    const i64 m = g->m;
    for (i64 i = t; i < (m + 2) * (m + 1); i += MPK_WAVE_THREADS) L.HH[i] = g->HH[i];
    for (i64 i = t; i < (m + 1) * (m + 1); i += MPK_WAVE_THREADS) L.HES[i] = g->HES[i];
    for (i64 i = t; i <= m; i += MPK_WAVE_THREADS) {
      L.CC[i] = g->CC[i];
      L.SS[i] = g->SS[i];
      L.LHH[i] = g->LHH[i];
    }
    for (i64 i = t; i < m + 2; i += MPK_WAVE_THREADS) L.RS[i] = g->RS[i];
  }
  if (t == 0) {
    L.ks = *g;
    if (arrays) {
      L.ks.HH = L.HH, L.ks.HES = L.HES, L.ks.CC = L.CC;
      L.ks.SS = L.SS, L.ks.RS = L.RS, L.ks.LHH = L.LHH;
    }
  }
  __syncthreads();
  return &L.ks;
}
// The LDS state back into g (g keeps its own array pointers).  Every thread
// calls it.
template <class LD>
__device__ __forceinline__ void small_state_out(KspState *g, LD &L, bool arrays) {
  __syncthreads();
  const int t = (int)threadIdx.x;
  if (arrays) {
    const i64 m = L.ks.m;
    for (i64 i = t; i < (m + 2) * (m + 1); i += MPK_WAVE_THREADS) g->HH[i] = L.HH[i];
    for (i64 i = t; i < (m + 1) * (m + 1); i += MPK_WAVE_THREADS) g->HES[i] = L.HES[i];
    for (i64 i = t; i <= m; i += MPK_WAVE_THREADS) {
      g->CC[i] = L.CC[i];
      g->SS[i] = L.SS[i];
      g->LHH[i] = L.LHH[i];
    }
    for (i64 i = t; i < m + 2; i += MPK_WAVE_THREADS) g->RS[i] = L.RS[i];
  }
  if (t == 0) {
    KspState o = L.ks;
    o.HH = g->HH, o.HES = g->HES, o.CC = g->CC;
    o.SS = g->SS, o.RS = g->RS, o.LHH = g->LHH;
    *g = o;
  }
}

struct SmallKsp {
  SmallCG c;  // the matrix, the PC, the state, c.X the solution
  int side, flexible;
  i64 m, orthog, cgstype, guess_zero;
  double *vv, *prevec, *TEMP, *TMP;  // GMRES / FGMRES: the bases (leading dimension n), work
  const double *B;
  double *R, *RP, *V, *T, *Sv, *P;  // BiCGStab
};

__device__ __forceinline__ bool small_flag(const i64 *f) { return *reinterpret_cast<const volatile i64 *>(f) != 0; }

// A pair operation (Op(e, n) on elements e, e + 1) over the vector, as k_ew2
// applies it; no barrier.
template <class Op>
__device__ __forceinline__ void small_ew(const Op &op, i64 n) {
  for (i64 e = 2 * (i64)threadIdx.x; e < n; e += 2 * MPK_WAVE_THREADS) op(e, n);
}
__device__ __forceinline__ void small_copy(const double *x, double *y, i64 n) {
  if (x == y) return;
  for (i64 i = threadIdx.x; i < n; i += MPK_WAVE_THREADS) y[i] = x[i];
}
// y = alpha x + y, VecAXPY's fma(alpha, x, y)
__device__ __forceinline__ void small_axpy(double *y, double alpha, const double *x, i64 n) {
  for (i64 i = threadIdx.x; i < n; i += MPK_WAVE_THREADS) y[i] = fma(alpha, x[i], y[i]);
}

// KSP_PCApplyBAorAB (ksp.mojo:246-265) for the left and right sides, w the
// work vector; ends with a barrier.
template <int PCT>
__device__ __forceinline__ void small_ba(const SmallKsp &a, const double *x, double *y, double *w, SmallKspLds &L) {
  if (a.side == PC_RIGHT) {
    small_pc<PCT>(a.c, x, w, L);
    small_spmv(a.c, w, y);
  } else {
    small_spmv(a.c, x, w);
    __syncthreads();
    small_pc<PCT>(a.c, w, y, L);
  }
  __syncthreads();
}

// The tile sums of x . V_j for G vectors from v0 (leading dimension ld)
// into L.mw[j0..]: k_partials_mdot's per-thread pairs (x's pairs loaded once
// for the G vectors) and wave trees.
// `h` is the half of the workgroup (256 threads, a block of the general
// path) that takes the tiles: both halves when they split the tiles, one
// when they split the groups.
template <int G, class LD>
__device__ __forceinline__ void small_mdot_tiles(const double *x, const double *v0, i64 ld, int j0, i64 n, int T,
                                                 int h0, int hs, LD &L) {
  const int t = (int)threadIdx.x, tl = t & 255, lane = t & 63, wv = (t >> 6) & 3;
  const bool al = aligned16(x) && aligned16(v0) && ld % 2 == 0;
  for (int j = (t >> 8) - h0; j >= 0 && j < T; j += hs) {
    const i64 base = (i64)j * MPK_TILE + 2 * tl;
    D2 xs[4];
#pragma unroll
    for (int q = 0; q < 4; q++) xs[q] = ld2(x, base + 512 * q, n, al);
    double u[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
      const double *v = v0 + (i64)g * ld;
      double p[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const D2 vv = ld2(v, base + 512 * q, n, al);
        p[q] = fma(xs[q].a, vv.a, xs[q].b * vv.b);
      }
      u[g] = (p[0] + p[1]) + (p[2] + p[3]);
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
      const double w = wave_tree(u[g]);
      if (lane == 0) L.mw[j0 + g][j][wv] = w;
    }
  }
}
// small_mdot_tiles for n <= 1024 (one tile; a thread's pairs at base + 1024
// and base + 1536 lie past n, their products +0.0, added as (0.0 + 0.0)),
// for the cnt <= 8 vectors vec(j0), ..: every vector's loads issued before
// any product (past cnt the last vector is reloaded, unused).  Called by
// the half of the workgroup that takes the group.
template <class VF, class LD>
__device__ __forceinline__ void small_mdot_tile1k(const double *x, VF vec, bool al, int j0, int cnt, i64 n,
                                                  LD &L) {
  const int t = (int)threadIdx.x, tl = t & 255, lane = t & 63, wv = (t >> 6) & 3;
  const i64 base = 2 * tl;
  const D2 x0 = ld2(x, base, n, al), x1 = ld2(x, base + 512, n, al);
  D2 v0[8], v1[8];
#pragma unroll
  for (int g = 0; g < 8; g++) {
    const double *v = vec(j0 + (g < cnt ? g : cnt - 1));
    v0[g] = ld2(v, base, n, al);
    v1[g] = ld2(v, base + 512, n, al);
  }
  double u[8];
#pragma unroll
  for (int g = 0; g < 8; g++) {
    const double p0 = fma(x0.a, v0[g].a, x0.b * v0[g].b), p1 = fma(x1.a, v1[g].a, x1.b * v1[g].b);
    u[g] = (p0 + p1) + (0.0 + 0.0);
  }
#pragma unroll
  for (int g = 0; g < 8; g++) {
    const double w = wave_tree(u[g]);
    if (lane == 0 && g < cnt) L.mw[j0 + g][0][wv] = w;
  }
}

// x . V_j, j < nv, into L.mres, n <= MPK_SMALL_PLAIN (at most two tiles):
// every sum's tile partials (block_tree's (W0 + W2) + (W1 + W3) of its wave
// sums), then k_finalize's fold of the T partials -- thread t < 256 holds
// P_t (0.0 past T), each wave sums its lanes by the halving tree, the result
// is (F0 + F2) + (F1 + F3).  With T <= 64 the waves past the first sum
// zeros (F1 = F2 = F3 = +0.0), so a sum's fold is one wave's tree of
// P_0, .., P_{T-1} and (F0 + 0.0) + (0.0 + 0.0): the same arithmetic, done
// by one wave per sum, the waves taking different sums.  Ends with a barrier.
// V_j for j < KL (a multiple of 4) is read from Vl (leading dimension
// MPK_SMALL_GM_ROWS: the basis vectors a kernel keeps in LDS), the others
// from V (leading dimension ld).
template <int KL, class LD>
__device__ __forceinline__ void small_mdot(const double *x, const double *V, i64 ld, int nv, i64 n, LD &L,
                                           const double *Vl) {
  constexpr int kl = KL;
  const int T = (int)((n + MPK_TILE - 1) / MPK_TILE);
  // one tile: the halves take alternate groups of four (of eight when n <=
  // 1024); two: one tile each
  const int half = (int)threadIdx.x >> 8;
  auto vec = [&](int j, i64 &ldj) -> const double * {
    ldj = j < kl ? (i64)MPK_SMALL_GM_ROWS : ld;
    return j < kl ? Vl + (i64)j * MPK_SMALL_GM_ROWS : V + (i64)j * ld;
  };
  bool one_k = false;  // the LDS kernels' one-tile form (n <= 1024)
  if constexpr (KL > 0) {
    if (n <= 1024) {
      auto vj = [&](int j) { return j < kl ? Vl + (i64)j * MPK_SMALL_GM_ROWS : V + (i64)j * ld; };
      const bool al = aligned16(x) && aligned16(V) && ld % 2 == 0 && aligned16(Vl);
      for (int j0 = 0, gi = 0; j0 < nv; j0 += 8, gi++)
        if ((gi & 1) == half) small_mdot_tile1k(x, vj, al, j0, nv - j0 < 8 ? nv - j0 : 8, n, L);
      one_k = true;
    }
  }
  if (!one_k) {
    int j0 = 0;
    for (int gi = 0; j0 + 4 <= nv; j0 += 4, gi++)
      if (T == 2 || (gi & 1) == half) {
        i64 ldj;
        const double *vj = vec(j0, ldj);
        small_mdot_tiles<4>(x, vj, ldj, j0, n, T, T == 2 ? 0 : half, T == 2 ? 2 : 1, L);
      }
    i64 ld0;
    const double *v0 = vec(j0, ld0);
    const int gl = nv / 4;  // the tail's group index
    if (T == 2 || (gl & 1) == half) {
      const int h0 = T == 2 ? 0 : half, hs = T == 2 ? 2 : 1;
      if (nv - j0 == 3)
        small_mdot_tiles<3>(x, v0, ld0, j0, n, T, h0, hs, L);
      else if (nv - j0 == 2)
        small_mdot_tiles<2>(x, v0, ld0, j0, n, T, h0, hs, L);
      else if (nv - j0 == 1)
        small_mdot_tiles<1>(x, v0, ld0, j0, n, T, h0, hs, L);
    }
  }
  __syncthreads();
  if (T == 1) {
    // one partial: the wave tree adds +0.0 (the other lanes' sums) to it at
    // each of its six steps, so one thread a sum does the same additions
    for (int g = (int)threadIdx.x; g < nv; g += MPK_WAVE_THREADS) {
      const double u = (L.mw[g][0][0] + L.mw[g][0][2]) + (L.mw[g][0][1] + L.mw[g][0][3]);
      const double f0 = (((((u + 0.0) + 0.0) + 0.0) + 0.0) + 0.0) + 0.0;
      L.mres[g] = (f0 + 0.0) + (0.0 + 0.0);
    }
  } else {
    const int lane = (int)threadIdx.x & 63;
    for (int g = (int)threadIdx.x >> 6; g < nv; g += MPK_WAVE_THREADS / 64) {
      const double u =
          lane < T ? (L.mw[g][lane][0] + L.mw[g][lane][2]) + (L.mw[g][lane][1] + L.mw[g][lane][3]) : 0.0;
      const double f0 = wave_tree(u);
      if (lane == 0) L.mres[g] = (f0 + 0.0) + (0.0 + 0.0);
    }
  }
  __syncthreads();
}

// One Arnoldi orthogonalisation of vnew against v_0..v_it (orthogonalize's
// steps); ends with a barrier.  The CGS sums and MAXPY read v_j, j < KL, from
// Vl (small_mdot's LDS copy of the basis).
template <int KL = 0, class LD>
__device__ __forceinline__ void small_orthogonalize(const SmallKsp &a, KspState *s, i64 it, double *vnew, LD &L,
                                                    const double *Vl = nullptr) {
  constexpr int kl = KL;
  const i64 n = a.c.n;
  const int t = (int)threadIdx.x;
  if (a.orthog == 1) {  // MGS
    for (i64 k = 0; k <= it; k++) {
      const double *vk = a.vv + k * n;
      small_reduce<1>(RedSet<1>{{{vnew, vk, RK_DOT}}}, n, L);
      if (t == 0) ksp_epilogue(s, EPI_GM_MGS, (int)k, L.res);
      __syncthreads();
      small_ew(OpAxpyDev{vnew, vk, s->LHH + k, aligned16(vnew) && aligned16(vk)}, n);
      __syncthreads();
    }
    return;
  }
  const int nv = (int)(it + 1);
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1 && small_flag(&s->skip_ref)) break;
    small_mdot<KL>(vnew, a.vv, n, nv, n, L, Vl);
    // EPI_GM_CGS_A / EPI_GM_CGS_B with each entry j <= it on thread j: the
    // same operations per entry (HH(j) = 0.0 - LHH(j) in the first pass, as
    // zeroing then subtracting gives), then, for refine-if-needed, the
    // serial norm in thread 0.  Nothing below reads an entry's LHH, HH or
    // HES before the next barrier but its own thread and that norm (MAXPY
    // takes LHH(j) = -mres(j) from mres), so only the norm waits for them.
    {
      const i64 itc = s->it, m = s->m;
      if (t <= itc) {
        const double l = -L.mres[t];
        s->LHH[t] = l;
        double *hh = s->HH + itc * (m + 2) + t, *he = s->HES + itc * (m + 1) + t;
        if (pass == 0) {
          *hh = 0.0 - l;
          *he = 0.0 - l;
        } else {
          *hh = *hh - l;
          *he = *he - l;
        }
      }
      if (s->cgstype == 1) {  // uniform
        __syncthreads();
        if (t == 0 && pass == 0) {
          double hnrm = 0.0;
          for (i64 j = 0; j <= itc; j++) hnrm = fma(s->LHH[j], s->LHH[j], hnrm);
          s->hnrm = sqrt(hnrm);
          s->skip_ref = 1;  // decided after the norm of the projected vector
        }
        __syncthreads();
      } else if (t == 0 && pass == 0) {
        s->skip_ref = s->cgstype == 2 ? 0 : 1;  // read after MAXPY's barrier
      }
    }
    // vmaxpy_dev(vnew, vv, n, LHH, nv, 1.0): each element's chain over the nv
    // vectors, eight vectors' loads issued before their multiply-adds
    // (rows i0 and i0 + 512 together; a group's loads past the last vector
    // reload the last one, unused, so that every load is unconditional)
    const double *mres = L.mres;  // LHH(j) = -mres(j)
    auto vec = [&](int j) { return j < kl ? Vl + (i64)j * MPK_SMALL_GM_ROWS : a.vv + (i64)j * n; };
    for (i64 i0 = t; i0 < n; i0 += 2 * MPK_WAVE_THREADS) {
      const i64 i1 = i0 + MPK_WAVE_THREADS < n ? i0 + MPK_WAVE_THREADS : i0;
      double acc0 = vnew[i0], acc1 = vnew[i1];
      for (int k = 0; k < nv; k += 8) {
        double v0[8], v1[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const double *vp = vec(k + u < nv ? k + u : nv - 1);
          v0[u] = vp[i0];
          v1[u] = vp[i1];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {  // past the last vector: acc kept
          const bool in = k + u < nv;
          const double c = 1.0 * -mres[in ? k + u : nv - 1];
          const double r0 = fma(c, v0[u], acc0), r1 = fma(c, v1[u], acc1);
          acc0 = in ? r0 : acc0;
          acc1 = in ? r1 : acc1;
        }
      }
      vnew[i0] = acc0;
      if (i1 != i0) vnew[i1] = acc1;
    }
    __syncthreads();
    if (pass == 0 && a.cgstype == 1) {
      small_reduce<1>(RedSet<1>{{{vnew, vnew, RK_SUMSQ}}}, n, L);
      if (t == 0) ksp_epilogue(s, EPI_GM_CGS_NORM, 0, L.res);
      __syncthreads();
    }
    if (pass == 0 && a.cgstype == 0) break;
  }
}

// KSPSolve_GMRES's restart cycles (mp_ksp_solve_gmres's launch sequence).
template <int PCT>
__global__ void __launch_bounds__(MPK_WAVE_THREADS) k_gmres_small(SmallKsp a) {
  __shared__ SmallKspLds L;
  KspState *s = small_state_in(a.c.s, L, true);
  const i64 n = a.c.n, m = a.m;
  const int t = (int)threadIdx.x;
  double *X = a.c.X, *vv = a.vv, *TEMP = a.TEMP, *TMP = a.TMP;
  i64 guess_zero = a.guess_zero;
  for (i64 cycle = 0;; cycle++) {
    // KSPInitialResidual(guess_zero, X, TEMP, TMP, vv, B)
    if (!guess_zero) {
      small_spmv(a.c, X, TEMP);
      small_copy(a.B, TMP, n);
      __syncthreads();
      small_axpy(TMP, -1.0, TEMP, n);
    } else {
      small_copy(a.B, TMP, n);
    }
    __syncthreads();
    if (a.side == PC_LEFT) {
      small_pc<PCT>(a.c, TMP, vv, L);
    } else {
      small_copy(TMP, vv, n);
      __syncthreads();
    }
    small_reduce<1>(RedSet<1>{{{vv, vv, RK_SUMSQ}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_GM_START, 0, L.res);
    __syncthreads();
    if (!small_flag(&s->stop)) small_ew(OpDivDev{vv, &s->res, aligned16(vv)}, n);
    __syncthreads();
    for (i64 k = 0; k < m; k++) {
      if (small_flag(&s->inner_stop)) break;
      double *vk = vv + k * n, *vnew = vv + (k + 1) * n;
      small_ba<PCT>(a, vk, vnew, TMP, L);
      small_orthogonalize(a, s, k, vnew, L);
      small_reduce<1>(RedSet<1>{{{vnew, vnew, RK_SUMSQ}}}, n, L);
      if (t == 0) ksp_epilogue(s, EPI_GM_STEP, 0, L.res);
      __syncthreads();
      if (!small_flag(&s->inner_stop)) small_ew(OpDivDev{vnew, &s->tt, aligned16(vnew)}, n);
      __syncthreads();
    }
    // KSPGMRESBuildSoln
    if (t == 0 && !small_flag(&s->stop)) gmres_backsub(s, 0);
    __syncthreads();
    if (!small_flag(&s->skip_build)) {
      small_ew(OpBuild{TEMP, vv, n, s, aligned16(TEMP) && aligned16(vv) && n % 2 == 0}, n);
      __syncthreads();
      // KSPUnwindPreconditioner(TEMP, TMP)
      small_copy(TEMP, TMP, n);
      __syncthreads();
      if (a.side == PC_RIGHT) {
        small_pc<PCT>(a.c, TEMP, TMP, L);
        small_copy(TMP, TEMP, n);
        __syncthreads();
      }
      small_axpy(X, 1.0, TEMP, n);
      __syncthreads();
    }
    if (t == 0) gmres_cycle_end(s, 0);
    __syncthreads();
    if (small_flag(&s->stop)) break;
    guess_zero = 0;
    if (cycle > a.c.max_it + 2) break;
  }
  small_state_out(a.c.s, L, true);
}

// k_gmres_small for systems of at most MPK_SMALL_GM_ROWS rows without sweeps:
// the Arnoldi step's vectors -- v_k, the new vector and the work vector --
// in LDS and the matrix rows in registers, so the product, the PC, the norm
// and the scaling touch no global memory; each new basis vector is written
// to vv once scaled, and the first MPK_SMALL_GM_BAS of them to LDS as well:
// the Gram-Schmidt sums and MAXPY read those there, the rest (and BuildSoln
// all of them) from vv.  The same operations on the same values.
// No sweeps: the reductions' and GMRES's LDS without the sweeps' arrays.
struct SmallGmresLds : SmallRedLds, SmallGmFields {
  double loc[3][MPK_SMALL_GM_ROWS] __attribute__((aligned(16)));
  double bas[MPK_SMALL_GM_BAS][MPK_SMALL_GM_ROWS] __attribute__((aligned(16)));
};

template <int PCT>
__global__ void __launch_bounds__(MPK_WAVE_THREADS) k_gmres_small_lds(SmallKsp a) {
  __shared__ SmallGmresLds L;
  SmallRows<MPK_SMALL_GM_ROWS / MPK_WAVE_THREADS> rows;
  const bool fits = small_rows_load(a.c, rows);
  KspState *s = small_state_in(a.c.s, L, true);
  const i64 n = a.c.n, m = a.m;
  const int t = (int)threadIdx.x;
  double *X = a.c.X, *vv = a.vv, *TEMP = a.TEMP, *TMP = a.TMP;
  const double *jac = a.c.jac;
  auto spmv = [&](const double *x, double *y) {
    if (fits)
      small_spmv_rows(rows, n, x, y);
    else
      small_spmv(a.c, x, y);
  };
  auto pc = [&](const double *x, double *y) {  // small_pc's arithmetic, then its barrier
    for (i64 i = t; i < n; i += MPK_WAVE_THREADS) y[i] = PCT == PCT_JACOBI ? x[i] * jac[i] : x[i];
    __syncthreads();
  };
  i64 guess_zero = a.guess_zero;
  for (i64 cycle = 0;; cycle++) {
    // KSPInitialResidual(guess_zero, X, TEMP, TMP, vv, B)
    if (!guess_zero) {
      small_spmv(a.c, X, TEMP);
      small_copy(a.B, TMP, n);
      __syncthreads();
      small_axpy(TMP, -1.0, TEMP, n);
    } else {
      small_copy(a.B, TMP, n);
    }
    __syncthreads();
    if (a.side == PC_LEFT) {
      small_pc<PCT>(a.c, TMP, vv, L);
    } else {
      small_copy(TMP, vv, n);
      __syncthreads();
    }
    small_reduce<1>(RedSet<1>{{{vv, vv, RK_SUMSQ}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_GM_START, 0, L.res);
    __syncthreads();
    if (!small_flag(&s->stop)) small_ew(OpDivDev{vv, &s->res, aligned16(vv)}, n);
    __syncthreads();
    double *vk = L.loc[0], *vnew = L.loc[1], *wk = L.loc[2];
    for (i64 i = t; i < n; i += MPK_WAVE_THREADS) {
      vk[i] = vv[i];
      L.bas[0][i] = vv[i];
    }
    __syncthreads();
    for (i64 k = 0; k < m; k++) {
      if (small_flag(&s->inner_stop)) break;
      if (a.side == PC_RIGHT) {  // small_ba
        pc(vk, wk);
        spmv(wk, vnew);
      } else {
        spmv(vk, wk);
        __syncthreads();
        pc(wk, vnew);
      }
      __syncthreads();
      small_orthogonalize<MPK_SMALL_GM_BAS>(a, s, k, vnew, L, &L.bas[0][0]);
      small_reduce<1>(RedSet<1>{{{vnew, vnew, RK_SUMSQ}}}, n, L);
      if (t == 0) ksp_epilogue(s, EPI_GM_STEP, 0, L.res);
      __syncthreads();
      if (!small_flag(&s->inner_stop)) small_ew(OpDivDev{vnew, &s->tt, true}, n);
      __syncthreads();
      double *g = vv + (k + 1) * n;
      for (i64 i = t; i < n; i += MPK_WAVE_THREADS) g[i] = vnew[i];
      if (k + 1 < MPK_SMALL_GM_BAS)  // uniform
        for (i64 i = t; i < n; i += MPK_WAVE_THREADS) L.bas[k + 1][i] = vnew[i];
      __syncthreads();
      double *q = vk;
      vk = vnew;
      vnew = q;
    }
    // KSPGMRESBuildSoln
    if (t == 0 && !small_flag(&s->stop)) gmres_backsub(s, 0);
    __syncthreads();
    if (!small_flag(&s->skip_build)) {
      small_ew(OpBuild{TEMP, vv, n, s, aligned16(TEMP) && aligned16(vv) && n % 2 == 0}, n);
      __syncthreads();
      // KSPUnwindPreconditioner(TEMP, TMP)
      small_copy(TEMP, TMP, n);
      __syncthreads();
      if (a.side == PC_RIGHT) {
        small_pc<PCT>(a.c, TEMP, TMP, L);
        small_copy(TMP, TEMP, n);
        __syncthreads();
      }
      small_axpy(X, 1.0, TEMP, n);
      __syncthreads();
    }
    if (t == 0) gmres_cycle_end(s, 0);
    __syncthreads();
    if (small_flag(&s->stop)) break;
    guess_zero = 0;
    if (cycle > a.c.max_it + 2) break;
  }
  small_state_out(a.c.s, L, true);
}

// KSPSolve_FGMRES's cycles after the first residual (mp_ksp_solve_fgmres's
// launch sequence).
template <int PCT>
__global__ void __launch_bounds__(MPK_WAVE_THREADS) k_fgmres_small(SmallKsp a) {
  __shared__ SmallKspLds L;
  KspState *s = small_state_in(a.c.s, L, true);
  const i64 n = a.c.n, m = a.m;
  const int t = (int)threadIdx.x;
  double *X = a.c.X, *vv = a.vv, *TEMP = a.TEMP;
  for (i64 cycle = 0;; cycle++) {
    small_reduce<1>(RedSet<1>{{{vv, vv, RK_SUMSQ}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_FGM_START, 0, L.res);
    __syncthreads();
    if (!small_flag(&s->inner_stop)) small_ew(OpScaleRecipDev{vv, &s->res, nullptr, aligned16(vv)}, n);
    __syncthreads();
    for (i64 k = 0; k < m; k++) {
      if (small_flag(&s->inner_stop)) break;
      double *vk = vv + k * n, *vnew = vv + (k + 1) * n, *pk = a.prevec + k * n;
      small_pc<PCT>(a.c, vk, pk, L);
      small_spmv(a.c, pk, vnew);
      __syncthreads();
      small_orthogonalize(a, s, k, vnew, L);
      small_reduce<1>(RedSet<1>{{{vnew, vnew, RK_SUMSQ}}}, n, L);
      if (t == 0) ksp_epilogue(s, EPI_FGM_STEP, 0, L.res);
      __syncthreads();
      if (!small_flag(&s->inner_stop)) small_ew(OpScaleRecipDev{vnew, &s->tt, &s->hapend, aligned16(vnew)}, n);
      __syncthreads();
    }
    if (t == 0 && !small_flag(&s->stop)) gmres_backsub(s, 1);
    __syncthreads();
    if (!small_flag(&s->skip_build)) {
      small_ew(OpBuild{TEMP, a.prevec, n, s, aligned16(TEMP) && aligned16(a.prevec) && n % 2 == 0}, n);
      __syncthreads();
      small_axpy(X, 1.0, TEMP, n);
      __syncthreads();
    }
    if (t == 0) gmres_cycle_end(s, 1);
    __syncthreads();
    if (small_flag(&s->stop)) break;
    // the next cycle's residual: vv = -1 (A x) + b
    small_spmv(a.c, X, TEMP);
    __syncthreads();
    for (i64 i = t; i < n; i += MPK_WAVE_THREADS) vv[i] = fma(-1.0, TEMP[i], a.B[i]);
    __syncthreads();
    if (cycle > a.c.max_it + 2) break;
  }
  small_state_out(a.c.s, L, true);
}

// k_fgmres_small in the form of k_gmres_small_lds: v_k, the preconditioned
// vector and the new vector in LDS, each preconditioned vector written to
// prevec and each scaled new vector to vv (BuildSoln and the Gram-Schmidt
// sums read them there).
template <int PCT>
__global__ void __launch_bounds__(MPK_WAVE_THREADS) k_fgmres_small_lds(SmallKsp a) {
  __shared__ SmallGmresLds L;
  SmallRows<MPK_SMALL_GM_ROWS / MPK_WAVE_THREADS> rows;
  const bool fits = small_rows_load(a.c, rows);
  KspState *s = small_state_in(a.c.s, L, true);
  const i64 n = a.c.n, m = a.m;
  const int t = (int)threadIdx.x;
  double *X = a.c.X, *vv = a.vv, *TEMP = a.TEMP;
  const double *jac = a.c.jac;
  for (i64 cycle = 0;; cycle++) {
    small_reduce<1>(RedSet<1>{{{vv, vv, RK_SUMSQ}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_FGM_START, 0, L.res);
    __syncthreads();
    if (!small_flag(&s->inner_stop)) small_ew(OpScaleRecipDev{vv, &s->res, nullptr, aligned16(vv)}, n);
    __syncthreads();
    double *vk = L.loc[0], *vnew = L.loc[1], *pk = L.loc[2];
    for (i64 i = t; i < n; i += MPK_WAVE_THREADS) {
      vk[i] = vv[i];
      L.bas[0][i] = vv[i];
    }
    __syncthreads();
    for (i64 k = 0; k < m; k++) {
      if (small_flag(&s->inner_stop)) break;
      double *gp = a.prevec + k * n;
      for (i64 i = t; i < n; i += MPK_WAVE_THREADS) {  // small_pc
        const double z = PCT == PCT_JACOBI ? vk[i] * jac[i] : vk[i];
        pk[i] = z;
        gp[i] = z;
      }
      __syncthreads();
      if (fits)
        small_spmv_rows(rows, n, pk, vnew);
      else
        small_spmv(a.c, pk, vnew);
      __syncthreads();
      small_orthogonalize<MPK_SMALL_GM_BAS>(a, s, k, vnew, L, &L.bas[0][0]);
      small_reduce<1>(RedSet<1>{{{vnew, vnew, RK_SUMSQ}}}, n, L);
      if (t == 0) ksp_epilogue(s, EPI_FGM_STEP, 0, L.res);
      __syncthreads();
      if (!small_flag(&s->inner_stop)) small_ew(OpScaleRecipDev{vnew, &s->tt, &s->hapend, true}, n);
      __syncthreads();
      double *g = vv + (k + 1) * n;
      for (i64 i = t; i < n; i += MPK_WAVE_THREADS) g[i] = vnew[i];
      if (k + 1 < MPK_SMALL_GM_BAS)  // uniform
        for (i64 i = t; i < n; i += MPK_WAVE_THREADS) L.bas[k + 1][i] = vnew[i];
      __syncthreads();
      double *q = vk;
      vk = vnew;
      vnew = q;
    }
    if (t == 0 && !small_flag(&s->stop)) gmres_backsub(s, 1);
    __syncthreads();
    if (!small_flag(&s->skip_build)) {
      small_ew(OpBuild{TEMP, a.prevec, n, s, aligned16(TEMP) && aligned16(a.prevec) && n % 2 == 0}, n);
      __syncthreads();
      small_axpy(X, 1.0, TEMP, n);
      __syncthreads();
    }
    if (t == 0) gmres_cycle_end(s, 1);
    __syncthreads();
    if (small_flag(&s->stop)) break;
    // the next cycle's residual: vv = -1 (A x) + b
    small_spmv(a.c, X, TEMP);
    __syncthreads();
    for (i64 i = t; i < n; i += MPK_WAVE_THREADS) vv[i] = fma(-1.0, TEMP[i], a.B[i]);
    __syncthreads();
    if (cycle > a.c.max_it + 2) break;
  }
  small_state_out(a.c.s, L, true);
}

// KSPSolve_BCGS's iterations (mp_ksp_solve_bcgs's loop).
template <int PCT>
__global__ void __launch_bounds__(MPK_WAVE_THREADS) k_bcgs_small(SmallKsp a) {
  __shared__ SmallKspLds L;
  KspState *s = small_state_in(a.c.s, L, false);
  const i64 n = a.c.n;
  const int t = (int)threadIdx.x;
  double *X = a.c.X, *R = a.R, *RP = a.RP, *V = a.V, *T = a.T, *Sv = a.Sv, *P = a.P;
  for (i64 it = 0; it <= a.c.max_it + 1; it++) {
    if (small_flag(&s->stop)) break;
    small_reduce<1>(RedSet<1>{{{R, RP, RK_DOT}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_BCGS_RHO, 0, L.res);
    __syncthreads();
    small_ew(OpBcgsP{P, R, V, s, aligned16(P) && aligned16(R) && aligned16(V)}, n);
    __syncthreads();
    small_ba<PCT>(a, P, V, T, L);
    small_reduce<1>(RedSet<1>{{{V, RP, RK_DOT}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_BCGS_D1, 0, L.res);
    __syncthreads();
    if (small_flag(&s->stop)) break;
    small_ew(OpWaxpyDev{Sv, V, R, &s->alpha, aligned16(Sv) && aligned16(V) && aligned16(R)}, n);
    __syncthreads();
    small_ba<PCT>(a, Sv, T, R, L);
    small_reduce<3>(RedSet<3>{{{Sv, T, RK_DOT}, {T, T, RK_SUMSQ}, {Sv, Sv, RK_SUMSQ}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_BCGS_D2, 0, L.res);
    __syncthreads();
    if (small_flag(&s->stop)) break;
    small_ew(OpBcgsX{X, P, Sv, s, aligned16(X) && aligned16(P) && aligned16(Sv)}, n);
    small_ew(OpWaxpyDev{R, T, Sv, &s->omega, aligned16(R) && aligned16(T) && aligned16(Sv)}, n);
    __syncthreads();
    small_reduce<1>(RedSet<1>{{{R, R, RK_SUMSQ}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_BCGS_DP, 0, L.res);
    __syncthreads();
  }
  small_state_out(a.c.s, L, false);
}

// k_bcgs_small with its vectors in LDS (systems of at most
// MPK_SMALL_LDS_ROWS rows without sweeps) and the matrix rows in registers,
// as k_cg_small_lds: the same steps on LDS copies of X, R, RP, V, T, S and P.
struct SmallBcgsLds {
  double vec[7][MPK_SMALL_LDS_ROWS] __attribute__((aligned(16)));
  double wsum[8][MPK_SMALL_TILES][4];
  double fsum[8][4];
  double res[8];
  KspState ks;
};

template <int PCT>
__global__ void __launch_bounds__(MPK_WAVE_THREADS) k_bcgs_small_lds(SmallKsp g) {
  __shared__ SmallBcgsLds L;
  const i64 n = g.c.n;
  const int t = (int)threadIdx.x;
  SmallRows<> rows;
  const bool fits = small_rows_load(g.c, rows);
  double *gv[7] = {g.c.X, g.R, g.RP, g.V, g.T, g.Sv, g.P};
  for (int k = 0; k < 7; k++)
    for (i64 i = t; i < n; i += MPK_WAVE_THREADS) L.vec[k][i] = gv[k][i];
  if (t == 0) L.ks = *g.c.s;
  __syncthreads();
  KspState *s = &L.ks;
  double *X = L.vec[0], *R = L.vec[1], *RP = L.vec[2], *V = L.vec[3], *T = L.vec[4], *Sv = L.vec[5], *P = L.vec[6];
  const double *jac = g.c.jac;
  auto spmv = [&](const double *x, double *y) {
    if (fits)
      small_spmv_rows(rows, n, x, y);
    else
      small_spmv(g.c, x, y);
  };
  auto pc = [&](const double *x, double *y) {  // small_pc's arithmetic, then its barrier
    for (i64 i = t; i < n; i += MPK_WAVE_THREADS) y[i] = PCT == PCT_JACOBI ? x[i] * jac[i] : x[i];
    __syncthreads();
  };
  auto ba = [&](const double *x, double *y, double *w) {  // small_ba
    if (g.side == PC_RIGHT) {
      pc(x, w);
      spmv(w, y);
    } else {
      spmv(x, w);
      __syncthreads();
      pc(w, y);
    }
    __syncthreads();
  };
  for (i64 it = 0; it <= g.c.max_it + 1; it++) {
    if (small_flag(&s->stop)) break;
    small_reduce<1>(RedSet<1>{{{R, RP, RK_DOT}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_BCGS_RHO, 0, L.res);
    __syncthreads();
    small_ew(OpBcgsP{P, R, V, s, true}, n);
    __syncthreads();
    ba(P, V, T);
    small_reduce<1>(RedSet<1>{{{V, RP, RK_DOT}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_BCGS_D1, 0, L.res);
    __syncthreads();
    if (small_flag(&s->stop)) break;
    small_ew(OpWaxpyDev{Sv, V, R, &s->alpha, true}, n);
    __syncthreads();
    ba(Sv, T, R);
    small_reduce<3>(RedSet<3>{{{Sv, T, RK_DOT}, {T, T, RK_SUMSQ}, {Sv, Sv, RK_SUMSQ}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_BCGS_D2, 0, L.res);
    __syncthreads();
    if (small_flag(&s->stop)) break;
    small_ew(OpBcgsX{X, P, Sv, s, true}, n);
    small_ew(OpWaxpyDev{R, T, Sv, &s->omega, true}, n);
    __syncthreads();
    small_reduce<1>(RedSet<1>{{{R, R, RK_SUMSQ}}}, n, L);
    if (t == 0) ksp_epilogue(s, EPI_BCGS_DP, 0, L.res);
    __syncthreads();
  }
  __syncthreads();
  for (int k = 0; k < 7; k++)
    for (i64 i = t; i < n; i += MPK_WAVE_THREADS) gv[k][i] = L.vec[k][i];
  if (t == 0) *g.c.s = L.ks;
}

// Whether a solve goes to the one-workgroup drivers: the tile-tree order (not
// MPK_REDUCTION=ref), n <= MPK_SMALL_ROWS (MPK_SMALL_ROWS=k lowers it; 0
// switches the drivers off), and a PC with two wave plans, or a PC without
// sweeps and n <= MPK_SMALL_PLAIN (past it the general path's fused
// launches, spread over the chip, are faster).
static std::atomic<int64_t> g_small_solves{0};
static bool small_ok(const PCView &pc, i64 n, bool ref, int *f0, int *f1) {
  const char *e = std::getenv("MPK_SMALL_ROWS");
  const i64 lim = e ? std::strtoll(e, nullptr, 10) : MPK_SMALL_ROWS;
  if (ref || n <= 0 || n > lim || n > MPK_SMALL_ROWS) return false;
  *f0 = *f1 = -1;
  if (pc.type == PCT_NONE || pc.type == PCT_JACOBI) return n <= MPK_SMALL_PLAIN;
  const char *plan = pc.type == PCT_ILU ? pc.ilu.plan : pc.type == PCT_ICC ? pc.icc.plan : nullptr;
  if (!plan) return false;
  *f0 = plan_wave_form(plan, 0);
  *f1 = plan_wave_form(plan, 1);
  return *f0 >= 0 && *f1 >= 0;
}

// Whether a GMRES / FGMRES (gmres) or BiCGStab solve goes to its
// one-workgroup driver: small_ok's rules, at most MPK_SMALL_PLAIN rows (past
// them one workgroup's share of the chip's bandwidth makes the Gram-Schmidt
// sums slower than the general path's launches), the left or right side, and
// for GMRES a restart of at most MPK_SMALL_M.
static bool small_ksp_ok(const PCView &pc, const KspState &h, i64 n, bool gmres, int *f0, int *f1) {
  if (n > MPK_SMALL_PLAIN || h.side == PC_SYMMETRIC) return false;
  if (gmres && (h.m < 1 || h.m > MPK_SMALL_M)) return false;
  return small_ok(pc, n, red_ref(), f0, f1);
}

// The matrix, PC and state part of a one-workgroup solve (the sweeps'
// scratch for ILU / IC); false when that scratch cannot be had.
static bool small_args(SmallCG &a, i64 n, i64 max_it, const DevCSR &A, const PCView &pc, KspState *d, double *X,
                       int f0, int f1) {
  a = SmallCG{n, max_it, A.rp, A.ci, A.v, X, nullptr, nullptr, nullptr, nullptr, d, pc.jac_diag, pc.ilu, pc.icc,
              f0, f1, nullptr, nullptr};
  if (pc.type == PCT_ILU || pc.type == PCT_ICC) {
    a.S = pc_sweep_scratch(n);
    if (!a.S) return false;
    a.st = static_cast<Staged *>(pc_staged_rows(a.S, n));
  }
  return true;
}

#define MPK_SMALL_LAUNCH(kern, pct, arg)                           \
  do {                                                             \
    switch (pct) {                                                 \
      case PCT_JACOBI:                                             \
        MPK_LAUNCH(kern<PCT_JACOBI>, 1, MPK_WAVE_THREADS, arg);    \
        break;                                                     \
      case PCT_ILU:                                                \
        MPK_LAUNCH(kern<PCT_ILU>, 1, MPK_WAVE_THREADS, arg);       \
        break;                                                     \
      case PCT_ICC:                                                \
        MPK_LAUNCH(kern<PCT_ICC>, 1, MPK_WAVE_THREADS, arg);       \
        break;                                                     \
      default:                                                     \
        MPK_LAUNCH(kern<PCT_NONE>, 1, MPK_WAVE_THREADS, arg);      \
    }                                                              \
  } while (0)

// Tiles up to which CG + Jacobi folds its reductions' partials in every block
// of the kernels that use them: 2048 (4,194,304 rows); MPK_CG_FOLD_TILES=k
// overrides it (read at each solve; tuning and tests).
static unsigned cg_fold_tiles() {
  const char *e = std::getenv("MPK_CG_FOLD_TILES");
  return e ? (unsigned)std::strtoul(e, nullptr, 10) : 2048u;
}

// Iterations queued between two reads of the stop flag.
static int chunk_for(i64 n) { return n < 65536 ? 64 : (n < 1048576 ? 16 : 8); }

}  // namespace mpk

using namespace mpk;

#define P64(a) reinterpret_cast<const int64_t *>(a)

extern "C" {

int64_t mpk_small_solves(void) { return g_small_solves.load(); }

// average ms of `reps` launches of k_cg_wdot (w = A p over the sliced copy,
// with the tile partials of p.w), HIP events on the library stream.  The
// kernel reads its stop flag with its first loads, so it gets a device flag
// holding 0 (kept for the process).
double mpk_time_cg_wdot(int64_t p, int64_t w, int64_t irow, int64_t nrow, int64_t soff, int64_t sv, int64_t sc,
                        int64_t reps) {
  Ctx *c = ctx();
  if (!c || reps <= 0 || nrow <= 0 || !soff) return -1.0;
  static std::mutex mu;
  static int64_t *no_stop = nullptr;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (!no_stop) {
      int64_t *f = nullptr;
      HIPCHK(hipMalloc(&f, sizeof(int64_t)));
      if (failed() || !f) return -1.0;
      HIPCHK(hipMemset(f, 0, sizeof(int64_t)));
      HIPCHK(hipDeviceSynchronize());
      if (failed()) return -1.0;
      no_stop = f;
    }
  }
  const unsigned T = ntiles(nrow);
  double *Pp = partials(T);
  if (!Pp) return -1.0;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int64_t r = 0; r < reps; r++) {
    if (sell_nt(nrow))
      MPK_LAUNCH(k_cg_wdot<true>, T, 4 * MPK_NT, (i64)nrow, (const int *)irow, (const int *)soff,
                 (const double *)sv, (const int *)sc, (const double *)p, (double *)w, Pp, (const int64_t *)no_stop);
    else
      MPK_LAUNCH(k_cg_wdot<false>, T, 4 * MPK_NT, (i64)nrow, (const int *)irow, (const int *)soff,
                 (const double *)sv, (const int *)sc, (const double *)p, (double *)w, Pp, (const int64_t *)no_stop);
  }
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  return failed() ? -1.0 : (double)ms / (double)reps;
}

// ksp.mojo:271-389 KSPSolve_CG
void mp_ksp_solve_cg(int64_t n, int64_t irow, int64_t icol, int64_t data, int64_t pci, int64_t pcf,
                     int64_t pstmp, int64_t work, int64_t cfg, int64_t cfgi, int64_t b, int64_t x) {
  (void)pstmp;
  Solve S;
  if (!ctx() || !S.start(n, P64(irow), P64(icol), P64(data), P64(pci), (const double *)pcf,
                         (const double *)cfg, P64(cfgi)))
    return;
  double *X = (double *)x;
  const double *B = (const double *)b;
  double *R = (double *)work, *Z = R + n, *P = R + 2 * n, *W = Z;
  const int64_t *st = S.stop;
  if (!S.h.guess_zero) {
    spmv(S.A, X, R, 0, nullptr);
    vaypx(R, -1.0, B, n, nullptr);
  } else {
    vcopy(B, R, n);
  }
  S.pc.apply(R, Z, nullptr);
  S.reduce<3>(RedSet<3>{{{Z, Z, RK_SUMSQ}, {R, R, RK_SUMSQ}, {Z, R, RK_DOT}}}, EPI_CG_INIT, 0, nullptr);
  int f0, f1;
  if (small_ok(S.pc, n, red_ref(), &f0, &f1)) {
    // the iterations in one launch of one workgroup
    SmallCG a{n, S.h.max_it, S.A.rp, S.A.ci, S.A.v, X, R, Z, P, W, S.d, S.pc.jac_diag, S.pc.ilu, S.pc.icc, f0, f1,
              nullptr, nullptr};
    if (S.pc.type == PCT_ILU || S.pc.type == PCT_ICC) {
      a.S = pc_sweep_scratch(n);
      if (!a.S) return;
      a.st = static_cast<Staged *>(pc_staged_rows(a.S, n));
    }
    const bool lds = W == Z && n <= MPK_SMALL_LDS_ROWS;
    switch (S.pc.type) {
      case PCT_JACOBI:
        if (lds)
          MPK_LAUNCH(k_cg_small_lds<PCT_JACOBI>, 1, MPK_WAVE_THREADS, a);
        else
          MPK_LAUNCH(k_cg_small<PCT_JACOBI>, 1, MPK_WAVE_THREADS, a);
        break;
      case PCT_ILU:
        MPK_LAUNCH(k_cg_small<PCT_ILU>, 1, MPK_WAVE_THREADS, a);
        break;
      case PCT_ICC:
        MPK_LAUNCH(k_cg_small<PCT_ICC>, 1, MPK_WAVE_THREADS, a);
        break;
      default:
        if (lds)
          MPK_LAUNCH(k_cg_small_lds<PCT_NONE>, 1, MPK_WAVE_THREADS, a);
        else
          MPK_LAUNCH(k_cg_small<PCT_NONE>, 1, MPK_WAVE_THREADS, a);
    }
    g_small_solves++;
    S.finish_out((double *)cfg, (int64_t *)cfgi);
    return;
  }
  const int chunk = chunk_for(n);
  int slot = 0;
  bool pending = false;
  // CG + Jacobi over the sliced copy: three launches an iteration, the p.w
  // and norm steps folded into the kernels that use their results.  Every
  // block folding every partial costs blocks x partials loads, so past
  // cg_fold_tiles() tiles the steps run in finalizes instead (five launches).
  // (the fused kernels reduce in the tile-tree order: the reference's order
  // runs the product, the updates and the PC apply as separate launches)
  const bool ref = red_ref();
  const bool jac = !ref && S.A.sell_off && n > 0 && S.pc.type == PCT_JACOBI;
  const bool fused = jac && ntiles(n) <= cg_fold_tiles();
  const unsigned nt = fused ? ntiles(n) : 0;
  double *Pw = fused ? partials((size_t)4 * nt) : nullptr, *Pn = fused ? Pw + nt : nullptr;
  if (fused && !Pw) return;
  for (i64 queued = 0;; queued += chunk) {
    for (int c = 0; c < chunk; c++) {
      if (fused) {
        const i64 it = queued + c;
        const bool alp = aligned16(P) && aligned16(Z);
        if (it == 0)
          ew2(n, OpCgP{P, Z, S.d, alp}, st);
        else if (alp)
          MPK_LAUNCH(k_cg_pz<true>, nt, MPK_NT, n, P, (const double *)Z, (const double *)Pn, (int)nt, S.d, it - 1,
                     st);
        else
          MPK_LAUNCH(k_cg_pz<false>, nt, MPK_NT, n, P, (const double *)Z, (const double *)Pn, (int)nt, S.d, it - 1,
                     st);
        if (sell_nt(n))
          MPK_LAUNCH(k_cg_wdot<true>, nt, 4 * MPK_NT, n, S.A.rp, S.A.sell_off, S.A.sell_v, S.A.sell_c,
                     (const double *)P, W, Pw, st);
        else
          MPK_LAUNCH(k_cg_wdot<false>, nt, 4 * MPK_NT, n, S.A.rp, S.A.sell_off, S.A.sell_v, S.A.sell_c,
                     (const double *)P, W, Pw, st);
        const bool al = aligned16(X) && aligned16(R) && aligned16(P) && aligned16(W) && aligned16(Z) &&
                        aligned16(S.pc.jac_diag);
        if (al)
          MPK_LAUNCH(k_cg_xrz<true>, nt, MPK_NT, n, X, R, Z, (const double *)P, (const double *)W, S.pc.jac_diag,
                     (const double *)Pw, (int)nt, S.d, Pn, st);
        else
          MPK_LAUNCH(k_cg_xrz<false>, nt, MPK_NT, n, X, R, Z, (const double *)P, (const double *)W, S.pc.jac_diag,
                     (const double *)Pw, (int)nt, S.d, Pn, st);
        continue;
      }
      ew2(n, OpCgP{P, Z, S.d, aligned16(P) && aligned16(Z)}, st);
      if (jac) {
        // five launches: AYPX, the product with its partial, the p.w step,
        // the update with its partials, the norm step
        const unsigned T = ntiles(n);
        double *Pp = partials((size_t)3 * T);
        if (sell_nt(n))
          MPK_LAUNCH(k_cg_wdot<true>, T, 4 * MPK_NT, n, S.A.rp, S.A.sell_off, S.A.sell_v, S.A.sell_c,
                     (const double *)P, W, Pp, st);
        else
          MPK_LAUNCH(k_cg_wdot<false>, T, 4 * MPK_NT, n, S.A.rp, S.A.sell_off, S.A.sell_v, S.A.sell_c,
                     (const double *)P, W, Pp, st);
        MPK_LAUNCH(k_finalize<KspEpi>, 1, MPK_NT, (const double *)Pp, (int)T, 1, (double *)nullptr,
                   KspEpi{S.d, EPI_CG_DPI, 0}, st);
        const bool al = aligned16(X) && aligned16(R) && aligned16(P) && aligned16(W) && aligned16(Z) &&
                        aligned16(S.pc.jac_diag);
        if (al)
          MPK_LAUNCH(k_cg_xrz<true>, T, MPK_NT, n, X, R, Z, (const double *)P, (const double *)W, S.pc.jac_diag,
                     (const double *)nullptr, (int)T, S.d, Pp, st);
        else
          MPK_LAUNCH(k_cg_xrz<false>, T, MPK_NT, n, X, R, Z, (const double *)P, (const double *)W, S.pc.jac_diag,
                     (const double *)nullptr, (int)T, S.d, Pp, st);
        MPK_LAUNCH(k_finalize<KspEpi>, 1, MPK_NT, (const double *)Pp, (int)T, 3, (double *)nullptr,
                   KspEpi{S.d, EPI_CG_NORM, 0}, st);
        continue;
      }
      if (!ref && S.A.sell_off && n > 0) {
        // the product and its reduction in one pass (k_cg_wdot)
        const unsigned T = ntiles(n);
        double *Pp = partials(T);
        if (sell_nt(n))
          MPK_LAUNCH(k_cg_wdot<true>, T, 4 * MPK_NT, n, S.A.rp, S.A.sell_off, S.A.sell_v, S.A.sell_c,
                     (const double *)P, W, Pp, st);
        else
          MPK_LAUNCH(k_cg_wdot<false>, T, 4 * MPK_NT, n, S.A.rp, S.A.sell_off, S.A.sell_v, S.A.sell_c,
                     (const double *)P, W, Pp, st);
        MPK_LAUNCH(k_finalize<KspEpi>, 1, MPK_NT, (const double *)Pp, (int)T, 1, (double *)nullptr,
                   KspEpi{S.d, EPI_CG_DPI, 0}, st);
      } else {
        spmv(S.A, P, W, 0, st);
        S.reduce<1>(RedSet<1>{{{P, W, RK_DOT}}}, EPI_CG_DPI, 0, st);
      }
      const bool al = aligned16(X) && aligned16(R) && aligned16(P) && aligned16(W);
      ew2(n, OpCgXR{X, R, P, W, S.d, al}, st);
      S.pc.apply(R, Z, st);
      S.reduce<3>(RedSet<3>{{{Z, Z, RK_SUMSQ}, {R, R, RK_SUMSQ}, {Z, R, RK_DOT}}}, EPI_CG_NORM, 0, st);
    }
    S.mark(slot);
    if (pending && S.done(slot ^ 1)) break;
    pending = true;
    slot ^= 1;
    if (queued > S.h.max_it + 2 * chunk) break;  // the device has stopped by now
  }
  S.finish_out((double *)cfg, (int64_t *)cfgi);
}

// ksp.mojo:395-505 KSPSolve_BCGS
void mp_ksp_solve_bcgs(int64_t n, int64_t irow, int64_t icol, int64_t data, int64_t pci, int64_t pcf,
                       int64_t pstmp, int64_t work, int64_t cfg, int64_t cfgi, int64_t b, int64_t x) {
  (void)pstmp;
  Solve S;
  if (!ctx() || !S.start(n, P64(irow), P64(icol), P64(data), P64(pci), (const double *)pcf,
                         (const double *)cfg, P64(cfgi)))
    return;
  double *X = (double *)x, *WK = (double *)work;
  const double *B = (const double *)b;
  double *R = WK, *RP = WK + n, *V = WK + 2 * n, *T = WK + 3 * n, *Sv = WK + 4 * n, *P = WK + 5 * n,
         *GUESS = WK + 6 * n;
  const int64_t *st = S.stop;
  const bool has_guess = S.h.side == PC_RIGHT && !S.h.guess_zero;
  S.initial_residual(S.h.guess_zero, X, V, T, R, B, nullptr);
  if (has_guess) {
    vcopy(X, GUESS, n);
    vset(X, 0.0, n, nullptr);
  }
  S.reduce<1>(RedSet<1>{{{R, R, RK_SUMSQ}}}, EPI_BCGS_INIT, 0, nullptr);
  // the set up of the loop runs unless the initial test stopped the solve
  S.copy_if(R, RP, st);
  vset(P, 0.0, n, st);
  vset(V, 0.0, n, st);
  const int chunk = chunk_for(n);
  int slot = 0;
  bool pending = false;
  int f0, f1;
  SmallKsp a{};
  if (small_ksp_ok(S.pc, S.h, n, false, &f0, &f1) && small_args(a.c, n, S.h.max_it, S.A, S.pc, S.d, X, f0, f1)) {
    // the iterations in one launch of one workgroup
    a.side = (int)S.h.side;
    a.R = R, a.RP = RP, a.V = V, a.T = T, a.Sv = Sv, a.P = P;
    if (n <= MPK_SMALL_LDS_ROWS && S.pc.type == PCT_JACOBI)
      MPK_LAUNCH(k_bcgs_small_lds<PCT_JACOBI>, 1, MPK_WAVE_THREADS, a);
    else if (n <= MPK_SMALL_LDS_ROWS && S.pc.type == PCT_NONE)
      MPK_LAUNCH(k_bcgs_small_lds<PCT_NONE>, 1, MPK_WAVE_THREADS, a);
    else
      MPK_SMALL_LAUNCH(k_bcgs_small, S.pc.type, a);
    g_small_solves++;
  } else
  for (i64 queued = 0;; queued += chunk) {
    for (int c = 0; c < chunk; c++) {
      S.reduce<1>(RedSet<1>{{{R, RP, RK_DOT}}}, EPI_BCGS_RHO, 0, st);
      ew2(n, OpBcgsP{P, R, V, S.d, aligned16(P) && aligned16(R) && aligned16(V)}, st);
      S.ba_or_ab(P, V, T, st);
      S.reduce<1>(RedSet<1>{{{V, RP, RK_DOT}}}, EPI_BCGS_D1, 0, st);
      ew2(n, OpWaxpyDev{Sv, V, R, &S.d->alpha, aligned16(Sv) && aligned16(V) && aligned16(R)}, st);
      S.ba_or_ab(Sv, T, R, st);
      S.reduce<3>(RedSet<3>{{{Sv, T, RK_DOT}, {T, T, RK_SUMSQ}, {Sv, Sv, RK_SUMSQ}}}, EPI_BCGS_D2, 0, st);
      ew2(n, OpBcgsX{X, P, Sv, S.d, aligned16(X) && aligned16(P) && aligned16(Sv)}, st);
      ew2(n, OpWaxpyDev{R, T, Sv, &S.d->omega, aligned16(R) && aligned16(T) && aligned16(Sv)}, st);
      S.reduce<1>(RedSet<1>{{{R, R, RK_SUMSQ}}}, EPI_BCGS_DP, 0, st);
    }
    S.mark(slot);
    if (pending && S.done(slot ^ 1)) break;
    pending = true;
    slot ^= 1;
    if (queued > S.h.max_it + 2 * chunk) break;
  }
  sync();
  ew2(n, OpBcgsSpecial{X, P, S.d, aligned16(X) && aligned16(P)}, nullptr);
  KspState hs;
  copy_dtoh(&hs, S.d, sizeof hs);
  if (!hs.stopped_at_init) S.unwind(X, T, nullptr);
  if (has_guess) vaxpy(X, 1.0, GUESS, n, nullptr);
  S.finish_out((double *)cfg, (int64_t *)cfgi);
}

}  // extern "C"

namespace mpk {

// One GMRES / FGMRES Arnoldi orthogonalisation of v_new against v_0..v_it.
static void orthogonalize(Solve &S, double *vv, i64 it, double *vnew) {
  const i64 n = S.n;
  const int64_t *inner = S.inner;
  if (S.h.orthog == 1) {  // MGS
    for (i64 k = 0; k <= it; k++) {
      S.reduce<1>(RedSet<1>{{{vnew, vv + k * n, RK_DOT}}}, EPI_GM_MGS, (int)k, inner);
      ew2(n, OpAxpyDev{vnew, vv + k * n, S.h.LHH + k, aligned16(vnew) && aligned16(vv + k * n)}, inner);
    }
    return;
  }
  // CGS: MDot, negate, MAXPY, subtract; refine when asked
  const int nv = (int)(it + 1);
  const bool ref = red_ref();
  // the reference's order: one partial per sum, from k_ref_mdot
  unsigned T = ref ? 1u : ntiles(n);
  for (int pass = 0; pass < 2; pass++) {
    const int64_t *stp = pass == 0 ? inner : S.skipref;
    // past MPK_FIN_LDS sums the finalize folds into the nv doubles after the partials
    double *Pp = partials((size_t)nv * T + (size_t)nv);
    if (!Pp) return;
    double *fout = nv > MPK_FIN_LDS ? Pp + (size_t)nv * T : nullptr;
    if (ref) {
      MPK_LAUNCH(k_ref_mdot, (nv + 2) / 3, MPK_REF_NT, (const double *)vnew, (const double *)vv, (i64)n, nv,
                 (i64)n, Pp, stp);
    } else {
      for (int g0 = 0; g0 < nv; g0 += 8) {
        int g = nv - g0 < 8 ? nv - g0 : 8;
        MPK_LAUNCH(k_partials_mdot<8>, T, MPK_NT, (const double *)vnew, (const double *)(vv + (i64)g0 * n),
                   (i64)n, g, (i64)n, Pp + (size_t)g0 * T, stp);
      }
    }
    MPK_LAUNCH(k_finalize<KspEpi>, 1, MPK_NT, (const double *)Pp, (int)T, nv, fout,
               KspEpi{S.d, pass == 0 ? EPI_GM_CGS_A : EPI_GM_CGS_B, 0}, stp);
    vmaxpy_dev(vnew, vv, n, S.h.LHH, nv, 1.0, n, false, stp);
    if (pass == 0 && S.h.cgstype == 1)
      S.reduce<1>(RedSet<1>{{{vnew, vnew, RK_SUMSQ}}}, EPI_GM_CGS_NORM, 0, inner);
    if (pass == 0 && S.h.cgstype == 0) break;
  }
}

}  // namespace mpk

extern "C" {

// ksp.mojo:667-769 KSPSolve_GMRES
void mp_ksp_solve_gmres(int64_t n, int64_t irow, int64_t icol, int64_t data, int64_t pci, int64_t pcf,
                        int64_t pstmp, int64_t work, int64_t vv_, int64_t hs, int64_t cfg, int64_t cfgi, int64_t b,
                        int64_t x) {
  (void)pstmp, (void)hs;
  Solve S;
  if (!ctx() || !S.start(n, P64(irow), P64(icol), P64(data), P64(pci), (const double *)pcf,
                         (const double *)cfg, P64(cfgi)))
    return;
  double *X = (double *)x, *vv = (double *)vv_;
  const double *B = (const double *)b;
  double *TEMP = (double *)work, *TEMP_MATOP = TEMP + n;
  const i64 m = S.h.m;
  const int64_t *st = S.stop, *inner = S.inner;
  i64 guess_zero = S.h.guess_zero;
  int f0, f1;
  SmallKsp a{};
  if (small_ksp_ok(S.pc, S.h, n, true, &f0, &f1) && small_args(a.c, n, S.h.max_it, S.A, S.pc, S.d, X, f0, f1)) {
    // the restart cycles in one launch of one workgroup
    a.side = (int)S.h.side;
    a.m = m, a.orthog = S.h.orthog, a.cgstype = S.h.cgstype, a.guess_zero = guess_zero;
    a.vv = vv, a.TEMP = TEMP, a.TMP = TEMP_MATOP, a.B = B;
    if (n <= MPK_SMALL_GM_ROWS && S.pc.type == PCT_JACOBI)
      MPK_LAUNCH(k_gmres_small_lds<PCT_JACOBI>, 1, MPK_WAVE_THREADS, a);
    else if (n <= MPK_SMALL_GM_ROWS && S.pc.type == PCT_NONE)
      MPK_LAUNCH(k_gmres_small_lds<PCT_NONE>, 1, MPK_WAVE_THREADS, a);
    else
      MPK_SMALL_LAUNCH(k_gmres_small, S.pc.type, a);
    g_small_solves++;
    S.finish_out((double *)cfg, (int64_t *)cfgi);
    return;
  }
  const int chunk = chunk_for(n);
  int islot = 0;
  bool ipending = false;
  for (i64 cycle = 0;; cycle++) {
    S.initial_residual(guess_zero, X, TEMP, TEMP_MATOP, vv, B, st);
    S.reduce<1>(RedSet<1>{{{vv, vv, RK_SUMSQ}}}, EPI_GM_START, 0, st);
    ew2(n, OpDivDev{vv, &S.d->res, aligned16(vv)}, st);
    for (i64 s = 0; s < m; s++) {
      double *vs = vv + s * n, *vnew = vv + (s + 1) * n;
      S.ba_or_ab(vs, vnew, TEMP_MATOP, inner);
      orthogonalize(S, vv, s, vnew);
      S.reduce<1>(RedSet<1>{{{vnew, vnew, RK_SUMSQ}}}, EPI_GM_STEP, 0, inner);
      ew2(n, OpDivDev{vnew, &S.d->tt, aligned16(vnew)}, inner);
      if (S.inner_over(s, m, chunk, islot, ipending)) break;
    }
    // KSPGMRESBuildSoln(RS, HH, vv, TEMP, TEMP_MATOP, x, x, m, it - 1)
    MPK_LAUNCH(k_gmres_backsub, 1, 1, S.d, 0, st);
    ew2(n, OpBuild{TEMP, vv, n, S.d, aligned16(TEMP) && aligned16(vv) && n % 2 == 0}, S.skipbuild);
    S.unwind(TEMP, TEMP_MATOP, S.skipbuild);
    vaxpy(X, 1.0, TEMP, n, S.skipbuild);
    MPK_LAUNCH(k_gmres_cycle_end, 1, 1, S.d, 0);
    S.mark(0);
    if (S.done(0)) break;
    guess_zero = 0;
    if (cycle > S.h.max_it + 2) break;
  }
  S.finish_out((double *)cfg, (int64_t *)cfgi);
}

// ksp.mojo:775-868 KSPSolve_FGMRES
void mp_ksp_solve_fgmres(int64_t n, int64_t irow, int64_t icol, int64_t data, int64_t pci, int64_t pcf,
                         int64_t pstmp, int64_t work, int64_t vv_, int64_t prevec_, int64_t hs, int64_t cfg,
                         int64_t cfgi, int64_t b, int64_t x) {
  (void)pstmp, (void)hs;
  Solve S;
  if (!ctx() || !S.start(n, P64(irow), P64(icol), P64(data), P64(pci), (const double *)pcf,
                         (const double *)cfg, P64(cfgi)))
    return;
  double *X = (double *)x, *vv = (double *)vv_, *prevec = (double *)prevec_;
  const double *B = (const double *)b;
  double *TEMP = (double *)work;
  const i64 m = S.h.m;
  const int64_t *st = S.stop, *inner = S.inner;
  // _fgmres_residual: vv0 = -1 (A x) + b
  if (!S.h.guess_zero) {
    spmv(S.A, X, TEMP, 0, nullptr);
    vwaxpy(vv, -1.0, TEMP, B, n, nullptr);
  } else {
    vcopy(B, vv, n);
  }
  int f0, f1;
  SmallKsp a{};
  if (small_ksp_ok(S.pc, S.h, n, true, &f0, &f1) && small_args(a.c, n, S.h.max_it, S.A, S.pc, S.d, X, f0, f1)) {
    // the restart cycles in one launch of one workgroup
    a.side = (int)S.h.side, a.flexible = 1;
    a.m = m, a.orthog = S.h.orthog, a.cgstype = S.h.cgstype;
    a.vv = vv, a.prevec = prevec, a.TEMP = TEMP, a.B = B;
    if (n <= MPK_SMALL_GM_ROWS && S.pc.type == PCT_JACOBI)
      MPK_LAUNCH(k_fgmres_small_lds<PCT_JACOBI>, 1, MPK_WAVE_THREADS, a);
    else if (n <= MPK_SMALL_GM_ROWS && S.pc.type == PCT_NONE)
      MPK_LAUNCH(k_fgmres_small_lds<PCT_NONE>, 1, MPK_WAVE_THREADS, a);
    else
      MPK_SMALL_LAUNCH(k_fgmres_small, S.pc.type, a);
    g_small_solves++;
    S.finish_out((double *)cfg, (int64_t *)cfgi);
    return;
  }
  const int chunk = chunk_for(n);
  int islot = 0;
  bool ipending = false;
  for (i64 cycle = 0;; cycle++) {
    S.reduce<1>(RedSet<1>{{{vv, vv, RK_SUMSQ}}}, EPI_FGM_START, 0, st);
    ew2(n, OpScaleRecipDev{vv, &S.d->res, nullptr, aligned16(vv)}, inner);
    for (i64 s = 0; s < m; s++) {
      double *vs = vv + s * n, *vnew = vv + (s + 1) * n, *ps = prevec + s * n;
      S.pc.apply(vs, ps, inner);
      spmv(S.A, ps, vnew, 0, inner);
      orthogonalize(S, vv, s, vnew);
      S.reduce<1>(RedSet<1>{{{vnew, vnew, RK_SUMSQ}}}, EPI_FGM_STEP, 0, inner);
      ew2(n, OpScaleRecipDev{vnew, &S.d->tt, &S.d->hapend, aligned16(vnew)}, inner);
      if (S.inner_over(s, m, chunk, islot, ipending)) break;
    }
    MPK_LAUNCH(k_gmres_backsub, 1, 1, S.d, 1, st);
    ew2(n, OpBuild{TEMP, prevec, n, S.d, aligned16(TEMP) && aligned16(prevec) && n % 2 == 0}, S.skipbuild);
    vaxpy(X, 1.0, TEMP, n, S.skipbuild);
    MPK_LAUNCH(k_gmres_cycle_end, 1, 1, S.d, 1);
    S.mark(0);
    if (S.done(0)) break;
    spmv(S.A, X, TEMP, 0, nullptr);
    vwaxpy(vv, -1.0, TEMP, B, n, nullptr);
    if (cycle > S.h.max_it + 2) break;
  }
  S.finish_out((double *)cfg, (int64_t *)cfgi);
}

}  // extern "C"
