// mpk_progs.h -- the row programs of the preconditioners' sweeps (the
// preconditioner kernels and the one-workgroup Krylov drivers share them).
//
// Each provides, for sweep row i (mpk_sweep.h runs them):
//   stage(i) -> Staged, begin(i)/end(i), col(j)/coef(j), fin(i, acc, staged),
//   stale(k).
#pragma once

#include "mpk_sweep.h"

namespace mpk {

struct IluFwd {  // tmp_i = b[r_i] - sum_{L row i} ba_j tmp[bj_j]
  const int *bi, *bj, *r;
  const double *ba, *b;
  __device__ Staged stage(i64 i) const { return Staged{b[r[i]], 0.0, 0.0, 0}; }
  __device__ int begin(i64 i) const { return bi[i]; }
  __device__ int end(i64 i) const { return bi[i + 1]; }
  __device__ i64 col(int j) const { return bj[j]; }
  __device__ double coef(int j) const { return -ba[j]; }
  __device__ double fin(i64, double acc, const Staged &) const { return acc; }
  __device__ double stale(i64) const { return 0.0; }
};
struct IluBwd {  // sum = tmp_i - sum_{U row i} ba tmp; x[c_i] = sum * ba[bdiag_i]
  const int *bj, *bdiag, *c;
  const double *ba, *F;
  double *x;
  __device__ Staged stage(i64 i) const { return Staged{F[i], ba[bdiag[i]], 0.0, c[i]}; }
  __device__ int begin(i64 i) const { return bdiag[i + 1] + 1; }
  __device__ int end(i64 i) const { return bdiag[i]; }
  __device__ i64 col(int j) const { return bj[j]; }
  __device__ double coef(int j) const { return -ba[j]; }
  __device__ double fin(i64, double acc, const Staged &st) const {
    double r = acc * st.s1;
    x[st.o] = r;
    return r;
  }
  __device__ double stale(i64) const { return 0.0; }
};
// IC forward: the reference scatters s_k += ua_ik y_i for i ascending; here
// row k gathers over the transpose in the same ascending i, up to the first
// row after k (the rest arrive after row k is final: k_icc_late).
struct IccFwd {
  const int *trp, *trow, *tpos, *tmid, *udiag, *c;
  const double *ua, *b;
  double *F;  // scaled values for the backward sweep
  __device__ Staged stage(i64 k) const { return Staged{b[c[k]], ua[udiag[k]], 0.0, 0}; }
  __device__ int begin(i64 k) const { return trp[k]; }
  __device__ int end(i64 k) const { return tmid[k]; }
  __device__ i64 col(int j) const { return trow[j]; }
  __device__ double coef(int j) const { return ua[tpos[j]]; }
  __device__ double fin(i64 k, double acc, const Staged &st) const {
    F[k] = acc * st.s1;
    return acc;
  }
  __device__ double stale(i64) const { return 0.0; }
};
// IC backward (ported.mojo:902-952): rows n-2 down to 0 add their strict
// part; the last row is copied as it is, so a strict part it holds (late
// entries, columns below the diagonal) is not applied.
struct IccBwd {
  const int *ui, *uj, *udiag, *c;
  const double *ua, *F;
  double *x;
  i64 last;  // n - 1
  __device__ Staged stage(i64 i) const { return Staged{F[i], 0.0, 0.0, c[i]}; }
  __device__ int begin(i64 i) const { return ui[i]; }
  __device__ int end(i64 i) const { return i == last ? ui[i] : udiag[i]; }
  __device__ i64 col(int j) const { return uj[j]; }
  __device__ double coef(int j) const { return ua[j]; }
  __device__ double fin(i64, double acc, const Staged &st) const {
    x[st.o] = acc;
    return acc;
  }
  // a column below the diagonal: the backward loop has not reached it yet,
  // so the reference reads its forward value
  __device__ double stale(i64 k) const { return F[k]; }
};

// SOR, zero initial guess (ported.mojo:655-683)
struct SorFwd {
  const int *rp, *ci, *dmark;
  const double *a, *x, *idiag;
  double *t, *yout;  // yout: final y when no backward sweep follows (else null)
  const double *yin;
  __device__ Staged stage(i64 i) const { return Staged{x[i], idiag[i], 0.0, 0}; }
  __device__ int begin(i64 i) const { return rp[i]; }
  __device__ int end(i64 i) const { return dmark[i]; }
  __device__ i64 col(int j) const { return ci[j]; }
  __device__ double coef(int j) const { return -a[j]; }
  __device__ double fin(i64 i, double acc, const Staged &st) const {
    t[i] = acc;
    double r = acc * st.s1;
    if (yout) yout[i] = r;
    return r;
  }
  __device__ double stale(i64 k) const { return yin[k]; }
};
struct SorBwd {
  const int *rp, *ci, *dmark;
  const double *a, *src, *idiag, *yf;  // src: t after a forward sweep, else x
  double *y;
  double omega;
  int use_t;
  __device__ Staged stage(i64 i) const { return Staged{src[i], idiag[i], use_t ? yf[i] : 0.0, 0}; }
  __device__ int begin(i64 i) const { return dmark[i] + 1; }
  __device__ int end(i64 i) const { return rp[i + 1]; }
  __device__ i64 col(int j) const { return ci[j]; }
  __device__ double coef(int j) const { return -a[j]; }
  __device__ double fin(i64 i, double acc, const Staged &st) const {
    double r = use_t ? fma(1.0 - omega, st.s2, acc * st.s1) : acc * st.s1;
    y[i] = r;
    return r;
  }
  __device__ double stale(i64 k) const { return y[k]; }
};

// The late forward contributions: F[k] += ua_ik y_i for the rows i > k whose
// strict part holds k, in ascending i, after row k was scaled.  A row whose
// strict part holds its own column (trow == k, which the reference's staging
// can produce) adds to itself before the scaling overwrites it: no effect.
__device__ __forceinline__ void icc_late_row(i64 k, const int *trp, const int *trow, const int *tpos,
                                             const int *tmid, const double *ua, const double *Y, double *F) {
  int j0 = tmid[k], j1 = trp[k + 1];
  if (j0 == j1) return;
  double f = F[k];
  for (int j = j0; j < j1; j++)
    if (trow[j] != k) f = fma(ua[tpos[j]], Y[trow[j]], f);
  F[k] = f;
}

}  // namespace mpk
