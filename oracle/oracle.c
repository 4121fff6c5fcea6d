/*
 * oracle.c -- CPU restatement of the reference's sequential Krylov core.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: the tests, the
 * smoke check and the CPU-baseline leg of bench.py load it as a checker.  The
 * product (the HIP library under mojo-petsc4py_amd/) never links, loads or
 * calls it.
 *
 * Every function restates a routine of lee101/mojo-petsc4py (file:line cited
 * beside it) with the same branch structure, loop order and expression shapes.
 * Built with -ffp-contract=off; every fused multiply-add is an explicit fma()
 * call, placed where LLVM contracts the reference's expression (`s += a * b`
 * becomes fma(a, b, s); `a * b + c * d` becomes fma(a, b, c * d); `s -= a * b`
 * becomes fma(-a, b, s)).  With contraction the oracle reproduces the
 * reference's published README result bit for bit (its = 1,
 * ||b - Ax|| = 1.4432899320127035e-15, README.md:50-94); without it the last
 * bits differ.
 *
 * Reductions come in two orders, selected by or_set_reduction_mode():
 *   0 (REF)  the reference's SIMD order with SW = 4, BLK = 16
 *            (src/simdvec.mojo:332-473, src/simd.mojo:18,32-33): four
 *            4-lane accumulators over 16-element blocks, lane-wise
 *            a0+a1+a2+a3, SW-steps, scalar tail s, result s + reduce_add(acc);
 *            reduce_add over 4 lanes is the halving tree (a0+a2)+(a1+a3).
 *            _mdot_basis (src/ksp.mojo:956-1062) keeps its own order.
 *   1 (DEV)  the fixed tile-tree order the GPU kernels implement: tiles of
 *            2048 elements; thread t of 256 owns the pairs at 512*j + 2*t
 *            (j < 4); a pair is v0 + v1 (fma(x0, y0, x1*y1) for products);
 *            the thread sums its pairs as (p0+p1)+(p2+p3); each 64-lane wave
 *            sums its lanes by the halving tree u[i] += u[i+w], w = 32..1;
 *            the four wave sums combine as (W0+W2)+(W1+W3).  Tile partials
 *            P_k are folded by thread t as P_t + P_{t+256} + ... (ascending)
 *            and the 256 thread values go through the same block tree.
 *            Leaves past n are +0.0.
 * Everything that is not a reduction is identical in both modes.
 *
 * Index arrays are int64 here, as in the reference's Python layer
 * (python/mojopetsc4py/PETSc/Mat.py:127-129).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t I;

#define SW 4
#define BLK 16

/* per thread, so that solves in both orders can run side by side */
static _Thread_local int g_mode = 0;

void or_set_reduction_mode(I mode) { g_mode = (int)mode; }
I or_get_reduction_mode(void) { return g_mode; }

/* ===================================================== reductions, REF */

/* Mojo SIMD.reduce_add over 4 lanes: split into halves and add, twice. */
static double red4(const double a[4]) { return (a[0] + a[2]) + (a[1] + a[3]); }

/* simdvec.mojo:332-353 dot */
static double ref_dot(const double *x, const double *y, I n) {
  double a0[4] = {0, 0, 0, 0}, a1[4] = {0, 0, 0, 0}, a2[4] = {0, 0, 0, 0},
         a3[4] = {0, 0, 0, 0};
  I i = 0;
  while (i + BLK <= n) {
    for (int l = 0; l < SW; l++) {
      a0[l] = fma(x[i + l], y[i + l], a0[l]);
      a1[l] = fma(x[i + SW + l], y[i + SW + l], a1[l]);
      a2[l] = fma(x[i + 2 * SW + l], y[i + 2 * SW + l], a2[l]);
      a3[l] = fma(x[i + 3 * SW + l], y[i + 3 * SW + l], a3[l]);
    }
    i += BLK;
  }
  double acc[4];
  for (int l = 0; l < SW; l++) acc[l] = ((a0[l] + a1[l]) + a2[l]) + a3[l];
  while (i + SW <= n) {
    for (int l = 0; l < SW; l++) acc[l] = fma(x[i + l], y[i + l], acc[l]);
    i += SW;
  }
  double s = 0.0;
  while (i < n) {
    s = fma(x[i], y[i], s);
    i += 1;
  }
  return s + red4(acc);
}

/* kind: 0 sum, 1 abs-sum, 2 sum of squares -- simdvec.mojo:399-473 */
static double ref_unary(const double *x, I n, int kind) {
  double a[4][4];
  memset(a, 0, sizeof a);
  I i = 0;
  /* a sum of squares accumulates as fma(v, v, acc); the others add */
#define ACC(acc, v) ((acc) = kind == 0 ? (acc) + (v) : kind == 1 ? (acc) + fabs(v) : fma((v), (v), (acc)))
  while (i + BLK <= n) {
    for (int q = 0; q < 4; q++)
      for (int l = 0; l < SW; l++) ACC(a[q][l], x[i + q * SW + l]);
    i += BLK;
  }
  double acc[4];
  for (int l = 0; l < SW; l++) acc[l] = ((a[0][l] + a[1][l]) + a[2][l]) + a[3][l];
  while (i + SW <= n) {
    for (int l = 0; l < SW; l++) ACC(acc[l], x[i + l]);
    i += SW;
  }
  double s = 0.0;
  while (i < n) {
    ACC(s, x[i]);
    i += 1;
  }
#undef ACC
  return s + red4(acc);
}

/* simdvec.mojo:356-396 dot_norm2: (x.y, x.x) in one pass */
static void ref_dot_norm2(const double *x, const double *y, I n, double *d,
                          double *q) {
  double dd[4][4], qq[4][4];
  memset(dd, 0, sizeof dd);
  memset(qq, 0, sizeof qq);
  I i = 0;
  while (i + BLK <= n) {
    for (int b = 0; b < 4; b++)
      for (int l = 0; l < SW; l++) {
        double v = x[i + b * SW + l];
        dd[b][l] = fma(v, y[i + b * SW + l], dd[b][l]);
        qq[b][l] = fma(v, v, qq[b][l]);
      }
    i += BLK;
  }
  double dacc[4], nacc[4];
  for (int l = 0; l < SW; l++) {
    dacc[l] = ((dd[0][l] + dd[1][l]) + dd[2][l]) + dd[3][l];
    nacc[l] = ((qq[0][l] + qq[1][l]) + qq[2][l]) + qq[3][l];
  }
  while (i + SW <= n) {
    for (int l = 0; l < SW; l++) {
      double v = x[i + l];
      dacc[l] = fma(v, y[i + l], dacc[l]);
      nacc[l] = fma(v, v, nacc[l]);
    }
    i += SW;
  }
  double ds = 0.0, qs = 0.0;
  while (i < n) {
    double v = x[i];
    ds = fma(v, y[i], ds);
    qs = fma(v, v, qs);
    i += 1;
  }
  *d = ds + red4(dacc);
  *q = qs + red4(nacc);
}

/* ksp.mojo:956-1062 _mdot_basis: res[j] = x . y_j, rows in groups of four
   sharing each x load; no SW-step loop; result reduce_add(acc) + tail. */
static void ref_mdot_basis(const double *x, const double *y, I nv, I n,
                           double *res) {
  I j = 0;
  while (j + 4 <= nv) {
    double a[4][4][4]; /* [row][acc][lane] */
    memset(a, 0, sizeof a);
    I i = 0;
    while (i + BLK <= n) {
      for (int r = 0; r < 4; r++) {
        const double *u = y + (j + r) * n;
        for (int q = 0; q < 4; q++)
          for (int l = 0; l < SW; l++)
            a[r][q][l] = fma(x[i + q * SW + l], u[i + q * SW + l], a[r][q][l]);
      }
      i += BLK;
    }
    double t[4] = {0, 0, 0, 0};
    while (i < n) {
      double xv = x[i];
      for (int r = 0; r < 4; r++) t[r] = fma(xv, y[(j + r) * n + i], t[r]);
      i += 1;
    }
    for (int r = 0; r < 4; r++) {
      double acc[4];
      for (int l = 0; l < SW; l++)
        acc[l] = ((a[r][0][l] + a[r][1][l]) + a[r][2][l]) + a[r][3][l];
      res[j + r] = red4(acc) + t[r];
    }
    j += 4;
  }
  while (j < nv) {
    const double *u = y + j * n;
    double b[4][4];
    memset(b, 0, sizeof b);
    I i = 0;
    while (i + BLK <= n) {
      for (int q = 0; q < 4; q++)
        for (int l = 0; l < SW; l++) b[q][l] = fma(x[i + q * SW + l], u[i + q * SW + l], b[q][l]);
      i += BLK;
    }
    double tb = 0.0;
    while (i < n) {
      tb = fma(x[i], u[i], tb);
      i += 1;
    }
    double acc[4];
    for (int l = 0; l < SW; l++) acc[l] = ((b[0][l] + b[1][l]) + b[2][l]) + b[3][l];
    res[j] = red4(acc) + tb;
    j += 1;
  }
}

/* ===================================================== reductions, DEV */

#define DEV_TILE 2048
#define DEV_NT 256

static double dev_block_tree(double *u) {
  double W[4];
  for (int w = 0; w < 4; w++) {
    double *l = u + 64 * w;
    for (int wd = 32; wd >= 1; wd >>= 1)
      for (int i = 0; i < wd; i++) l[i] = l[i] + l[i + wd];
    W[w] = l[0];
  }
  return (W[0] + W[2]) + (W[1] + W[3]);
}

/* The pair (e, e+1) owned by one thread in one 16-byte slot, combined as
   v0 + v1; for products the first one is fused: fma(x0, y0, x1*y1).
   kind: 0 sum, 1 abs-sum, 2 sum of squares, 3 dot(x, y).  Past n: +0.0. */
static double dev_pair(int kind, const double *x, const double *y, I e, I n) {
  double x0 = e < n ? x[e] : 0.0, x1 = e + 1 < n ? x[e + 1] : 0.0;
  switch (kind) {
  case 0: return x0 + x1;
  case 1: return fabs(x0) + fabs(x1);
  case 2: return fma(x0, x0, x1 * x1);
  default: {
    double y0 = e < n ? y[e] : 0.0, y1 = e + 1 < n ? y[e + 1] : 0.0;
    return fma(x0, y0, x1 * y1);
  }
  }
}

static double dev_reduce(int kind, const double *x, const double *y, I n) {
  I T = (n + DEV_TILE - 1) / DEV_TILE;
  if (T == 0) return 0.0;
  double *P = (double *)malloc(sizeof(double) * (size_t)T);
  double u[DEV_NT];
  for (I k = 0; k < T; k++) {
    I base = k * DEV_TILE;
    for (int t = 0; t < DEV_NT; t++) {
      double p[4];
      for (int j = 0; j < 4; j++) {
        I e = base + 512 * j + 2 * t;
        p[j] = dev_pair(kind, x, y, e, n);
      }
      u[t] = (p[0] + p[1]) + (p[2] + p[3]);
    }
    P[k] = dev_block_tree(u);
  }
  for (int t = 0; t < DEV_NT; t++) {
    double a = t < T ? P[t] : 0.0;
    for (I k = t + DEV_NT; k < T; k += DEV_NT) a = a + P[k];
    u[t] = a;
  }
  double r = dev_block_tree(u);
  free(P);
  return r;
}

/* ===================================================== reductions, dispatch */

static double r_dot(const double *x, const double *y, I n) {
  return g_mode ? dev_reduce(3, x, y, n) : ref_dot(x, y, n);
}
static double r_sumsq(const double *x, I n) {
  return g_mode ? dev_reduce(2, x, 0, n) : ref_unary(x, n, 2);
}
static double r_asum(const double *x, I n) {
  return g_mode ? dev_reduce(1, x, 0, n) : ref_unary(x, n, 1);
}
static double r_total(const double *x, I n) {
  return g_mode ? dev_reduce(0, x, 0, n) : ref_unary(x, n, 0);
}
static void r_dot_norm2(const double *x, const double *y, I n, double *d, double *q) {
  if (g_mode) {
    *d = dev_reduce(3, x, y, n);
    *q = dev_reduce(2, x, 0, n);
  } else {
    ref_dot_norm2(x, y, n, d, q);
  }
}
static void r_mdot_basis(const double *x, const double *y, I nv, I n, double *res) {
  if (g_mode) {
    for (I j = 0; j < nv; j++) res[j] = dev_reduce(3, x, y + j * n, n);
  } else {
    ref_mdot_basis(x, y, nv, n, res);
  }
}

/* exported raw reductions, for tests */
double or_red_dot(I x, I y, I n) { return r_dot((double *)x, (double *)y, n); }
double or_red_sumsq(I x, I n) { return r_sumsq((double *)x, n); }
double or_red_asum(I x, I n) { return r_asum((double *)x, n); }
double or_red_total(I x, I n) { return r_total((double *)x, n); }
void or_red_mdot(I x, I y, I nv, I n, I res) {
  r_mdot_basis((double *)x, (double *)y, nv, n, (double *)res);
}

/* ===================================================== Vec (ported.mojo) */

/* ported.mojo:75-98 pointwise family; simdvec.mojo:158-213 */
void or_vec_pointwise_mult(I w, I x, I y, I n) {
  double *W = (double *)w, *X = (double *)x, *Y = (double *)y;
  for (I i = 0; i < n; i++) W[i] = X[i] * Y[i];
}
void or_vec_pointwise_divide(I w, I x, I y, I n) {
  double *W = (double *)w, *X = (double *)x, *Y = (double *)y;
  for (I i = 0; i < n; i++) W[i] = X[i] / Y[i];
}
/* Mojo min/max on Float64: a < b ? a : b and a > b ? a : b */
static double mn(double a, double b) { return a < b ? a : b; }
static double mx(double a, double b) { return a > b ? a : b; }
void or_vec_pointwise_min(I w, I x, I y, I n) {
  double *W = (double *)w, *X = (double *)x, *Y = (double *)y;
  for (I i = 0; i < n; i++) W[i] = mn(X[i], Y[i]);
}
void or_vec_pointwise_max(I w, I x, I y, I n) {
  double *W = (double *)w, *X = (double *)x, *Y = (double *)y;
  for (I i = 0; i < n; i++) W[i] = mx(X[i], Y[i]);
}
void or_vec_pointwise_max_abs(I w, I x, I y, I n) {
  double *W = (double *)w, *X = (double *)x, *Y = (double *)y;
  for (I i = 0; i < n; i++) W[i] = mx(fabs(X[i]), fabs(Y[i]));
}

/* ported.mojo:100-105 VecNorm2 */
static double vec_norm2(const double *x, I n) {
  if (!n) return 0.0;
  return sqrt(r_sumsq(x, n));
}

/* simdvec.mojo:476-488 norm_infinity: first NaN adopted, scan stops */
static double norm_inf(const double *x, I n) {
  double m = 0.0;
  for (I i = 0; i < n; i++) {
    double t = fabs(x[i]);
    if (t > m || t != t) {
      m = t;
      if (t != t) break;
    }
  }
  return m;
}

/* ported.mojo:113-135 VecNorm_Seq */
void or_vec_norm(I x, I ntype, I z, I n) {
  const double *X = (const double *)x;
  double *Z = (double *)z;
  double z0 = 0.0, z1 = 0.0;
  if (n) {
    if (ntype == 1 || ntype == 2 || ntype == 4) {
      if (ntype == 4) {
        z0 = r_asum(X, n);
        z1 = vec_norm2(X, n);
      } else {
        z0 = vec_norm2(X, n);
      }
    } else if (ntype == 3) {
      z0 = norm_inf(X, n);
    } else if (ntype == 0) {
      z0 = r_asum(X, n);
    }
  }
  Z[0] = z0;
  if (ntype == 4) Z[1] = z1;
}

void or_vec_dot(I x, I y, I val, I n) { *(double *)val = r_dot((double *)x, (double *)y, n); }
void or_vec_tdot(I x, I y, I val, I n) { *(double *)val = r_dot((double *)y, (double *)x, n); }
void or_vec_dot_norm2(I x, I y, I sval, I rval, I n) {
  r_dot_norm2((double *)x, (double *)y, n, (double *)sval, (double *)rval);
}
void or_vec_sum(I x, I val, I n) { *(double *)val = r_total((double *)x, n); }
void or_vec_mean(I x, I val, I n) { *(double *)val = r_total((double *)x, n) / (double)n; }

/* ported.mojo:170-208 VecMinMax_Seq */
static void vec_minmax(const double *x, I *idx, double *z, double seed, int is_max, I n) {
  I j = -1;
  double mm = seed;
  if (n) {
    j = 0;
    mm = x[0];
    for (I i = 1; i < n; i++) {
      double t = x[i];
      if (is_max) {
        if (t > mm) { j = i; mm = t; }
      } else {
        if (t < mm) { j = i; mm = t; }
      }
    }
  }
  *z = mm;
  *idx = j;
}
void or_vec_max(I x, I idx, I z, I n) {
  vec_minmax((double *)x, (I *)idx, (double *)z, -1.7976931348623157e308, 1, n);
}
void or_vec_min(I x, I idx, I z, I n) {
  vec_minmax((double *)x, (I *)idx, (double *)z, 1.7976931348623157e308, 0, n);
}

/* ported.mojo:200-205 VecNormalize_Seq */
void or_vec_normalize(I x, I v, I n) {
  double *X = (double *)x;
  double z = vec_norm2(X, n);
  if (z > 0.0)
    for (I i = 0; i < n; i++) X[i] = X[i] / z;
  *(double *)v = z;
}

/* simdvec.mojo:226-244 reciprocal_inplace */
void or_vec_reciprocal(I x, I n) {
  double *X = (double *)x;
  for (I i = 0; i < n; i++) X[i] = X[i] == 0.0 ? 0.0 : 1.0 / X[i];
}
void or_vec_exp(I x, I n) { double *X = (double *)x; for (I i = 0; i < n; i++) X[i] = exp(X[i]); }
void or_vec_log(I x, I n) { double *X = (double *)x; for (I i = 0; i < n; i++) X[i] = log(X[i]); }
void or_vec_sqrtabs(I x, I n) { double *X = (double *)x; for (I i = 0; i < n; i++) X[i] = sqrt(fabs(X[i])); }
/* simdvec.mojo:247-258 chop_inplace: strict < tol */
void or_vec_chop(I x, I n, double tol) {
  double *X = (double *)x;
  for (I i = 0; i < n; i++)
    if (fabs(X[i]) < tol) X[i] = 0.0;
}
void or_vec_abs(I x, I n) { double *X = (double *)x; for (I i = 0; i < n; i++) X[i] = fabs(X[i]); }
void or_vec_conjugate(I x, I n) { (void)x; (void)n; }
void or_vec_copy(I x, I y, I n) { memmove((void *)y, (void *)x, sizeof(double) * (size_t)n); }
void or_vec_swap(I x, I y, I n) {
  double *X = (double *)x, *Y = (double *)y;
  for (I i = 0; i < n; i++) { double t = X[i]; X[i] = Y[i]; Y[i] = t; }
}
void or_vec_zero_entries(I x, I n) { double *X = (double *)x; for (I i = 0; i < n; i++) X[i] = 0.0; }
/* ported.mojo:251-258 VecSet_Seq */
void or_vec_set(I x, double alpha, I n) {
  double *X = (double *)x;
  if (alpha == 0.0) { or_vec_zero_entries(x, n); return; }
  for (I i = 0; i < n; i++) X[i] = alpha;
}
void or_vec_scale(I x, double alpha, I n) { double *X = (double *)x; for (I i = 0; i < n; i++) X[i] = X[i] * alpha; }
void or_vec_shift(I x, double alpha, I n) { double *X = (double *)x; for (I i = 0; i < n; i++) X[i] = X[i] + alpha; }

/* ported.mojo:278-282 VecAXPY_Seq -> simdvec.axpy: y + alpha*x */
void or_vec_axpy(I y, double alpha, I x, I n) {
  double *Y = (double *)y, *X = (double *)x;
  for (I i = 0; i < n; i++) Y[i] = fma(alpha, X[i], Y[i]);
}
/* ported.mojo:285-294 VecAYPX_Seq */
void or_vec_aypx(I y, double alpha, I x, I n) {
  double *Y = (double *)y, *X = (double *)x;
  if (alpha == 0.0) {
    or_vec_copy(x, y, n);
  } else if (alpha == 1.0) {
    or_vec_axpy(y, alpha, x, n);
  } else {
    for (I i = 0; i < n; i++) Y[i] = fma(alpha, Y[i], X[i]);
  }
}
/* ported.mojo:297-308 VecAXPBY_Seq */
void or_vec_axpby(I y, double alpha, double beta, I x, I n) {
  double *Y = (double *)y, *X = (double *)x;
  if (beta == 0.0) {
    for (I i = 0; i < n; i++) Y[i] = X[i] * alpha;
  } else if (alpha == 0.0) {
    or_vec_scale(y, beta, n);
  } else {
    for (I i = 0; i < n; i++) Y[i] = fma(alpha, X[i], beta * Y[i]);
  }
}
/* ported.mojo:311-315 VecAXPBYPCZ_Seq: (alpha x + beta y) + gamma z */
void or_vec_axpbypcz(I z, double alpha, double beta, double gamma, I x, I y, I n) {
  double *Z = (double *)z, *X = (double *)x, *Y = (double *)y;
  for (I i = 0; i < n; i++) Z[i] = fma(gamma, Z[i], fma(alpha, X[i], beta * Y[i]));
}
/* ported.mojo:318-320 VecWAXPY_Seq: alpha x + y */
void or_vec_waxpy(I w, double alpha, I x, I y, I n) {
  double *W = (double *)w, *X = (double *)x, *Y = (double *)y;
  for (I i = 0; i < n; i++) W[i] = fma(alpha, X[i], Y[i]);
}
/* ported.mojo:329-335 VecMAXPY_Seq: k sequential axpy sweeps */
void or_vec_maxpy(I y, I nv, I alpha, I x, I n) {
  const double *A = (const double *)alpha;
  const I *XS = (const I *)x;
  for (I k = 0; k < nv; k++) or_vec_axpy(y, A[k], XS[k], n);
}
/* simdvec.mojo:491-500 max_pointwise_divide */
void or_vec_max_pointwise_divide(I x, I y, I rval, I n) {
  const double *X = (const double *)x, *Y = (const double *)y;
  double m = 0.0;
  for (I i = 0; i < n; i++) {
    double yi = Y[i];
    double v = fabs(yi == 0.0 ? X[i] : X[i] / yi);
    m = v > m ? v : m;
  }
  *(double *)rval = m;
}

/* ===================================================== Mat (ported.mojo) */

/* ported.mojo:354-377 MatGetDiagonalMarkers_SeqAIJ */
I or_mat_get_diagonal_markers(I irow, I icol, I diag, I n) {
  const I *R = (const I *)irow, *C = (const I *)icol;
  I *D = (I *)diag;
  I dense = 1;
  for (I i = 0; i < n; i++) {
    I bl = R[i], be = R[i + 1];
    D[i] = be;
    for (I j = bl; j < be; j++)
      if (C[j] == i) { D[i] = j; break; }
    if (D[i] == be) dense = 0;
  }
  return dense;
}

/* ported.mojo:380-391 MatMult_SeqAIJ */
void or_mat_mult(I x, I y, I irow, I icol, I data, I nrow) {
  const double *X = (const double *)x, *A = (const double *)data;
  double *Y = (double *)y;
  const I *R = (const I *)irow, *C = (const I *)icol;
  for (I i = 0; i < nrow; i++) {
    double sum = 0.0;
    for (I j = R[i]; j < R[i + 1]; j++) sum = fma(A[j], X[C[j]], sum);
    Y[i] = sum;
  }
}
/* ported.mojo:394-405 MatMultAdd_SeqAIJ */
void or_mat_mult_add(I x, I y, I irow, I icol, I data, I nrow) {
  const double *X = (const double *)x, *A = (const double *)data;
  double *Y = (double *)y;
  const I *R = (const I *)irow, *C = (const I *)icol;
  for (I i = 0; i < nrow; i++) {
    double sum = Y[i];
    for (I j = R[i]; j < R[i + 1]; j++) sum = fma(A[j], X[C[j]], sum);
    Y[i] = sum;
  }
}
/* ported.mojo:408-431 MatMultTranspose_SeqAIJ */
void or_mat_mult_transpose(I x, I y, I irow, I icol, I data, I nrow, I ncols) {
  const double *X = (const double *)x, *A = (const double *)data;
  double *Y = (double *)y;
  const I *R = (const I *)irow, *C = (const I *)icol;
  for (I j = 0; j < ncols; j++) Y[j] = 0.0;
  for (I i = 0; i < nrow; i++) {
    double xi = X[i];
    for (I jb = R[i]; jb < R[i + 1]; jb++) {
      I c = C[jb];
      Y[c] = fma(A[jb], xi, Y[c]);
    }
  }
}
/* ported.mojo:434-447 MatGetDiagonal_SeqAIJ */
void or_mat_get_diagonal(I diag, I irow, I icol, I data, I nrow) {
  double *D = (double *)diag;
  const double *A = (const double *)data;
  const I *R = (const I *)irow, *C = (const I *)icol;
  for (I i = 0; i < nrow; i++) {
    double v = 0.0;
    for (I j = R[i]; j < R[i + 1]; j++)
      if (C[j] == i) { v = A[j]; break; }
    D[i] = v;
  }
}
void or_mat_scale(I data, I nnz, double alpha) {
  double *A = (double *)data;
  for (I j = 0; j < nnz; j++) A[j] = A[j] * alpha;
}
void or_mat_zero_entries(I data, I nnz) {
  double *A = (double *)data;
  for (I j = 0; j < nnz; j++) A[j] = 0.0;
}

/* ===================================================== PC (ported.mojo) */

/* ported.mojo:570-607 row reductions */
static void row_sum(double *d, const I *R, const double *A, I n) {
  for (I i = 0; i < n; i++) {
    double s = 0.0;
    for (I j = R[i]; j < R[i + 1]; j++) s += A[j];
    d[i] = s;
  }
}
static void row_sum_abs(double *d, const I *R, const double *A, I n) {
  for (I i = 0; i < n; i++) {
    double s = 0.0;
    for (I j = R[i]; j < R[i + 1]; j++) s += fabs(A[j]);
    d[i] = s;
  }
}
static void row_max_abs(double *d, const I *R, const double *A, I n) {
  for (I i = 0; i < n; i++) {
    double s = 0.0;
    for (I j = R[i]; j < R[i + 1]; j++) {
      double v = fabs(A[j]);
      s = v > s ? v : s;
    }
    d[i] = s;
  }
}
/* ported.mojo:487-499 _JacobiFill */
static void jacobi_fill(double *dst, I jtype, I arow, I acol, I adata, I n) {
  if (jtype == 0)
    or_mat_get_diagonal((I)dst, arow, acol, adata, n);
  else if (jtype == 1)
    row_max_abs(dst, (const I *)arow, (const double *)adata, n);
  else if (jtype == 2)
    row_sum_abs(dst, (const I *)arow, (const double *)adata, n);
  else
    row_sum(dst, (const I *)arow, (const double *)adata, n);
}
/* ported.mojo:502-541 PCSetUp_Jacobi */
void or_pc_jacobi_setup(I diag, I jtype, I useabs, I fixdiag, double scale, I irow,
                        I icol, I data, I nrow, I didx) {
  double *D = (double *)diag, *DX = (double *)didx;
  jacobi_fill(D, jtype, irow, icol, data, nrow);
  if (jtype == 2) {
    or_mat_get_diagonal(didx, irow, icol, data, nrow);
    for (I i = 0; i < nrow; i++)
      if (DX[i] < 0.0) D[i] = -D[i];
    if (scale != 1.0)
      for (I i = 0; i < nrow; i++) D[i] = fma(scale, D[i] - DX[i], DX[i]);
  }
  or_vec_reciprocal(diag, nrow);
  if (useabs) or_vec_abs(diag, nrow);
  if (fixdiag)
    for (I i = 0; i < nrow; i++)
      if (D[i] == 0.0) D[i] = 1.0;
}
/* ported.mojo:544-562 PCSetUp_Jacobi_Symmetric */
void or_pc_jacobi_setup_symmetric(I diagsqrt, I jtype, I irow, I icol, I data, I nrow) {
  double *D = (double *)diagsqrt;
  jacobi_fill(D, jtype, irow, icol, data, nrow);
  for (I i = 0; i < nrow; i++) {
    double d = D[i];
    if (d < 0.0)
      D[i] = 1.0 / sqrt(fabs(-d));
    else if (d > 0.0)
      D[i] = 1.0 / sqrt(fabs(d));
    else
      D[i] = 1.0;
  }
}
/* ported.mojo:565-567 PCApply_Jacobi */
void or_pc_jacobi_apply(I y, I x, I diag, I n) { or_vec_pointwise_mult(y, x, diag, n); }

/* ported.mojo:610-630 MatInvertDiagonalForSOR_SeqAIJ */
static void sor_invert_diag(double *idiag, double *mdiag, const I *dmark, const double *A,
                            double omega, double fshift, I n) {
  if (omega == 1.0 && fshift <= 0.0) {
    for (I i = 0; i < n; i++) {
      mdiag[i] = A[dmark[i]];
      idiag[i] = 1.0 / A[dmark[i]];
    }
  } else {
    for (I i = 0; i < n; i++) {
      mdiag[i] = A[dmark[i]];
      idiag[i] = omega / (fshift + A[dmark[i]]);
    }
  }
}
/* ported.mojo:633-748 PCApply_SOR (x = rhs, y = solution) */
static void pc_apply_sor(const double *x, double *y, const I *R, const I *C, const double *A,
                         const I *dmark, double *idiag, double *mdiag, double *t, double omega,
                         double fshift, I its, I lits, I flag, I n) {
  I it = its * lits;
  sor_invert_diag(idiag, mdiag, dmark, A, omega, fshift, n);
  int use_t = 0;
  if (flag & (1 | 4)) {
    for (I i = 0; i < n; i++) {
      double sum = x[i];
      for (I j = R[i]; j < dmark[i]; j++) sum = fma(-A[j], y[C[j]], sum);
      t[i] = sum;
      y[i] = sum * idiag[i];
    }
    use_t = 1;
  }
  if (flag & (2 | 8)) {
    for (I i = n - 1; i >= 0; i--) {
      double sum = (use_t ? t : x)[i];
      for (I j = dmark[i] + 1; j < R[i + 1]; j++) sum = fma(-A[j], y[C[j]], sum);
      if (use_t)
        y[i] = fma(1.0 - omega, y[i], sum * idiag[i]);
      else
        y[i] = sum * idiag[i];
    }
  }
  it -= 1;
  while (it > 0) {
    it -= 1;
    if (flag & (1 | 4)) {
      for (I i = 0; i < n; i++) {
        double sum = x[i];
        for (I j = R[i]; j < dmark[i]; j++) sum = fma(-A[j], y[C[j]], sum);
        t[i] = sum;
        for (I j = dmark[i] + 1; j < R[i + 1]; j++) sum = fma(-A[j], y[C[j]], sum);
        y[i] = fma(1.0 - omega, y[i], sum * idiag[i]);
      }
      use_t = 1;
    }
    if (flag & (2 | 8)) {
      for (I i = n - 1; i >= 0; i--) {
        double sum = (use_t ? t : x)[i];
        if (use_t) {
          for (I j = dmark[i] + 1; j < R[i + 1]; j++) sum = fma(-A[j], y[C[j]], sum);
          y[i] = fma(1.0 - omega, y[i], sum * idiag[i]);
        } else {
          for (I j = R[i]; j < R[i + 1]; j++) sum = fma(-A[j], y[C[j]], sum);
          y[i] = fma(1.0 - omega, y[i], fma(mdiag[i], y[i], sum) * idiag[i]);
        }
      }
    }
  }
}
/* capi.mojo:272-280 mp_pc_sor_apply: its = lits = 1 */
void or_pc_sor_apply(I x, I y, I irow, I icol, I data, I dmark, I idiag, I mdiag, I t,
                     double omega, double fshift, I flag, I nrow) {
  pc_apply_sor((const double *)x, (double *)y, (const I *)irow, (const I *)icol,
               (const double *)data, (const I *)dmark, (double *)idiag, (double *)mdiag,
               (double *)t, omega, fshift, 1, 1, flag, nrow);
}
void or_pc_sor_apply_its(I x, I y, I irow, I icol, I data, I dmark, I idiag, I mdiag, I t,
                         double omega, double fshift, I its, I lits, I flag, I nrow) {
  pc_apply_sor((const double *)x, (double *)y, (const I *)irow, (const I *)icol,
               (const double *)data, (const I *)dmark, (double *)idiag, (double *)mdiag,
               (double *)t, omega, fshift, its, lits, flag, nrow);
}

/* ------------------------------------------------ incomplete-LL machinery */

/* ported.mojo:759-786 LLInsertLocation; returns (nlnk + 1, entry) */
static void ll_insert(I assume_sorted, I k, I idx_start, I entry, I *nlnk, I *lnkdata, I *lnk) {
  I ld = *lnkdata;
  if (!assume_sorted && k != 0 && entry < ld) ld = idx_start;
  I location = ld;
  for (;;) {
    ld = lnk[location];
    if (entry > ld)
      location = ld;
    else
      break;
  }
  lnk[location] = entry;
  lnk[entry] = ld;
  *nlnk += 1;
  *lnkdata = entry;
}
/* ported.mojo:789-799 _BTLookupSet */
static I bt_lookup_set(I *bt, I entry) {
  I old = bt[entry];
  bt[entry] = 1;
  return old;
}
/* ported.mojo:802-818 IncompleteLLInit */
static I ll_init(I nidx, const I *idx, I idx_start, const I *perm, I *lnk, I *lnklvl, I *bt) {
  I nlnk = 0, lnkdata = idx_start;
  for (I k = 0; k < nidx; k++) {
    I entry = perm[idx[k]];
    if (bt_lookup_set(bt, entry) == 0) {
      ll_insert(0, k, idx_start, entry, &nlnk, &lnkdata, lnk);
      lnklvl[entry] = 0;
    }
  }
  return nlnk;
}
/* ported.mojo:821-844 IncompleteLLAdd */
static I ll_add(I nidx, const I *idx, double level, const I *idxlvl, I idx_start, I *lnk,
                I *lnklvl, I *bt, I prow_offset, I assume_sorted) {
  I nlnk = 0, lnkdata = idx_start;
  for (I k = 0; k < nidx; k++) {
    I incrlev = idxlvl[k] + prow_offset + 1;
    if ((double)incrlev <= level) {
      I entry = idx[k];
      if (bt_lookup_set(bt, entry) == 0) {
        ll_insert(assume_sorted, k, idx_start, entry, &nlnk, &lnkdata, lnk);
        lnklvl[entry] = incrlev;
      } else if (lnklvl[entry] > incrlev) {
        lnklvl[entry] = incrlev;
      }
    }
  }
  return nlnk;
}
/* ported.mojo:847-861 IncompleteLLClean */
static void ll_clean(I idx_start, I lnk_max, I nlnk, I *lnk, I *lnklvl, I *indices,
                     I *indiceslvl, I *bt) {
  I idx = idx_start;
  for (I j = 0; j < nlnk; j++) {
    idx = lnk[idx];
    indices[j] = idx;
    indiceslvl[j] = lnklvl[idx];
    lnklvl[idx] = -1;
    bt[idx] = 0;
  }
  lnk[idx_start] = lnk_max;
}

/* ported.mojo:864-899 ILUApplyFactor_SeqAIJ */
void or_pc_ilu_apply(I b, I x, I tmp, I bi, I bj, I ba, I bdiag, I r, I c, I nrow) {
  const double *B = (const double *)b, *BA = (const double *)ba;
  double *X = (double *)x, *T = (double *)tmp;
  const I *BI = (const I *)bi, *BJ = (const I *)bj, *BD = (const I *)bdiag;
  const I *Rr = (const I *)r, *Cc = (const I *)c;
  if (!nrow) return;
  T[0] = B[Rr[0]];
  for (I i = 1; i < nrow; i++) {
    double sum = B[Rr[i]];
    for (I j = BI[i]; j < BI[i + 1]; j++) sum = fma(-BA[j], T[BJ[j]], sum);
    T[i] = sum;
  }
  for (I i = nrow - 1; i >= 0; i--) {
    I v = BD[i + 1] + 1;
    I nz = BD[i] - BD[i + 1] - 1;
    double sum = T[i];
    for (I j = 0; j < nz; j++) sum = fma(-BA[v + j], T[BJ[v + j]], sum);
    sum = sum * BA[v + nz];
    T[i] = sum;
    X[Cc[i]] = sum;
  }
}

/* ported.mojo:902-952 ICCApplyFactor_SeqAIJ */
I or_pc_icc_apply(I b, I x, I ui, I uj, I ua, I udiag, I c, I scratch, I nrow) {
  const double *B = (const double *)b, *UA = (const double *)ua;
  double *X = (double *)x, *S = (double *)scratch;
  const I *UI = (const I *)ui, *UJ = (const I *)uj, *UD = (const I *)udiag, *Cc = (const I *)c;
  if (!nrow) return 0;
  for (I i = 0; i < nrow; i++) S[i] = B[Cc[i]];
  for (I i = 0; i < nrow; i++) {
    double xi = S[i];
    for (I j = UI[i]; j < UD[i]; j++) {
      I k = UJ[j];
      S[k] = fma(UA[j], xi, S[k]);
    }
    S[i] = xi * UA[UD[i]];
  }
  for (I i = nrow - 2; i >= 0; i--) {
    double xi = S[i];
    for (I j = UI[i]; j < UD[i]; j++) xi = fma(UA[j], S[UJ[j]], xi);
    S[i] = xi;
    X[Cc[i]] = xi;
  }
  X[Cc[nrow - 1]] = S[nrow - 1];
  return 0;
}

/* ported.mojo:957-1037 MatILUFactorNumeric_SeqAIJ */
I or_pc_ilu_numeric(I ba, I rtmp, I irow, I icol, I data, I bi, I bj, I bdiag, I r, I ic,
                    I nrow) {
  double *BA = (double *)ba, *RT = (double *)rtmp;
  const I *AR = (const I *)irow, *AC = (const I *)icol;
  const double *AD = (const double *)data;
  const I *BI = (const I *)bi, *BJ = (const I *)bj, *BD = (const I *)bdiag;
  const I *Rr = (const I *)r, *IC = (const I *)ic;
  for (I i = 0; i < nrow; i++) {
    I jend = BI[i + 1] - BI[i], jj = BI[i];
    for (I j = 0; j < jend; j++) RT[BJ[jj++]] = 0.0;
    jend = BD[i] - BD[i + 1];
    jj = BD[i + 1] + 1;
    for (I j = 0; j < jend; j++) RT[BJ[jj++]] = 0.0;
    jend = AR[Rr[i] + 1] - AR[Rr[i]];
    jj = AR[Rr[i]];
    for (I j = 0; j < jend; j++) {
      RT[IC[AC[jj]]] = AD[jj];
      jj++;
    }
    jj = BI[i];
    I row = BJ[jj];
    jj++;
    I nzL = BI[i + 1] - BI[i];
    for (I k = 0; k < nzL; k++) {
      double pc = RT[row];
      if (pc != 0.0) {
        double pv = BA[BD[row]];
        double multiplier = pc * pv;
        RT[row] = multiplier;
        I pj = BD[row + 1] + 1;
        I nz = BD[row] - BD[row + 1] - 1;
        for (I j = 0; j < nz; j++) {
          I cc = BJ[pj + j];
          RT[cc] = fma(-multiplier, BA[pj + j], RT[cc]);
        }
      }
      row = BJ[jj];
      jj++;
    }
    jend = BI[i + 1] - BI[i];
    jj = BI[i];
    for (I j = 0; j < jend; j++) {
      BA[jj] = RT[BJ[jj]];
      jj++;
    }
    jend = BD[i] - BD[i + 1] - 1;
    jj = BD[i + 1] + 1;
    for (I j = 0; j < jend; j++) {
      BA[jj] = RT[BJ[jj]];
      jj++;
    }
    BA[BD[i]] = 1.0 / RT[i];
  }
  return BD[0] + 1;
}

/* ported.mojo:1040-1124 MatICCFactorNumeric_SeqAIJ */
void or_pc_icc_numeric(I ua, I rtmp, I il, I c2r, I irow, I icol, I data, I ui, I uj, I udiag,
                       I r, I riip, I nrow) {
  double *BA = (double *)ua, *RT = (double *)rtmp;
  I *IL = (I *)il, *C2R = (I *)c2r;
  const I *AR = (const I *)irow, *AC = (const I *)icol;
  const double *AD = (const double *)data;
  const I *BI = (const I *)ui, *BJ = (const I *)uj, *BD = (const I *)udiag;
  const I *Rr = (const I *)r, *RI = (const I *)riip;
  for (I i = 0; i < nrow; i++) C2R[i] = nrow;
  if (nrow) IL[0] = 0;
  for (I k = 0; k < nrow; k++) {
    I jend = BI[k + 1] - BI[k], jj = BI[k];
    for (I j = 0; j < jend; j++) RT[BJ[jj++]] = 0.0;
    I bval = BI[k];
    for (I j = AR[Rr[k]]; j < AR[Rr[k] + 1]; j++) {
      I col = RI[AC[j]];
      if (col >= k) {
        RT[col] = AD[j];
        BA[bval] = 0.0;
        bval += 1;
      }
    }
    double dk = RT[k];
    I i = C2R[k];
    while (i < k) {
      I nexti = C2R[i];
      I ili = IL[i];
      double uikdi = -BA[ili] * BA[BD[i]];
      dk = fma(uikdi, BA[ili], dk);
      BA[ili] = uikdi;
      I jmin = ili + 1, jmax = BI[i + 1];
      if (jmin < jmax) {
        for (I j = jmin; j < jmax; j++) {
          I cc = BJ[j];
          RT[cc] = fma(uikdi, BA[j], RT[cc]);
        }
        IL[i] = jmin;
        I cj = BJ[jmin];
        C2R[i] = C2R[cj];
        C2R[cj] = i;
      }
      i = nexti;
    }
    I jmin = BI[k], jmax = BI[k + 1] - 1;
    if (jmin < jmax) {
      for (I j = jmin; j < jmax; j++) BA[j] = RT[BJ[j]];
      IL[k] = jmin;
      I cj = BJ[jmin];
      C2R[k] = C2R[cj];
      C2R[cj] = k;
    }
    BA[BD[k]] = 1.0 / dk;
  }
}

/* ===================================================== factor.mojo */

/* factor.mojo:31-67 MatILUFactorSymbolic_SeqAIJ_ilu0 */
I or_ilu_symbolic_ilu0(I bi, I bj, I bdiag, I adiag, I irow, I icol, I nrow) {
  I *BI = (I *)bi, *BJ = (I *)bj, *BD = (I *)bdiag;
  const I *AD = (const I *)adiag, *AR = (const I *)irow, *AC = (const I *)icol;
  I k = 0;
  BI[0] = 0;
  for (I i = 0; i < nrow; i++) {
    I nz = AD[i] - AR[i];
    BI[i + 1] = BI[i] + nz;
    I src = AR[i];
    for (I j = 0; j < nz; j++) BJ[k++] = AC[src + j];
  }
  BD[nrow] = BI[nrow] - 1;
  for (I i = nrow - 1; i >= 0; i--) {
    I nz = AR[i + 1] - AD[i] - 1;
    I src = AD[i] + 1;
    for (I j = 0; j < nz; j++) BJ[k++] = AC[src + j];
    BJ[k++] = i;
    BD[i] = BD[i + 1] + nz + 1;
  }
  return k;
}

/* factor.mojo:70-122 MatILUFactorSymbolic_SeqAIJ (level of fill) */
I or_ilu_symbolic(I bi, I bj, I bjlvl, I bdiag, I irow, I icol, I r, I ic, I levels, I cap,
                  I scratch, I nrow) {
  I *BI = (I *)bi, *BJ = (I *)bj, *BJL = (I *)bjlvl, *BD = (I *)bdiag;
  const I *AR = (const I *)irow, *AC = (const I *)icol, *Rr = (const I *)r, *IC = (const I *)ic;
  I nl = nrow + 1;
  I *lnk = (I *)scratch, *lnklvl = lnk + nl, *bt = lnk + 2 * nl;
  lnk[nrow] = nrow;
  BI[0] = 0;
  BD[0] = 0;
  for (I i = 0; i < nrow; i++) {
    I ri = Rr[i];
    I nzi = 0;
    I nnz = AR[ri + 1] - AR[ri];
    const I *cols = AC + AR[ri];
    lnk[i] = -1;
    nzi += ll_init(nnz, cols, nrow, IC, lnk, lnklvl, bt);
    I nzbd = 0;
    I prow = lnk[nrow];
    while (prow < i) {
      I nb = BD[prow];
      const I *pcols = BJ + BI[prow] + nb + 1;
      const I *pcolslvl = BJL + BI[prow] + nb + 1;
      I lnz = BI[prow + 1] - BI[prow] - nb - 1;
      nzi += ll_add(lnz, pcols, (double)levels, pcolslvl, nrow, lnk, lnklvl, bt, lnklvl[prow], 1);
      prow = lnk[prow];
      nzbd += 1;
    }
    BD[i] = nzbd;
    if (BI[i] + nzi > cap) return -(BI[i] + nzi);
    BI[i + 1] = BI[i] + nzi;
    ll_clean(nrow, nrow, nzi, lnk, lnklvl, BJ + BI[i], BJL + BI[i], bt);
  }
  return BI[nrow];
}

/* factor.mojo:125-273 MatICCFactorSymbolic_SeqAIJ (level of fill).  The pivot
   rows are read from the staging block at offset 0, which holds the most
   recently drained row (factor.mojo:204-206 against :235-238); restated as is. */
I or_icc_symbolic(I ui, I uj, I ujlvl, I udiag, I irow, I icol, I rip, I riip, I levels, I cap,
                  I scratch, I nrow, I scratchlen) {
  I *UI = (I *)ui, *UJ = (I *)uj, *UJL = (I *)ujlvl, *UD = (I *)udiag;
  const I *AR = (const I *)irow, *AC = (const I *)icol, *RP = (const I *)rip,
          *RI = (const I *)riip;
  I nl = nrow + 1;
  if (6 * nl + 2 * cap > scratchlen) return -(6 * nl + 2 * cap);
  I *S = (I *)scratch;
  I *lnk = S, *lnklvl = S + nl, *bt = S + 2 * nl, *jl = S + 3 * nl, *il = S + 4 * nl,
    *ajtmp = S + 5 * nl, *ujtmp = S + 6 * nl, *ujtmp_lvl = S + 6 * nl + cap;
  for (I i = 0; i < nrow; i++) {
    jl[i] = nrow;
    il[i] = 0;
  }
  lnk[nrow] = nrow;
  UI[0] = 0;
  for (I k = 0; k < nrow; k++) {
    I nzk = 0;
    I rk = RP[k];
    I ncols = AR[rk + 1] - AR[rk];
    I ncols_upper = 0;
    for (I j = 0; j < ncols; j++) {
      I col = AC[AR[rk] + j];
      if (RI[col] > k) ajtmp[ncols_upper++] = col;
    }
    nzk += ll_init(ncols_upper, ajtmp, nrow, RI, lnk, lnklvl, bt);
    if (bt[rk] == 0) {
      I lnkdata = nrow;
      ll_insert(0, ncols_upper, nrow, rk, &nzk, &lnkdata, lnk);
      lnklvl[rk] = 0;
      bt[rk] = 1;
    }
    I uk = UI[k];
    I prow = jl[k];
    while (prow < k) {
      I nextprow = jl[prow];
      I jmin = il[prow] + 1;
      I jmax = UI[prow + 1];
      I ncols2 = jmax - jmin;
      I pbase = jmin - UI[prow];
      const I *cols = ujtmp + pbase;
      const I *colslvl = ujtmp_lvl + pbase;
      if (ncols2 > 0)
        nzk += ll_add(ncols2, cols, (double)levels, colslvl, nrow, lnk, lnklvl, bt, colslvl[-1], 1);
      if (jmin < jmax) {
        il[prow] = jmin;
        I cj = cols[0];
        jl[prow] = jl[cj];
        jl[cj] = prow;
      }
      prow = nextprow;
    }
    if (uk + nzk > cap) return -(uk + nzk);
    ll_clean(nrow, nrow, nzk, lnk, lnklvl, ujtmp, ujtmp_lvl, bt);
    if (nzk > 1) {
      I cj = ujtmp[1];
      jl[k] = jl[cj];
      jl[cj] = k;
      il[k] = uk + 1;
    }
    I diag_at = -1;
    for (I c = 0; c < nzk; c++)
      if (ujtmp[c] == rk) { diag_at = c; break; }
    I w = 0;
    for (I c = 0; c < nzk; c++)
      if (c != diag_at) {
        UJ[uk + w] = ujtmp[c];
        UJL[uk + w] = ujtmp_lvl[c];
        w += 1;
      }
    for (I c = 1; c < w; c++) {
      I key = UJ[uk + c];
      I kk = c - 1;
      while (kk >= 0 && UJ[uk + kk] > key) {
        UJ[uk + kk + 1] = UJ[uk + kk];
        kk -= 1;
      }
      UJ[uk + kk + 1] = key;
    }
    UJ[uk + nzk - 1] = rk;
    UJL[uk + nzk - 1] = ujtmp_lvl[diag_at];
    UD[k] = uk + nzk - 1;
    UI[k + 1] = uk + nzk;
  }
  return UI[nrow];
}

/* factor.mojo:279-332 MatGetOrdering_RCM (index-order neighbours) */
void or_ordering_rcm(I perm, I level, I visited, I queue, I irow, I icol, I nrow) {
  I *P = (I *)perm, *L = (I *)level, *V = (I *)visited, *Q = (I *)queue;
  const I *AR = (const I *)irow, *AC = (const I *)icol;
  for (I i = 0; i < nrow; i++) {
    L[i] = AR[i + 1] - AR[i];
    V[i] = 0;
  }
  I filled = 0;
  for (I start = 0; start < nrow; start++) {
    if (V[start] == 0) {
      V[start] = 1;
      Q[0] = start;
      I head = 0, tail = 1;
      while (head < tail) {
        I node = Q[head++];
        for (I j = AR[node]; j < AR[node + 1]; j++) {
          I c = AC[j];
          if (V[c] == 0) {
            V[c] = 1;
            Q[tail++] = c;
          }
        }
      }
      for (I i = 0; i < tail / 2; i++) {
        I t = Q[i];
        Q[i] = Q[tail - 1 - i];
        Q[tail - 1 - i] = t;
      }
      for (I i = 0; i < tail; i++) P[filled + i] = Q[i];
      filled += tail;
    }
  }
}

/* factor.mojo:335-381 MatILUFreeSpaceContiguousLU */
void or_ilu_contiguous_lu(I bjsrc, I bi, I bdiag, I bj, I n) {
  const I *SRC = (const I *)bjsrc;
  I *BI = (I *)bi, *BD = (I *)bdiag, *BJ = (I *)bj;
  I bi_temp = BI[n];
  I src = 0;
  I nnzL = n ? BD[0] : 0;
  for (I row = 0; row < n; row++) {
    I nnz = BI[row + 1] - BI[row];
    if (row == 0)
      BI[row] = 0;
    else
      BI[row] = BI[row - 1] + nnzL;
    nnzL = BD[row];
    I dst = BI[row];
    for (I k = 0; k < nnzL; k++) BJ[dst + k] = SRC[src + k];
    BD[row] = bi_temp - 1;
    BJ[bi_temp - 1] = row;
    I nnzU = nnz - nnzL;
    bi_temp -= nnzU;
    nnzU -= 1;
    dst = bi_temp;
    for (I k = 0; k < nnzU; k++) BJ[dst + k] = SRC[src + nnzL + 1 + k];
    src += nnz;
  }
  if (n) {
    BI[n] = BI[n - 1] + nnzL;
    BD[n] = BD[n - 1] - 1;
  }
}

/* factor.mojo:384-412 MatICCFreeSpaceContiguousCholesky */
void or_icc_contiguous_cholesky(I ujsrc, I ui, I udiag, I bi, I bdiag, I bj, I n) {
  const I *SRC = (const I *)ujsrc, *UI = (const I *)ui, *UD = (const I *)udiag;
  I *BI = (I *)bi, *BD = (I *)bdiag, *BJ = (I *)bj;
  for (I row = 0; row < n; row++)
    for (I j = UI[row]; j < UD[row] + 1; j++) BJ[j] = SRC[j];
  if (n) {
    BI[0] = 0;
    for (I row = 0; row < n; row++) BI[row + 1] = UD[row] + 1;
    BD[n] = BI[n];
  }
}

/* ===================================================== KSP (ksp.mojo) */

#define KSP_CONVERGED_RTOL 2
#define KSP_CONVERGED_ATOL 3
#define KSP_CONVERGED_ITS 4
#define KSP_CONVERGED_HAPPY_BREAKDOWN 7
#define KSP_DIVERGED_NULL (-2)
#define KSP_DIVERGED_ITS (-3)
#define KSP_DIVERGED_DTOL (-4)
#define KSP_DIVERGED_BREAKDOWN (-5)
#define KSP_DIVERGED_INDEFINITE_PC (-8)
#define KSP_DIVERGED_NANORINF (-9)
#define KSP_DIVERGED_INDEFINITE_MAT (-10)
#define NORM_NONE 1
#define NORM_PRECONDITIONED 2
#define NORM_UNPRECONDITIONED 3
#define NORM_NATURAL 4
#define PC_LEFT 0
#define PC_RIGHT 1
#define PC_SYMMETRIC 2
#define PCT_JACOBI 2
#define PCT_SOR 4
#define PCT_ILU 5
#define PCT_ICC 6

/* ksp.mojo:96-152 KSPConvergedDefault */
static I converged_default(I n, double rnorm, double *cfg, I *cfgi) {
  I normtype = cfgi[0], min_it = cfgi[3], chknorm = cfgi[4], convmaxits = cfgi[5],
    max_it = cfgi[6];
  if (convmaxits && n >= max_it) return KSP_CONVERGED_ITS;
  if (normtype == NORM_NONE) return 0;
  double rtol = cfg[0], atol = cfg[1], divtol = cfg[2];
  if (!n && cfg[14] == 0.0) {
    double rnorm0 = rnorm;
    if (cfgi[12] != 0) {
      double snorm = cfg[12];
      if (!snorm) snorm = rnorm;
      rnorm0 = snorm;
    }
    cfg[5] = rnorm0;
    double ttol = rtol * rnorm0;
    cfg[6] = ttol > atol ? ttol : atol;
    cfg[14] = 1.0;
  }
  if (n <= chknorm) return 0;
  /* the reference compares against MAX_REAL (ksp.mojo:85,138), so an actual
     infinity falls through to the divergence test */
  if (rnorm != rnorm || fabs(rnorm) == 1.7976931348623157e308)
    return KSP_DIVERGED_NANORINF;
  if (n < min_it) return 0;
  double rnorm0 = cfg[5], ttol = cfg[6];
  if (rnorm <= ttol) {
    if (rnorm < atol) return KSP_CONVERGED_ATOL;
    return KSP_CONVERGED_RTOL;
  }
  if (rnorm >= divtol * rnorm0) return KSP_DIVERGED_DTOL;
  return 0;
}

I or_ksp_converged_default(I n, double rnorm, I cfg, I cfgi) {
  return converged_default(n, rnorm, (double *)cfg, (I *)cfgi);
}

/* ksp.mojo:155-198 KSP_PCApply */
static void ksp_pcapply(const I *pci, const double *pcf, double *pstmp, I n, double *x, double *y) {
  I kind = pci[0];
  if (kind == PCT_JACOBI) {
    or_pc_jacobi_apply((I)y, (I)x, pci[1], n);
  } else if (kind == PCT_ILU) {
    or_pc_ilu_apply((I)x, (I)y, (I)pstmp, pci[1], pci[2], pci[3], pci[4], pci[5], pci[6], n);
  } else if (kind == PCT_ICC) {
    or_pc_icc_apply((I)x, (I)y, pci[1], pci[2], pci[3], pci[4], pci[5], (I)pstmp, n);
  } else if (kind == PCT_SOR) {
    pc_apply_sor(x, y, (const I *)pci[1], (const I *)pci[2], (const double *)pci[3],
                 (const I *)pci[4], (double *)pci[5], (double *)pci[6], (double *)pci[7], pcf[0],
                 pcf[1], 1, 1, pci[10], n);
  } else {
    or_vec_copy((I)x, (I)y, n);
  }
}
/* ksp.mojo:201-208 PCApplySymmetric (reads the Jacobi square root from pci[5]) */
static void pc_apply_symmetric(const I *pci, const double *pcf, double *pstmp, I n, double *x,
                               double *y) {
  if (pci[0] == PCT_JACOBI)
    or_pc_jacobi_apply((I)y, (I)x, pci[5], n);
  else
    ksp_pcapply(pci, pcf, pstmp, n, x, y);
}
/* ksp.mojo:211-218 KSPUnwindPreconditioner */
static void unwind_pc(const I *pci, const double *pcf, double *pstmp, I side, I n, double *vsoln,
                      double *vt1) {
  or_vec_copy((I)vsoln, (I)vt1, n);
  if (side == PC_RIGHT) {
    ksp_pcapply(pci, pcf, pstmp, n, vsoln, vt1);
    or_vec_copy((I)vt1, (I)vsoln, n);
  }
}

typedef struct {
  const I *R, *C;
  const double *A;
  I n;
} Op;

static void op_mult(const Op *op, double *x, double *y) {
  or_mat_mult((I)x, (I)y, (I)op->R, (I)op->C, (I)op->A, op->n);
}

/* ksp.mojo:221-243 KSPInitialResidual */
static void initial_residual(const I *pci, const double *pcf, double *pstmp, I side, I guess_zero,
                             const Op *op, double *vsoln, double *vt1, double *vt2, double *vres,
                             double *vb) {
  I n = op->n;
  if (!guess_zero) {
    op_mult(op, vsoln, vt1);
    or_vec_copy((I)vb, (I)vt2, n);
    or_vec_axpy((I)vt2, -1.0, (I)vt1, n);
  } else {
    or_vec_copy((I)vb, (I)vt2, n);
  }
  if (side == PC_SYMMETRIC)
    pc_apply_symmetric(pci, pcf, pstmp, n, vt2, vres);
  else if (side == PC_LEFT)
    ksp_pcapply(pci, pcf, pstmp, n, vt2, vres);
  else
    or_vec_copy((I)vt2, (I)vres, n);
}
/* ksp.mojo:246-265 KSP_PCApplyBAorAB */
static void apply_ba_or_ab(const I *pci, const double *pcf, double *pstmp, I side, const Op *op,
                           double *x, double *y, double *w) {
  I n = op->n;
  if (side == PC_RIGHT) {
    ksp_pcapply(pci, pcf, pstmp, n, x, w);
    op_mult(op, w, y);
  } else if (side == PC_LEFT) {
    op_mult(op, x, w);
    ksp_pcapply(pci, pcf, pstmp, n, w, y);
  } else {
    pc_apply_symmetric(pci, pcf, pstmp, n, x, w);
    op_mult(op, w, y);
    or_vec_copy((I)y, (I)w, n);
    pc_apply_symmetric(pci, pcf, pstmp, n, w, y);
  }
}

static void finish(double *cfg, I *cfgi, double rnorm, I its, I reason) {
  cfg[7] = rnorm;
  cfgi[10] = its;
  cfgi[11] = reason;
}
static I sign_real(double x) { return x < 0.0 ? -1 : (x == 0.0 ? 0 : 1); }

/* ksp.mojo:271-389 KSPSolve_CG */
void or_ksp_solve_cg(I n, I irow, I icol, I data, I pci_, I pcf_, I pstmp_, I work, I cfg_,
                     I cfgi_, I b, I x) {
  const I *pci = (const I *)pci_;
  const double *pcf = (const double *)pcf_;
  double *pstmp = (double *)pstmp_, *cfg = (double *)cfg_;
  I *cfgi = (I *)cfgi_;
  Op op = {(const I *)irow, (const I *)icol, (const double *)data, n};
  I normtype = cfgi[0], chknorm = cfgi[4], max_it = cfgi[6], guess_zero = cfgi[2];
  double *X = (double *)x, *B = (double *)b;
  double *R = (double *)work, *Z = R + n, *P = R + 2 * n, *W = Z;
  double a = 1.0, beta = 0.0, betaold = 1.0, bcoef = 0.0, dpi = 0.0, dpiold = 0.0, dp = 0.0;
  I its = 0, reason = 0;
  (void)a;
  if (!guess_zero) {
    op_mult(&op, X, R);
    or_vec_aypx((I)R, -1.0, (I)B, n);
  } else {
    or_vec_copy((I)B, (I)R, n);
  }
  if (normtype == NORM_PRECONDITIONED) {
    ksp_pcapply(pci, pcf, pstmp, n, R, Z);
    dp = vec_norm2(Z, n);
  } else if (normtype == NORM_UNPRECONDITIONED) {
    dp = vec_norm2(R, n);
  } else if (normtype == NORM_NATURAL) {
    ksp_pcapply(pci, pcf, pstmp, n, R, Z);
    beta = r_dot(Z, R, n);
    dp = sqrt(fabs(beta));
  } else {
    dp = 0.0;
  }
  cfg[9] = dp;
  reason = converged_default(0, dp, cfg, cfgi);
  if (reason != 0) {
    finish(cfg, cfgi, dp, its, reason);
    return;
  }
  if (normtype != NORM_PRECONDITIONED && normtype != NORM_NATURAL)
    ksp_pcapply(pci, pcf, pstmp, n, R, Z);
  if (normtype != NORM_NATURAL) beta = r_dot(Z, R, n);
  I i = 0;
  for (;;) {
    its = i + 1;
    if (beta == 0.0) {
      reason = KSP_CONVERGED_ATOL;
      break;
    } else if (i > 0 && beta * betaold < 0.0) {
      reason = KSP_DIVERGED_INDEFINITE_PC;
      break;
    }
    if (!i) {
      or_vec_copy((I)Z, (I)P, n);
      bcoef = 0.0;
    } else {
      bcoef = beta / betaold;
      or_vec_aypx((I)P, bcoef, (I)Z, n);
    }
    dpiold = dpi;
    op_mult(&op, P, W);
    dpi = r_dot(P, W, n);
    betaold = beta;
    if (dpi == 0.0 || (i > 0 && sign_real(dpi) * sign_real(dpiold) < 0)) {
      reason = KSP_DIVERGED_INDEFINITE_MAT;
      break;
    }
    a = beta / dpi;
    or_vec_axpy((I)X, a, (I)P, n);
    or_vec_axpy((I)R, -a, (I)W, n);
    if (normtype == NORM_PRECONDITIONED && chknorm < i + 2) {
      ksp_pcapply(pci, pcf, pstmp, n, R, Z);
      dp = vec_norm2(Z, n);
    } else if (normtype == NORM_UNPRECONDITIONED && chknorm < i + 2) {
      dp = vec_norm2(R, n);
    } else if (normtype == NORM_NATURAL) {
      ksp_pcapply(pci, pcf, pstmp, n, R, Z);
      beta = r_dot(Z, R, n);
      dp = sqrt(fabs(beta));
    } else {
      dp = 0.0;
    }
    cfg[9 + its] = dp;
    reason = converged_default(its, dp, cfg, cfgi);
    if (reason != 0) break;
    if ((normtype != NORM_PRECONDITIONED && normtype != NORM_NATURAL) || chknorm >= i + 2)
      ksp_pcapply(pci, pcf, pstmp, n, R, Z);
    if (normtype != NORM_NATURAL || chknorm >= i + 2) beta = r_dot(Z, R, n);
    i += 1;
    if (i >= max_it) break;
  }
  if (reason == 0) reason = KSP_DIVERGED_ITS;
  finish(cfg, cfgi, dp, its, reason);
}

/* ksp.mojo:395-505 KSPSolve_BCGS */
void or_ksp_solve_bcgs(I n, I irow, I icol, I data, I pci_, I pcf_, I pstmp_, I work, I cfg_,
                       I cfgi_, I b, I x) {
  const I *pci = (const I *)pci_;
  const double *pcf = (const double *)pcf_;
  double *pstmp = (double *)pstmp_, *cfg = (double *)cfg_;
  I *cfgi = (I *)cfgi_;
  Op op = {(const I *)irow, (const I *)icol, (const double *)data, n};
  I normtype = cfgi[0], side = cfgi[1], guess_zero = cfgi[2], chknorm = cfgi[4],
    max_it = cfgi[6];
  double *X = (double *)x, *B = (double *)b, *WK = (double *)work;
  double *R = WK, *RP = WK + n, *V = WK + 2 * n, *T = WK + 3 * n, *S = WK + 4 * n,
         *P = WK + 5 * n, *GUESS = WK + 6 * n;
  I has_guess = 0;
  initial_residual(pci, pcf, pstmp, side, guess_zero, &op, X, V, T, R, B);
  if (side == PC_RIGHT && !guess_zero) {
    or_vec_copy((I)X, (I)GUESS, n);
    or_vec_set((I)X, 0.0, n);
    has_guess = 1;
  }
  double dp = normtype != NORM_NONE ? vec_norm2(R, n) : 0.0;
  I its = 0;
  cfg[9] = dp;
  I reason = converged_default(0, dp, cfg, cfgi);
  if (reason != 0) {
    if (has_guess) or_vec_axpy((I)X, 1.0, (I)GUESS, n);
    finish(cfg, cfgi, dp, its, reason);
    return;
  }
  or_vec_copy((I)R, (I)RP, n);
  double rhoold = 1.0, alpha = 1.0, omegaold = 1.0, omega = 1.0, rho = 1.0, d1 = 0.0, d2 = 0.0;
  or_vec_set((I)P, 0.0, n);
  or_vec_set((I)V, 0.0, n);
  I i = 0;
  for (;;) {
    rho = r_dot(R, RP, n);
    double beta = (rho / rhoold) * (alpha / omegaold);
    or_vec_axpbypcz((I)P, 1.0, -omegaold * beta, beta, (I)R, (I)V, n);
    apply_ba_or_ab(pci, pcf, pstmp, side, &op, P, V, T);
    d1 = r_dot(V, RP, n);
    if (d1 == 0.0) {
      reason = KSP_DIVERGED_BREAKDOWN;
      break;
    }
    alpha = rho / d1;
    or_vec_waxpy((I)S, -alpha, (I)V, (I)R, n);
    apply_ba_or_ab(pci, pcf, pstmp, side, &op, S, T, R);
    d1 = r_dot(S, T, n);
    d2 = r_sumsq(T, n);
    if (d2 == 0.0) {
      d1 = r_dot(S, S, n);
      if (d1 != 0.0) {
        reason = KSP_DIVERGED_BREAKDOWN;
        break;
      }
      or_vec_axpy((I)X, alpha, (I)P, n);
      its += 1;
      dp = 0.0;
      reason = KSP_CONVERGED_RTOL;
      break;
    }
    omega = d1 / d2;
    or_vec_axpbypcz((I)X, alpha, omega, 1.0, (I)P, (I)S, n);
    or_vec_waxpy((I)R, -omega, (I)T, (I)S, n);
    if (normtype != NORM_NONE && chknorm < i + 2) dp = vec_norm2(R, n);
    rhoold = rho;
    omegaold = omega;
    its += 1;
    cfg[9 + its] = dp;
    reason = converged_default(its, dp, cfg, cfgi);
    if (reason != 0) break;
    if (rho == 0.0) {
      reason = KSP_DIVERGED_BREAKDOWN;
      break;
    }
    i += 1;
    if (i >= max_it) break;
  }
  if (reason == 0) reason = KSP_DIVERGED_ITS;
  unwind_pc(pci, pcf, pstmp, side, n, X, T);
  if (has_guess) or_vec_axpy((I)X, 1.0, (I)GUESS, n);
  finish(cfg, cfgi, dp, its, reason);
}

/* ---------------------------------------------------------------- GMRES */

#define HH_(a, b) (hh[(b) * (m + 2) + (a)])
#define HES_(a, b) (hes[(b) * (m + 1) + (a)])

/* ksp.mojo:511-530 MGS */
static void gmres_mgs(double *vv, double *hh, double *hes, I m, I it, I n) {
  double *vnew = vv + (it + 1) * n;
  for (I k = 0; k <= it; k++) {
    double d = r_dot(vnew, vv + k * n, n);
    HH_(k, it) = d;
    HES_(k, it) = d;
    or_vec_axpy((I)vnew, -d, (I)(vv + k * n), n);
  }
}
/* simdvec.mojo:503-518 maxpy over a flat basis */
static void maxpy_basis(double *y, const double *alpha, const double *base, I nv, I n) {
  for (I k = 0; k < nv; k++) {
    const double *xk = base + k * n;
    double a = alpha[k];
    for (I i = 0; i < n; i++) y[i] = fma(a, xk[i], y[i]);
  }
}
/* ksp.mojo:533-586 CGS with refinement */
static void gmres_cgs(double *vv, double *hh, double *hes, double *lhh, I m, I it, I n,
                      I cgstype) {
  int refine = cgstype == 2 ? 1 : 0;
  double *vnew = vv + (it + 1) * n;
  for (I j = 0; j <= it; j++) {
    HH_(j, it) = 0.0;
    HES_(j, it) = 0.0;
  }
  r_mdot_basis(vnew, vv, it + 1, n, lhh);
  for (I j = 0; j <= it; j++) lhh[j] = -lhh[j];
  maxpy_basis(vnew, lhh, vv, it + 1, n);
  for (I j = 0; j <= it; j++) {
    HH_(j, it) = HH_(j, it) - lhh[j];
    HES_(j, it) = HES_(j, it) - lhh[j];
  }
  if (cgstype == 1) {
    double hnrm = 0.0;
    for (I j = 0; j <= it; j++) hnrm = fma(lhh[j], lhh[j], hnrm);
    hnrm = sqrt(hnrm);
    if (vec_norm2(vnew, n) < hnrm) refine = 1;
  }
  if (refine) {
    r_mdot_basis(vnew, vv, it + 1, n, lhh);
    for (I j = 0; j <= it; j++) lhh[j] = -lhh[j];
    maxpy_basis(vnew, lhh, vv, it + 1, n);
    for (I j = 0; j <= it; j++) {
      HH_(j, it) = HH_(j, it) - lhh[j];
      HES_(j, it) = HES_(j, it) - lhh[j];
    }
  }
}
/* ksp.mojo:589-629 KSPGMRESUpdateHessenberg; returns reason, writes *res */
static I gmres_update_hessenberg(double *hh, double *cc, double *ss, double *rs, I it, I m,
                                 I hapend, double *res) {
  for (I j = 1; j <= it; j++) {
    I idx = it * (m + 2) + (j - 1);
    double tt = hh[idx];
    hh[idx] = fma(cc[j - 1], tt, ss[j - 1] * hh[idx + 1]);
    hh[idx + 1] = fma(cc[j - 1], hh[idx + 1], -(ss[j - 1] * tt));
  }
  if (!hapend) {
    I base = it * (m + 2) + it;
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
/* ksp.mojo:1084-1087 _maxpy_vv_n: zero y then k sequential sweeps */
static void maxpy_zero(double *y, const double *alpha, const double *base, I nv, I n) {
  for (I i = 0; i < n; i++) y[i] = 0.0;
  maxpy_basis(y, alpha, base, nv, n);
}
/* ksp.mojo:632-664 KSPGMRESBuildSoln */
static I gmres_build_soln(double *rs, double *hh, double *vv, double *temp, double *wtmp,
                          double *vs, double *vdest, I m, I it, I n, const I *pci,
                          const double *pcf, double *pstmp, I side) {
  if (it < 0) {
    or_vec_copy((I)vs, (I)vdest, n);
    return 0;
  }
  if (hh[it * (m + 2) + it] != 0.0)
    rs[it] = rs[it] / hh[it * (m + 2) + it];
  else
    return KSP_DIVERGED_BREAKDOWN;
  for (I ii = 1; ii <= it; ii++) {
    I k = it - ii;
    double tt = rs[k];
    for (I j = k + 1; j <= it; j++) tt = fma(-hh[j * (m + 2) + k], rs[j], tt);
    if (hh[k * (m + 2) + k] == 0.0) return KSP_DIVERGED_BREAKDOWN;
    rs[k] = tt / hh[k * (m + 2) + k];
  }
  maxpy_zero(temp, rs, vv, it + 1, n);
  unwind_pc(pci, pcf, pstmp, side, n, temp, wtmp);
  or_vec_axpy((I)vdest, 1.0, (I)temp, n);
  return 0;
}
static double normalize(double *x, I n) {
  double z = vec_norm2(x, n);
  if (z > 0.0)
    for (I i = 0; i < n; i++) x[i] = x[i] / z;
  return z;
}
static void hs_split(double *hs, I m, double **HH, double **HES, double **CC, double **SS,
                     double **RS, double **OR) {
  I hhn = (m + 2) * (m + 1), hesn = (m + 1) * (m + 1), ccn = m + 1, rsn = m + 2;
  *HH = hs;
  *HES = hs + hhn;
  *CC = hs + hhn + hesn;
  *SS = hs + hhn + hesn + ccn;
  *RS = hs + hhn + hesn + 2 * ccn;
  *OR = hs + hhn + hesn + 2 * ccn + rsn;
}

/* ksp.mojo:667-769 KSPSolve_GMRES */
void or_ksp_solve_gmres(I n, I irow, I icol, I data, I pci_, I pcf_, I pstmp_, I work, I vv_,
                        I hs_, I cfg_, I cfgi_, I b, I x) {
  const I *pci = (const I *)pci_;
  const double *pcf = (const double *)pcf_;
  double *pstmp = (double *)pstmp_, *cfg = (double *)cfg_, *vv = (double *)vv_;
  I *cfgi = (I *)cfgi_;
  Op op = {(const I *)irow, (const I *)icol, (const double *)data, n};
  I normtype = cfgi[0], side = cfgi[1], guess_zero = cfgi[2], max_it = cfgi[6], m = cfgi[7],
    orthog = cfgi[8], cgstype = cfgi[9];
  double haptol = cfg[3], breakdowntol = cfg[4];
  double *HH, *HES, *CC, *SS, *RS, *ORTH;
  hs_split((double *)hs_, m, &HH, &HES, &CC, &SS, &RS, &ORTH);
  double *TEMP = (double *)work, *TEMP_MATOP = TEMP + n;
  double *X = (double *)x, *B = (double *)b;
  I its = 0, itcount = 0, reason = 0;
  double res = 0.0, prev_rnorm = -1.0;
  while (reason == 0) {
    initial_residual(pci, pcf, pstmp, side, guess_zero, &op, X, TEMP, TEMP_MATOP, vv, B);
    I it = 0;
    res = normalize(vv, n);
    if (prev_rnorm > 0.0 && fabs(res - prev_rnorm) > breakdowntol * cfg[8]) {
      reason = KSP_DIVERGED_BREAKDOWN;
      finish(cfg, cfgi, res, its, reason);
      return;
    }
    RS[0] = res;
    cfg[8] = res;
    cfg[9 + its] = res;
    if (!res)
      reason = KSP_CONVERGED_ATOL;
    else
      reason = converged_default(its, res, cfg, cfgi);
    I hapend = 0;
    while (reason == 0 && it < m && its < max_it) {
      apply_ba_or_ab(pci, pcf, pstmp, side, &op, vv + it * n, vv + (it + 1) * n, TEMP_MATOP);
      if (orthog == 1)
        gmres_mgs(vv, HH, HES, m, it, n);
      else
        gmres_cgs(vv, HH, HES, ORTH, m, it, n, cgstype);
      double tt = normalize(vv + (it + 1) * n, n);
      HH[it * (m + 2) + it + 1] = tt;
      HES[it * (m + 1) + it + 1] = tt;
      double hapbnd = fabs(tt / RS[it]);
      hapbnd = hapbnd > haptol ? haptol : hapbnd;
      hapend = tt < hapbnd ? 1 : 0;
      reason = gmres_update_hessenberg(HH, CC, SS, RS, it, m, hapend, &res);
      it += 1;
      its += 1;
      cfg[9 + its] = res;
      if (reason != 0) break;
      reason = converged_default(its, res, cfg, cfgi);
      if (hapend) {
        if (normtype == NORM_NONE)
          reason = KSP_CONVERGED_HAPPY_BREAKDOWN;
        else if (reason == 0) {
          reason = KSP_DIVERGED_BREAKDOWN;
          break;
        }
      }
    }
    itcount += it;
    I breason = gmres_build_soln(RS, HH, vv, TEMP, TEMP_MATOP, X, X, m, it - 1, n, pci, pcf,
                                 pstmp, side);
    if (reason == 0 && breason != 0) reason = breason;
    if (itcount >= max_it) {
      if (reason == 0) reason = KSP_DIVERGED_ITS;
      break;
    }
    prev_rnorm = res;
    guess_zero = 0;
  }
  finish(cfg, cfgi, res, its, reason);
}

/* ksp.mojo:886-903 _fgmres_buildsoln */
static void fgmres_build_soln(double *rs, double *hh, double *prevec, double *temp, I m, I it,
                              I n, double *vdest) {
  if (it < 0) return;
  if (hh[it * (m + 2) + it] != 0.0)
    rs[it] = rs[it] / hh[it * (m + 2) + it];
  else
    rs[it] = 0.0;
  for (I ii = 1; ii <= it; ii++) {
    I k = it - ii;
    double tt = rs[k];
    for (I j = k + 1; j <= it; j++) tt = fma(-hh[j * (m + 2) + k], rs[j], tt);
    rs[k] = tt / hh[k * (m + 2) + k];
  }
  maxpy_zero(temp, rs, prevec, it + 1, n);
  or_vec_axpy((I)vdest, 1.0, (I)temp, n);
}
/* ksp.mojo:874-880 _fgmres_residual */
static void fgmres_residual(const Op *op, double *x, double *temp, double *vv, double *b) {
  op_mult(op, x, temp);
  or_vec_waxpy((I)vv, -1.0, (I)temp, (I)b, op->n);
}

/* ksp.mojo:775-868 KSPSolve_FGMRES */
void or_ksp_solve_fgmres(I n, I irow, I icol, I data, I pci_, I pcf_, I pstmp_, I work, I vv_,
                         I prevec_, I hs_, I cfg_, I cfgi_, I b, I x) {
  const I *pci = (const I *)pci_;
  const double *pcf = (const double *)pcf_;
  double *pstmp = (double *)pstmp_, *cfg = (double *)cfg_, *vv = (double *)vv_,
         *prevec = (double *)prevec_;
  I *cfgi = (I *)cfgi_;
  Op op = {(const I *)irow, (const I *)icol, (const double *)data, n};
  I guess_zero = cfgi[2], max_it = cfgi[6], m = cfgi[7], orthog = cfgi[8], cgstype = cfgi[9];
  double haptol = cfg[3];
  double *HH, *HES, *CC, *SS, *RS, *ORTH;
  hs_split((double *)hs_, m, &HH, &HES, &CC, &SS, &RS, &ORTH);
  double *TEMP = (double *)work;
  double *X = (double *)x, *B = (double *)b;
  I its = 0, reason = 0;
  double res_norm = 0.0;
  if (!guess_zero)
    fgmres_residual(&op, X, TEMP, vv, B);
  else
    or_vec_copy((I)B, (I)vv, n);
  for (;;) {
    I loc_it = 0;
    res_norm = vec_norm2(vv, n);
    RS[0] = res_norm;
    cfg[9 + its] = res_norm;
    reason = converged_default(its, res_norm, cfg, cfgi);
    if (reason == 0) {
      or_vec_scale((I)vv, 1.0 / res_norm, n);
      I hapend = 0;
      while (reason == 0 && loc_it < m && its < max_it) {
        ksp_pcapply(pci, pcf, pstmp, n, vv + loc_it * n, prevec + loc_it * n);
        op_mult(&op, prevec + loc_it * n, vv + (loc_it + 1) * n);
        if (orthog == 1)
          gmres_mgs(vv, HH, HES, m, loc_it, n);
        else
          gmres_cgs(vv, HH, HES, ORTH, m, loc_it, n, cgstype);
        double tt = vec_norm2(vv + (loc_it + 1) * n, n);
        HH[loc_it * (m + 2) + loc_it + 1] = tt;
        HES[loc_it * (m + 1) + loc_it + 1] = tt;
        double hapbnd = fabs(tt / RS[loc_it]);
        hapbnd = hapbnd > haptol ? haptol : hapbnd;
        if (tt > hapbnd) {
          or_vec_scale((I)(vv + (loc_it + 1) * n), 1.0 / tt, n);
        } else {
          or_vec_set((I)TEMP, 0.0, n);
          hapend = 1;
        }
        reason = gmres_update_hessenberg(HH, CC, SS, RS, loc_it, m, hapend, &res_norm);
        if (reason != 0) break;
        loc_it += 1;
        its += 1;
        cfg[9 + its] = res_norm;
        reason = converged_default(its, res_norm, cfg, cfgi);
        if (hapend && reason == 0) {
          reason = KSP_DIVERGED_BREAKDOWN;
          break;
        }
      }
    }
    fgmres_build_soln(RS, HH, prevec, TEMP, m, loc_it - 1, n, X);
    if (its >= max_it) {
      if (reason == 0) reason = KSP_DIVERGED_ITS;
      break;
    }
    if (reason != 0) break;
    fgmres_residual(&op, X, TEMP, vv, B);
  }
  finish(cfg, cfgi, res_norm, its, reason);
}
