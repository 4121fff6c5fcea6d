// mpk_ops.h -- host-side launchers shared between the library's source files.
// `stop` (may be null) is a device flag; a kernel launched with it returns at
// once when the flag is nonzero, which is how a solver's queued iterations
// become no-ops after it has finished.
#pragma once

#include "mpk_internal.h"

namespace mpk {

// ---- Vec (mpk_vec.hip)
void vaxpy(double *y, double alpha, const double *x, i64 n, const int64_t *stop);
void vaypx(double *y, double alpha, const double *x, i64 n, const int64_t *stop);
void vaypx_raw(double *y, double alpha, const double *x, i64 n, const int64_t *stop);
void vaxpby_raw(double *y, double alpha, double beta, const double *x, i64 n, const int64_t *stop);
void vaxpbypcz(double *z, double alpha, double beta, double gamma, const double *x, const double *y,
               i64 n, const int64_t *stop);
void vwaxpy(double *w, double alpha, const double *x, const double *y, i64 n, const int64_t *stop);
void vscale(double *x, double alpha, i64 n, const int64_t *stop);
void vdivide(double *x, double z, i64 n, const int64_t *stop);
void vset(double *x, double alpha, i64 n, const int64_t *stop);
void vcopy(const double *x, double *y, i64 n);
void vpointwise_mult(double *w, const double *x, const double *y, i64 n, const int64_t *stop);
void vmaxpy(double *y, int nv, const double *alpha_host, const double *const *xs, i64 n,
            bool zero_first, const int64_t *stop);
// y = (zero_first ? 0 : y) + sum_k (sign * coef_dev[k]) * (base + k ld), k < nv
void vmaxpy_dev(double *y, const double *base, i64 ld, const double *coef_dev, int nv, double sign,
                i64 n, bool zero_first, const int64_t *stop);
double dot_host(const double *x, const double *y, i64 n);
double norm2_host(const double *x, i64 n);
double reduce1_host(const double *x, i64 n, int kind);

// ---- Mat (mpk_mat.hip)
struct DevCSR {
  i64 nrow, ncol;
  const int *rp;    // nrow + 1
  const int *ci;    // nnz
  const double *v;  // nnz
  // optional sliced copy (64-row slices, column-major within a slice)
  const int *sell_off = nullptr;
  const double *sell_v = nullptr;
  const int *sell_c = nullptr;
};
// y = A x (add = 0) or y = y + A x (add = 1); serial row sums in CSR order,
// over the sliced copy when the matrix has one.
void spmv(const DevCSR &A, const double *x, double *y, int add, const int64_t *stop);
// Whether the sliced products of an nrow-row matrix load the matrix stream
// non-temporally: from 3,000,000 rows on (a 7-point operator there is about
// 270 MB, past the 256 MB last-level cache); MPK_SELL_NT=0|1 overrides
// (read at each call).
bool sell_nt(i64 nrow);
// The most padded slots a sliced copy can hold: its slice offsets and slot
// indices are int32, and its last offset may read -1 - count.
constexpr i64 kSellMaxSlots = 2147483647;
inline bool sell_fits(i64 padded) { return padded >= 0 && padded <= kSellMaxSlots; }
// The padded slot count of a sliced copy from its last slice offset, which
// holds -1 - that count when the copy carries 16-bit column offsets.
inline i64 sell_slots(int last) { return last < 0 ? -1 - (i64)last : (i64)last; }

// Builds a sliced copy of A for the lifetime of the object and attaches it
// to A (no-op when A already has one).
struct SellTemp {
  int *off = nullptr, *c = nullptr;
  double *v = nullptr;
  explicit SellTemp(DevCSR &A);
  ~SellTemp();
  SellTemp(const SellTemp &) = delete;
  SellTemp &operator=(const SellTemp &) = delete;
};

}  // namespace mpk
