// mpk_mat.hip -- SEQAIJ kernels and the mp_mat_* exports.
//
// The device matrix is CSR with int32 row pointers and column indices and
// float64 values.  Every row sum runs serially in CSR order from 0.0 (or from
// y_i for MultAdd), one lane per row, with fma(a_ij, x_j, sum) -- the
// arithmetic of MatMult_SeqAIJ (src/ported.mojo:380-405) -- so a row gives
// the same bits as the CPU path.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <tuple>
#include <vector>

#include "mpk_internal.h"
#include "mpk_ops.h"
#include "../../include/mpk.h"

namespace mpk {

// One lane per row.  Row i's values and indices are contiguous, so a wave
// reads a contiguous span of the value and index streams spread over its
// rows; the lines are reused across the row loop from L1/L2.
__global__ void __launch_bounds__(256) k_spmv_row(i64 nrow, const int *__restrict__ rp,
                                                  const int *__restrict__ ci,
                                                  const double *__restrict__ v,
                                                  const double *__restrict__ x, double *__restrict__ y,
                                                  int add, const int64_t *stop) {
  if (stop && *stop) return;
  i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= nrow) return;
  int j = rp[i], je = rp[i + 1];
  double s = add ? y[i] : 0.0;
  for (; j < je; j++) s = fma(v[j], x[ci[j]], s);
  y[i] = s;
}

// ------------------------------------------------------------ sliced ELL
//
// The same matrix in slices of 64 rows (one wave): slice s stores entry k of
// its row l at soff[s] + 64 k + l, so each load instruction of a wave reads
// 512 contiguous bytes of values (256 of indices).  Row i still runs its
// entries in CSR order, serially, from 0.0 -- the row sum is the CSR one bit
// for bit; a lane stops at its own row length, and the padding slots of
// shorter rows are never read.  Blocks are remapped so that each of the 8
// XCDs sweeps one contiguous range of rows, which keeps the x entries that
// neighbouring rows share inside that XCD's L2.

template <bool NT>
__global__ void __launch_bounds__(256) k_spmv_sell(i64 nrow, const int *__restrict__ rp,
                                                   const int *__restrict__ soff, const double *__restrict__ sv,
                                                   const int *__restrict__ sc, const double *__restrict__ x,
                                                   double *__restrict__ y, int add, const int64_t *stop) {
  if (stop && *stop) return;
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const i64 i = (i64)b * 256 + threadIdx.x;
  if (i >= nrow) return;
  const short *sd = sell_offsets16(soff, sc, nrow);
  const int len = rp[i + 1] - rp[i];
  const int base = soff[i >> 6] + (int)(i & 63);
  const double s0 = add ? y[i] : 0.0;
  st_stream<NT>(y + i, sd ? sell_row<NT>(sv, SellCols16{sd}, x, base, len, s0, i)
                          : sell_row<NT>(sv, SellCols32{sc}, x, base, len, s0, i));
}

// scatter CSR into the slice layout (one lane per row)
__global__ void k_sell_fill(i64 nrow, const int *rp, const int *ci, const double *v, const int *soff, double *sv,
                            int *sc) {
  i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= nrow) return;
  const int j0 = rp[i], len = rp[i + 1] - j0;
  const int base = soff[i >> 6] + (int)(i & 63);
  for (int k = 0; k < len; k++) {
    sv[base + 64 * k] = v[j0 + k];
    sc[base + 64 * k] = ci[j0 + k];
  }
}

// The 16-bit column offsets of the slots (after the int32 columns, at sc +
// padded slots; the same slot positions).
__global__ void k_sell_fill16(i64 nrow, const int *rp, const int *ci, const int *soff, short *sd) {
  i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= nrow) return;
  const int j0 = rp[i], len = rp[i + 1] - j0;
  const int base = soff[i >> 6] + (int)(i & 63);
  for (int k = 0; k < len; k++) sd[base + 64 * k] = (short)(ci[j0 + k] - (int)i);
}

// The largest |column - row| over a CSR matrix's entries, into *out (which
// starts at 0): a wave's maximum, then one atomic per wave.
__global__ void k_csr_span(i64 nrow, const int *rp, const int *ci, int *out) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  long long m = 0;
  if (i < nrow)
    for (int j = rp[i]; j < rp[i + 1]; j++) {
      const long long d = (long long)ci[j] - i;
      m = d < 0 ? (-d > m ? -d : m) : (d > m ? d : m);
    }
  int mi = m > 0x7fffffff ? 0x7fffffff : (int)m;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const int o = __shfl_down(mi, w, 64);
    mi = o > mi ? o : mi;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, mi);
}

// Fills a sliced copy (soff: the host offsets, copied to soff_d; sv; sc with
// room for 6 bytes a slot): values and int32 columns, then, when every entry
// lies within 32767 columns of its row, the 16-bit offsets, marked in the
// last offset.  False when a launch failed.
static bool sell_fill(const int *rp, const int *ci, const double *v, i64 nrow, std::vector<int> &soff,
                      int *soff_d, double *sv, int *sc) {
  const i64 ns = (nrow + 63) / 64, total = soff[ns];
  int *span = reinterpret_cast<int *>(ctx()->dres);
  const int zero = 0;
  copy_htod(span, &zero, sizeof zero);
  MPK_LAUNCH(k_csr_span, nblocks(nrow, 256), 256, nrow, rp, ci, span);
  int width = 0x7fffffff;
  copy_dtoh(&width, span, sizeof width);
  if (failed()) return false;
  const char *cols = std::getenv("MPK_SELL_COLS");  // "32": int32 columns only (comparisons)
  if (cols && cols[0] == '3' && cols[1] == '2' && !cols[2]) width = 0x7fffffff;
  if (width <= 32767) soff[ns] = (int)(-1 - total);
  copy_htod(soff_d, soff.data(), soff.size() * sizeof(int));
  MPK_LAUNCH(k_sell_fill, nblocks(nrow, 256), 256, nrow, rp, ci, v, (const int *)soff_d, sv, sc);
  if (width <= 32767)
    MPK_LAUNCH(k_sell_fill16, nblocks(nrow, 256), 256, nrow, rp, ci, (const int *)soff_d,
               reinterpret_cast<short *>(sc + total));
  return !failed();
}

// slice offsets from the row pointers; returns the padded entry count.  The
// offsets and every slot index are int32 (the last offset also encodes the
// column form as -1 - count), so past INT32_MAX padded slots the offsets are
// not filled in: callers check the count (sell_fits) and keep the row kernel.
template <class Ix>
static i64 sell_offsets_from(const Ix *rp, i64 nrow, std::vector<int> &soff) {
  i64 ns = (nrow + 63) / 64;
  soff.assign(ns + 1, 0);
  i64 total = 0;
  for (i64 s = 0; s < ns; s++) {
    i64 w = 0;
    for (i64 i = s * 64; i < nrow && i < s * 64 + 64; i++)
      w = (i64)rp[i + 1] - (i64)rp[i] > w ? (i64)rp[i + 1] - (i64)rp[i] : w;
    if (sell_fits(total)) soff[s] = (int)total;
    total += 64 * w;
  }
  if (sell_fits(total)) soff[ns] = (int)total;
  return total;
}
i64 sell_offsets(const int *rp_d, i64 nrow, std::vector<int> &soff) {
  std::vector<int> rp(nrow + 1);
  copy_dtoh(rp.data(), rp_d, (nrow + 1) * sizeof(int));
  return sell_offsets_from(rp.data(), nrow, soff);
}

bool sell_nt(i64 nrow) {
  const char *e = std::getenv("MPK_SELL_NT");  // read at each launch (tuning and tests)
  if (e && (e[0] == '0' || e[0] == '1') && !e[1]) return e[0] == '1';
  return nrow >= 3000000;
}

void spmv(const DevCSR &A, const double *x, double *y, int add, const int64_t *stop) {
  if (A.nrow <= 0 || !ctx()) return;
  if (A.sell_off) {
    if (sell_nt(A.nrow))
      MPK_LAUNCH(k_spmv_sell<true>, nblocks(A.nrow, 256), 256, A.nrow, A.rp, A.sell_off, A.sell_v, A.sell_c, x, y,
                 add, stop);
    else
      MPK_LAUNCH(k_spmv_sell<false>, nblocks(A.nrow, 256), 256, A.nrow, A.rp, A.sell_off, A.sell_v, A.sell_c, x, y,
                 add, stop);
    return;
  }
  MPK_LAUNCH(k_spmv_row, nblocks(A.nrow, 256), 256, A.nrow, A.rp, A.ci, A.v, x, y, add, stop);
}

// A temporary slice copy of a CSR matrix, owned by the object.
SellTemp::SellTemp(DevCSR &A) {
  if (A.nrow <= 0 || A.sell_off || !ctx()) return;
  std::vector<int> soff;
  i64 total = sell_offsets(A.rp, A.nrow, soff);
  if (!sell_fits(total)) return;  // int32 offsets: keep the row kernel
  HIPCHK(hipMalloc(&off, soff.size() * sizeof(int)));
  HIPCHK(hipMalloc(&v, (total > 0 ? total : 1) * sizeof(double)));
  HIPCHK(hipMalloc(&c, (total > 0 ? 6 * total : 4)));  // int32 columns, then 16-bit offsets
  if (failed() || !sell_fill(A.rp, A.ci, A.v, A.nrow, soff, off, v, c)) return;
  A.sell_off = off;
  A.sell_v = v;
  A.sell_c = c;
}
SellTemp::~SellTemp() {
  if (!off) return;
  HIPCHK(hipStreamSynchronize(stream()));
  HIPCHK(hipFree(off));
  HIPCHK(hipFree(v));
  HIPCHK(hipFree(c));
}

// ------------------------------------------------------------ transpose

// y = A^T x as a gather over the transpose: column c's entries are visited in
// increasing row, which is the order the reference's scatter adds them in
// (ported.mojo:408-431).  tp[k] is the position of the entry in A's arrays.
__global__ void __launch_bounds__(256) k_spmv_t(i64 ncol, const int *trp, const int *trow,
                                                const int *tp, const double *v, const double *x,
                                                double *y) {
  i64 c = (i64)blockIdx.x * 256 + threadIdx.x;
  if (c >= ncol) return;
  double s = 0.0;
  for (int k = trp[c]; k < trp[c + 1]; k++) s = fma(v[tp[k]], x[trow[k]], s);
  y[c] = s;
}

// Builds the transpose map of a CSR pattern into caller-owned device arrays:
// trp[ncol + 1], trow[nnz] (source row), tpos[nnz] (position in A's arrays),
// each column's entries in increasing row.
bool csr_transpose_map(const int *rp, const int *ci, i64 nrow, i64 ncol, int *trp_d, int *trow_d, int *tpos_d) {
  Ctx *c = ctx();
  if (!c) return false;
  std::vector<int> hrp(nrow + 1);
  HIPCHK(hipMemcpyAsync(hrp.data(), rp, (nrow + 1) * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  i64 nnz = hrp[nrow];
  std::vector<int> hci(nnz > 0 ? nnz : 1);
  if (nnz > 0) {
    HIPCHK(hipMemcpyAsync(hci.data(), ci, nnz * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  std::vector<int> trp(ncol + 1, 0), trow(nnz > 0 ? nnz : 1), tp(nnz > 0 ? nnz : 1);
  for (i64 k = 0; k < nnz; k++) trp[hci[k] + 1]++;
  for (i64 col = 0; col < ncol; col++) trp[col + 1] += trp[col];
  std::vector<int> fill(trp.begin(), trp.end() - 1);
  for (i64 i = 0; i < nrow; i++)
    for (int k = hrp[i]; k < hrp[i + 1]; k++) {
      int pos = fill[hci[k]]++;
      trow[pos] = (int)i;
      tp[pos] = k;
    }
  HIPCHK(hipMemcpyAsync(trp_d, trp.data(), (ncol + 1) * sizeof(int), hipMemcpyHostToDevice, c->stream));
  if (nnz > 0) {
    HIPCHK(hipMemcpyAsync(trow_d, trow.data(), nnz * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(tpos_d, tp.data(), nnz * sizeof(int), hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return !failed();
}

static i64 csr_nnz(const int *rp, i64 nrow) {
  int h = 0;
  copy_dtoh(&h, rp + nrow, sizeof(int));
  return h;
}

// ------------------------------------------------------------ row scans

__global__ void k_get_diagonal(i64 nrow, const int *rp, const int *ci, const double *v, double *d) {
  i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= nrow) return;
  double r = 0.0;
  for (int j = rp[i]; j < rp[i + 1]; j++)
    if (ci[j] == i) {
      r = v[j];
      break;
    }
  d[i] = r;
}

// diag[i] = position of a_ii in row i, or the row end; *dense cleared to 0
// by any row without a diagonal (ported.mojo:354-377)
__global__ void k_diag_markers(i64 nrow, const int *rp, const int *ci, int *diag, int *dense) {
  i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= nrow) return;
  int be = rp[i + 1], m = be;
  for (int j = rp[i]; j < be; j++)
    if (ci[j] == i) {
      m = j;
      break;
    }
  diag[i] = m;
  if (m == be) atomicAnd(dense, 0);
}

__global__ void k_scale_vals(i64 nnz, double *v, double alpha) {
  i64 j = (i64)blockIdx.x * 256 + threadIdx.x;
  if (j < nnz) v[j] = v[j] * alpha;
}

}  // namespace mpk

using namespace mpk;

extern "C" {

void mp_mat_mult(int64_t x, int64_t y, int64_t irow, int64_t icol, int64_t data, int64_t nrow) {
  DevCSR A{nrow, 0, (const int *)irow, (const int *)icol, (const double *)data};
  spmv(A, (const double *)x, (double *)y, 0, nullptr);
}

void mp_mat_mult_add(int64_t x, int64_t y, int64_t irow, int64_t icol, int64_t data, int64_t nrow) {
  DevCSR A{nrow, 0, (const int *)irow, (const int *)icol, (const double *)data};
  spmv(A, (const double *)x, (double *)y, 1, nullptr);
}

void mp_mat_mult_transpose(int64_t x, int64_t y, int64_t irow, int64_t icol, int64_t data, int64_t nrow,
                           int64_t ncols) {
  // the reference signature carries no transpose, so one is built for the
  // call; mpk_csr_transpose + mpk_mat_mult_transpose_t keep it instead
  if (ncols <= 0 || !ctx()) return;
  if (nrow <= 0) {
    vset((double *)y, 0.0, ncols, nullptr);
    return;
  }
  i64 nnz = csr_nnz((const int *)irow, nrow);
  int *trp = nullptr, *trow = nullptr, *tpos = nullptr;
  HIPCHK(hipMalloc(&trp, (ncols + 1) * sizeof(int)));
  HIPCHK(hipMalloc(&trow, (nnz > 0 ? nnz : 1) * sizeof(int)));
  HIPCHK(hipMalloc(&tpos, (nnz > 0 ? nnz : 1) * sizeof(int)));
  if (csr_transpose_map((const int *)irow, (const int *)icol, nrow, ncols, trp, trow, tpos))
    MPK_LAUNCH(k_spmv_t, nblocks(ncols, 256), 256, (i64)ncols, (const int *)trp, (const int *)trow,
               (const int *)tpos, (const double *)data, (const double *)x, (double *)y);
  HIPCHK(hipStreamSynchronize(stream()));
  HIPCHK(hipFree(trp));
  HIPCHK(hipFree(trow));
  HIPCHK(hipFree(tpos));
}

// The transpose map of a CSR pattern, into caller-owned [D] int32 arrays
// trp[ncols + 1], trow[nnz], tpos[nnz].  Returns 0 on success.
int64_t mpk_csr_transpose(int64_t irow, int64_t icol, int64_t nrow, int64_t ncols, int64_t trp, int64_t trow,
                          int64_t tpos) {
  if (nrow < 0 || ncols <= 0) return 0;
  return csr_transpose_map((const int *)irow, (const int *)icol, nrow, ncols, (int *)trp, (int *)trow,
                           (int *)tpos)
             ? 0
             : -1;
}

// y = A^T x over a map made by mpk_csr_transpose.
void mpk_mat_mult_transpose_t(int64_t x, int64_t y, int64_t data, int64_t ncols, int64_t trp, int64_t trow,
                              int64_t tpos) {
  if (ncols <= 0 || !ctx()) return;
  MPK_LAUNCH(k_spmv_t, nblocks(ncols, 256), 256, (i64)ncols, (const int *)trp, (const int *)trow,
             (const int *)tpos, (const double *)data, (const double *)x, (double *)y);
}

void mp_mat_get_diagonal(int64_t diag, int64_t irow, int64_t icol, int64_t data, int64_t nrow) {
  if (nrow <= 0 || !ctx()) return;
  MPK_LAUNCH(k_get_diagonal, nblocks(nrow, 256), 256, (i64)nrow, (const int *)irow, (const int *)icol,
             (const double *)data, (double *)diag);
}

void mp_mat_scale(int64_t data, int64_t nnz, double alpha) {
  if (nnz <= 0 || !ctx()) return;
  MPK_LAUNCH(k_scale_vals, nblocks(nnz, 256), 256, (i64)nnz, (double *)data, alpha);
}

void mp_mat_zero_entries(int64_t data, int64_t nnz) {
  if (nnz <= 0 || !ctx()) return;
  vset((double *)data, 0.0, nnz, nullptr);
}

int64_t mp_mat_get_diagonal_markers(int64_t irow, int64_t icol, int64_t diag, int64_t nrow) {
  Ctx *c = ctx();
  if (!c) return 0;
  if (nrow <= 0) return 1;
  int *dense = reinterpret_cast<int *>(c->dres);
  int one = 1;
  HIPCHK(hipMemcpyAsync(dense, &one, sizeof(int), hipMemcpyHostToDevice, c->stream));
  MPK_LAUNCH(k_diag_markers, nblocks(nrow, 256), 256, (i64)nrow, (const int *)irow, (const int *)icol,
             (int *)diag, dense);
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, dense, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return h;
}

}  // extern "C"

extern "C" {

// The sliced copy of a CSR matrix, in caller-owned [D] arrays: soff[nslices + 1]
// (nslices = ceil(nrow / 64)), sv[padded], sc[padded]; mpk_sell_size returns
// the padded entry count.
int64_t mpk_sell_size(int64_t irow, int64_t nrow) {
  if (nrow <= 0 || !ctx()) return 0;
  std::vector<int> soff;
  return sell_offsets((const int *)irow, nrow, soff);
}

int64_t mpk_sell_size_host(int64_t irow, int64_t nrow) {
  if (nrow <= 0) return 0;
  std::vector<int> soff;
  const i64 total = sell_offsets_from((const int64_t *)irow, nrow, soff);
  return sell_fits(total) ? total : -1;
}

int64_t mpk_sell_build(int64_t irow, int64_t icol, int64_t data, int64_t nrow, int64_t soff, int64_t sv,
                       int64_t sc) {
  if (nrow <= 0 || !ctx()) return 0;
  std::vector<int> off;
  const i64 total = sell_offsets((const int *)irow, nrow, off);
  if (!sell_fits(total)) {
    set_error(4, "sliced copy of %lld padded slots: past the int32 slot indices (at most %lld)", (long long)total,
              (long long)kSellMaxSlots);
    return -1;
  }
  return sell_fill((const int *)irow, (const int *)icol, (const double *)data, nrow, off, (int *)soff,
                   (double *)sv, (int *)sc)
             ? 0
             : -1;
}

// y = A x (add = 0) or y = y + A x (add = 1) over the sliced copy.
void mpk_sell_mult(int64_t x, int64_t y, int64_t irow, int64_t soff, int64_t sv, int64_t sc, int64_t nrow,
                   int64_t add) {
  if (nrow <= 0 || !ctx()) return;
  DevCSR A{nrow, 0, (const int *)irow, nullptr, nullptr, (const int *)soff, (const double *)sv,
           (const int *)sc};
  spmv(A, (const double *)x, (double *)y, (int)add, nullptr);
}

// Average duration in ms of `reps` back-to-back SpMV launches, timed with HIP
// events on the library's stream (the stream the kernel runs on).
double mpk_time_spmv(int64_t x, int64_t y, int64_t irow, int64_t icol, int64_t data, int64_t nrow, int64_t soff,
                     int64_t sv, int64_t sc, int64_t reps) {
  Ctx *c = ctx();
  if (!c || reps <= 0) return -1.0;
  DevCSR A{nrow, 0, (const int *)irow, (const int *)icol, (const double *)data, (const int *)soff,
           (const double *)sv, (const int *)sc};
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int64_t r = 0; r < reps; r++) spmv(A, (const double *)x, (double *)y, 0, nullptr);
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  return (double)ms / (double)reps;
}

}  // extern "C"
