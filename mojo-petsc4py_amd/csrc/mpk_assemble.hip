// mpk_assemble.hip -- Mat assembly on the device: the fold of a log of entry
// writes into CSR arrays (Mat._materialize, the reference's
// python/mojopetsc4py/PETSc/Mat.py:383-439 semantics), and the CSR arrays of
// a dense array's nonzeros (from_dense, Mat.py:555-589).
//
// Fold: every write is (row, column, value, add).  The entry's result is the
// per-entry dictionary's: an insert replaces the value, an add accumulates
// onto the value so far (0.0 when it has none); every entry ever written
// stays stored, columns ascending within a row.  On the device the writes are
// keyed by row * ncols + column and radix-sorted with their call positions
// (stable: call order kept within a key); each run of equal keys is folded in
// that order by one thread, acc = add ? acc + value : value from acc = 0.0,
// so the sums are the host fold's, bit for bit.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>

#include "mpk_internal.h"

namespace mpk {

// Keys and call positions; flag bit 0: a row or column outside the matrix.
__global__ void k_fold_keys(i64 nw, const int64_t *rows, const int64_t *cols, i64 nrows, i64 ncols,
                            unsigned long long *key, int *seq, unsigned *flag) {
  const i64 k = (i64)blockIdx.x * 256 + threadIdx.x;
  if (k >= nw) return;
  const int64_t r = rows[k], c = cols[k];
  const bool bad = r < 0 || r >= nrows || c < 0 || c >= ncols;
  key[k] = bad ? 0ull : (unsigned long long)r * (unsigned long long)ncols + (unsigned long long)c;
  seq[k] = (int)k;
  if (bad) atomicOr(flag, 1u);
}

// head[k] = 1 where a run of equal keys starts (head[nw] = 0 for the scan).
__global__ void k_fold_heads(i64 nw, const unsigned long long *skey, int *head) {
  const i64 k = (i64)blockIdx.x * 256 + threadIdx.x;
  if (k < nw) head[k] = k == 0 || skey[k] != skey[k - 1];
  if (k == nw) head[k] = 0;
}

// The runs' first positions: starts[run] = k; starts[m] = nw.
__global__ void k_fold_starts(i64 nw, const int *head, const int *run, int *starts) {
  const i64 k = (i64)blockIdx.x * 256 + threadIdx.x;
  if (k < nw && head[k]) starts[run[k]] = (int)k;
  if (k == nw) starts[run[nw]] = (int)nw;
}

// One thread per run: the entry's column and folded value, and its row's
// count.
__global__ void k_fold_runs(i64 m, i64 ncols, const int *starts, const unsigned long long *skey, const int *sseq,
                            const double *vals, const unsigned char *adds, int *ci, double *v, int *rowcnt) {
  const i64 s = (i64)blockIdx.x * 256 + threadIdx.x;
  if (s >= m) return;
  const int k0 = starts[s], k1 = starts[s + 1];
  const unsigned long long key = skey[k0];
  const i64 row = (i64)(key / (unsigned long long)ncols), col = (i64)(key % (unsigned long long)ncols);
  double acc = 0.0;
  for (int k = k0; k < k1; k++) {
    const int j = sseq[k];
    acc = adds[j] ? acc + vals[j] : vals[j];
  }
  ci[s] = (int)col;
  v[s] = acc;
  atomicAdd(rowcnt + row, 1);
}

// Dense to CSR: one wave per row, 64 columns at a time; a nonzero is any
// value != 0.0 (NaN included, -0.0 excluded), as numpy's nonzero counts.
__global__ void __launch_bounds__(64) k_dense_count(i64 n, i64 m, const double *a, int *cnt) {
  const i64 r = blockIdx.x;
  const int lane = (int)threadIdx.x;
  int c = 0;
  for (i64 j = lane; j < m; j += 64) c += a[r * m + j] != 0.0;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) c += __shfl_down(c, w, 64);
  if (lane == 0) cnt[r] = c;
  if (r == 0 && lane == 0) cnt[n] = 0;
}
__global__ void __launch_bounds__(64) k_dense_fill(i64 m, const double *a, const int *rp, int *ci, double *v) {
  const i64 r = blockIdx.x;
  const int lane = (int)threadIdx.x;
  int at = rp[r];
  for (i64 j0 = 0; j0 < m; j0 += 64) {
    const i64 j = j0 + lane;
    const double x = j < m ? a[r * m + j] : 0.0;
    const bool nz = x != 0.0;
    const unsigned long long ball = __ballot(nz);
    if (nz) {
      const int q = at + __popcll(ball & ((1ull << lane) - 1ull));
      ci[q] = (int)j;
      v[q] = x;
    }
    at += __popcll(ball);
  }
}

// Grown scratch of the fold (device), kept.
struct FoldScratch {
  int64_t *rows = nullptr, *cols = nullptr;
  double *vals = nullptr;
  unsigned char *adds = nullptr;
  unsigned long long *key = nullptr, *skey = nullptr;
  int *seq = nullptr, *sseq = nullptr, *head = nullptr, *run = nullptr, *starts = nullptr;
  unsigned *flag = nullptr;
  void *tmp = nullptr;
  size_t tmp_cap = 0;
  i64 cap = 0;
  bool ensure(i64 nw) {
    if (nw > cap) {
      for (void *q : {(void *)rows, (void *)cols, (void *)vals, (void *)adds, (void *)key, (void *)skey, (void *)seq,
                      (void *)sseq, (void *)head, (void *)run, (void *)starts})
        if (q) HIPCHK(hipFree(q));
      HIPCHK(hipMalloc(&rows, nw * sizeof(int64_t)));
      HIPCHK(hipMalloc(&cols, nw * sizeof(int64_t)));
      HIPCHK(hipMalloc(&vals, nw * sizeof(double)));
      HIPCHK(hipMalloc(&adds, nw));
      HIPCHK(hipMalloc(&key, nw * sizeof(unsigned long long)));
      HIPCHK(hipMalloc(&skey, nw * sizeof(unsigned long long)));
      HIPCHK(hipMalloc(&seq, nw * sizeof(int)));
      HIPCHK(hipMalloc(&sseq, nw * sizeof(int)));
      for (int **q : {&head, &run, &starts}) HIPCHK(hipMalloc(q, (nw + 1) * sizeof(int)));
      cap = nw;
    }
    if (!flag) HIPCHK(hipMalloc(&flag, 4 * sizeof(unsigned)));
    return !failed();
  }
  bool ensure_tmp(size_t bytes) {
    if (bytes <= tmp_cap && tmp) return true;
    if (tmp) HIPCHK(hipFree(tmp));
    HIPCHK(hipMalloc(&tmp, bytes > 0 ? bytes : 16));
    tmp_cap = bytes;
    return !failed();
  }
};
static FoldScratch &fold_scratch() {
  static FoldScratch s;
  return s;
}

}  // namespace mpk

using namespace mpk;

extern "C" {

int64_t mpk_mat_fold(int64_t rows, int64_t cols, int64_t vals, int64_t adds, int64_t nw, int64_t nrows,
                     int64_t ncols, int64_t rp, int64_t ci, int64_t v) {
  if (!ctx()) return -1;
  if (nw < 0 || nw >= ((int64_t)1 << 31) || nrows < 0 || ncols < 0 || nrows >= ((int64_t)1 << 31) ||
      ncols >= ((int64_t)1 << 31)) {
    set_error(4, "mpk_mat_fold: sizes outside the int32 index range");
    return -1;
  }
  int *RP = (int *)rp;
  if (nw == 0 || nrows == 0 || ncols == 0) {
    if (nw > 0) {
      set_error(4, "mpk_mat_fold: a write lies outside the matrix");
      return -1;
    }
    HIPCHK(hipMemsetAsync(RP, 0, (nrows + 1) * sizeof(int), stream()));
    return failed() ? -1 : 0;
  }
  FoldScratch &S = fold_scratch();
  if (!S.ensure(nw)) return -1;
  copy_htod(S.rows, (const void *)rows, nw * sizeof(int64_t));
  copy_htod(S.cols, (const void *)cols, nw * sizeof(int64_t));
  copy_htod(S.vals, (const void *)vals, nw * sizeof(double));
  copy_htod(S.adds, (const void *)adds, nw);
  HIPCHK(hipMemsetAsync(S.flag, 0, sizeof(unsigned), stream()));
  MPK_LAUNCH(k_fold_keys, nblocks(nw, 256), 256, nw, (const int64_t *)S.rows, (const int64_t *)S.cols, nrows, ncols,
             S.key, S.seq, S.flag);
  // sort by key, call position kept within a key
  const unsigned long long top = (unsigned long long)nrows * (unsigned long long)ncols - 1ull;
  int bits = 1;
  while (bits < 64 && (top >> bits) != 0ull) bits++;
  size_t tb = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, S.key, S.skey, S.seq, S.sseq, (int)nw, 0, bits, stream()));
  size_t sb = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, S.head, S.run, (int)(nw + 1), stream()));
  size_t rb = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, rb, RP, RP, (int)(nrows + 1), stream()));
  if (!S.ensure_tmp(std::max(tb, std::max(sb, rb)))) return -1;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(S.tmp, tb, S.key, S.skey, S.seq, S.sseq, (int)nw, 0, bits, stream()));
  // runs of equal keys
  MPK_LAUNCH(k_fold_heads, nblocks(nw + 1, 256), 256, nw, (const unsigned long long *)S.skey, S.head);
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(S.tmp, sb, S.head, S.run, (int)(nw + 1), stream()));
  MPK_LAUNCH(k_fold_starts, nblocks(nw + 1, 256), 256, nw, (const int *)S.head, (const int *)S.run, S.starts);
  int m = 0;
  copy_dtoh(&m, S.run + nw, sizeof m);
  unsigned flag = 0;
  copy_dtoh(&flag, S.flag, sizeof flag);
  if (failed()) return -1;
  if (flag & 1u) {
    set_error(4, "mpk_mat_fold: a write lies outside the matrix");
    return -1;
  }
  // the entries, their rows' counts, then the row starts
  HIPCHK(hipMemsetAsync(RP, 0, (nrows + 1) * sizeof(int), stream()));
  MPK_LAUNCH(k_fold_runs, nblocks(m, 256), 256, (i64)m, ncols, (const int *)S.starts,
             (const unsigned long long *)S.skey, (const int *)S.sseq, (const double *)S.vals,
             (const unsigned char *)S.adds, (int *)ci, (double *)v, RP);
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(S.tmp, rb, RP, RP, (int)(nrows + 1), stream()));
  return failed() ? -1 : (int64_t)m;
}

int64_t mpk_dense_to_csr(int64_t dense, int64_t nrows, int64_t ncols, int64_t rp, int64_t ci, int64_t v,
                         int64_t cap) {
  if (!ctx()) return -1;
  if (nrows < 0 || ncols < 0 || nrows * ncols >= ((int64_t)1 << 31)) {
    set_error(4, "mpk_dense_to_csr: more than 2^31 - 1 elements (int32 entry offsets)");
    return -1;
  }
  int *RP = (int *)rp;
  if (nrows == 0) {
    HIPCHK(hipMemsetAsync(RP, 0, sizeof(int), stream()));
    return failed() ? -1 : 0;
  }
  const double *A = (const double *)dense;
  MPK_LAUNCH(k_dense_count, nrows, 64, nrows, ncols, A, RP);
  size_t tb = 0;
  FoldScratch &S = fold_scratch();
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, RP, RP, (int)(nrows + 1), stream()));
  if (!S.ensure(1) || !S.ensure_tmp(tb)) return -1;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(S.tmp, tb, RP, RP, (int)(nrows + 1), stream()));
  int nnz = 0;
  copy_dtoh(&nnz, RP + nrows, sizeof nnz);
  if (failed()) return -1;
  if (!ci || !v || cap < nnz) return -(int64_t)nnz;
  MPK_LAUNCH(k_dense_fill, nrows, 64, ncols, A, (const int *)RP, (int *)ci, (double *)v);
  return failed() ? -1 : (int64_t)nnz;
}

}  // extern "C"
