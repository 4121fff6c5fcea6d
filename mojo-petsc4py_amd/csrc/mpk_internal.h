// mpk_internal.h -- shared pieces of the MI355X Krylov library (libmpk.so).
//
// Conventions used by every source file:
//  * one HIP context per process: the current device and one non-blocking
//    stream at default priority; every kernel goes on that stream, so kernels
//    never overlap and each one sees everything the previous one wrote;
//  * vectors, matrices and factors live in device memory; scalar results and
//    control blocks cross the C ABI through host memory;
//  * index arrays on the device are int32 (row pointers and column indices);
//  * all arithmetic is IEEE binary64 with explicit fma() where the reference
//    expression contracts (the library is built with -ffp-contract=off), so an
//    elementwise kernel or a serial row sum gives the same bits as the CPU;
//  * reductions use the fixed tile-tree order below, which depends only on n.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

typedef int64_t i64;

#define MPK_TILE 2048  // elements per reduction tile
#define MPK_NT 256     // threads per block in tile kernels

namespace mpk {

struct Ctx {
  int device = -1;
  hipStream_t stream = nullptr;
  double *partials = nullptr;  // reduction partials, grown on demand
  size_t partials_cap = 0;     // in doubles
  double *dres = nullptr;      // device scalar results (1024 doubles)
  double *hres = nullptr;      // pinned host mirror of dres
  int *flags = nullptr;        // per-row ready flags for the sync-free sweeps
  size_t flags_cap = 0;        // in ints
  unsigned *tickets = nullptr; // block tickets for the sync-free sweeps (64)
  double *dtmp = nullptr;      // device scratch vector, grown on demand
  size_t dtmp_cap = 0;
  double *fixed = nullptr;     // the fixed scratch array of the no-allocation entry point
  size_t fixed_cap = 0;        // in doubles
  unsigned *watch = nullptr;   // fine-grained host words of the sweep waits: [0] abort, [1] fault
};

// The sweep waits' watch: its host words, the polls a wait makes before it
// gives up, and a generation that changes whenever either changes (each
// source file that launches polling kernels copies it to its device side
// when the generation moved; mpk_sweep.h).
void sweep_watch_state(unsigned **word, unsigned *limit, unsigned *gen);

// The context if it is initialised already, else nullptr; never initialises
// (so never allocates).
Ctx *ctx_ready();

// The process context, initialised on first use on the current device.
// Returns nullptr (and records the error) when no HIP device is usable.
Ctx *ctx();
hipStream_t stream();

// Error state: the C ABI cannot throw, so failures are recorded here and read
// back through mpk_status()/mpk_last_error().
void set_error(int code, const char *fmt, ...);
bool failed();

double *partials(size_t ndoubles);
// Host <-> device copies ordered after every kernel queued on the library's
// stream (a non-blocking stream, which plain hipMemcpy does not wait for);
// they return when the copy is done.
void copy_dtoh(void *dst, const void *src, size_t bytes);
void copy_htod(void *dst, const void *src, size_t bytes);
int *flag_buffer(size_t nints);
double *scratch_vector(size_t ndoubles);

inline unsigned nblocks(i64 n, i64 per) { return (unsigned)((n + per - 1) / per); }
inline unsigned ntiles(i64 n) { return nblocks(n, MPK_TILE); }

// Synchronise the stream and record any asynchronous error.
void sync();

// Reduction order: false (the default) for the tile-tree order below, true
// for the reference's SIMD order (mpk_reduce.h).  Set by
// mpk_set_reduction_order(); MPK_REDUCTION=ref|dev gives the initial value.
bool red_ref();

}  // namespace mpk

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      ::mpk::set_error(1, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

#define MPK_LAUNCH(kernel, grid, block, ...)                                        \
  do {                                                                              \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, ::mpk::stream(), __VA_ARGS__); \
    HIPCHK(hipGetLastError());                                                      \
  } while (0)

// ------------------------------------------------------------------ device side

namespace mpk {

// Two consecutive doubles; loaded as one 16-byte access when aligned.
struct D2 {
  double a, b;
};

__device__ __forceinline__ D2 ld2(const double *p, i64 e, i64 n, bool al) {
  D2 r;
  if (e + 1 < n) {
    if (al) {
      double2 v = *reinterpret_cast<const double2 *>(p + e);
      r.a = v.x;
      r.b = v.y;
    } else {
      r.a = p[e];
      r.b = p[e + 1];
    }
  } else {
    // e == n - 1 or past the end.  A relaxed atomic load (a plain load at
    // workgroup scope), so that the compiler cannot merge it with the other
    // branch's load of p[e]: the merge would split the aligned 16-byte access
    // in two.
    r.a = e < n ? __hip_atomic_load(p + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0.0;
    r.b = 0.0;
  }
  return r;
}

__device__ __forceinline__ void st2(double *p, i64 e, i64 n, bool al, D2 v) {
  if (e + 1 < n) {
    if (al) {
      *reinterpret_cast<double2 *>(p + e) = make_double2(v.a, v.b);
    } else {
      p[e] = v.a;
      p[e + 1] = v.b;
    }
  } else if (e < n) {
    p[e] = v.a;
  }
}

__host__ __device__ __forceinline__ bool aligned16(const void *p) {
  return (reinterpret_cast<uintptr_t>(p) & 15) == 0;
}

// The hardware hands block b of a launch to XCD b % 8; this maps block b to a
// tile index so that each XCD sweeps one contiguous run of tiles (neighbouring
// tiles share data through that XCD's L2).
#define MPK_XCDS 8
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
  unsigned per = nb / MPK_XCDS, rem = nb % MPK_XCDS, x = b % MPK_XCDS, k = b / MPK_XCDS;
  return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

// Loads of the sliced matrix stream and stores of the product: non-temporal
// (NT) for matrices well past the last-level cache, so neither stream pushes
// out of L2 the lines of x that rows a grid plane apart read again; ordinary
// otherwise, so a matrix that fits the cache stays resident across products.
template <bool NT, class T>
__device__ __forceinline__ T ld_stream(const T *p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

template <bool NT, class T>
__device__ __forceinline__ void st_stream(T *p, T v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// Columns of a sliced copy.  The column array holds the int32 column of every
// slot; when every entry lies within 32767 columns of its row it is followed
// by the same slots as 16-bit offsets from the row, and the last slice offset
// soff[nslices] holds -1 - (padded slots) instead of the slot count.  The
// products then read the offsets: 2 bytes a slot instead of 4.
struct SellCols32 {
  const int *c;
  template <bool NT>
  __device__ __forceinline__ int col(int q, int) const {
    return ld_stream<NT>(c + q);
  }
};
struct SellCols16 {
  const short *d;
  template <bool NT>
  __device__ __forceinline__ int col(int q, int i) const {
    return i + (int)ld_stream<NT>(d + q);
  }
};
// The 16-bit offsets of a sliced copy with nrow rows, or null (int32 columns).
__device__ __forceinline__ const short *sell_offsets16(const int *soff, const int *sc, i64 nrow) {
  const int last = soff[(nrow + 63) >> 6];
  return last < 0 ? reinterpret_cast<const short *>(sc + (-1 - (i64)last)) : nullptr;
}

// One row (row i) of a sliced product: entries base + 64 k, k < len, summed in
// order onto s (fma(v, x[c], s), k = 0, 1, ...).  The values and columns of
// four entries are loaded before their gathers of x, so each lane keeps
// several loads in flight; the arithmetic order is the serial one.
template <bool NT, class Cols>
__device__ __forceinline__ double sell_row(const double *__restrict__ sv, const Cols &cols,
                                           const double *__restrict__ x, int base, int len, double s, i64 i) {
  for (int k = 0; k < len; k += 4) {
    int c[4];
    double v[4], xv[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (k + j < len) {
        c[j] = cols.template col<NT>(base + 64 * (k + j), (int)i);
        v[j] = ld_stream<NT>(sv + base + 64 * (k + j));
      }
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (k + j < len) xv[j] = x[c[j]];
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (k + j < len) s = fma(v[j], xv[j], s);
  }
  return s;
}

// Lane i + n's value within lane i's row of 16 (DPP row_shl:n; lanes whose
// i + n leaves the row read 0.0), both halves of the double.
template <int N>
__device__ __forceinline__ double row_shl_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x100 + N, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x100 + N, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Halving tree over the 64 lanes of a wave: u[i] += u[i + w], w = 32..1.
// Lane 0 ends with the wave's sum; the other lanes hold partial values.  At
// step w only lanes i < 2w still feed lane 0, so for w <= 8 (i + w < 16) a
// row shift gives every lane that matters its operand: the lane-crossing
// permute is needed for w = 32 and 16 only.
__device__ __forceinline__ double wave_tree(double u) {
  u = u + __shfl_down(u, 32, 64);
  u = u + __shfl_down(u, 16, 64);
  u = u + row_shl_d<8>(u);
  u = u + row_shl_d<4>(u);
  u = u + row_shl_d<2>(u);
  u = u + row_shl_d<1>(u);
  return u;
}

// Block tree over 256 threads: wave trees, then (W0 + W2) + (W1 + W3).
// Returns the sum in thread 0; `sh` is 4 doubles of shared memory.
__device__ __forceinline__ double block_tree(double u, double *sh) {
  u = wave_tree(u);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = u;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (sh[0] + sh[2]) + (sh[1] + sh[3]);
  __syncthreads();
  return r;
}

}  // namespace mpk
