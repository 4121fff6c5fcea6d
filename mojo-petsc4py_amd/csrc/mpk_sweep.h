// mpk_sweep.h -- sync-free triangular sweeps.
//
// A sweep computes rows in dependency order (forward: row i needs rows < i;
// backward: rows > i).  One wave owns 64 consecutive rows of the sweep order;
// each lane computes one row as a serial sum over the row's entries in stored
// order, so the arithmetic of every row is the CPU's.  Rows publish their
// value in `pub`, which is filled with a signalling-NaN sentinel before the
// sweep: one 64-bit relaxed atomic load both tests readiness and fetches the
// value, so no fence orders a flag against the data.  A dependency inside the
// lane's own wave is read from LDS instead.  Blocks take their position in the
// sweep from a ticket counter, so every block a block waits on has already
// started -- there is no deadlock however the hardware orders dispatch.
#pragma once

#include "mpk_internal.h"

namespace mpk {

#define MPK_SENTINEL 0x7FF4D5A17B0E0000ull  // signalling NaN; arithmetic never yields it
#define MPK_QNAN 0x7FF8000000000000ull

__device__ __forceinline__ unsigned long long load_pub(const double *pub, i64 k) {
  return __hip_atomic_load(reinterpret_cast<const unsigned long long *>(pub + k), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_pub(double *pub, i64 k, double v) {
  unsigned long long b = __double_as_longlong(v);
  if (b == MPK_SENTINEL) b = MPK_QNAN;  // a NaN keeps being a NaN
  __hip_atomic_store(reinterpret_cast<unsigned long long *>(pub + k), b, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// Row programs.  Each provides, for sweep row i:
//   init(i) -> accumulator start
//   begin(i), end(i) -> entry range
//   col(j), coef(j) -> dependency row and coefficient; acc = fma(coef, v, acc)
//   finish(i, acc) -> published value (and any side outputs)
//   stale(k) -> value of a row that is not earlier in the sweep (malformed input)

// Generic driver: 64-thread blocks (one wave).
template <class Prog>
__global__ void __launch_bounds__(64) k_sweep(Prog prog, i64 n, int forward, unsigned *ticket,
                                              unsigned ticket_base, double *pub, const int64_t *stop) {
  __shared__ unsigned blk_sh;
  __shared__ double lv[64];
  if (threadIdx.x == 0) blk_sh = atomicAdd(ticket, 1u) - ticket_base;
  __syncthreads();
  if (stop && *stop) return;
  const i64 L0 = (i64)blk_sh * 64;
  const int lane = threadIdx.x;
  const i64 L = L0 + lane;
  const bool valid = L < n;
  const i64 i = valid ? (forward ? L : n - 1 - L) : 0;
  bool done = !valid;
  int j = 0, je = 0;
  double acc = 0.0;
  if (valid) {
    acc = prog.init(i);
    j = prog.begin(i);
    je = prog.end(i);
  }
  unsigned long long mask = 0;  // lanes of this wave that have published
  for (;;) {
    bool moved = false;
    if (!done) {
      const int j_in = j;
      while (j < je) {
        const i64 k = prog.col(j);
        const i64 Lk = forward ? k : n - 1 - k;
        double v;
        if (Lk >= L) {
          // not earlier in the sweep: only a malformed matrix (an SOR row
          // without its diagonal) reaches this; it reads the pre-sweep value
          v = prog.stale(k);
        } else if (Lk >= L0) {
          if (!((mask >> (Lk - L0)) & 1ull)) break;
          v = lv[Lk - L0];
        } else {
          unsigned long long b = load_pub(pub, k);
          if (b == MPK_SENTINEL) break;
          v = __longlong_as_double((long long)b);
        }
        acc = fma(prog.coef(j), v, acc);
        j++;
      }
      if (j >= je) {
        double r = prog.finish(i, acc);
        lv[lane] = r;
        store_pub(pub, i, r);
        done = true;
      }
      moved = j != j_in || done;
    }
    __syncthreads();  // one wave: orders the LDS writes before the next reads
    const unsigned long long newmask = __ballot(done);
    if (newmask == ~0ull) break;
    // nothing new in this wave: the rows it waits on belong to earlier waves
    if (newmask == mask && __ballot(moved) == 0ull) __builtin_amdgcn_s_sleep(2);
    mask = newmask;
  }
}

}  // namespace mpk
