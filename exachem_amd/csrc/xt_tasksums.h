// Exact per-task energy sums of the persistent kernels (k_spatial_fp64 in
// xt_kernels.hip, k_spatial_mixed in xt_mixed.hip): thread 0 of a workgroup
// adds its boxes' partials into running sums in LDS (xt_fixed.h's modular
// addition) and adds them into the task's global pair when the workgroup
// moves to another task or leaves.  The same bits as one global add per box.
#pragma once
#include <hip/hip_runtime.h>

#include "xt_device.h"
#include "xt_fixed.h"

namespace xt {

// A partial that is not finite or beyond 2^62 Eh sets *R.err instead, and the
// host reports the run as failed.
__device__ __forceinline__ void sums_add(const RunDev& R, unsigned long long* s, double e4,
                                         double e5) {
  unsigned long long lo, hi;
  if(fx_split(e4, lo, hi)) fx_add_local(&s[0], lo, hi);
  else atomicOr(R.err, 1ull);
  if(fx_split(e5, lo, hi)) fx_add_local(&s[2], lo, hi);
  else atomicOr(R.err, 1ull);
}

// Adds a split pair into a task's global pair (the second atomic carries
// what the first overflowed).
__device__ __forceinline__ void fx_add_global(unsigned long long* acc, unsigned long long lo,
                                              unsigned long long hi) {
  const unsigned long long old   = atomicAdd(&acc[0], lo);
  const unsigned long long carry = (old + lo < old) ? 1ull : 0ull;
  atomicAdd(&acc[1], hi + carry);
}

// Adds the running sums s (E4 lo, hi, E5 lo, hi) of `task` into its global
// accumulators and clears them.
__device__ __forceinline__ void sums_flush(const RunDev& R, int task, unsigned long long* s) {
  if(task >= 0 && (s[0] | s[1] | s[2] | s[3])) {
    fx_add_global(&R.acc[4 * (size_t) task], s[0], s[1]);
    fx_add_global(&R.acc[4 * (size_t) task + 2], s[2], s[3]);
  }
  s[0] = s[1] = s[2] = s[3] = 0ull;
}

}  // namespace xt
