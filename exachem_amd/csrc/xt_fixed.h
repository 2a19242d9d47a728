// Exact 128-bit fixed-point energy sums (scale 2^64 per Eh) shared by the
// persistent kernels and the host: a per-box FP64 partial is split into a
// two's-complement (lo, hi) pair whose sum with carry over any order of
// additions is exact, so a task's energy does not depend on which workgroup
// finished first.  Plain C++ so that the host tests compile it with g++.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define XT_HD __host__ __device__
#else
#define XT_HD
#endif

namespace xt {

// Magnitude bound of one partial: the integer part must fit the signed high
// word with room for carries; the FP64 -> integer conversions below are only
// defined for finite values in range.
constexpr double kFxLimit = 4611686018427387904.0;  // 2^62

// Splits p into (lo, hi) with p = hi + lo * 2^-64 (two's complement, hi
// signed).  Returns false, leaving lo = hi = 0, for a NaN, an infinity or
// |p| >= 2^62; the caller flags the run instead of adding.
XT_HD inline bool fx_split(double p, unsigned long long& lo, unsigned long long& hi) {
  const double a = p < 0.0 ? -p : p;
  if(!(a < kFxLimit)) {  // false for NaN too
    lo = hi = 0ull;
    return false;
  }
  hi                = (unsigned long long) a;  // floor
  const double frac = a - (double) hi;         // exact
  lo                = (unsigned long long) (frac * 18446744073709551616.0);
  if(p < 0.0) {
    lo = ~lo + 1ull;
    hi = ~hi + (lo == 0ull ? 1ull : 0ull);
  }
  return true;
}

// Adds the pair (lo, hi) into w[0..1] modulo 2^128, with the carry rule of
// the kernels' global atomics.  A workgroup sums the partials of its
// consecutive boxes of one task this way and adds the sum globally once: the
// same bits as adding each partial globally, since modular addition is exact
// and associative.
XT_HD inline void fx_add_local(unsigned long long* w, unsigned long long lo, unsigned long long hi) {
  const unsigned long long l = w[0] + lo;
  w[1] += hi + (l < lo ? 1ull : 0ull);
  w[0] = l;
}

// Value of an accumulator pair (lo, hi) in Eh, converted through its
// magnitude so that the result is within a few ulp of the exact sum also for
// small negative values (hi = -1 with lo near 2^64 would otherwise cancel).
XT_HD inline double fx_value(const unsigned long long* acc) {
  unsigned long long lo = acc[0], hi = acc[1];
  const bool         neg = (long long) hi < 0;
  if(neg) {
    lo = ~lo + 1ull;
    hi = ~hi + (lo == 0ull ? 1ull : 0ull);
  }
  const double m = (double) hi + (double) lo * 5.421010862427522170037264e-20;  // 2^-64
  return neg ? -m : m;
}

}  // namespace xt
