// Exact, order-independent accumulation of FP64 terms in two int64 words.
//
// The sparse products scatter many terms into few accumulators from many
// threads (LDS atomics), so the order of FP64 additions, and with it the
// last bits of the sums, would change from run to run.  Integer addition is
// associative: each term is mapped exactly to a fixed-point integer
// V = v 2^S and added as two int64 words, the result is the same whatever
// the order, and it is converted to FP64 once at the end.
//
// For an accumulator that receives fewer than 2^h terms, all bounded by
// |v| < 2^e:  B = 63 - h, S = 2B - 1 - e, so |V| < 2^(2B-1);  V = H 2^B + L
// with H = trunc(V / 2^B) (|H| < 2^(B-1)) and |L| < 2^B (same sign as V).
// Fewer than 2^h terms keep |sum H| < 2^62 and |sum L| < 2^63: no overflow.
// Bits of V below 2^0 are dropped (truncation toward zero), i.e. each term
// keeps 2B - 1 >= 2 (63 - h) - 1 bits below the largest term's magnitude:
// more than 80 bits for any h <= 20, far beyond the 53 of an FP64 sum.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace skema {

struct FxScale {
  int S;        // scale exponent: V = v * 2^S
  int B;        // split: V = H * 2^B + L
  int finite;   // 0 when a term is inf or NaN (the sums are then NaN)
};

// Scale for accumulators receiving fewer than 2^h terms with h from `terms`
// (the number of terms, >= 0), every term bounded by maxabs.
__host__ __device__ inline FxScale fx_scale(double maxabs, int64_t terms) {
  FxScale f;
  f.S = 0;
  f.finite = 1;
  int h = 1;
  while (h < 40 && ((int64_t)1 << h) <= terms) ++h;
  f.B = 63 - h;
  if (!(maxabs <= 1.7976931348623157e308)) {  // inf or NaN among the terms
    f.finite = 0;
    return f;
  }
  if (maxabs == 0.0) {
    f.S = 0;
    return f;
  }
  int e = 0;
  (void)frexp(maxabs, &e);  // maxabs = m 2^e, 0.5 <= m < 1, so maxabs < 2^e
  f.S = 2 * f.B - 1 - e;
  return f;
}

// Exact split of v * 2^S (|v * 2^S| < 2^(2B-1)).
__device__ inline void fx_split(double v, const FxScale& f, long long& H, long long& L) {
  const double V = ldexp(v, f.S);
  const double Hd = trunc(ldexp(V, -f.B));
  const double Ld = V - ldexp(Hd, f.B);  // exact: |Ld| < 2^B, its bits are a subset of V's
  H = (long long)Hd;
  L = (long long)Ld;  // drops the bits of V below 2^0
}

// FP64 value of an accumulated (sum H, sum L) pair.
__device__ inline double fx_value(long long H, long long L, const FxScale& f) {
  if (!f.finite) return __builtin_nan("");
  return ldexp((double)H, f.B - f.S) + ldexp((double)L, -f.S);
}

// Atomic add of (h, l) into the LDS pair at p[0], p[1].
__device__ inline void fx_add_lds(unsigned long long* p, long long h, long long l) {
  if (h) atomicAdd(p, (unsigned long long)h);
  if (l) atomicAdd(p + 1, (unsigned long long)l);
}

}  // namespace skema
