// Deterministic, order-independent accumulation of FP64 terms in one int64.
//
// The sparse products scatter many terms into few accumulators from many
// threads (LDS atomics), so the order of FP64 additions, and with it the
// last bits of the sums, would change from run to run.  Integer addition is
// associative: each term is mapped to a fixed-point integer V = trunc(v 2^S)
// and added as one int64, the sum is the same whatever the order, and it is
// converted to FP64 once at the end.
//
// For an accumulator that receives fewer than 2^h terms, all bounded by
// |v| < 2^e:  S = 62 - h - e, so |V| < 2^(62 - h) and |sum V| < 2^62: no
// overflow.  Each term keeps 62 - h bits below the largest term's magnitude
// (truncation toward zero drops the rest): at least 50 bits for up to 4096
// terms, the resolution of an FP64 sum whose largest term is that large.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace skema {

struct FxScale {
  int S;        // scale exponent: V = trunc(v * 2^S)
  int finite;   // 0 when a term is inf or NaN (the sums are then NaN)
};

// Scale for accumulators receiving `terms` (>= 0) terms or fewer, every term
// bounded by maxabs.
__host__ __device__ inline FxScale fx_scale(double maxabs, int64_t terms) {
  FxScale f;
  f.S = 0;
  f.finite = 1;
  int h = 1;
  while (h < 40 && ((int64_t)1 << h) <= terms) ++h;  // terms < 2^h
  if (!(maxabs <= 1.7976931348623157e308)) {  // inf or NaN among the terms
    f.finite = 0;
    return f;
  }
  if (maxabs == 0.0) return f;
  int e = 0;
  (void)frexp(maxabs, &e);  // maxabs = m 2^e, 0.5 <= m < 1, so maxabs < 2^e
  f.S = 62 - h - e;
  return f;
}

// trunc(v * 2^S); |v| <= maxabs of the scale, so the result is below 2^(62-h).
__device__ inline long long fx_term(double v, const FxScale& f) { return (long long)trunc(ldexp(v, f.S)); }

// FP64 value of an accumulated sum.
__device__ inline double fx_value(long long V, const FxScale& f) {
  if (!f.finite) return __builtin_nan("");
  return ldexp((double)V, -f.S);
}

// Atomic add of V into the LDS accumulator at p.
__device__ inline void fx_add_lds(unsigned long long* p, long long V) {
  if (V) atomicAdd(p, (unsigned long long)V);
}

}  // namespace skema
