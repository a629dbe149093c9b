// FP64 division by a denominator whose reciprocal is known ahead, with the
// exact bits of the compiler's IEEE division sequence (v_div_scale, v_rcp,
// two Newton steps, v_div_fmas, v_div_fixup) whenever that sequence does not
// rescale; callers check the bounds below and fall back to '/' otherwise.
#pragma once

#include <hip/hip_runtime.h>

namespace npbx {

// The reciprocal that the compiler's IEEE FP64 division sequence computes
// from an unscaled denominator: the hardware estimate refined by two Newton
// steps.  With it the sequence ends m = n y, r = fma(-den, m, n),
// q = fma(r, y, m) (its div_fmas and div_fixup steps return their operand
// when nothing was rescaled and the result is a normal number).
__device__ __forceinline__ double division_reciprocal(double den) {
  double y = __builtin_amdgcn_rcp(den);
  double e = __builtin_fma(-den, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-den, y, 1.0);
  return __builtin_fma(y, e, y);
}

__device__ __forceinline__ double short_division(double n, double den, double y, double& m) {
  m = n * y;
  const double r = __builtin_fma(-den, m, n);
  return __builtin_fma(r, y, m);
}

__device__ __forceinline__ int exponent_of(double x) {
  return (__double2hiint(x) >> 20) & 0x7ff;
}

// den's side of the no-rescale condition: a normal number whose reciprocal
// is normal, with room for the numerator bounds below.
__device__ __forceinline__ bool den_in_range(double den) {
  const int e = exponent_of(den);
  return e >= 100 && e <= 1900;
}

// The numerator's side, for dens in range of magnitude dmin or more: the
// division sequence rescales when n is zero, tiny (biased exponent <= 53),
// not finite, at least 2^768 times den (in exponent), or when the quotient
// is subnormal.  With m = n y, the quotient to within a few units of its
// last place, lo <= |m| < hi excludes all of them:
//   hi = 2^765: |n / den| < 2^766, so the exponents differ by at most 766;
//   lo = max(2^-1020, 2^-960 / dmin): the quotient is normal, and
//        |n| = |m den| >= 2^-960 (biased exponent above 53).
// A NaN fails both comparisons, an infinite n gives an infinite m.
__device__ __forceinline__ void quotient_bounds(double dmin, double& lo, double& hi) {
  lo = fmax(0x1p-1020, 0x1p-960 / dmin);
  hi = 0x1p765;
}

__device__ __forceinline__ bool unscaled_division(double m, double lo, double hi) {
  const double a = fabs(m);
  return a >= lo && a < hi;
}

// A wave-uniform double (the value of the first active lane), so that values
// every lane computes alike live in scalar registers.
__device__ __forceinline__ double uniform_double(double x) {
  return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(x)),
                          __builtin_amdgcn_readfirstlane(__double2loint(x)));
}

// A uniform divisor and its reciprocal.  prepare_divisor narrows [lo, hi),
// the |m| range in which the short form is exact for every divisor prepared
// into it (empty when a divisor is outside den_in_range).
struct Divisor {
  double d, y;
};

__device__ __forceinline__ Divisor prepare_divisor(double d, double& lo, double& hi) {
  double dlo, dhi;
  quotient_bounds(fabs(d), dlo, dhi);
  const bool fast = den_in_range(d);
  lo = fmax(lo, fast ? dlo : INFINITY);
  hi = fmin(hi, dhi);
  return {d, uniform_double(fast ? division_reciprocal(d) : 0.0)};
}

// Division policies for code that computes a chunk of quotients at once.
// ShortDiv is the short form alone and tracks the smallest and largest |m|
// it saw; FullDiv is the division operator.
struct ShortDiv {
  double mn = INFINITY, mx = 0.0;
  __device__ __forceinline__ double operator()(double n, const Divisor& v) {
    double m;
    const double q = short_division(n, v.d, v.y, m);
    mn = fmin(mn, fabs(m));
    mx = fmax(mx, fabs(m));
    return q;
  }
};
struct FullDiv {
  __device__ __forceinline__ double operator()(double n, const Divisor& v) { return n / v.d; }
};

// out[c] = f(c, div) for c < K with the short division; a chunk whose |m|
// left [dv.lo, dv.hi) anywhere is redone with the division operator.  With
// no branch on the common path the chunk's loads issue together.
template <int K, class D, class F>
__device__ __forceinline__ void chunk_eval(const D& dv, double (&out)[K], F&& f) {
  ShortDiv sd;
#pragma unroll
  for (int c = 0; c < K; ++c) out[c] = f(c, sd);
  if (!(sd.mn >= dv.lo && sd.mx < dv.hi)) {
    FullDiv fd;
#pragma unroll
    for (int c = 0; c < K; ++c) out[c] = f(c, fd);
  }
}

}  // namespace npbx
