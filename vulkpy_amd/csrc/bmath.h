// Small math helpers shared by device kernels and their host-side checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace vkh {

// sin and cos of x in [0, 8): reduction r = x - q pi/2 as one FP64 fused
// multiply-add (double(pi/2) is within 6e-17 of pi/2, so r keeps its full
// relative precision even next to the zeros of sin and cos, where a
// float-only Cody-Waite reduction cancels), then the Taylor series of sin (to
// r^9) and cos (to r^10) in FP32 on |r| <= pi/4, whose truncation errors are
// below 2^-28 relative, and the quadrant's sign / swap.  Used for
// Box-Muller's angle in [0, 2 pi), where a general-argument sincos spends
// most of its work on range reduction that is never needed.
__host__ __device__ inline void sincos_small(float x, float& s, float& c) {
  const float q = rintf(x * 0.636619772f);  // 2 / pi
  const float r = (float)fma(-(double)q, 1.5707963267948966, (double)x);
  const float r2 = r * r;
  float ps = fmaf(r2, 2.75573192e-6f, -1.98412698e-4f);  // 1/9!, -1/7!
  ps = fmaf(r2, ps, 8.33333333e-3f);                     // 1/5!
  ps = fmaf(r2, ps, -1.66666667e-1f);                    // -1/3!
  const float sr = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, -2.75573192e-7f, 2.48015873e-5f);  // -1/10!, 1/8!
  pc = fmaf(r2, pc, -1.38888889e-3f);                    // -1/6!
  pc = fmaf(r2, pc, 4.16666667e-2f);                     // 1/4!
  const float cr = fmaf(r2 * r2, pc, fmaf(r2, -0.5f, 1.0f));
  const int iq = (int)q & 3;
  const float a = (iq & 1) ? cr : sr, b = (iq & 1) ? sr : cr;
  s = (iq & 2) ? -a : a;
  c = ((iq + 1) & 2) ? -b : b;
}

// The FP32 functions below evaluate in FP64 with a relative error below 2^-40
// and round once, so they are within 0.5 ulp + 2^-16 ulp of the exact value,
// like a full FP64 libm call rounded to FP32, at a fraction of its cost (the
// FP64 tanh and log of the device library take 4x and 2x the time of
// streaming the data on MI355X).

// tanh: |x| < 2^-12: x (1 - x^2/3), whose omitted terms are below 2^-49
// relative; otherwise (e - 1) / (e + 1) with e = exp(2|x|), where e - 1
// (>= 2^-11) keeps all but 11 bits of e's precision; 1 from |x| >= 9.5 on,
// where 1 - tanh(x) < 2^-26 rounds to 1.  NaN stays NaN; the sign of zero
// is kept.
__host__ __device__ inline float tanh_f32(float x) {
  const double d = (double)x, a = fabs(d);
  if (!(a < 9.5)) return a == a ? copysignf(1.0f, x) : x;
  if (a < 0x1p-12) return (float)(d * fma(-d * d, 1.0 / 3.0, 1.0));
  const double e = exp(2.0 * a);
  return copysignf((float)((e - 1.0) / (e + 1.0)), x);
}

// atanh(s) = s + s^3/3 + ... + s^15/15 for |s| <= 0.1716, whose omitted
// terms are below 2^-45 relative
__host__ __device__ inline double atanh_series(double s) {
  const double z = s * s;
  double p = fma(z, 1.0 / 15.0, 1.0 / 13.0);
  p = fma(z, p, 1.0 / 11.0);
  p = fma(z, p, 1.0 / 9.0);
  p = fma(z, p, 1.0 / 7.0);
  p = fma(z, p, 1.0 / 5.0);
  p = fma(z, p, 1.0 / 3.0);
  return fma(s * z, p, s);
}

// Natural log before rounding: with x = m 2^k, m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s) for s = (m - 1) / (m + 1), |s| <= 0.1716.  k_out
// receives k and the result is log m; x = 1 gives exactly 0.  x must be
// positive and finite.
__host__ __device__ inline double log_mant(double x, int& k_out) {
  int k;
  double m = frexp(x, &k);  // m in [0.5, 1)
  if (m < 0.70710678118654752) {
    m *= 2.0;
    --k;
  }
  k_out = k;
  return 2.0 * atanh_series((m - 1.0) / (m + 1.0));
}

// log of a positive finite double, FP64 error below 2^-44 relative
__host__ __device__ inline double log_pos(double x) {
  int k;
  const double lm = log_mant(x, k);
  return fma((double)k, 0.69314718055994531, lm);
}

// log and log2 of FP32 x through log_mant: k ln 2 + log m and k + log m /
// ln 2, each with one rounding in FP64 before the FP32 rounding; exact at
// x = 1 (and log2 at every power of two).  0 -> -inf, negative -> NaN, +inf
// and NaN pass through.
__host__ __device__ inline float log_f32(float x) {
  if (!(x > 0.0f) || x == INFINITY) return x == 0.0f ? -INFINITY : (x < 0.0f ? NAN : x);
  int k;
  const double lm = log_mant(x, k);
  return (float)fma((double)k, 0.69314718055994531, lm);
}

__host__ __device__ inline float log2_f32(float x) {
  if (!(x > 0.0f) || x == INFINITY) return x == 0.0f ? -INFINITY : (x < 0.0f ? NAN : x);
  int k;
  const double lm = log_mant(x, k);
  return (float)fma(lm, 1.4426950408889634, (double)k);
}

// a ** b (C pow semantics) for FP32 operands.  A finite positive base with a
// finite exponent takes exp(b log a) in FP64: log a from log_mant (relative
// error below 2^-45) gives b log a with an absolute error below 2^-37 for
// every |b log a| < 150, beyond which the FP32 result is 0 or inf either way,
// so the result is within 0.5 ulp + 2^-13 ulp.  Every other case (zero,
// negative or infinite base, non-finite exponent) is the FP64 pow rounded.
__host__ __device__ inline float pow_f32(float a, float b) {
  if (a > 0.0f && a < INFINITY && fabsf(b) < INFINITY) {
    int k;
    const double lm = log_mant(a, k);
    return (float)exp((double)b * fma((double)k, 0.69314718055994531, lm));
  }
  return (float)pow((double)a, (double)b);
}

// sinh: |x| < 2^-5: x (1 + x^2/6 + x^4/120 + x^6/5040), omitted terms below
// 2^-45 relative; otherwise (e - 1/e) / 2 with e = exp(|x|), where the
// difference (>= 2^-4) keeps all but 5 bits of e's precision.  Overflow to
// inf, NaN and the sign of zero follow the FP64 function.
__host__ __device__ inline float sinh_f32(float x) {
  const double d = (double)x, a = fabs(d);
  if (a < 0x1p-5) {
    const double z = d * d;
    return (float)(d * fma(z, fma(z, fma(z, 1.0 / 5040.0, 1.0 / 120.0), 1.0 / 6.0), 1.0));
  }
  const double e = exp(a);
  return copysignf((float)(0.5 * (e - 1.0 / e)), x);
}

// cosh: (e + 1/e) / 2 with e = exp(|x|) (no cancellation)
__host__ __device__ inline float cosh_f32(float x) {
  const double e = exp(fabs((double)x));
  return (float)(0.5 * (e + 1.0 / e));
}

// asinh: |x| < 2^-5: x (1 - x^2/6 + 3 x^4/40 - 5 x^6/112), omitted terms
// below 2^-44 relative; otherwise log(|x| + sqrt(x^2 + 1)) (x^2 fits FP64
// for every FP32 x), whose argument is at least 1 + 2^-5.
__host__ __device__ inline float asinh_f32(float x) {
  const double d = (double)x, a = fabs(d);
  if (a < 0x1p-5) {
    const double z = d * d;
    return (float)(d * fma(z, fma(z, fma(z, -5.0 / 112.0, 3.0 / 40.0), -1.0 / 6.0), 1.0));
  }
  if (!(a < INFINITY)) return x;  // inf, NaN
  return copysignf((float)log_pos(a + sqrt(fma(a, a, 1.0))), x);
}

// acosh (x >= 1): log(x + sqrt(x^2 - 1)); x^2 - 1 is exact in FP64 for FP32
// x, so near x = 1 the result keeps its relative precision.  x < 1 and NaN
// give NaN, inf gives inf.
__host__ __device__ inline float acosh_f32(float x) {
  const double d = (double)x;
  if (!(d >= 1.0)) return NAN;
  if (d == INFINITY) return x;
  return (float)log_pos(d + sqrt(fma(d, d, -1.0)));
}

// atanh (|x| <= 1): |x| <= 0.1716: the series; otherwise
// log((1 + |x|) / (1 - |x|)) / 2, where 1 +- |x| are exact in FP64 and the
// log is at least 0.34.  +-1 give +-inf, |x| > 1 and NaN give NaN.
__host__ __device__ inline float atanh_f32(float x) {
  const double d = (double)x, a = fabs(d);
  if (a <= 0.1716) return (float)atanh_series(d);
  if (!(a <= 1.0)) return NAN;
  if (a == 1.0) return copysignf(INFINITY, x);
  return copysignf((float)(0.5 * log_pos((1.0 + a) / (1.0 - a))), x);
}

}  // namespace vkh
