// Small math helpers shared by device kernels and their host-side checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

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
// like a full FP64 libm call rounded to FP32, at a fraction of its cost: none
// calls an FP64 library transcendental on its common path (the device
// library's FP64 tan, asin, tanh or log take 2x to 4x the time of streaming
// the data on MI355X).  The cores are polynomials in FP64 after an exact or
// near-exact argument reduction; every constant is either 1/n!, 1/(2n+1), a
// value of pi / 2 / ln 2 / atan split into pieces, or the coefficients of a
// Chebyshev fit made for this file (asin below).

// ---- pieces of constants ----------------------------------------------------
// pi/2 = PIO2_1 + PIO2_2 + PIO2_3 to 2^-120: the first two have 33
// significant bits, so q * PIO2_1 and q * PIO2_2 are exact for |q| < 2^20.
constexpr double PIO2_1 = 0x1.921fb54400000p+0;
constexpr double PIO2_2 = 0x1.0b4611a600000p-34;
constexpr double PIO2_3 = 0x1.3198a2e037073p-69;
constexpr double PIO2 = 1.5707963267948966, PI = 3.141592653589793;
constexpr double TWO_OVER_PI = 0.6366197723675814;
// ln 2 = LN2_HI + LN2_LO, LN2_HI with 32 significant bits (k * LN2_HI exact
// for |k| < 2^21)
constexpr double LN2_HI = 0x1.62e42fee00000p-1;
constexpr double LN2_LO = 0x1.a39ef35793c76p-33;
constexpr double LN2 = 0.6931471805599453, LOG2E = 1.4426950408889634;

// ---- reciprocals -------------------------------------------------------------
// 1/b for a nonzero finite b: the hardware's approximate FP64 reciprocal
// (device) then two Newton steps, each squaring the relative error, which
// leaves it within an FP64 rounding of 1/b; the host computes the quotient
// directly.  Cheaper on the device than the correctly rounded FP64 division
// sequence, and accurate far beyond what the FP32 results need.
__device__ inline double recip_seed(double b) { return __builtin_amdgcn_rcp(b); }
__host__ inline double recip_seed(double b) { return 1.0 / b; }

__host__ __device__ inline double recip(double b) {
  double r = recip_seed(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  return fma(r, e, r);
}

// 1/sqrt(x) for a positive normal double the same way: approximate
// reciprocal square root, two Newton steps y += y (1 - x y^2) / 2
__device__ inline double rsqrt_seed(double x) { return __builtin_amdgcn_rsq(x); }
__host__ inline double rsqrt_seed(double x) { return 1.0 / sqrt(x); }

__host__ __device__ inline double rsqrt_pos(double x) {
  double y = rsqrt_seed(x);
  double e = fma(-(x * y), y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-(x * y), y, 1.0);
  return fma(0.5 * y, e, y);
}

// sqrt x for finite x that is 0 or not below the smallest normal double:
// x * rsqrt_pos(x) (relative error below 2^-49), x for +-0, NaN for negative
// x and NaN.  The FP64 square root the compiler emits for sqrt() costs twice
// the instructions (scaling for denormals and its own refinement); the
// callers below never pass a positive denormal or an infinity.
__host__ __device__ inline double sqrt_nonneg(double x) {
  return x > 0.0 ? x * rsqrt_pos(x) : (x == 0.0 ? x : NAN);
}

// ---- FP64 fused multiply-adds with a constant operand ---------------------------
// A 64-bit constant is not an inline operand of a vector instruction, so the
// compiler copies it into vector registers, two moves per use, which cost as
// much issue time as the FMA itself.  These name it as an operand held in
// scalar registers instead (filled by scalar moves, which issue beside the
// vector instructions): one v_fma_f64 (v_mul_f64) in all, with the IEEE result
// of fma(a, b, c) (a * b).  The operand marked k must be the same in every
// lane (a literal constant).  Host: the standard functions.
__device__ inline double fma_kc(double a, double b, double c) {  // c constant
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
  return r;
}
__host__ inline double fma_kc(double a, double b, double c) { return fma(a, b, c); }
__device__ inline double fma_kb(double a, double b, double c) {  // b constant
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
  return r;
}
__host__ inline double fma_kb(double a, double b, double c) { return fma(a, b, c); }
// -a * b + c, b constant
__device__ inline double fnma_kb(double a, double b, double c) {
  double r;
  asm("v_fma_f64 %0, -%1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
  return r;
}
__host__ inline double fnma_kb(double a, double b, double c) { return fma(-a, b, c); }
__device__ inline double mul_kb(double a, double b) {  // b constant
  double r;
  asm("v_mul_f64 %0, %1, %2" : "=v"(r) : "v"(a), "s"(b));
  return r;
}
__host__ inline double mul_kb(double a, double b) { return a * b; }

// ---- exp ----------------------------------------------------------------------
// e^t for |t| <= 700: t = k ln 2 + r with k = rint(t / ln 2) and |r| <= ln 2 / 2
// (two-piece ln 2: r is exact to 2^-85 relative), then the Taylor polynomial
// of e^r to r^11 (truncation below 2^-47 relative) and the exact scaling by
// 2^k.  Measured on MI355X against the alternatives for 2^27 floats: as fast
// as the FP64 library exp, and faster than a 2^(j/64) table with a degree-5
// polynomial (whose per-lane table reads cost more than the terms saved).
__host__ __device__ inline double exp_core(double t) {
  const double k = rint(mul_kb(t, LOG2E));
  double r = fnma_kb(k, LN2_HI, t);
  r = fnma_kb(k, LN2_LO, r);
  double p = fma_kb(r, 1.0 / 39916800.0, 1.0 / 3628800.0);  // 1/11!, 1/10!
  p = fma_kc(p, r, 1.0 / 362880.0);
  p = fma_kc(p, r, 1.0 / 40320.0);
  p = fma_kc(p, r, 1.0 / 5040.0);
  p = fma_kc(p, r, 1.0 / 720.0);
  p = fma_kc(p, r, 1.0 / 120.0);
  p = fma_kc(p, r, 1.0 / 24.0);
  p = fma_kc(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

// e^t rounded to FP32: 0 below -104 (e^-104 < 2^-150, half the smallest
// denormal), inf above 89 (e^89 > FP32 max); NaN stays NaN
__host__ __device__ inline float exp_to_f32(double t) {
  if (!(t <= 89.0)) return t != t ? (float)t : INFINITY;
  if (t < -104.0) return 0.0f;
  return (float)exp_core(t);
}

__host__ __device__ inline float exp_f32(float x) { return exp_to_f32((double)x); }

// 2^x: x ln 2 in FP64 carries an absolute error below 2^-46 for |x| < 150
__host__ __device__ inline float exp2_f32(float x) { return exp_to_f32((double)x * LN2); }

// ---- log ----------------------------------------------------------------------
// atanh(s) = s + s^3/3 + ... + s^15/15 for |s| <= 0.1716, whose omitted
// terms are below 2^-45 relative
__host__ __device__ inline double atanh_series(double s) {
  const double z = s * s;
  double p = fma_kb(z, 1.0 / 15.0, 1.0 / 13.0);
  p = fma_kc(z, p, 1.0 / 11.0);
  p = fma_kc(z, p, 1.0 / 9.0);
  p = fma_kc(z, p, 1.0 / 7.0);
  p = fma_kc(z, p, 1.0 / 5.0);
  p = fma_kc(z, p, 1.0 / 3.0);
  return fma(s * z, p, s);
}

// Natural log before rounding: x = m 2^k with m in [sqrt(1/2), sqrt(2)) read
// from the bit pattern, log m = 2 atanh(s) for s = (m - 1) / (m + 1),
// |s| <= 0.1716.  k_out receives k and the result is log m; x = 1 gives
// exactly 0.  x must be a positive normal double (every positive finite FP32
// value is).
__host__ __device__ inline double log_mant(double x, int& k_out) {
  uint64_t bits;
  memcpy(&bits, &x, 8);
  int k = (int)(bits >> 52) - 1023;
  bits = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;  // m in [1, 2)
  double m;
  memcpy(&m, &bits, 8);
  if (m > 1.4142135623730951) {
    m *= 0.5;
    ++k;
  }
  k_out = k;
  return 2.0 * atanh_series((m - 1.0) * recip(m + 1.0));
}

// log of a positive normal double, FP64 error below 2^-44 relative
__host__ __device__ inline double log_pos(double x) {
  int k;
  const double lm = log_mant(x, k);
  return fma_kb((double)k, LN2_HI, fma_kb((double)k, LN2_LO, lm));
}

// log and log2 of FP32 x through log_mant: k ln 2 + log m and k + log m /
// ln 2, each rounded in FP64 before the FP32 rounding; exact at x = 1 (and
// log2 at every power of two).  0 -> -inf, negative -> NaN, +inf and NaN pass
// through.
__host__ __device__ inline float log_f32(float x) {
  if (!(x > 0.0f) || x == INFINITY) return x == 0.0f ? -INFINITY : (x < 0.0f ? NAN : x);
  return (float)log_pos((double)x);
}

__host__ __device__ inline float log2_f32(float x) {
  if (!(x > 0.0f) || x == INFINITY) return x == 0.0f ? -INFINITY : (x < 0.0f ? NAN : x);
  int k;
  const double lm = log_mant(x, k);
  return (float)fma_kb(lm, LOG2E, (double)k);
}

// a ** b (C pow semantics) for FP32 operands.  A finite positive base with a
// finite exponent takes exp(b log a) in FP64: log a with a relative error
// below 2^-44 gives b log a with an absolute error below 2^-36 for every
// |b log a| < 105, beyond which the FP32 result is 0 or inf either way, so
// the result is within 0.5 ulp + 2^-12 ulp.  Every other case (zero,
// negative or infinite base, non-finite exponent) is the FP64 pow rounded.
__host__ __device__ inline float pow_f32(float a, float b) {
  if (a > 0.0f && a < INFINITY && fabsf(b) < INFINITY) return exp_to_f32((double)b * log_pos((double)a));
  return (float)pow((double)a, (double)b);
}

// ---- hyperbolic functions ---------------------------------------------------
// tanh: |x| < 2^-12: x (1 - x^2/3), whose omitted terms are below 2^-49
// relative; otherwise (e - 1) / (e + 1) with e = exp(2|x|), where e - 1
// (>= 2^-11) keeps all but 11 bits of e's precision; 1 from |x| >= 9.5 on,
// where 1 - tanh(x) < 2^-26 rounds to 1.  NaN stays NaN; the sign of zero
// is kept.
__host__ __device__ inline float tanh_f32(float x) {
  const double d = (double)x, a = fabs(d);
  if (!(a < 9.5)) return a == a ? copysignf(1.0f, x) : x;
  if (a < 0x1p-12) return (float)(d * fma(-d * d, 1.0 / 3.0, 1.0));
  const double e = exp_core(2.0 * a);
  return copysignf((float)((e - 1.0) * recip(e + 1.0)), x);
}

// sinh: |x| < 2^-5: x (1 + x^2/6 + x^4/120 + x^6/5040), omitted terms below
// 2^-45 relative; otherwise (e - 1/e) / 2 with e = exp(|x|), where the
// difference (>= 2^-4) keeps all but 5 bits of e's precision; +-inf beyond
// |x| = 90 (sinh(90) > FP32 max).  NaN and the sign of zero are kept.
__host__ __device__ inline float sinh_f32(float x) {
  const double d = (double)x, a = fabs(d);
  if (a < 0x1p-5) {
    const double z = d * d;
    return (float)(d * fma(z, fma(z, fma(z, 1.0 / 5040.0, 1.0 / 120.0), 1.0 / 6.0), 1.0));
  }
  if (!(a < 90.0)) return a == a ? copysignf(INFINITY, x) : x;
  const double e = exp_core(a);
  return copysignf((float)(0.5 * (e - recip(e))), x);
}

// cosh: (e + 1/e) / 2 with e = exp(|x|) (no cancellation); inf beyond 90
__host__ __device__ inline float cosh_f32(float x) {
  const double a = fabs((double)x);
  if (!(a < 90.0)) return a == a ? INFINITY : x;
  const double e = exp_core(a);
  return (float)(0.5 * (e + recip(e)));
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
  return copysignf((float)log_pos(a + sqrt_nonneg(fma(a, a, 1.0))), x);
}

// acosh (x >= 1): log(x + sqrt(x^2 - 1)); x^2 - 1 is exact in FP64 for FP32
// x, so near x = 1 the result keeps its relative precision.  x < 1 and NaN
// give NaN, inf gives inf.
__host__ __device__ inline float acosh_f32(float x) {
  const double d = (double)x;
  if (!(d >= 1.0)) return NAN;
  if (d == INFINITY) return x;
  return (float)log_pos(d + sqrt_nonneg(fma(d, d, -1.0)));
}

// atanh (|x| <= 1): |x| <= 0.1716: the series; otherwise
// log((1 + |x|) / (1 - |x|)) / 2, where 1 +- |x| are exact in FP64 and the
// log is at least 0.34.  +-1 give +-inf, |x| > 1 and NaN give NaN.
__host__ __device__ inline float atanh_f32(float x) {
  const double d = (double)x, a = fabs(d);
  if (a <= 0.1716) return (float)atanh_series(d);
  if (!(a <= 1.0)) return NAN;
  if (a == 1.0) return copysignf(INFINITY, x);
  return copysignf((float)(0.5 * log_pos((1.0 + a) * recip(1.0 - a))), x);
}

// ---- circular functions -------------------------------------------------------
// Below 2^-13 in magnitude sin, tan, asin and atan differ from x by less than
// x^2/3 < 2^-26 relative, half an FP32 ulp at most, and cos from 1 by less
// than 2^-27: those round to x and to 1 (the sign of zero is kept).
constexpr double TINY = 0x1p-13;

// sin of r, |r| <= pi/2: r + r^3 S(r^2) with S the degree-6 polynomial of a
// Chebyshev fit (made for this file) to (sin(sqrt z) - sqrt z) / z^1.5 on
// z in [0, (pi/2)^2]: relative error below 2^-50
__host__ __device__ inline double sin_pio2(double r) {
  const double z = r * r;
  double p = fma_kb(z, -0x1.a1095286ba17ap-41, 0x1.60f6f0d66f8acp-33);
  p = fma_kc(p, z, -0x1.ae63bdfa8502ap-26);
  p = fma_kc(p, z, 0x1.71de394e5888cp-19);
  p = fma_kc(p, z, -0x1.a01a019e69965p-13);
  p = fma_kc(p, z, 0x1.1111111110067p-7);
  p = fma_kc(p, z, -0x1.555555555554fp-3);
  return fma(r * z, p, r);
}

// sin and cos of r, |r| <= pi/4, for tan: fits of the same kind on
// z in [0, (pi/4)^2], relative errors below 2^-45 and 2^-49
__host__ __device__ inline double sin_pio4(double r) {
  const double z = r * r;
  double p = fma_kb(z, -0x1.aa285788aaa42p-26, 0x1.71d9a9f41c5a9p-19);
  p = fma_kc(p, z, -0x1.a019fd9b35ee5p-13);
  p = fma_kc(p, z, 0x1.1111110fd3d43p-7);
  p = fma_kc(p, z, -0x1.555555555516bp-3);
  return fma(r * z, p, r);
}

__host__ __device__ inline double cos_pio4(double r) {
  const double z = r * r;
  double p = fma_kb(z, 0x1.1c81c3531fff2p-29, -0x1.27e25f4bb4e6fp-22);
  p = fma_kc(p, z, 0x1.a019ff53a6a1cp-16);
  p = fma_kc(p, z, -0x1.6c16c16b614fcp-10);
  p = fma_kc(p, z, 0x1.5555555555437p-5);
  return fma(z * z, p, fma(z, -0.5, 1.0));
}

// x = (m/2) pi + r for an integer m of a given parity, |r| <= pi/2 (sin, cos:
// m even, m = 2q) or |r| <= pi/4 (tan: any m), for |x| < 2^20: m from FP64
// rounding, r from the three pieces of pi/2 (the first two products exact
// for |m| < 2^20), accurate to 2^-70 absolute, far below r itself next to any
// multiple of pi/2 an FP32 value can come near.
__host__ __device__ inline double reduce_pio2(double x, double m) {
  double t = fnma_kb(m, PIO2_1, x);
  t = fnma_kb(m, PIO2_2, t);
  return fnma_kb(m, PIO2_3, t);
}

constexpr double REDUCE_MAX = 0x1p20;  // beyond: the FP64 library functions

// sin x = (-1)^q sin r with x = q pi + r; cos x = sin(x + pi/2) = (-1)^q sin r
// with x = (q - 1/2) pi + r.  One polynomial for every lane.
__host__ __device__ inline float sin_f32(float x) {
  const double d = (double)x, a = fabs(d);
  if (a < TINY) return x;
  if (!(a < REDUCE_MAX)) return (float)sin(d);  // huge, inf, NaN
  const double q = rint(mul_kb(d, 1.0 / PI));
  const double v = sin_pio2(reduce_pio2(d, 2.0 * q));
  return (float)(((int)q & 1) ? -v : v);
}

__host__ __device__ inline float cos_f32(float x) {
  const double d = (double)x, a = fabs(d);
  if (a < TINY) return 1.0f;
  if (!(a < REDUCE_MAX)) return (float)cos(d);
  const double q = rint(fma_kb(d, 1.0 / PI, 0.5));
  const double v = sin_pio2(reduce_pio2(d, 2.0 * q - 1.0));
  return (float)(((int)q & 1) ? -v : v);
}

// tan = sin / cos of the reduced argument, or -cos / sin in odd quadrants
__host__ __device__ inline float tan_f32(float x) {
  const double d = (double)x, a = fabs(d);
  if (a < TINY) return x;
  if (!(a < REDUCE_MAX)) return (float)tan(d);
  const double m = rint(mul_kb(d, TWO_OVER_PI));
  const double r = reduce_pio2(d, m);
  const double s = sin_pio4(r), c = cos_pio4(r);
  return (float)(((int)m & 1) ? -c * recip(s) : s * recip(c));
}

// atan takes the FP64 library function, which measured faster on MI355X
// than a table-and-series core with its divisions (235 against 326 us for
// 2^27 floats)
__host__ __device__ inline float atan_f32(float x) { return (float)atan((double)x); }

// asin(s) = s + s^3 P(s^2) for |s| <= 1/2, with P the degree-9 polynomial of
// a Chebyshev fit (made for this file) to (asin(sqrt z) - sqrt z) / z^1.5 on
// z in [0, 1/4]: asin's relative error below 2^-46.
__host__ __device__ inline double asin_half(double s) {
  const double z = s * s;
  double p = fma_kb(z, 0x1.c93a92d53b4f1p-6, -0x1.815314c864b09p-9);
  p = fma_kc(p, z, 0x1.00d47e7966d94p-6);
  p = fma_kc(p, z, 0x1.b02442413f6bap-7);
  p = fma_kc(p, z, 0x1.1dc2ef640046fp-6);
  p = fma_kc(p, z, 0x1.6e72146fda29ep-6);
  p = fma_kc(p, z, 0x1.f1c81c59ea536p-6);
  p = fma_kc(p, z, 0x1.6db6d8e71341bp-5);
  p = fma_kc(p, z, 0x1.33333335a9cd6p-4);
  p = fma_kc(p, z, 0x1.5555555554f05p-3);
  return fma(s * z, p, s);
}

// asin: |x| <= 1/2 directly; above, asin|x| = pi/2 - 2 asin(sqrt((1 - |x|) / 2)),
// where 1 - |x| is exact in FP64 and the result is at least pi/6.  The
// polynomial's argument is selected per lane, so every lane evaluates it
// once.  |x| > 1 and NaN give NaN.
__host__ __device__ inline float asin_f32(float x) {
  const double d = (double)x, a = fabs(d);
  if (a < TINY) return x;
  const bool inner = a <= 0.5;
  const double p = asin_half(inner ? a : sqrt_nonneg(0.5 * (1.0 - a)));
  const double v = inner ? p : PIO2 - 2.0 * p;
  return copysignf((float)v, x);
}

// acos: |x| <= 1/2: pi/2 - asin(x) (at least pi/3); x > 1/2:
// 2 asin(sqrt((1 - x) / 2)); x < -1/2: pi - 2 asin(sqrt((1 + x) / 2)).
__host__ __device__ inline float acos_f32(float x) {
  const double d = (double)x, a = fabs(d);
  const bool inner = a <= 0.5;
  const double p = asin_half(inner ? d : sqrt_nonneg(0.5 * (1.0 - a)));
  return (float)(inner ? PIO2 - p : (d > 0.0 ? 2.0 * p : PI - 2.0 * p));
}

// 1 / sqrt(x): positive normal FP32 x through rsqrt_pos (relative error
// below 2^-50 before the FP32 rounding); 0, denormals, negatives, inf and
// NaN through the FP64 square root and division (0 -> +inf, negative -> NaN,
// inf -> 0)
__host__ __device__ inline float invsqrt_f32(float x) {
  if (x >= 0x1p-126f && x < INFINITY) return (float)rsqrt_pos((double)x);
  return (float)(1.0 / sqrt((double)x));
}

}  // namespace vkh
