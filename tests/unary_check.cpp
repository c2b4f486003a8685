// Host check of the FP32 transcendentals of vulkpy_amd/csrc/bmath.h against
// the FP64 library functions: every FP32 input whose bit pattern is a
// multiple of the stride (argv[1]; 1 = every input) in the range where each
// function computes (both signs where the function is defined for both), and
// for pow every positive finite base at 64 times the stride against 31
// exponents.  Prints the largest error of each function in units of the last
// place of the FP64 reference rounded to float, in the order of the table
// below (pow fourth), then the number of inputs checked and the number of
// special-value mismatches (signed zeros, negatives, infinities, NaN, 1,
// powers of two, domain edges).
#include "../vulkpy_amd/csrc/bmath.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

static float from_bits(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

static double ulp_err(float got, double ref) {
  if (std::isinf(ref) || std::isinf(got)) return got == (float)ref ? 0.0 : 1e9;
  const float rf = (float)ref;
  const float a = std::fabs(rf);
  const float u = a == 0.0f ? 1.4e-45f : std::nextafter(a, INFINITY) - a;
  return std::fabs((double)got - ref) / (double)u;
}

static bool same(float a, float b) {
  return (std::isnan(a) && std::isnan(b)) || (a == b && std::signbit(a) == std::signbit(b));
}

struct Fn {
  float (*f)(float);
  double (*ref)(double);
  uint32_t lo, hi;  // bit patterns [lo, hi) of the magnitudes checked
  bool both_signs;
};

static double d_tanh(double x) { return std::tanh(x); }
static double d_log(double x) { return std::log(x); }
static double d_log2(double x) { return std::log2(x); }
static double d_sinh(double x) { return std::sinh(x); }
static double d_cosh(double x) { return std::cosh(x); }
static double d_asinh(double x) { return std::asinh(x); }
static double d_acosh(double x) { return std::acosh(x); }
static double d_atanh(double x) { return std::atanh(x); }
static double d_sin(double x) { return std::sin(x); }
static double d_cos(double x) { return std::cos(x); }
static double d_tan(double x) { return std::tan(x); }
static double d_asin(double x) { return std::asin(x); }
static double d_acos(double x) { return std::acos(x); }
static double d_atan(double x) { return std::atan(x); }
static double d_exp(double x) { return std::exp(x); }
static double d_exp2(double x) { return std::exp2(x); }
static double d_invsqrt(double x) { return 1.0 / std::sqrt(x); }

int main(int argc, char** argv) {
  const uint64_t stride = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1;
  const Fn fns[] = {
      {vkh::tanh_f32, d_tanh, 0u, 0x41180000u, true},        // |x| < 9.5
      {vkh::log_f32, d_log, 1u, 0x7f800000u, false},         // positive finite
      {vkh::log2_f32, d_log2, 1u, 0x7f800000u, false},
      {nullptr, nullptr, 0u, 0u, false},                      // pow, below
      {vkh::sinh_f32, d_sinh, 0u, 0x42b30000u, true},        // |x| < 89.5
      {vkh::cosh_f32, d_cosh, 0u, 0x42b30000u, true},
      {vkh::asinh_f32, d_asinh, 0u, 0x7f800000u, true},      // all finite
      {vkh::acosh_f32, d_acosh, 0x3f800000u, 0x7f800000u, false},  // [1, max]
      {vkh::atanh_f32, d_atanh, 0u, 0x3f800000u, true},      // |x| < 1
      {vkh::sin_f32, d_sin, 0u, 0x49800000u, true},          // |x| < 2^20 (reduced here)
      {vkh::cos_f32, d_cos, 0u, 0x49800000u, true},
      {vkh::tan_f32, d_tan, 0u, 0x49800000u, true},
      {vkh::asin_f32, d_asin, 0u, 0x3f800001u, true},        // |x| <= 1
      {vkh::acos_f32, d_acos, 0u, 0x3f800001u, true},
      {vkh::atan_f32, d_atan, 0u, 0x7f800001u, true},        // all, inf included
      {vkh::exp_f32, d_exp, 0u, 0x42d00000u, true},          // |x| < 104
      {vkh::exp2_f32, d_exp2, 0u, 0x43160000u, true},        // |x| < 150
      {vkh::invsqrt_f32, d_invsqrt, 1u, 0x7f800000u, false}, // positive finite
  };
  constexpr int NF = sizeof(fns) / sizeof(fns[0]);
  double err[NF] = {};
  uint64_t n = 0;
  for (int i = 0; i < NF; ++i) {
    const Fn& F = fns[i];
    if (!F.f) continue;
    double e = 0;
    uint64_t c = 0;
#pragma omp parallel for reduction(max : e) reduction(+ : c) schedule(static)
    for (int64_t b = F.lo; b < (int64_t)F.hi; b += (int64_t)stride) {
      const float x = from_bits((uint32_t)b);
      const double r = F.ref((double)x);
      e = std::fmax(e, ulp_err(F.f(x), r));
      ++c;
      if (F.both_signs) {
        e = std::fmax(e, ulp_err(F.f(-x), F.ref(-(double)x)));
        ++c;
      }
    }
    err[i] = e;
    n += c;
  }
  // pow: positive finite bases at 64 times the stride against a fixed set of
  // exponents
  const float exps[] = {0.0f, 1.0f, -1.0f, 0.5f, -0.5f, 2.0f, 3.0f, -2.0f, 0.37f, 1.1f, 2.7f, -3.3f,
                        7.25f, -7.25f, 12.5f, 31.0f, -31.0f, 1e-3f, -1e-3f, 1e-7f, 100.0f, -100.0f,
                        150.0f, -150.0f, 0.999f, 1.001f, 23.4f, -23.4f, 64.0f, -64.0f, 1e4f};
  {
    double e = 0;
    uint64_t c = 0;
#pragma omp parallel for reduction(max : e) reduction(+ : c) schedule(static)
    for (int64_t b = 1; b < 0x7f800000; b += 64 * (int64_t)stride) {
      const float a = from_bits((uint32_t)b);
      for (float x : exps) {
        e = std::fmax(e, ulp_err(vkh::pow_f32(a, x), std::pow((double)a, (double)x)));
        ++c;
      }
    }
    err[3] = e;
    n += c;
  }
  int bad = 0;
  const float specials[] = {0.0f, -0.0f, -1.0f, 1.0f, -INFINITY, INFINITY, NAN, 1e-45f, -1e-45f,
                            3.4e38f, -3.4e38f, 100.0f, -100.0f, 0.99999994f, 1.0000001f, 2.0f};
  for (const Fn& F : fns) {
    if (!F.f) continue;
    for (float x : specials) bad += !same(F.f(x), (float)F.ref((double)x));
  }
  for (int k = -149; k < 128; ++k) {
    const float x = std::ldexp(1.0f, k);
    bad += vkh::log2_f32(x) != (float)k;
  }
  bad += vkh::log_f32(1.0f) != 0.0f;
  const float pbases[] = {0.0f, -0.0f, -2.0f, -0.5f, 1.0f, INFINITY, -INFINITY, NAN, 2.0f};
  const float pexps[] = {0.0f, -0.0f, 1.0f, 2.0f, -3.0f, 0.5f, INFINITY, -INFINITY, NAN};
  for (float a : pbases)
    for (float x : pexps) bad += !same(vkh::pow_f32(a, x), (float)std::pow((double)a, (double)x));
  for (int i = 0; i < NF; ++i) std::printf("%.4f ", err[i]);
  std::printf("%llu %d\n", (unsigned long long)n, bad);
  return 0;
}
