// Host check of tanh_f32, log_f32 and log2_f32 (vulkpy_amd/csrc/bmath.h):
// every FP32 input whose bit pattern is a multiple of the stride (argv[1];
// 1 = every input) in the ranges where the functions compute -- tanh on
// [0, 9.5) and both signs, log / log2 on all positive finite floats -- against
// the FP64 library functions.  Prints the largest error of each in units of
// the last place of the FP64 reference rounded to float, the number of
// inputs, and the number of special-value mismatches (0, -0, negative, inf,
// NaN, 1, powers of two).
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

int main(int argc, char** argv) {
  const uint64_t stride = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1;
  double et = 0, el = 0, e2 = 0;
  uint64_t n = 0;
  const uint32_t tanh_end = 0x41180000u;  // 9.5f
  for (uint64_t b = 0; b < tanh_end; b += stride, ++n) {
    const float x = from_bits((uint32_t)b);
    const double r = std::tanh((double)x);
    et = std::fmax(et, ulp_err(vkh::tanh_f32(x), r));
    et = std::fmax(et, ulp_err(vkh::tanh_f32(-x), -r));
  }
  for (uint64_t b = 1; b < 0x7f800000u; b += stride, ++n) {
    const float x = from_bits((uint32_t)b);
    el = std::fmax(el, ulp_err(vkh::log_f32(x), std::log((double)x)));
    e2 = std::fmax(e2, ulp_err(vkh::log2_f32(x), std::log2((double)x)));
  }
  int bad = 0;
  const float specials[] = {0.0f, -0.0f, -1.0f, -INFINITY, INFINITY, NAN, 1.0f, 1e-45f, 3.4e38f};
  for (float x : specials) {
    bad += !same(vkh::tanh_f32(x), (float)std::tanh((double)x));
    bad += !same(vkh::log_f32(x), (float)std::log((double)x));
    bad += !same(vkh::log2_f32(x), (float)std::log2((double)x));
  }
  for (int k = -149; k < 128; ++k) {
    const float x = std::ldexp(1.0f, k);
    bad += vkh::log2_f32(x) != (float)k;
  }
  bad += vkh::log_f32(1.0f) != 0.0f;
  // pow: positive finite bases at 64 times the stride against a fixed set of
  // exponents, and the special cases against the FP64 pow rounded
  const float exps[] = {0.0f, 1.0f, -1.0f, 0.5f, -0.5f, 2.0f, 3.0f, -2.0f, 0.37f, 1.1f, 2.7f, -3.3f,
                        7.25f, -7.25f, 12.5f, 31.0f, -31.0f, 1e-3f, -1e-3f, 1e-7f, 100.0f, -100.0f,
                        150.0f, -150.0f, 0.999f, 1.001f, 23.4f, -23.4f, 64.0f, -64.0f, 1e4f};
  double ep = 0;
  for (uint64_t b = 1; b < 0x7f800000u; b += 64 * stride) {
    const float a = from_bits((uint32_t)b);
    for (float e : exps) {
      ep = std::fmax(ep, ulp_err(vkh::pow_f32(a, e), std::pow((double)a, (double)e)));
      ++n;
    }
  }
  const float pbases[] = {0.0f, -0.0f, -2.0f, -0.5f, 1.0f, INFINITY, -INFINITY, NAN, 2.0f};
  const float pexps[] = {0.0f, -0.0f, 1.0f, 2.0f, -3.0f, 0.5f, INFINITY, -INFINITY, NAN};
  for (float a : pbases)
    for (float e : pexps) bad += !same(vkh::pow_f32(a, e), (float)std::pow((double)a, (double)e));
  std::printf("%.4f %.4f %.4f %.4f %llu %d\n", et, el, e2, ep, (unsigned long long)n, bad);
  return 0;
}
