// Host check of sincos_small (vulkpy_amd/csrc/bmath.h) at every Box-Muller
// angle the generator can produce: x = float(6.28318530718f * u) for
// u = k * 2^-23, k < 2^23 (prng_box_muller.comp:26 with the float uniforms of
// prng_xoshiro128pp_float.comp:33).  Prints the largest errors of sin and cos
// in units of the last place of the FP64 reference rounded to float, and the
// number of angles checked.
#include "../vulkpy_amd/csrc/bmath.h"

#include <cmath>
#include <cstdio>

static double ulp_err(float got, double ref) {
  const float rf = (float)ref;
  const float u = std::nextafter(std::fabs(rf), INFINITY) - std::fabs(rf);
  return std::fabs((double)got - ref) / (double)u;
}

int main() {
  double ms = 0, mc = 0;
  long n = 0;
  for (long k = 0; k < (1L << 23); ++k, ++n) {
    const volatile float u = (float)k * (1.0f / 8388608.0f);
    const float x = 6.28318530718f * u;
    float s, c;
    vkh::sincos_small(x, s, c);
    const double es = ulp_err(s, std::sin((double)x)), ec = ulp_err(c, std::cos((double)x));
    if (es > ms) ms = es;
    if (ec > mc) mc = ec;
  }
  std::printf("%.4f %.4f %ld\n", ms, mc, n);
  return 0;
}
