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

}  // namespace vkh
