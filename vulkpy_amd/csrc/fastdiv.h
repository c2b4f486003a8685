// Division of 32-bit unsigned integers n < 2^31 by a runtime-constant divisor
// d >= 1 with one multiply-high, an add and a shift: q = (mulhi(n, m) + n) >> s
// where s = ceil(log2 d) and m = floor(2^32 (2^s - d) / d) + 1 (the
// round-up method of Granlund & Montgomery, "Division by invariant integers
// using multiplication", PLDI 1994).  Usable from host and device code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vkh {

struct FastDiv {
  uint32_t d = 1, m = 0, shift = 0;
  FastDiv() = default;
  explicit FastDiv(uint32_t div) : d(div) {
    for (shift = 0; shift < 32; ++shift)
      if ((1u << shift) >= div) break;
    const uint64_t one = 1;
    const uint64_t magic = ((one << 32) * ((one << shift) - div)) / div + 1;
    m = (uint32_t)magic;
  }
  __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const {
    const uint32_t t = (uint32_t)(((uint64_t)n * m) >> 32);
    return (t + n) >> shift;
  }
};

}  // namespace vkh
