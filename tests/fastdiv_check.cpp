// Host check of vulkpy_amd/csrc/fastdiv.h: exhaustive over small n and d,
// then edge cases and pseudo-random pairs up to n < 2^31.  Prints the number
// of mismatches.
#include "../vulkpy_amd/csrc/fastdiv.h"

#include <cstdio>

int main() {
  unsigned long long bad = 0, checked = 0;
  for (uint32_t d = 1; d <= 2048; ++d) {
    vkh::FastDiv f(d);
    for (uint32_t n = 0; n < 70000; ++n, ++checked) bad += f.div(n) != n / d;
  }
  const uint32_t top = 0x7fffffffu;
  uint64_t x = 88172645463325252ull;
  for (int it = 0; it < 2000000; ++it) {
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    uint32_t d = (uint32_t)(x % top) + 1;
    if (it % 3 == 0) d = (uint32_t)(x % 70000) + 1;
    if (it % 5 == 0) d = 1u << (x % 31);
    vkh::FastDiv f(d);
    uint32_t ns[6] = {(uint32_t)((x >> 17) % (top + 1ull)), top, top - 1, d - 1, d, d * 2 < top ? d * 2 - 1 : top};
    for (uint32_t n : ns) { if (n > top) continue; ++checked; bad += f.div(n) != n / d; }
  }
  std::printf("%llu %llu\n", bad, checked);
  return bad != 0;
}
