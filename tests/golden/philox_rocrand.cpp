// Known-answer vectors for Philox4x32-10 from AMD's open-source rocRAND
// engine (host path of rocrand_philox4x32_10.h), used to pin the oracle's
// Philox.  Engine state mapping: seed = key1:key0, subsequence = ctr3:ctr2,
// offset = 4 * (ctr1:ctr0) selects the 128-bit block with counter ctr
// (so ctr1 must stay below 2^30 for the offset to fit in 64 bits).
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

int main() {
  const uint32_t cases[][6] = {
      {0u, 0u, 0u, 0u, 0u, 0u},
      {0xffffffffu, 0x3fffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
      {0x243f6a88u, 0x05a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
      {0u, 0u, 0u, 0u, 42u, 0u},
      {17u, 3u, 9u, 0u, 42u, 1u},
      {999u, 127u, 5u, 4095u, 42u, 7u},
      {123456u, 65535u, 0u, 12u, 7u, 3u},
  };
  printf("[\n");
  const int n = sizeof(cases) / sizeof(cases[0]);
  for (int i = 0; i < n; ++i) {
    const uint32_t* c = cases[i];
    const unsigned long long seed = ((unsigned long long)c[5] << 32) | c[4];
    const unsigned long long sub = ((unsigned long long)c[3] << 32) | c[2];
    const unsigned long long off = (((unsigned long long)c[1] << 32) | c[0]) * 4ull;
    rocrand_device::philox4x32_10_engine eng(seed, sub, off);
    uint32_t r[4];
    for (int k = 0; k < 4; ++k) r[k] = eng();
    printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n", c[0], c[1], c[2],
           c[3], c[4], c[5], r[0], r[1], r[2], r[3], i + 1 < n ? "," : "");
  }
  printf("]\n");
  return 0;
}
