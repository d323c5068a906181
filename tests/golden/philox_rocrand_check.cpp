// Cross-check of the Philox4x32-10 known-answer vectors against ROCm's
// open-source rocRAND implementation (host side). Prints one JSON object per
// vector: {"ctr": [...], "key": [...], "out": [...]}.
// Build: hipcc -O1 -I/opt/rocm/include philox_rocrand_check.cpp -o check
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>

struct Probe : rocrand_device::philox4x32_10_engine {
  uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
  const unsigned v[][6] = {
      {0u, 0u, 0u, 0u, 0u, 0u},
      {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
      {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
      {1u, 2u, 3u, 4u, 5u, 6u},
      {0x00000007u, 0x02000003u, 0x00000011u, 0x00000005u, 0x0357ba20u, 0x00000000u},
  };
  Probe p;
  for (auto &t : v) {
    uint4 o = p.run(make_uint4(t[0], t[1], t[2], t[3]), make_uint2(t[4], t[5]));
    std::printf("{\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}\n",
                t[0], t[1], t[2], t[3], t[4], t[5], o.x, o.y, o.z, o.w);
  }
  return 0;
}
