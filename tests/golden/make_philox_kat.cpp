// Writes tests/golden/philox_kat.json: Philox4x32-10 outputs from rocRAND's
// own implementation (rocrand_philox4x32_10.h, ten_rounds) for a fixed set of
// counters/keys, to cross-check the oracle's and the library's Philox.
// Build/run (host only):  hipcc --offload-arch=gfx950 -O1 make_philox_kat.cpp -o /tmp/kat && /tmp/kat
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

struct Exposed : rocrand_device::philox4x32_10_engine {
    using rocrand_device::philox4x32_10_engine::ten_rounds;
};

int main() {
    const uint32_t cases[][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {1u, 2u, 3u, 4u, 5u, 6u},
        {12345u, 7u, (3u << 8) | 2u, 0u, 1u, 0u},
        {0u, 99u, 1u, 0u, 0xdeadbeefu, 0x01234567u},
    };
    Exposed eng;
    std::printf("[\n");
    const int n = sizeof(cases) / sizeof(cases[0]);
    for (int i = 0; i < n; ++i) {
        const uint32_t* c = cases[i];
        uint4 ctr{c[0], c[1], c[2], c[3]};
        uint2 key{c[4], c[5]};
        uint4 r = eng.ten_rounds(ctr, key);
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    c[0], c[1], c[2], c[3], c[4], c[5], r.x, r.y, r.z, r.w, i + 1 < n ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
