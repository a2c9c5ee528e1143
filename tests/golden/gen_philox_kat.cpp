// Writes Philox4x32-10 known-answer vectors from rocRAND's host-callable
// engine (rocrand_philox4x32_10.h) to stdout as JSON. The engine seeded with
// (seed, subsequence, offset = 4*c) holds counter (c lo, c hi, subsequence lo,
// subsequence hi) and key (seed lo, seed hi); its first four outputs are the
// ten-round block of that counter. Build and run:
//   hipcc -O1 -o /tmp/gen_philox_kat tests/golden/gen_philox_kat.cpp && /tmp/gen_philox_kat
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdint>
#include <cstdio>

int main() {
    struct V { uint32_t c[4]; uint32_t k[2]; };
    const V vs[] = {
        {{0u, 0u, 0u, 0u}, {0u, 0u}},
        {{1u, 0u, 0u, 0u}, {0u, 0u}},
        {{0u, 0u, 1u, 0u}, {0u, 0u}},
        {{12345u, 7u, 99u, 3u}, {123u, 0u}},
        {{4000000u, 0u, 0u, 0u}, {0xdeadbeefu, 0xcafef00du}},
        {{0xffffffffu, 0x3fffffffu, 0xffffffffu, 0xffffffffu}, {0xffffffffu, 0xffffffffu}},
        {{17u, 1000u, 0u, 0u}, {42u, 0u}},
    };
    std::printf("[\n");
    const int nv = sizeof(vs) / sizeof(vs[0]);
    for (int q = 0; q < nv; ++q) {
        const V& v = vs[q];
        const unsigned long long seed = (unsigned long long)v.k[0] | ((unsigned long long)v.k[1] << 32);
        const unsigned long long sub = (unsigned long long)v.c[2] | ((unsigned long long)v.c[3] << 32);
        const unsigned long long off = 4ull * ((unsigned long long)v.c[0] | ((unsigned long long)v.c[1] << 32));
        rocrand_device::philox4x32_10_engine e(seed, sub, off);
        uint32_t o[4];
        for (int i = 0; i < 4; ++i) o[i] = e.next();
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n", v.c[0], v.c[1],
                    v.c[2], v.c[3], v.k[0], v.k[1], o[0], o[1], o[2], o[3], q + 1 < nv ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
