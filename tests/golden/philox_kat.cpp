// Writes Philox4x32-10 known-answer vectors computed by rocRAND's
// philox4x32_10_engine::ten_rounds, an independent implementation used to
// pin the engine's and the oracle's generator.  Host-only program.
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>

// ten_rounds is protected; a subclass exposes it unchanged.
struct Exposed : rocrand_device::philox4x32_10_engine {
    uint4 rounds(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main()
{
    const unsigned ctrs[][4] = {{0, 0, 0, 0},
                                {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
                                {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u},
                                {1, 2, 3, 4},
                                {7, 123456, 999, (1u << 24) | (3u << 16)}};
    const unsigned keys[][2] = {{0, 0}, {0xffffffffu, 0xffffffffu}, {0xa4093822u, 0x299f31d0u}, {5, 0}, {5, 0}};
    Exposed eng;
    printf("[\n");
    for (int i = 0; i < 5; i++) {
        uint4 c = {ctrs[i][0], ctrs[i][1], ctrs[i][2], ctrs[i][3]};
        uint2 k = {keys[i][0], keys[i][1]};
        uint4 o = eng.rounds(c, k);
        printf(" {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n", c.x, c.y, c.z, c.w,
               k.x, k.y, o.x, o.y, o.z, o.w, i < 4 ? "," : "");
    }
    printf("]\n");
    return 0;
}
