// Writes Philox4x32-10 golden vectors computed by rocRAND's philox4x32_10_engine
// (a library routine, called on the host) to philox_rocrand.json.
//
// Build and run (host code only; no GPU needed):
//   hipcc -O1 -o /tmp/gen_philox tests/golden/gen_philox_vectors.cpp && /tmp/gen_philox > tests/golden/philox_rocrand.json
//
// rocRAND's engine(seed, subsequence, 0) starts at counter
// (0, 0, subsequence lo, subsequence hi) with key (seed lo, seed hi); discard(4n)
// then advances the 64-bit low half (c0, c1) by n.  A block (c0, c1, c2, c3) is
// reached with subsequence = c2 | c3 << 32 and discards summing to c0 | c1 << 32
// (split in quarters so that every discard argument fits in 64 bits).
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdint>
#include <cstdio>

static void block(uint64_t seed, const uint32_t c[4], uint32_t out[4])
{
    unsigned long long subseq = (unsigned long long)c[2] | ((unsigned long long)c[3] << 32);
    unsigned long long lowhalf = (unsigned long long)c[0] | ((unsigned long long)c[1] << 32);
    rocrand_device::philox4x32_10_engine e(seed, subseq, 0ull);
    unsigned long long quarter = lowhalf >> 2;
    for (int q = 0; q < 4; ++q) e.discard(4ull * quarter);
    e.discard(4ull * (lowhalf & 3ull));
    for (int i = 0; i < 4; ++i) out[i] = e.next();
}

int main()
{
    struct Case { uint64_t seed; uint32_t c[4]; };
    Case cases[] = {
        {0ull, {0u, 0u, 0u, 0u}},
        {0xffffffffffffffffull, {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu}},
        {0x299f31d0a4093822ull, {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u}},
        {1ull, {0u, 0u, 0u, 3u << 28}},
        {7ull, {5u, 123u, 2u, (1u << 28) | (1u << 24)}},
        {4095ull, {77u, 5000u, 7u, 1u << 28}},
        {0x123456789abcdefull, {1u, 2u, 3u, (2u << 28) | (4u << 16)}},
        {42ull, {0x7fffffffu, 0x80000000u, 0u, 0u}},
    };
    std::printf("[\n");
    const int n = sizeof(cases) / sizeof(cases[0]);
    for (int i = 0; i < n; ++i) {
        uint32_t out[4];
        block(cases[i].seed, cases[i].c, out);
        std::printf("  {\"seed\": %llu, \"ctr\": [%u, %u, %u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    (unsigned long long)cases[i].seed, cases[i].c[0], cases[i].c[1], cases[i].c[2],
                    cases[i].c[3], out[0], out[1], out[2], out[3], i + 1 < n ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
