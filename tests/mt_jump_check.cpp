// Checks the host jump-ahead of mt_jump.cpp against brute force: for a seed
// the characteristic polynomial was not derived from, the words J = c * step
// draws ahead (c = 1..3, step = 624 * 256) equal the XOR of the windows at
// the jump polynomial's set bits, for windows starting at several offsets
// (word 0 of a window counts by its top bit only).  Prints "ok" and exits 0.
#include <cstdio>
#include <vector>

#include "mt_jump.hpp"

static std::vector<uint32_t> words(uint32_t seed, size_t n)
{
    std::vector<uint32_t> x(n + 624);
    x[0] = seed;
    for (int i = 1; i < 624; ++i) x[i] = 1812433253u * (x[i - 1] ^ (x[i - 1] >> 30)) + (uint32_t)i;
    for (size_t j = 624; j < n + 624; ++j) {
        const uint32_t y = (x[j - 624] & 0x80000000u) | (x[j - 623] & 0x7fffffffu);
        x[j] = x[j - 227] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    x.resize(n);
    return x;
}

int main()
{
    const uint64_t step = 624ull * 256;
    std::vector<uint64_t> g;
    if (!fgl::mt_jump_polys(step, 3, g)) {
        std::printf("no characteristic polynomial of degree 19937\n");
        return 1;
    }
    const std::vector<uint32_t> x = words(12345u, 3 * step + 60000);
    int bad = 0, dense = 0;
    for (int c = 1; c <= 3; ++c) {
        const uint64_t *p = g.data() + (size_t)(c - 1) * fgl::kMtPolyWords;
        int weight = 0;
        for (int w = 0; w < fgl::kMtPolyWords; ++w) weight += __builtin_popcountll(p[w]);
        dense += weight > 1000;
        for (size_t n : {1ul, 5ul, 624ul, 1000ul, 20000ul})
            for (size_t k : {0ul, 1ul, 300ul, 623ul}) {
                uint32_t acc = 0;
                for (int i = 0; i < fgl::kMtDegree; ++i)
                    if ((p[i >> 6] >> (i & 63)) & 1u) acc ^= x[n + i + k];
                const uint32_t mask = k == 0 ? 0x80000000u : 0xffffffffu;
                if ((acc ^ x[n + c * step + k]) & mask) ++bad;
            }
    }
    if (bad || dense != 3) {
        std::printf("bad %d dense %d\n", bad, dense);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
