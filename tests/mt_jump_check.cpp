// Checks the host jump-ahead of mt_jump.cpp against brute force: for a seed
// the characteristic polynomial was not derived from, the words J = c * step
// draws ahead equal the XOR of the windows at the jump polynomial's set
// bits, for windows starting at several offsets (word 0 of a window counts
// by its top bit only).  Two steps (624 * 256 and 624 * 512) are asked for
// in turn, the first again with more polynomials after the second, as the
// context's two chunk lengths do.  Prints "ok" and exits 0.
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

// Polynomials c = 1..n of `step` against the words; returns the mismatches,
// and counts the polynomials of more than 1000 terms in dense.
static int check(uint64_t step, int n, const std::vector<uint32_t> &x, int &dense)
{
    std::vector<uint64_t> g;
    if (!fgl::mt_jump_polys(step, n, g) || (int)(g.size() / fgl::kMtPolyWords) != n) return 1000;
    int bad = 0;
    for (int c = 1; c <= n; ++c) {
        const uint64_t *p = g.data() + (size_t)(c - 1) * fgl::kMtPolyWords;
        int weight = 0;
        for (int w = 0; w < fgl::kMtPolyWords; ++w) weight += __builtin_popcountll(p[w]);
        dense += weight > 1000;
        for (size_t m : {1ul, 5ul, 624ul, 1000ul, 20000ul})
            for (size_t k : {0ul, 1ul, 300ul, 623ul}) {
                uint32_t acc = 0;
                for (int i = 0; i < fgl::kMtDegree; ++i)
                    if ((p[i >> 6] >> (i & 63)) & 1u) acc ^= x[m + i + k];
                const uint32_t mask = k == 0 ? 0x80000000u : 0xffffffffu;
                if ((acc ^ x[m + c * step + k]) & mask) ++bad;
            }
    }
    return bad;
}

int main()
{
    const uint64_t s1 = 624ull * 256, s2 = 624ull * 512;
    const std::vector<uint32_t> x = words(12345u, 4 * s1 + 60000);
    int dense = 0;
    const int bad = check(s1, 3, x, dense) + check(s2, 2, x, dense) + check(s1, 4, x, dense);
    if (bad || dense != 9) {
        std::printf("bad %d dense %d\n", bad, dense);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
