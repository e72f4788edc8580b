// mt_jump.cpp -- jump polynomials for std::mt19937 (host code; built by the
// host compiler with the carry-less multiply instruction enabled per function).
//
// phi(t): Berlekamp-Massey on bit 0 of the untempered words x_1, x_2, ...
// (of any seed) gives the sequence's minimal polynomial, which for MT19937
// is its characteristic polynomial, of degree 19937.  Residues mod phi are
// multiplied with 64 x 64 carry-less products and reduced 64 bits at a time.
#include "mt_jump.hpp"

#include <immintrin.h>

#include <cstring>
#include <map>
#include <mutex>

namespace fgl {
namespace {

constexpr int kN = 624, kM = 397;
constexpr uint32_t kUpper = 0x80000000u, kLower = 0x7fffffffu, kMatrix = 0x9908b0dfu;
constexpr int kW = kMtPolyWords;
constexpr int kProdWords = 2 * kW + 2;   // products and the reduction's spill

// The untempered sequence x_0 .. x_{n-1} of std::mt19937(seed).
std::vector<uint32_t> mt_words(uint32_t seed, size_t n)
{
    std::vector<uint32_t> x(n + kN);
    x[0] = seed;
    for (int i = 1; i < kN; ++i) x[i] = 1812433253u * (x[i - 1] ^ (x[i - 1] >> 30)) + (uint32_t)i;
    for (size_t j = kN; j < n + kN; ++j) {
        const uint32_t y = (x[j - kN] & kUpper) | (x[j - kN + 1] & kLower);
        x[j] = x[j - kN + kM] ^ (y >> 1) ^ ((y & 1u) ? kMatrix : 0u);
    }
    x.resize(n);
    return x;
}

inline bool bit(const uint64_t *v, size_t i) { return (v[i >> 6] >> (i & 63)) & 1u; }

// 64 bits of v starting at bit i (v holds at least i / 64 + 2 words).
inline uint64_t bits64(const uint64_t *v, size_t i)
{
    const size_t q = i >> 6;
    const unsigned b = (unsigned)(i & 63);
    return b ? (v[q] >> b) | (v[q + 1] << (64 - b)) : v[q];
}

// v ^= h << i (v holds at least i / 64 + 2 words).
inline void xor_at(uint64_t *v, uint64_t h, size_t i)
{
    const size_t q = i >> 6;
    const unsigned b = (unsigned)(i & 63);
    v[q] ^= h << b;
    if (b) v[q + 1] ^= h >> (64 - b);
}

// Berlekamp-Massey over GF(2): the shortest connection polynomial
// C(x) = 1 + c_1 x + ... + c_L x^L with s_n = sum c_i s_{n-i}.  Returns L;
// C's coefficients in c (bit i = c_i).
int berlekamp_massey(const std::vector<uint8_t> &s, std::vector<uint64_t> &c)
{
    const size_t N = s.size(), words = 2 * (N / 64) + 8;
    // r: the sequence reversed (bit j = s_{N-1-j}), so the discrepancy at n is
    // parity(C & (r >> (N - 1 - n))) word by word (C has degree <= L <= n + 1)
    std::vector<uint64_t> r(N / 64 + 4, 0), b(words, 0), t(words, 0);
    c.assign(words, 0);
    for (size_t j = 0; j < N; ++j)
        if (s[N - 1 - j]) r[j >> 6] |= 1ull << (j & 63);
    c[0] = b[0] = 1;
    int L = 0;
    size_t m = 1;
    for (size_t n = 0; n < N; ++n) {
        const size_t o = N - 1 - n;
        uint64_t acc = 0;
        for (size_t w = 0; w <= (size_t)L / 64; ++w) acc ^= c[w] & bits64(r.data(), o + 64 * w);
        if (!(__builtin_popcountll(acc) & 1)) {
            ++m;
            continue;
        }
        // C ^= B x^m over the words B can occupy (its degree is at most L)
        const size_t top = (size_t)L / 64 + 2;
        if ((m >> 6) + top + 1 >= words) return -1;
        t = c;
        for (size_t w = 0; w < top; ++w) xor_at(c.data(), b[w], 64 * w + m);
        if (2 * (size_t)L <= n) {
            L = (int)(n + 1 - (size_t)L);
            b = t;
            m = 1;
        } else {
            ++m;
        }
    }
    return L;
}

struct Phi {
    uint64_t w[kW + 1];          // coefficients of degrees 0..19937
    std::vector<int> low;        // exponents e < 19937 with phi_e = 1
    std::vector<int> near;       // d = 19937 - e for the exponents e in [19937 - 63, 19936]
};

__attribute__((target("pclmul"))) inline void clmul_hw(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi)
{
    const __m128i r = _mm_clmulepi64_si128(_mm_cvtsi64_si128((long long)a), _mm_cvtsi64_si128((long long)b), 0);
    lo = (uint64_t)_mm_cvtsi128_si64(r);
    hi = (uint64_t)_mm_cvtsi128_si64(_mm_unpackhi_epi64(r, r));
}

inline void clmul_sw(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi)
{
    lo = hi = 0;
    for (int i = 0; i < 64; ++i)
        if ((a >> i) & 1u) {
            lo ^= b << i;
            if (i) hi ^= b >> (64 - i);
        }
}

__attribute__((target("pclmul"))) void mul_hw(const uint64_t *a, const uint64_t *b, uint64_t *p)
{
    std::memset(p, 0, kProdWords * sizeof(uint64_t));
    for (int i = 0; i < kW; ++i) {
        if (!a[i]) continue;
        for (int j = 0; j < kW; ++j) {
            uint64_t lo, hi;
            clmul_hw(a[i], b[j], lo, hi);
            p[i + j] ^= lo;
            p[i + j + 1] ^= hi;
        }
    }
}

void mul_sw(const uint64_t *a, const uint64_t *b, uint64_t *p)
{
    std::memset(p, 0, kProdWords * sizeof(uint64_t));
    for (int i = 0; i < kW; ++i) {
        if (!a[i]) continue;
        for (int j = 0; j < kW; ++j) {
            uint64_t lo, hi;
            clmul_sw(a[i], b[j], lo, hi);
            p[i + j] ^= lo;
            p[i + j + 1] ^= hi;
        }
    }
}

// p (degree < 2 * 19968) mod phi into out (kW words).  Windows of 64 bits at
// 19937 + 64 k, from the top: h with h * t^(64 k) * phi cancelling the window
// solves h ^ XOR_d (h >> d) = window (d over `near`: phi's terms that reach
// into the window), a triangular system taken from the top bit down.
void reduce(const Phi &phi, uint64_t *p, uint64_t *out)
{
    for (int k = (kProdWords * 64 - 64 - kMtDegree) / 64; k >= 0; --k) {
        const size_t at = (size_t)kMtDegree + 64 * (size_t)k;
        const uint64_t win = bits64(p, at);
        if (!win) continue;
        uint64_t h = 0;
        for (int j = 63; j >= 0; --j) {
            uint64_t v = (win >> j) & 1u;
            for (int d : phi.near)
                if (j + d < 64) v ^= (h >> (j + d)) & 1u;
            h |= v << j;
        }
        xor_at(p, h, at);
        for (int e : phi.low) xor_at(p, h, (size_t)e + 64 * (size_t)k);
    }
    std::memcpy(out, p, kW * sizeof(uint64_t));
    out[kW - 1] &= (1ull << (kMtDegree - 64 * (kW - 1))) - 1;
}

struct Series {
    std::vector<uint64_t> g1, polys;   // t^step mod phi; t^(c step) mod phi for c = 1..
};

struct Table {
    std::mutex mu;
    bool tried = false, ok = false, hw = false;
    Phi phi;
    std::map<uint64_t, Series> series;   // by step
};

Table &table()
{
    static Table t;
    return t;
}

void mulmod(Table &T, const uint64_t *a, const uint64_t *b, uint64_t *out)
{
    uint64_t p[kProdWords];
    if (T.hw) mul_hw(a, b, p);
    else mul_sw(a, b, p);
    reduce(T.phi, p, out);
}

bool init_phi(Table &T)
{
    // bit 0 of x_1 .. x_N: x_0 counts by its top bit only and is left out
    const size_t N = 2 * (size_t)kMtDegree + 256;
    const std::vector<uint32_t> x = mt_words(5489u, N + 1);
    std::vector<uint8_t> s(N);
    for (size_t n = 0; n < N; ++n) s[n] = (uint8_t)(x[n + 1] & 1u);
    std::vector<uint64_t> c;
    const int L = berlekamp_massey(s, c);
    if (L != kMtDegree) return false;
    // phi(t) = t^L C(1/t): phi_j = c_{L-j}
    std::memset(T.phi.w, 0, sizeof(T.phi.w));
    for (int j = 0; j <= L; ++j)
        if (bit(c.data(), (size_t)(L - j))) T.phi.w[j >> 6] |= 1ull << (j & 63);
    for (int e = 0; e < L; ++e)
        if (bit(T.phi.w, (size_t)e)) {
            T.phi.low.push_back(e);
            if (e >= L - 63) T.phi.near.push_back(L - e);
        }
    T.hw = __builtin_cpu_supports("pclmul");
    return true;
}

}  // namespace

bool mt_jump_polys(uint64_t step, int n, std::vector<uint64_t> &out)
{
    Table &T = table();
    std::lock_guard<std::mutex> lock(T.mu);
    if (!T.tried) {
        T.tried = true;
        T.ok = init_phi(T);
    }
    if (!T.ok) return false;
    Series &S = T.series[step];
    if (S.g1.empty()) {
        // t^step by square-and-multiply from the top bit of step
        S.g1.assign(kW, 0);
        S.g1[0] = 1;
        std::vector<uint64_t> t1(kW, 0);
        t1[0] = 2;   // t
        for (int b = 63; b >= 0; --b) {
            mulmod(T, S.g1.data(), S.g1.data(), S.g1.data());
            if ((step >> b) & 1u) mulmod(T, S.g1.data(), t1.data(), S.g1.data());
        }
        S.polys = S.g1;
    }
    while ((int)(S.polys.size() / kW) < n) {
        const size_t have = S.polys.size();
        S.polys.resize(have + kW);
        mulmod(T, S.polys.data() + have - kW, S.g1.data(), S.polys.data() + have);
    }
    out.assign(S.polys.begin(), S.polys.begin() + (size_t)n * kW);
    return true;
}

}  // namespace fgl
