// rng.hip -- the reference's random stream on gfx950.
//
// The reference draws every random float from one function-static
// std::mt19937 (default seed 5489) through uniform_real_distribution<float>
// (src/utility.h:90-103), and builds its random vectors by rejection
// (src/geometry.h:947-990).  Its consumers (Render::DoSSAO, the PCF / PCSS
// shadow filters) take the stream strictly in pixel (or fragment) order, so
// the device replays it in two steps: the generator's draws are produced in
// parallel chunks reached by jumping ahead, and the accepted vectors written
// in stream order; the per-pixel kernels then read the vectors of their
// pixel at a known offset.
//
// Arithmetic, fixed by the C++ library the reference is built with:
//  * MT19937: the standard twist and tempering on 624 words;
//  * generate_canonical<float, 24>: one 32-bit draw u, float(u) * 2^-32
//    (float(u) rounded to nearest), a result of 1 replaced by the largest
//    float below 1;
//  * Random(-1, 1) = -1 + 2 * r in float;
//  * Vector3f(Random(..), Random(..), Random(..)) evaluates its arguments
//    right to left (the g++ that builds the reference), so a unit-sphere try
//    takes z, y, x from three consecutive draws, a unit-disk try y, x from
//    two; a try is kept when (x*x + y*y) + z*z < 1.
#include "fgl_kernels.hpp"
#include "mt_jump.hpp"

namespace fgl {

constexpr uint32_t kMtN = 624, kMtM = 397;
constexpr uint32_t kMtUpper = 0x80000000u, kMtLower = 0x7fffffffu, kMtMatrix = 0x9908b0dfu;

void mt_seed_state(uint32_t seed, MtState &st)
{
    st.mt[0] = seed;
    for (uint32_t i = 1; i < kMtN; ++i) st.mt[i] = 1812433253u * (st.mt[i - 1] ^ (st.mt[i - 1] >> 30)) + i;
    st.idx = kMtN;
    st.pad = 0;
    st.draws = 0;
}

__device__ __forceinline__ uint32_t mt_twist_word(uint32_t cur, uint32_t next, uint32_t far)
{
    const uint32_t y = (cur & kMtUpper) | (next & kMtLower);
    return far ^ (y >> 1) ^ ((y & 1u) ? kMtMatrix : 0u);
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t z)
{
    z ^= z >> 11;
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= z >> 18;
    return z;
}

// Random(-1, 1) from one draw.
__device__ __forceinline__ float random_pm1(uint32_t u)
{
    float r = (float)u * 2.3283064365386963e-10f;   // 2^-32, exact scaling
    if (r >= 1.f) r = __uint_as_float(0x3f7fffffu);
    return -1.f + 2.f * r;
}

// The generator as one sequence: x_{n+624} = twist(x_n, x_{n+1}, x_{n+397})
// (a block's twist is this recurrence, new words used as they appear).  The
// 227 words x_n .. x_{n+226} read only words at least 227 back, so a
// workgroup makes them together, one barrier per step, keeping the last 1024
// words in an LDS ring (word n in slot n % 1024).  In step s thread t makes
// x_{624+j}, j = 227 s + t, from x_j and x_{j+1} (made two steps back or
// earlier, so read during the step before, while it computes) and x_{j+397}
// (the word the thread made in the step before).  A step writes the slots of
// words 1024 back, none of which it or the next step's reads touch, so the
// next step's reads are issued first and their latency overlaps the twist.
constexpr uint32_t kMtRing = 1024, kMtLag = kMtN - kMtM;

// ring holds x_0 .. x_623 (slots 0..623); the next `total` words x_624 ..
// go through emit(j, x_{624+j}) once each, in steps of 227.  Needs at least
// 227 threads; every thread of the workgroup calls it.  Past the end the
// last step computes and stores words nobody reads, and emits only j < total.
template <typename Emit>
__device__ __forceinline__ void mt_run(uint32_t *ring, uint32_t total, Emit emit)
{
    constexpr uint32_t R = kMtRing - 1;
    const uint32_t tid = threadIdx.x;
    const bool lane_on = tid < kMtLag;
    uint32_t cur = 0, nxt = 0, far = 0;
    if (lane_on) {
        cur = ring[tid];
        nxt = ring[tid + 1];
        far = ring[tid + kMtM];
    }
    for (uint32_t j0 = 0; j0 < total; j0 += kMtLag) {
        if (lane_on) {
            const uint32_t j = j0 + tid;
            const uint32_t ncur = ring[(j + kMtLag) & R], nnxt = ring[(j + kMtLag + 1) & R];
            const uint32_t v = mt_twist_word(cur, nxt, far);
            ring[(j + kMtN) & R] = v;
            if (j < total) emit(j, v);
            cur = ncur;
            nxt = nnxt;
            far = v;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ uint32_t mt_untemper(uint32_t y)
{
    y ^= y >> 18;
    y ^= (y << 15) & 0xefc60000u;
    uint32_t x = y;
#pragma unroll
    for (int i = 0; i < 5; ++i) x = y ^ ((x << 7) & 0x9d2c5680u);
    return x ^ (x >> 11) ^ (x >> 22);
}

// One try of D draws at p: the vector (x, y, z) and whether it is kept.
template <int D>
__device__ __forceinline__ bool mt_try(const uint32_t *__restrict__ p, float &x, float &y, float &z)
{
    z = 0.f;
    if (D == 3) {
        z = random_pm1(p[0]);
        y = random_pm1(p[1]);
        x = random_pm1(p[2]);
    } else {
        y = random_pm1(p[0]);
        x = random_pm1(p[1]);
    }
    return (x * x + y * y) + z * z < 1.f;
}

// ---- the stream in parallel: jump ahead, generate, keep ----
//
// A fill takes the stream's next draws in chunks of K blocks (624 K draws).
// Chunk 0 continues from the context's array; chunk c > 0 starts from the
// array c 624 K words further on, which the jump gives (mt_jump.hpp): the XOR
// of the windows x_{B+i} .. x_{B+i+623} of the base sequence over the set
// bits i of t^(c 624 K) mod phi, x_B .. being the context's array and the 32
// blocks after it.  The tries are then counted, scanned and kept in
// parallel, and a commit moves the context's stream past the tries a caller
// used.  Every window of a jump, x_{B + 19967 + 623} at most, lies in the
// base sequence.
static_assert(kMtBaseWords >= 64 * kMtPolyWords + 623, "the jump's windows stay in the base sequence");

// One workgroup: the context's array and the 32 blocks after it, untempered.
__global__ __launch_bounds__(256) void k_mt_base(const MtState *__restrict__ st, uint32_t *__restrict__ base)
{
    __shared__ uint32_t ring[kMtRing];
    const int tid = threadIdx.x;
    for (int i = tid; i < (int)kMtN; i += 256) {
        const uint32_t v = st->mt[i];
        ring[i] = v;
        base[i] = v;
    }
    __syncthreads();
    mt_run(ring, (uint32_t)(kMtBaseWords - (int)kMtN), [&](uint32_t j, uint32_t v) { base[kMtN + j] = v; });
}

// Workgroup c - 1 jumps to chunk c's array: states[c][k] = XOR of base[i + k]
// over the set bits i of polynomial c, the base sequence in LDS.  The set
// bits are split by i mod 4: for the bits i = 4 q + r, thread t gathers the
// 16-byte group base[4 (q + t) .. 4 (q + t) + 3], words k + r of the result
// for k + r in [4 t, 4 t + 4), so each bit costs a thread one 16-byte LDS
// read; three wavefronts per class, so each class's chain of dependent LDS
// reads runs beside the others; the four partial arrays are combined at the
// end (word k = XOR over r of partial r at k + r).  Each class's q values are listed in LDS (per-word
// counts scanned by one wavefront per class), padded to a multiple of eight
// with a q whose groups are zero, so every wavefront reads eight q values
// with one broadcast load.
constexpr int kJumpGroups = ((int)kMtN + 3 + 3) / 4;                       // 157: k + r < 627
constexpr int kJumpClassThreads = 192;                                     // three wavefronts per class
constexpr int kJumpThreads = 4 * kJumpClassThreads;
static_assert(kJumpGroups <= kJumpClassThreads, "a thread per 16-byte group of a class");
constexpr int kJumpZeroQ = kMtBaseWords / 4;                               // groups past the base: zero
constexpr int kJumpLdsWords = kMtBaseWords + 4 * kJumpGroups;              // base, then the zero groups
constexpr int kJumpListMax = kMtPolyWords * 64 + 4 * 7 + 8;                // every bit, padding, one read ahead
static_assert(kMtBaseWords % 4 == 0 && kJumpZeroQ < 65536, "16-bit q values; aligned zero groups");
static_assert(4 * (kMtDegree / 4 + kJumpGroups - 1) + 3 < kMtBaseWords, "the groups a bit reads lie in the base");

__global__ __launch_bounds__(kJumpThreads) void k_mt_jump(const uint32_t *__restrict__ base, const uint64_t *__restrict__ polys,
                                                 uint32_t *__restrict__ states)
{
    __shared__ __align__(16) uint32_t b[kJumpLdsWords];
    __shared__ __align__(16) uint16_t qs[kJumpListMax];
    __shared__ uint32_t off[4][kMtPolyWords];
    __shared__ uint32_t ncls[4], cls[5];                // class sizes; starts (multiples of 8), cls[4] = the end
    __shared__ uint32_t part[4][4 * kJumpGroups];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c = blockIdx.x + 1;
    const uint64_t *g = polys + (size_t)(c - 1) * kMtPolyWords;
    constexpr uint64_t kRes = 0x1111111111111111ull;   // bits i with i mod 4 = 0
    for (int i = tid; i < kMtBaseWords / 4; i += kJumpThreads) ((uint4 *)b)[i] = ((const uint4 *)base)[i];
    for (int i = tid; i < kJumpGroups; i += kJumpThreads) ((uint4 *)b)[kJumpZeroQ + i] = make_uint4(0u, 0u, 0u, 0u);
    for (int w = tid; w < kMtPolyWords; w += kJumpThreads) {
        const uint64_t v = g[w];
#pragma unroll
        for (int r = 0; r < 4; ++r) off[r][w] = (uint32_t)__popcll(v & (kRes << r));
    }
    __syncthreads();
    if (wave < 4) {
        // wavefront r: exclusive scan of class r's 312 counts, 5 words a lane
        constexpr int kPer = (kMtPolyWords + 63) / 64;
        const int r = wave;
        uint32_t v[kPer], sum = 0;
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int q = lane * kPer + e;
            v[e] = q < kMtPolyWords ? off[r][q] : 0u;
            sum += v[e];
        }
        uint32_t incl = sum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(incl, d);
            if (lane >= d) incl += o;
        }
        uint32_t run = incl - sum;
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int q = lane * kPer + e;
            if (q < kMtPolyWords) off[r][q] = run;
            run += v[e];
        }
        if (lane == 63) ncls[r] = incl;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t at = 0;
        for (int r = 0; r < 4; ++r) {
            cls[r] = at;
            at += (ncls[r] + 7u) & ~7u;
        }
        cls[4] = at;
    }
    __syncthreads();
    for (int w = tid; w < kMtPolyWords; w += kJumpThreads) {
        const uint64_t v = g[w];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            uint32_t pos = cls[r] + off[r][w];
            uint64_t bits = v & (kRes << r);
            while (bits) {
                qs[pos++] = (uint16_t)((w * 64 + __builtin_ctzll(bits)) >> 2);
                bits &= bits - 1;
            }
        }
    }
    if (tid < 4)   // pad class tid to its rounded length with the zero groups
        for (uint32_t e = cls[tid] + ncls[tid]; e < cls[tid + 1]; ++e) qs[e] = (uint16_t)kJumpZeroQ;
    __syncthreads();
    const int r = tid / kJumpClassThreads, t = tid - r * kJumpClassThreads;
    if (t < kJumpGroups) {
        const uint4 *b4 = (const uint4 *)b + t;
        {
            uint4 acc = make_uint4(0u, 0u, 0u, 0u);
            const uint4 *q8 = (const uint4 *)(qs + cls[r]);
            const uint32_t n8 = (cls[r + 1] - cls[r]) / 8;
            uint4 v = q8[0];
            for (uint32_t e = 0; e < n8; ++e) {
                const uint4 next = q8[e + 1];   // read ahead (past the class: unused)
                const uint4 a0 = b4[v.x & 0xffffu], a1 = b4[v.x >> 16];
                const uint4 a2 = b4[v.y & 0xffffu], a3 = b4[v.y >> 16];
                const uint4 a4 = b4[v.z & 0xffffu], a5 = b4[v.z >> 16];
                const uint4 a6 = b4[v.w & 0xffffu], a7 = b4[v.w >> 16];
                acc.x ^= ((a0.x ^ a1.x) ^ (a2.x ^ a3.x)) ^ ((a4.x ^ a5.x) ^ (a6.x ^ a7.x));
                acc.y ^= ((a0.y ^ a1.y) ^ (a2.y ^ a3.y)) ^ ((a4.y ^ a5.y) ^ (a6.y ^ a7.y));
                acc.z ^= ((a0.z ^ a1.z) ^ (a2.z ^ a3.z)) ^ ((a4.z ^ a5.z) ^ (a6.z ^ a7.z));
                acc.w ^= ((a0.w ^ a1.w) ^ (a2.w ^ a3.w)) ^ ((a4.w ^ a5.w) ^ (a6.w ^ a7.w));
                v = next;
            }
            ((uint4 *)part[r])[t] = acc;
        }
    }
    __syncthreads();
    for (int k = tid; k < (int)kMtN; k += kJumpThreads)
        states[(size_t)c * kMtN + k] = part[0][k] ^ part[1][k + 1] ^ part[2][k + 2] ^ part[3][k + 3];
}

// Workgroup c generates chunk c: K blocks of tempered draws at
// raw[(624 - idx) + c 624 K ..]; chunk 0 first writes the rest of the
// context's current block at raw[0 .. 624 - idx).
__global__ __launch_bounds__(256) void k_mt_gen(const MtState *__restrict__ st, const uint32_t *__restrict__ states,
                                                int K, uint32_t *__restrict__ raw)
{
    __shared__ uint32_t ring[kMtRing];
    const int tid = threadIdx.x, c = blockIdx.x;
    const uint32_t idx = st->idx;
    const uint32_t *src = c == 0 ? st->mt : states + (size_t)c * kMtN;
    for (int i = tid; i < (int)kMtN; i += 256) ring[i] = src[i];
    __syncthreads();
    const size_t a = kMtN - idx;
    if (c == 0)
        for (uint32_t i = tid; i < a; i += 256) raw[i] = mt_temper(ring[idx + i]);
    uint32_t *out = raw + a + (size_t)c * K * kMtN;
    mt_run(ring, (uint32_t)K * kMtN, [&](uint32_t j, uint32_t v) { out[j] = mt_temper(v); });
}

// Tries are taken 2048 to a workgroup, 8 consecutive ones to a thread (its
// 8 D draws read as 16-byte loads): k_mt_count writes each workgroup's count
// of kept tries, a scan of the counts gives each workgroup its first
// vector's index, and k_mt_take recounts its own tries and writes the kept
// vectors in order.  No per-try array is written.
constexpr uint32_t kTryItems = 8, kTryTile = kTryItems * 256;
static_assert(kTryTile == kMtTryTile, "tries per workgroup");

// The kept flags (bit k) of tries t0 .. t0 + 7 (those below T), their
// vectors in xyz[3 k ..].
template <int D>
__device__ __forceinline__ uint32_t mt_tries8(const uint32_t *__restrict__ raw, uint32_t T, uint32_t t0, float *xyz)
{
    uint32_t w[kTryItems * D];
    if (t0 + kTryItems <= T) {
        // D t0 words from a 16-byte aligned array: t0 is a multiple of 8
        const uint4 *p = (const uint4 *)(raw + (size_t)D * t0);
#pragma unroll
        for (int e = 0; e < 2 * D; ++e) {
            const uint4 q = p[e];
            w[4 * e] = q.x;
            w[4 * e + 1] = q.y;
            w[4 * e + 2] = q.z;
            w[4 * e + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int e = 0; e < (int)kTryItems * D; ++e) w[e] = t0 + (uint32_t)(e / D) < T ? raw[(size_t)D * t0 + e] : 0u;
    }
    uint32_t kept = 0;
#pragma unroll
    for (int k = 0; k < (int)kTryItems; ++k) {
        float x, y, z;
        const bool ok = mt_try<D>(w + D * k, x, y, z);
        xyz[3 * k] = x;
        xyz[3 * k + 1] = y;
        xyz[3 * k + 2] = z;
        if (ok && t0 + (uint32_t)k < T) kept |= 1u << k;
    }
    return kept;
}

// Exclusive scan of one value per thread over a 256-thread workgroup.
__device__ __forceinline__ uint32_t try_block_scan(uint32_t v, uint32_t *lds4, uint32_t &total)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) lds4[wave] = inc;
    __syncthreads();
    uint32_t prefix = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint32_t c = lds4[w];
        if (w < wave) prefix += c;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return prefix + inc - v;
}

template <int D>
__global__ __launch_bounds__(256) void k_mt_count(const uint32_t *__restrict__ raw, uint32_t T,
                                                  uint32_t *__restrict__ part)
{
    __shared__ uint32_t lds4[4];
    float xyz[3 * kTryItems];
    const uint32_t t0 = blockIdx.x * kTryTile + threadIdx.x * kTryItems;
    const uint32_t c = (uint32_t)__popc(mt_tries8<D>(raw, T, t0, xyz));
    uint32_t total;
    try_block_scan(c, lds4, total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// part: exclusive scan of the workgroups' counts (part[nb] = every kept
// try); the first n kept vectors to out in stream order.
template <int D>
__global__ __launch_bounds__(256) void k_mt_take(const uint32_t *__restrict__ raw, uint32_t T,
                                                 const uint32_t *__restrict__ part, uint32_t n, float *__restrict__ out)
{
    __shared__ uint32_t lds4[4];
    const uint32_t first = part[blockIdx.x];
    if (first >= n) return;   // the whole workgroup: its vectors are not asked for
    float xyz[3 * kTryItems];
    const uint32_t t0 = blockIdx.x * kTryTile + threadIdx.x * kTryItems;
    const uint32_t kept = mt_tries8<D>(raw, T, t0, xyz);
    uint32_t total;
    uint32_t p = first + try_block_scan((uint32_t)__popc(kept), lds4, total);
#pragma unroll
    for (int k = 0; k < (int)kTryItems; ++k)
        if ((kept >> k) & 1u) {
            if (p < n) {
                float *o = out + (size_t)p * D;
                o[0] = xyz[3 * k];
                o[1] = xyz[3 * k + 1];
                if (D == 3) o[2] = xyz[3 * k + 2];
            }
            ++p;
        }
}

// The stream moves past the try that kept the m-th vector (m >= 1, at most
// part[nb]): the draws up to it are consumed, and the array becomes the
// block holding the next draw (its words untempered from raw), idx its
// position there -- 624 when the next draw starts a new block, as the
// generator itself leaves it.  The workgroup holding the m-th kept try is
// found by a binary search over part, the try by recounting its tries.
template <int D>
__global__ __launch_bounds__(256) void k_mt_commit(MtState *__restrict__ st, const uint32_t *__restrict__ raw,
                                                   const uint32_t *__restrict__ part, uint32_t nb, uint32_t T,
                                                   uint32_t m)
{
    __shared__ uint32_t s_b, s_t;
    __shared__ uint32_t lds4[4];
    const int tid = threadIdx.x;
    const uint32_t idx0 = st->idx;
    if (tid == 0) {
        uint32_t lo = 0, hi = nb - 1;   // the smallest b with part[b + 1] >= m
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (part[mid + 1] >= m) hi = mid;
            else lo = mid + 1;
        }
        s_b = lo;
    }
    __syncthreads();
    {
        float xyz[3 * kTryItems];
        const uint32_t t0 = s_b * kTryTile + (uint32_t)tid * kTryItems;
        const uint32_t kept = mt_tries8<D>(raw, T, t0, xyz);
        const uint32_t cnt = (uint32_t)__popc(kept);
        uint32_t total;
        const uint32_t ex = part[s_b] + try_block_scan(cnt, lds4, total);
        if (ex < m && m <= ex + cnt) {
            // the (m - ex)-th kept try of this thread's eight
            uint32_t j = m - ex, k = 0;
            for (; k < kTryItems; ++k)
                if (((kept >> k) & 1u) && --j == 0) break;
            s_t = t0 + k;
        }
    }
    __syncthreads();
    const unsigned long long e = (unsigned long long)D * (s_t + 1ull);
    unsigned long long blk = (idx0 + e) / kMtN;
    uint32_t idx1 = (uint32_t)((idx0 + e) % kMtN);
    if (idx1 == 0) {
        blk -= 1;
        idx1 = kMtN;
    }
    if (blk > 0) {
        const uint32_t *src = raw + (kMtN - idx0) + (size_t)(blk - 1) * kMtN;
        for (int i = tid; i < (int)kMtN; i += 256) st->mt[i] = mt_untemper(src[i]);
    }
    __syncthreads();   // every thread has read idx before it changes
    if (tid == 0) {
        st->idx = idx1;
        st->draws += e;
    }
}

static inline unsigned blocks256(size_t n) { return (unsigned)((n + 255) / 256); }

void launch_mt_base(const MtState *st, uint32_t *base, hipStream_t s)
{
    hipLaunchKernelGGL(k_mt_base, dim3(1), dim3(256), 0, s, st, base);
}

void launch_mt_jump(const uint32_t *base, const uint64_t *polys, int n_jumps, uint32_t *states, hipStream_t s)
{
    if (n_jumps <= 0) return;
    hipLaunchKernelGGL(k_mt_jump, dim3(n_jumps), dim3(kJumpThreads), 0, s, base, polys, states);
}

void launch_mt_gen(const MtState *st, const uint32_t *states, int n_chunks, int K, uint32_t *raw, hipStream_t s)
{
    hipLaunchKernelGGL(k_mt_gen, dim3(n_chunks), dim3(256), 0, s, st, states, K, raw);
}

uint32_t mt_try_blocks(uint32_t T) { return (T + kTryTile - 1) / kTryTile; }

void launch_mt_count(int D, const uint32_t *raw, uint32_t T, uint32_t *part, hipStream_t s)
{
    if (T == 0) return;
    if (D == 3) hipLaunchKernelGGL(k_mt_count<3>, dim3(mt_try_blocks(T)), dim3(256), 0, s, raw, T, part);
    else hipLaunchKernelGGL(k_mt_count<2>, dim3(mt_try_blocks(T)), dim3(256), 0, s, raw, T, part);
}

void launch_mt_take(int D, const uint32_t *raw, uint32_t T, const uint32_t *part, uint32_t n, float *out,
                    hipStream_t s)
{
    if (T == 0 || n == 0) return;
    if (D == 3) hipLaunchKernelGGL(k_mt_take<3>, dim3(mt_try_blocks(T)), dim3(256), 0, s, raw, T, part, n, out);
    else hipLaunchKernelGGL(k_mt_take<2>, dim3(mt_try_blocks(T)), dim3(256), 0, s, raw, T, part, n, out);
}

void launch_mt_commit(int D, MtState *st, const uint32_t *raw, const uint32_t *part, uint32_t T, uint32_t m,
                      hipStream_t s)
{
    if (T == 0 || m == 0) return;
    const uint32_t nb = mt_try_blocks(T);
    if (D == 3) hipLaunchKernelGGL(k_mt_commit<3>, dim3(1), dim3(256), 0, s, st, raw, part, nb, T, m);
    else hipLaunchKernelGGL(k_mt_commit<2>, dim3(1), dim3(256), 0, s, st, raw, part, nb, T, m);
}

}  // namespace fgl
