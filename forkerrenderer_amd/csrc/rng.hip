// rng.hip -- the reference's random stream on gfx950.
//
// The reference draws every random float from one function-static
// std::mt19937 (default seed 5489) through uniform_real_distribution<float>
// (src/utility.h:90-103), and builds its random vectors by rejection
// (src/geometry.h:947-990).  Its consumers (Render::DoSSAO, the PCF / PCSS
// shadow filters) take the stream strictly in pixel order, so the device
// replays it in two steps: one workgroup runs the generator and writes the
// accepted vectors in stream order (k_mt_accept); the per-pixel kernels then
// read the vectors of their pixel at a known offset.
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

// The twist of std::mt19937 on 624 LDS words in three dependent phases:
// words [0,227) read old words only; [227,454) read the new [0,227);
// [454,624) read the new [227,397) (and word 623 the new word 0).  Each
// phase reads, waits for every read, then writes.
__device__ __forceinline__ void mt_twist(uint32_t *mt)
{
    const int tid = threadIdx.x;
    uint32_t v = 0;
    if (tid < 227) v = mt_twist_word(mt[tid], mt[tid + 1], mt[tid + kMtM]);
    __syncthreads();
    if (tid < 227) mt[tid] = v;
    __syncthreads();
    if (tid < 227) v = mt_twist_word(mt[227 + tid], mt[228 + tid], mt[tid]);
    __syncthreads();
    if (tid < 227) mt[227 + tid] = v;
    __syncthreads();
    if (tid < 170) {
        const int i = 454 + tid;
        v = mt_twist_word(mt[i], mt[(i + 1) % kMtN], mt[i - 227]);
    }
    __syncthreads();
    if (tid < 170) mt[454 + tid] = v;
    __syncthreads();
}

// One workgroup of 256 threads: draws tries of D draws (3: unit sphere,
// 2: unit disk) until n_accept are kept, writes the kept vectors (D floats
// each) to out in stream order, and leaves the stream right after the last
// kept try -- where the reference's next draw would come from.
template <int D>
__global__ __launch_bounds__(256) void k_mt_accept(MtState *__restrict__ st, float *__restrict__ out,
                                                   uint32_t n_accept)
{
    __shared__ uint32_t mt[kMtN];
    __shared__ uint32_t buf[kMtN + 4];
    __shared__ uint32_t wcnt[4];
    __shared__ uint32_t s_last;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < (int)kMtN; i += 256) mt[i] = st->mt[i];
    uint32_t idx = st->idx;
    unsigned long long draws = st->draws;
    uint32_t carry = 0, accepted = 0;
    __syncthreads();
    while (accepted < n_accept) {
        if (idx >= kMtN) {
            mt_twist(mt);
            idx = 0;
        }
        const uint32_t avail = kMtN - idx;
        for (uint32_t i = tid; i < avail; i += 256) buf[carry + i] = mt_temper(mt[idx + i]);
        __syncthreads();
        const uint32_t total = carry + avail, ntries = total / D, need = n_accept - accepted;
        uint32_t base = 0;
        bool done = false;
        for (uint32_t c0 = 0; c0 < ntries; c0 += 256) {
            const uint32_t t = c0 + tid;
            bool ok = false;
            float x = 0.f, y = 0.f, z = 0.f;
            if (t < ntries) {
                if (D == 3) {
                    z = random_pm1(buf[3 * t]);
                    y = random_pm1(buf[3 * t + 1]);
                    x = random_pm1(buf[3 * t + 2]);
                } else {
                    y = random_pm1(buf[2 * t]);
                    x = random_pm1(buf[2 * t + 1]);
                }
                ok = (x * x + y * y) + z * z < 1.f;
            }
            const unsigned long long m = __ballot(ok);
            if (lane == 0) wcnt[wave] = (uint32_t)__popcll(m);
            __syncthreads();
            uint32_t pre = 0, tot = 0;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const uint32_t cw = wcnt[w];
                if (w < wave) pre += cw;
                tot += cw;
            }
            const uint32_t rank = base + pre + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (ok && rank < need) {
                float *o = out + (size_t)(accepted + rank) * D;
                o[0] = x;
                o[1] = y;
                if (D == 3) o[2] = z;
                if (rank == need - 1) s_last = t;
            }
            __syncthreads();   // wcnt is rewritten by the next chunk; s_last is visible
            if (base + tot >= need) {
                done = true;
                break;
            }
            base += tot;
        }
        if (done) {
            const uint32_t used = (s_last + 1) * D - carry;   // draws of this block up to the last kept try
            idx += used;
            draws += used;
            accepted = n_accept;
            break;
        }
        accepted += base;
        draws += avail;
        idx = kMtN;
        const uint32_t rem = total - ntries * D;   // a try started here ends in the next block
        uint32_t keep = 0;
        if ((uint32_t)tid < rem) keep = buf[ntries * D + tid];
        __syncthreads();
        if ((uint32_t)tid < rem) buf[tid] = keep;
        carry = rem;
        __syncthreads();
    }
    __syncthreads();
    for (int i = tid; i < (int)kMtN; i += 256) st->mt[i] = mt[i];
    if (tid == 0) {
        st->idx = idx;
        st->draws = draws;
    }
}

void launch_rng_accept(int draws_per_try, MtState *st, float *out, uint32_t n_accept, hipStream_t s)
{
    if (n_accept == 0) return;
    if (draws_per_try == 3) hipLaunchKernelGGL(k_mt_accept<3>, dim3(1), dim3(256), 0, s, st, out, n_accept);
    else hipLaunchKernelGGL(k_mt_accept<2>, dim3(1), dim3(256), 0, s, st, out, n_accept);
}

}  // namespace fgl
