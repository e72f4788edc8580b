// shade.hip -- screen-space passes for gfx950: deferred lighting, kernel SSAO,
// AO blurs and the quantise + SSAA resolve.  All read/write SoA float planes
// in HBM with 64-wide rows per wavefront.
#include "shade_core.hpp"

namespace fgl {

template <bool STREAM>
__global__ __launch_bounds__(256) void k_lighting(FrameUniforms u, const float *__restrict__ g,
                                                  const float *__restrict__ ao, int ao_is_one, NetAo net,
                                                  const float *__restrict__ smap, ShadowSamples ss,
                                                  float *__restrict__ fb)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= u.W || y >= u.H) return;
    lighting_px<STREAM>(u, g, ao, ao_is_one, smap, ss, fb, x, y, net);
}

__global__ __launch_bounds__(256) void k_ssao_ref(FrameUniforms u, const float *__restrict__ samples,
                                                  const float *__restrict__ pow3,
                                                  const float *__restrict__ depth,
                                                  const float *__restrict__ g, float *__restrict__ ao)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= u.W || y >= u.H) return;
    ssao_ref_px(u, samples, pow3, depth, g, ao, x, y);
}

// ---- PCSS sample offsets for the stochastic filter (FGL_SHADOW_REF_PCSS) ----
//
// The lighting pass visits pixels in row order (the forward pass its passing
// fragments in raster order) and each shading point draws n_blocker
// unit-disk samples, plus n_pcf more when its blocker search finds a blocker
// (PCSS, shadow.cpp:93-106), so a point's offset into the stream depends on
// the outcome of every search before it.  Most searches are decided by the
// shadow map alone: k_pcss_classify checks every texel a sample of the point
// could reach (the square around the disk) and marks the point "no blocker
// for any sample set", "a blocker for any sample set" or "depends on the
// samples".  Only the last kind is walked in order (k_pcss_points,
// k_pcss_guess, k_pcss_band, k_pcss_walk_blocks, k_pcss_chain below), at
// offsets that count the decided points before it.  The points' inputs are planes of stride P:
// normal, world position, light NDC.

// Light 0's shadow coordinate and bias at point i, as shade_point computes them.
__device__ __forceinline__ void pcss_pixel(const FrameUniforms &u, const float *__restrict__ g, size_t P, size_t i,
                                           Vec3 &sc, float &bias)
{
    const Vec3 N = vec3(g[i], g[P + i], g[2 * P + i]);
    const Vec3 Pw = vec3(g[3 * P + i], g[4 * P + i], g[5 * P + i]);
    const Vec3 lndc = vec3(g[6 * P + i], g[7 * P + i], g[8 * P + i]);
    sc = shadow_coord(lndc);
    bias = shadow_bias(N, light_dir(u.lights[0], Pw));
}

// cnt[i] = n_blocker + n_pcf for "a blocker for any sample set", n_blocker
// otherwise; unc[i] = 1 for "depends on the samples".  Nothing is drawn with
// shadows off or without a light.
__global__ __launch_bounds__(256) void k_pcss_classify(FrameUniforms u, const float *__restrict__ g, size_t P,
                                                       const float *__restrict__ smap, uint32_t *__restrict__ cnt,
                                                       uint32_t *__restrict__ unc)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const DevShadowCfg &c = u.shadow;
    if (!(u.shadow_on && u.n_lights > 0)) {
        cnt[i] = 0;
        unc[i] = 0;
        return;
    }
    Vec3 sc;
    float bias;
    pcss_pixel(u, g, P, i, sc, bias);
    // every sample coordinate lies in [sc - F, sc + F], F = (float)blocker_filter
    // (|(float)(d * filter)| <= (float)filter for |d| < 1; float add and the
    // texel index are monotone)
    const float F = (float)c.blocker_filter;
    const float u0 = sc.x + (-F), u1 = sc.x + F, v0 = sc.y + (-F), v1 = sc.y + F;
    int cls = 2;
    if (u0 == u0 && u1 == u1 && v0 == v0 && v1 == v1 && sc.z == sc.z && bias == bias) {
        const bool inside = u0 >= 0.f && u1 <= 1.f && v0 >= 0.f && v1 <= 1.f;
        bool any_blk = false, any_clear = !inside;   // samples off the map never block
        if (u1 >= 0.f && u0 <= 1.f && v1 >= 0.f && v0 <= 1.f) {
            const int w = (int)((float)u.SW - 0.001f), h = (int)((float)u.SH - 0.001f);
            const int x0 = (int)((float)w * (u0 < 0.f ? 0.f : u0)), x1 = (int)((float)w * (u1 > 1.f ? 1.f : u1));
            const int y0 = (int)((float)h * (v0 < 0.f ? 0.f : v0)), y1 = (int)((float)h * (v1 > 1.f ? 1.f : v1));
            for (int y = y0; y <= y1 && !(any_blk && any_clear); ++y)
                for (int x = x0; x <= x1; ++x) {
                    const float d = smap[(size_t)x + (size_t)y * u.SW];
                    const float s = ((double)d < 0.001) ? 1.f : d;
                    const bool blk = sc.z > s + bias;
                    // a blocker depth within reach of 0.001 could average
                    // below it: not decided
                    if (blk && !((double)s > 0.0011)) any_clear = true;
                    any_blk |= blk;
                    any_clear |= !blk;
                }
        }
        if (!any_blk) cls = 0;
        else if (!any_clear) cls = 1;
    }
    cnt[i] = (uint32_t)c.n_blocker + (cls == 1 ? (uint32_t)c.n_pcf : 0u);
    unc[i] = cls == 2 ? 1u : 0u;
}

// list[upos[i]] = i for the undecided pixels (upos = exclusive scan of the
// undecided flags, total at upos[P]; a flag is upos[i + 1] - upos[i]).
__global__ __launch_bounds__(256) void k_pcss_compact(const uint32_t *__restrict__ upos, size_t P,
                                                      uint32_t *__restrict__ list)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P && upos[i + 1] != upos[i]) list[upos[i]] = (uint32_t)i;
}

// The undecided points' inputs that do not depend on the stream: light 0's
// shadow coordinate and bias, and the offset of their samples were no earlier
// undecided point to find a blocker (base of the point).
__global__ __launch_bounds__(256) void k_pcss_points(FrameUniforms u, const float *__restrict__ g, size_t P,
                                                     const uint32_t *__restrict__ base,
                                                     const uint32_t *__restrict__ list, uint32_t n,
                                                     float4 *__restrict__ pt, uint32_t *__restrict__ pbase)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const size_t i = list[k];
    Vec3 sc;
    float bias;
    pcss_pixel(u, g, P, i, sc, bias);
    pt[k] = make_float4(sc.x, sc.y, sc.z, bias);
    pbase[k] = base[i];
}

// The undecided points in order: point k draws from
// pbase[k] + n_pcf * (k - M_k), M_k the undecided points before it that find
// no blocker, so each outcome moves every later point's samples.  They are
// settled a segment [s0, s0 + L) at a time, M_{s0} = M0 known:
//  * a prediction of M_k: every point evaluated once at the current
//    prediction (k_pcss_guess), its misses scanned (kGuesses times, from a
//    straight line at the miss rate so far) -- exact wherever the outcomes
//    do not depend on the samples' offset, which is most points;
//  * k_pcss_band evaluates every point at the 63 values of M_k around its
//    prediction, a wavefront per point, the outcomes as a bit mask;
//  * k_pcss_walk_blocks and k_pcss_chain follow the actual M_k through the
//    masks in order, write found[k] (1: a blocker) and stop where M_k
//    leaves the band or the segment ends; the host starts the next segment
//    there.
// Every evaluation is the blocker search exactly as FindAverageBlockDepth
// runs it (shadow.cpp:65-91: the samples in order, blocker depths summed in
// that order).  Every segment settles its first point (its prediction is
// exact).
constexpr int kBandHalf = 31;   // M_k in [prediction - 31, prediction + 31]: a wavefront per point
constexpr int kGuesses = 2;     // prediction rounds per segment (more do not pay)

// The blocker search of an undecided point at sample offset o (the stream
// samples' loads grouped as in the shading, fgl_device.hpp).
__device__ __forceinline__ bool pcss_search(const FrameUniforms &u, const float *__restrict__ smap,
                                            const float *__restrict__ pairs, float4 q, uint32_t o)
{
    const float dB = pcss_blocker_stream(u, smap, vec3(q.x, q.y, q.z), q.w, pairs + 2 * (size_t)o, u.shadow.n_blocker);
    return !((double)dB < 0.001);
}

// pred == nullptr: the prediction M0 + round(i * rate); else M0 + pred[i]
// (the scanned misses of the previous guess).  Clamped to the possible
// [M0, M0 + i].
__device__ __forceinline__ uint32_t pcss_predict(const uint32_t *__restrict__ pred, uint32_t M0, float rate,
                                                 uint32_t i)
{
    const uint32_t e = pred ? pred[i] : (uint32_t)((float)i * rate + 0.5f);
    return M0 + (e < i ? e : i);
}

__global__ __launch_bounds__(256) void k_pcss_guess(FrameUniforms u, const float *__restrict__ smap,
                                                    const float *__restrict__ pairs, const float4 *__restrict__ pt,
                                                    const uint32_t *__restrict__ pbase, uint32_t s0, uint32_t L,
                                                    uint32_t M0, float rate, const uint32_t *__restrict__ pred,
                                                    uint32_t *__restrict__ miss)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= L) return;
    const uint32_t k = s0 + i, M = pcss_predict(pred, M0, rate, i);
    miss[i] = pcss_search(u, smap, pairs, pt[k], pbase[k] + (uint32_t)u.shadow.n_pcf * (k - M)) ? 0u : 1u;
}

__global__ __launch_bounds__(256) void k_pcss_band(FrameUniforms u, const float *__restrict__ smap,
                                                   const float *__restrict__ pairs, const float4 *__restrict__ pt,
                                                   const uint32_t *__restrict__ pbase, uint32_t s0, uint32_t L,
                                                   uint32_t M0, const uint32_t *__restrict__ pred,
                                                   unsigned long long *__restrict__ masks)
{
    const uint32_t i = (blockIdx.x * 256 + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (i >= L) return;   // whole wavefronts
    const int cand = (int)lane;
    const long long M = (long long)pcss_predict(pred, M0, 0.f, i) - kBandHalf + cand;
    bool f = false;
    if (cand <= 2 * kBandHalf && M >= (long long)M0 && M <= (long long)M0 + i) {
        const uint32_t k = s0 + i;
        f = pcss_search(u, smap, pairs, pt[k], pbase[k] + (uint32_t)u.shadow.n_pcf * (k - (uint32_t)M));
    }
    const unsigned long long b = __ballot(f);
    if (lane == 0) masks[i] = b;
}

// Following the actual M_k through the masks.  With P_i the prediction
// (misses before point i in the segment, pred[i]) and D_i = 31 + M_i - M0 -
// P_i the actual M_i's place in point i's band, D_{i+1} = D_i + (miss_i(D_i)
// - (P_{i+1} - P_i)): D stays put while the prediction holds.  So a
// wavefront walking 64 points tests them all at one D; a ballot finds the
// first point whose outcome breaks the prediction (equals the predicted
// step), the points up to it settle (every one before it with the opposite
// outcome), and D moves by one from there.
//
// The segment is cut into blocks of 1024 points, and every block is walked
// from each of the 63 places D can enter it with (k_pcss_walk_blocks: a
// workgroup per block and 16 places, a wavefront per place; block 0 only
// from 31), each walk writing its outcomes, the points it settled and the D
// it left with.  k_pcss_chain then follows the actual D from block to block
// through those results and copies the chosen walks' outcomes.  A walk stops
// where D leaves the band; so does the chain, at the first block whose walk
// stopped or which D enters from outside the band.
constexpr int kBandWidth = 2 * kBandHalf + 1;
constexpr int kChainWaves = 16;   // start places per workgroup
static_assert(kPcssChainBlock == 1024 && kPcssBandWidth == kBandWidth, "chain layout");

__global__ __launch_bounds__(64 * kChainWaves) void k_pcss_walk_blocks(const unsigned long long *__restrict__ masks,
                                                                      const uint32_t *__restrict__ pred, uint32_t L,
                                                                      unsigned long long *__restrict__ fbits,
                                                                      uint32_t *__restrict__ res)
{
    __shared__ unsigned long long sm[kPcssChainBlock];
    __shared__ int sdl[kPcssChainBlock];
    const uint32_t b = blockIdx.x, at = b * kPcssChainBlock;
    const uint32_t nb = L - at < kPcssChainBlock ? L - at : kPcssChainBlock;
    for (uint32_t j = threadIdx.x; j < nb; j += 64 * kChainWaves) {
        sm[j] = masks[at + j];
        sdl[j] = (int)(pred[at + j + 1] - pred[at + j]);   // pred[L]: the segment's predicted total
    }
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63), d0 = (int)blockIdx.y * kChainWaves + (int)(threadIdx.x >> 6);
    if (d0 >= kBandWidth || (b == 0 && d0 != kBandHalf)) return;   // whole wavefronts, after the only barrier
    unsigned long long *fo = fbits + ((size_t)b * kBandWidth + d0) * (kPcssChainBlock / 64);
    int D = d0;
    uint32_t settled = 0;
    for (uint32_t g = 0; g < nb; g += 64) {
        const uint32_t cnt = nb - g < 64 ? nb - g : 64;
        const unsigned long long valid = cnt == 64 ? ~0ull : ((1ull << cnt) - 1ull);
        // (slots past the block are never settled: valid masks them out)
        const unsigned long long m = sm[g + lane];
        const int dl = sdl[g + lane];
        const unsigned long long bm = dl ? m : ~m;
        const unsigned long long dlm = __ballot(dl != 0);
        uint32_t start = 0;
        unsigned long long brks = 0;
        bool stop = false;
        while (start < cnt) {
            if (D < 0 || D > 2 * kBandHalf) {
                stop = true;
                break;
            }
            const unsigned long long hit = __ballot((uint32_t)(bm >> D) & 1u) & valid & (~0ull << start);
            if (!hit) {
                start = cnt;
                break;
            }
            const uint32_t j = (uint32_t)__builtin_ctzll(hit);
            brks |= 1ull << j;
            D += ((dlm >> j) & 1ull) ? -1 : 1;
            start = j + 1;
        }
        if (lane == 0) fo[g / 64] = ~(dlm ^ brks);   // outcomes of the settled points
        settled += start;
        if (stop) break;
    }
    if (lane == 0) {
        uint32_t *r = res + 2 * ((size_t)b * kBandWidth + d0);
        r[0] = settled;
        r[1] = (uint32_t)D;
    }
}

// One workgroup: the chain through the blocks' walks (their results staged
// in LDS), then found[] for the settled points.  out = (first unsettled
// point, M there).
__global__ __launch_bounds__(256) void k_pcss_chain(const unsigned long long *__restrict__ fbits,
                                                    const uint32_t *__restrict__ res, const uint32_t *__restrict__ pred,
                                                    uint32_t s0, uint32_t L, uint32_t M0, uint32_t *__restrict__ found,
                                                    uint32_t *__restrict__ out)
{
    constexpr uint32_t kMaxBlocks = kPcssSegmentMax / kPcssChainBlock;
    __shared__ uint32_t sres[2 * kBandWidth * kMaxBlocks];
    __shared__ int pick[kMaxBlocks];
    __shared__ uint32_t s_settled;
    __shared__ int s_D;
    const uint32_t tid = threadIdx.x;
    const uint32_t nblk = (L + kPcssChainBlock - 1) / kPcssChainBlock;
    for (uint32_t i = tid; i < 2 * kBandWidth * nblk; i += 256) sres[i] = res[i];
    __syncthreads();
    if (tid == 0) {
        int D = kBandHalf;
        uint32_t settled = 0;
        for (uint32_t b = 0; b < nblk; ++b) {
            if (D < 0 || D > 2 * kBandHalf) break;
            const uint32_t len = L - b * kPcssChainBlock < kPcssChainBlock ? L - b * kPcssChainBlock : kPcssChainBlock;
            const uint32_t r = sres[2 * (b * kBandWidth + (uint32_t)D)];
            pick[b] = D;
            D = (int)sres[2 * (b * kBandWidth + (uint32_t)D) + 1];
            settled += r;
            if (r < len) break;
        }
        s_settled = settled;
        s_D = D;
    }
    __syncthreads();
    const uint32_t settled = s_settled;
    for (uint32_t k = tid; k < settled; k += 256) {
        const uint32_t b = k / kPcssChainBlock, kk = k - b * kPcssChainBlock;
        const unsigned long long w = fbits[((size_t)b * kBandWidth + (uint32_t)pick[b]) * (kPcssChainBlock / 64) + kk / 64];
        found[s0 + k] = (uint32_t)(w >> (kk & 63)) & 1u;
    }
    if (tid == 0) {
        out[0] = s0 + settled;
        out[1] = (uint32_t)((int)M0 + (int)pred[settled] + s_D - kBandHalf);
    }
}

// offs[i] = base[i] + n_pcf * fpos[upos[i]] for i = 0..P (offs[P] = total).
__global__ __launch_bounds__(256) void k_pcss_offsets(const uint32_t *__restrict__ base,
                                                      const uint32_t *__restrict__ upos,
                                                      const uint32_t *__restrict__ fpos, uint32_t n_pcf, size_t P,
                                                      uint32_t *__restrict__ offs)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= P) offs[i] = base[i] + n_pcf * fpos[upos[i]];
}

// Kernel SSAO over 64 x 16-pixel tiles (1024 threads, a wavefront per row):
// the workgroup first copies the depth plane around its tile, 32 pixels
// each side, into LDS (coalesced 16-byte loads), and taps that land there
// read LDS; the others read the plane.  Per tap it issues about 42 vector
// ALU operations (projection, reciprocal, bounds, address, compare).
constexpr int kSsaoTW = 64, kSsaoTH = 16, kSsaoHalo = 32;
static_assert(kSsaoTW * kSsaoTH == 1024 && kSsaoHalo % 4 == 0, "one thread per pixel; 16-byte staging");
constexpr int kSsaoLW = kSsaoTW + 2 * kSsaoHalo, kSsaoLH = kSsaoTH + 2 * kSsaoHalo;

__global__ __launch_bounds__(1024) void k_ssao_tiled(FrameUniforms u, SsaoArgs sa,
                                                     const float *__restrict__ depth,
                                                     const float *__restrict__ g,
                                                     float *__restrict__ ao)
{
    __shared__ float tile[kSsaoLH * kSsaoLW];
    const int x0 = blockIdx.x * kSsaoTW, y0 = blockIdx.y * kSsaoTH;
    const int lx0 = x0 - kSsaoHalo, ly0 = y0 - kSsaoHalo;
    const bool vec = (u.W & 3) == 0;   // rows start 16-byte aligned
    for (int k = threadIdx.x; k < kSsaoLH * kSsaoLW / 4; k += 1024) {
        const int row = k / (kSsaoLW / 4), c4 = k - row * (kSsaoLW / 4);
        const int gy = ly0 + row, gx = lx0 + 4 * c4;
        float v[4] = {0.f, 0.f, 0.f, 0.f};   // outside the plane: never read
        if (gy >= 0 && gy < u.H) {
            const float *src = depth + (size_t)gy * u.W;
            if (vec && gx >= 0 && gx + 3 < u.W) {
                const float4 q = *(const float4 *)(src + gx);
                v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (gx + e >= 0 && gx + e < u.W) v[e] = src[gx + e];
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) tile[row * kSsaoLW + 4 * c4 + e] = v[e];
    }
    __syncthreads();
    const int x = x0 + (int)(threadIdx.x % kSsaoTW), y = y0 + (int)(threadIdx.x / kSsaoTW);
    if (x >= u.W || y >= u.H) return;
    const uint32_t W = (uint32_t)u.W;
    // The plane is read through a global-address-space pointer: with two
    // generic pointers the compiler merges both reads into one flat load,
    // which sends the LDS taps down the vector-memory path as well.
    const __attribute__((address_space(1))) float *gdepth =
        (const __attribute__((address_space(1))) float *)depth;
    ssao_px_with(u, sa, depth, g, ao, x, y, [&](int sx, int sy) {
        const uint32_t tx = (uint32_t)(sx - lx0), ty = (uint32_t)(sy - ly0);
        const bool near = tx < (uint32_t)kSsaoLW && ty < (uint32_t)kSsaoLH;
        // Every lane reads LDS (a far lane at the tile's first word) and far
        // lanes also read the plane, into another register: the LDS read
        // does not wait for the plane's.
        const float dl = tile[near ? ty * kSsaoLW + tx : 0u];
        float dg = 0.f;
        if (!near) dg = gdepth[(uint32_t)sx + (uint32_t)sy * W];
        return near ? dl : dg;
    });
}

// Buffer1f::TwoPassGaussianBlurDenoised (src/buffer.cpp:59-98) is in place and
// order dependent: the taps left of (or below) a pixel read values already
// blurred.  One thread owns a whole row (then a whole column) and walks it in
// the reference's order, so every read sees exactly what the reference sees.
__constant__ float c_ref_blur_w[5] = {(float)0.227027, (float)0.1945946, (float)0.1216216,
                                      (float)0.054054, (float)0.016216};

// Buffer1f::TwoPassGaussianBlurDenoised (src/buffer.cpp:59-98) along one
// line of n samples at stride s, in place: each output is the reference's
// sum in its order -- the centre, then for i = 1..4 the sample i ahead
// (not yet blurred) and the sample i behind (already blurred), indices
// clamped to the line.  The last four outputs and the next four inputs stay
// in registers, and the inputs are loaded kPF steps ahead (samples ahead are
// never written before they are read), so a step waits on its own
// arithmetic, not on memory.
constexpr int kPF = 64;

__device__ __forceinline__ void blur_rec_line(float *__restrict__ p, int n, size_t s)
{
    const float w0 = c_ref_blur_w[0], w1 = c_ref_blur_w[1], w2 = c_ref_blur_w[2], w3 = c_ref_blur_w[3],
                w4 = c_ref_blur_w[4];
    auto at = [&](int x) { return p[(size_t)(x < n - 1 ? x : n - 1) * s]; };
    float a0 = at(0), a1 = at(1), a2 = at(2), a3 = at(3), a4 = at(4);
    // x = 0: every sample behind clamps to sample 0, not yet blurred
    float r = a0 * w0;
    r += a1 * w1;
    r += a0 * w1;
    r += a2 * w2;
    r += a0 * w2;
    r += a3 * w3;
    r += a0 * w3;
    r += a4 * w4;
    r += a0 * w4;
    p[0] = r;
    float o1 = r, o2 = r, o3 = r, o4 = r;   // from here on they clamp to the blurred sample 0
    a0 = a1;
    a1 = a2;
    a2 = a3;
    a3 = a4;
    float ring[kPF];
#pragma unroll
    for (int j = 0; j < kPF; ++j) ring[j] = at(5 + j);
    for (int x0 = 1; x0 < n; x0 += kPF) {
#pragma unroll
        for (int j = 0; j < kPF; ++j) {
            const int x = x0 + j;
            if (x < n) {
                a4 = ring[j];                   // sample x + 4
                ring[j] = at(x + 4 + kPF);      // for step x + kPF
                float q = a0 * w0;
                q += a1 * w1;
                q += o1 * w1;
                q += a2 * w2;
                q += o2 * w2;
                q += a3 * w3;
                q += o3 * w3;
                q += a4 * w4;
                q += o4 * w4;
                p[(size_t)x * s] = q;
                o4 = o3;
                o3 = o2;
                o2 = o1;
                o1 = q;
                a0 = a1;
                a1 = a2;
                a2 = a3;
                a3 = a4;
            }
        }
    }
}

__global__ __launch_bounds__(64) void k_blur_rec_rows(float *__restrict__ a, int W, int H)
{
    const int h = blockIdx.x * 64 + threadIdx.x;
    if (h >= H) return;
    blur_rec_line(a + (size_t)h * W, W, 1);
}

__global__ __launch_bounds__(64) void k_blur_rec_cols(float *__restrict__ a, int W, int H)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    if (x >= W) return;
    blur_rec_line(a + x, H, (size_t)W);
}

struct BlurWeights {
    float w[9];
    int R;
};

// Separable Gaussian in one pass: a workgroup stages a (16 + 2R) x (64 + 2R)
// block of the source in LDS (coordinates clamped to the image, as the taps
// are), runs the horizontal taps for its 16 + 2R rows (the rows the vertical
// taps of its 64 x 16 outputs read, clamped likewise) into a second LDS
// block, then the vertical taps; each tap sum in the reference loop's order
// (centre, then +i and -i for i = 1..R), so the result is the
// horizontal-then-vertical two-pass filter bit for bit, without an
// intermediate plane in HBM.  Writes dst (not src: neighbouring workgroups
// read src).
constexpr int kBlurTW = 64, kBlurTH = 16, kBlurMaxR = 8;

template <int R>
__global__ __launch_bounds__(256) void k_blur_fused(const float *__restrict__ src, float *__restrict__ dst,
                                                    int W, int H, BlurWeights bw)
{
    __shared__ float s[kBlurTH + 2 * kBlurMaxR][kBlurTW + 2 * kBlurMaxR];
    __shared__ float t[kBlurTH + 2 * kBlurMaxR][kBlurTW];
    constexpr int rows = kBlurTH + 2 * R, cols = kBlurTW + 2 * R;
    const int x0 = blockIdx.x * kBlurTW, y0 = blockIdx.y * kBlurTH;
    for (int k = threadIdx.x; k < rows * cols; k += 256) {
        const int j = k / cols, i = k - j * cols;
        const int gy = clampi(y0 - R + j, 0, H - 1), gx = clampi(x0 - R + i, 0, W - 1);
        s[j][i] = src[(size_t)gx + (size_t)gy * W];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < rows * kBlurTW; k += 256) {
        const int j = k / kBlurTW, x = k - j * kBlurTW;
        float r = s[j][x + R] * bw.w[0];
#pragma unroll
        for (int i = 1; i <= R; ++i) {
            r += s[j][x + R + i] * bw.w[i];
            r += s[j][x + R - i] * bw.w[i];
        }
        t[j][x] = r;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kBlurTH * kBlurTW; k += 256) {
        const int yy = k / kBlurTW, x = k - yy * kBlurTW;
        const int gx = x0 + x, gy = y0 + yy;
        if (gx >= W || gy >= H) continue;
        float r = t[yy + R][x] * bw.w[0];
#pragma unroll
        for (int i = 1; i <= R; ++i) {
            r += t[yy + R + i][x] * bw.w[i];
            r += t[yy + R - i][x] * bw.w[i];
        }
        dst[(size_t)gx + (size_t)gy * W] = r;
    }
}

__global__ __launch_bounds__(256) void k_resolve(const float *__restrict__ fb, int W, int H, int k,
                                                 uint8_t *__restrict__ img)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W / k || y >= H / k) return;
    resolve_px(fb, W, H, k, img, x, y);
}

// Plain 16-byte copy: the reference point for HBM byte counters.
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                              size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

void launch_copy(const void *src, void *dst, size_t bytes, hipStream_t s)
{
    size_t n = bytes / 16;
    if (n == 0) return;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const float4 *)src,
                       (float4 *)dst, n);
}

static inline dim3 grid2(int W, int H) { return dim3((W + 63) / 64, (H + 3) / 4); }

void launch_lighting(const FrameUniforms &u, const float *gbuf, const float *ao, int ao_is_one, const NetAo &net,
                     const float *shadow_map, const ShadowSamples &ss, float *fb, hipStream_t s)
{
    if (u.W <= 0 || u.H <= 0) return;
    if (ss.pairs)
        hipLaunchKernelGGL(k_lighting<true>, grid2(u.W, u.H), dim3(256), 0, s, u, gbuf, ao, ao_is_one, net,
                           shadow_map, ss, fb);
    else
        hipLaunchKernelGGL(k_lighting<false>, grid2(u.W, u.H), dim3(256), 0, s, u, gbuf, ao, ao_is_one, net,
                           shadow_map, ss, fb);
}

void launch_ssao_reference(const FrameUniforms &u, const float *samples, const float *pow3,
                           const float *depth, const float *gbuf, float *ao, hipStream_t s)
{
    if (u.W <= 0 || u.H <= 0) return;
    hipLaunchKernelGGL(k_ssao_ref, grid2(u.W, u.H), dim3(256), 0, s, u, samples, pow3, depth, gbuf, ao);
}

void launch_pcss_classify(const FrameUniforms &u, const float *g, size_t P, const float *shadow_map, uint32_t *cnt,
                          uint32_t *unc, hipStream_t s)
{
    if (P == 0) return;
    hipLaunchKernelGGL(k_pcss_classify, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, u, g, P, shadow_map, cnt,
                       unc);
}

void launch_pcss_compact(const uint32_t *upos, size_t P, uint32_t *list, hipStream_t s)
{
    if (P == 0) return;
    hipLaunchKernelGGL(k_pcss_compact, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, upos, P, list);
}

void launch_pcss_points(const FrameUniforms &u, const float *g, size_t P, const uint32_t *base, const uint32_t *list,
                        uint32_t n, float4 *pt, uint32_t *pbase, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_pcss_points, dim3((n + 255) / 256), dim3(256), 0, s, u, g, P, base, list, n, pt, pbase);
}

void launch_pcss_segment(const FrameUniforms &u, const float *shadow_map, const float *pairs, const float4 *pt,
                         const uint32_t *pbase, uint32_t s0, uint32_t L, uint32_t M0, float rate, uint32_t *pred,
                         uint32_t *scan_tmp, unsigned long long *masks, unsigned long long *fbits, uint32_t *res,
                         uint32_t *found, uint32_t *out, hipStream_t s)
{
    if (L == 0 || L > kPcssSegmentMax) return;
    // kGuesses guesses: from the straight line, then each from the misses of
    // the one before; the last lands in pred + (L + 1)
    uint32_t *bufs[2] = {pred + (L + 1), pred};
    const uint32_t *prev = nullptr;
    for (int q = 0; q < kGuesses; ++q) {
        uint32_t *dst = bufs[(kGuesses - 1 - q) & 1];
        hipLaunchKernelGGL(k_pcss_guess, dim3((L + 255) / 256), dim3(256), 0, s, u, shadow_map, pairs, pt, pbase, s0, L,
                           M0, rate, prev, dst);
        launch_exclusive_scan(dst, dst, L, scan_tmp, s);
        prev = dst;
    }
    const uint32_t *pred2 = prev;
    hipLaunchKernelGGL(k_pcss_band, dim3((L + 3) / 4), dim3(256), 0, s, u, shadow_map, pairs, pt, pbase, s0, L, M0,
                       (const uint32_t *)pred2, masks);
    const uint32_t nblk = (L + kPcssChainBlock - 1) / kPcssChainBlock;
    hipLaunchKernelGGL(k_pcss_walk_blocks, dim3(nblk, (kBandWidth + kChainWaves - 1) / kChainWaves),
                       dim3(64 * kChainWaves), 0, s, masks, pred2, L, fbits, res);
    hipLaunchKernelGGL(k_pcss_chain, dim3(1), dim3(256), 0, s, fbits, res, pred2, s0, L, M0, found, out);
}

void launch_pcss_offsets(const uint32_t *base, const uint32_t *upos, const uint32_t *fpos, uint32_t n_pcf, size_t P,
                         uint32_t *offs, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcss_offsets, dim3((unsigned)((P + 1 + 255) / 256)), dim3(256), 0, s, base, upos, fpos, n_pcf,
                       P, offs);
}

void launch_ssao(const FrameUniforms &u, const SsaoArgs &sa, const float *depth,
                 const float *gbuf, float *ao, hipStream_t s)
{
    if (u.W <= 0 || u.H <= 0) return;
    hipLaunchKernelGGL(k_ssao_tiled, dim3((u.W + kSsaoTW - 1) / kSsaoTW, (u.H + kSsaoTH - 1) / kSsaoTH), dim3(1024),
                       0, s, u, sa, depth, gbuf, ao);
}

void launch_blur_recursive(float *ao, int W, int H, hipStream_t s)
{
    if (W <= 0 || H <= 0) return;
    hipLaunchKernelGGL(k_blur_rec_rows, dim3((H + 63) / 64), dim3(64), 0, s, ao, W, H);
    hipLaunchKernelGGL(k_blur_rec_cols, dim3((W + 63) / 64), dim3(64), 0, s, ao, W, H);
}

void launch_blur_separable(const float *src, float *dst, int W, int H, const float *w, int R, hipStream_t s)
{
    if (W <= 0 || H <= 0) return;
    BlurWeights bw;
    for (int i = 0; i < 9; ++i) bw.w[i] = i <= R ? w[i] : 0.f;
    bw.R = R;
    const dim3 grid((W + kBlurTW - 1) / kBlurTW, (H + kBlurTH - 1) / kBlurTH);
    switch (R) {
#define FGL_BLUR_CASE(r) \
    case r: hipLaunchKernelGGL(k_blur_fused<r>, grid, dim3(256), 0, s, src, dst, W, H, bw); break;
        FGL_BLUR_CASE(1) FGL_BLUR_CASE(2) FGL_BLUR_CASE(3) FGL_BLUR_CASE(4)
        FGL_BLUR_CASE(5) FGL_BLUR_CASE(6) FGL_BLUR_CASE(7) FGL_BLUR_CASE(8)
#undef FGL_BLUR_CASE
    default: break;
    }
}

// recip(w) against the IEEE quotient 1.f / w for every non-NaN float w:
// counts[0] += floats compared, counts[1] += those that differ, counts[2] +=
// those that took the reciprocal-and-Newton path.
__global__ __launch_bounds__(256) void k_check_recip(unsigned long long *__restrict__ counts)
{
    __shared__ unsigned long long part[3][4];
    unsigned long long n = 0, bad = 0, fast = 0;
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < (1ull << 32); i += stride) {
        const float w = __uint_as_float((uint32_t)i);
        if (w != w) continue;
        ++n;
        const float a = fabsf(w);
        fast += (a >= 0x1p-125f && a <= 0x1p125f) ? 1u : 0u;
        bad += __float_as_uint(recip(w)) != __float_as_uint(1.f / w) ? 1u : 0u;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        n += __shfl_xor(n, o, 64);
        bad += __shfl_xor(bad, o, 64);
        fast += __shfl_xor(fast, o, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        part[0][wave] = n;
        part[1][wave] = bad;
        part[2][wave] = fast;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const unsigned long long v = part[threadIdx.x][0] + part[threadIdx.x][1] + part[threadIdx.x][2] +
                                     part[threadIdx.x][3];
        atomicAdd(&counts[threadIdx.x], v);
    }
}

void launch_check_recip(unsigned long long *counts, hipStream_t s)
{
    hipLaunchKernelGGL(k_check_recip, dim3(4096), dim3(256), 0, s, counts);
}

void launch_resolve(const float *fb, int W, int H, int k, uint8_t *img, hipStream_t s)
{
    int ow = W / k, oh = H / k;
    if (ow <= 0 || oh <= 0) return;
    hipLaunchKernelGGL(k_resolve, grid2(ow, oh), dim3(256), 0, s, fb, W, H, k, img);
}

}  // namespace fgl
