// shade.hip -- screen-space passes for gfx950: deferred lighting, kernel SSAO,
// AO blurs and the quantise + SSAA resolve.  All read/write SoA float planes
// in HBM with 64-wide rows per wavefront.
#include "shade_core.hpp"

namespace fgl {

__global__ __launch_bounds__(256) void k_lighting(FrameUniforms u, const float *__restrict__ g,
                                                  const float *__restrict__ ao, int ao_is_one,
                                                  const float *__restrict__ smap, ShadowSamples ss,
                                                  float *__restrict__ fb)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= u.W || y >= u.H) return;
    lighting_px(u, g, ao, ao_is_one, smap, ss, fb, x, y);
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

// The PCSS sample offsets: one wavefront walks the pixels in the lighting
// pass's order (row order); at each pixel lanes evaluate the blocker
// samples at the pixel's offset in parallel and lane order sums them as
// FindAverageBlockDepth does (shadow.cpp:65-91), giving the pixel's
// consumption: n_blocker pairs, plus n_pcf when a blocker is found
// (PCSS, shadow.cpp:93-106).  Only light 0 is shadowed; with shadows off or
// no light nothing is consumed.
__global__ __launch_bounds__(64) void k_pcss_walk(FrameUniforms u, const float *__restrict__ g,
                                                  const float *__restrict__ smap,
                                                  const float *__restrict__ pairs,
                                                  uint32_t *__restrict__ offs, uint32_t *__restrict__ total)
{
    const int lane = threadIdx.x;
    const size_t P = (size_t)u.W * u.H;
    const DevShadowCfg &c = u.shadow;
    const bool consume = u.shadow_on && u.n_lights > 0;
    uint32_t o = 0;
    for (size_t i = 0; i < P; ++i) {
        if (lane == 0) offs[i] = o;
        if (!consume) continue;
        const Vec3 N = vec3(g[i], g[P + i], g[2 * P + i]);
        const Vec3 Pw = vec3(g[3 * P + i], g[4 * P + i], g[5 * P + i]);
        const Vec3 lndc = vec3(g[6 * P + i], g[7 * P + i], g[8 * P + i]);
        const Vec3 L = light_dir(u.lights[0], Pw);
        const Vec3 sc = shadow_coord(lndc);
        const float bias = shadow_bias(N, L);
        float sv = 0.f;
        int blk = 0;
        if (lane < c.n_blocker) {
            const float *d = pairs + 2 * ((size_t)o + lane);
            const float su = sc.x + (float)((double)d[0] * c.blocker_filter);
            const float sw = sc.y + (float)((double)d[1] * c.blocker_filter);
            sv = sample_shadow(smap, u.SW, u.SH, su, sw);
            blk = sc.z > sv + bias;
        }
        float blocker = 0.f, nb = 0.f;
        for (int j = 0; j < c.n_blocker; ++j) {
            const float sj = __shfl(sv, j);
            if (__shfl(blk, j)) {
                blocker += sj;
                nb += 1.f;
            }
        }
        const float dB = (nb < 1.f) ? 0.f : blocker / nb;
        o += (uint32_t)c.n_blocker + (((double)dB < 0.001) ? 0u : (uint32_t)c.n_pcf);
    }
    if (lane == 0) *total = o;
}

__global__ __launch_bounds__(256) void k_ssao(FrameUniforms u, SsaoArgs sa,
                                              const float *__restrict__ depth,
                                              const float *__restrict__ g,
                                              float *__restrict__ ao)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= u.W || y >= u.H) return;
    ssao_px(u, sa, depth, g, ao, x, y);
}

// Buffer1f::TwoPassGaussianBlurDenoised (src/buffer.cpp:59-98) is in place and
// order dependent: the taps left of (or below) a pixel read values already
// blurred.  One thread owns a whole row (then a whole column) and walks it in
// the reference's order, so every read sees exactly what the reference sees.
__constant__ float c_ref_blur_w[5] = {(float)0.227027, (float)0.1945946, (float)0.1216216,
                                      (float)0.054054, (float)0.016216};

__global__ __launch_bounds__(64) void k_blur_rec_rows(float *__restrict__ a, int W, int H)
{
    const int h = blockIdx.x * 64 + threadIdx.x;
    if (h >= H) return;
    float *row = a + (size_t)h * W;
    for (int x = 0; x < W; ++x) {
        float r = row[x] * c_ref_blur_w[0];
#pragma unroll
        for (int i = 1; i < 5; ++i) {
            r += row[clampi(x + i, 0, W - 1)] * c_ref_blur_w[i];
            r += row[clampi(x - i, 0, W - 1)] * c_ref_blur_w[i];
        }
        row[x] = r;
    }
}

__global__ __launch_bounds__(64) void k_blur_rec_cols(float *__restrict__ a, int W, int H)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    if (x >= W) return;
    for (int h = 0; h < H; ++h) {
        float r = a[(size_t)x + (size_t)h * W] * c_ref_blur_w[0];
#pragma unroll
        for (int i = 1; i < 5; ++i) {
            r += a[(size_t)x + (size_t)clampi(h + i, 0, H - 1) * W] * c_ref_blur_w[i];
            r += a[(size_t)x + (size_t)clampi(h - i, 0, H - 1) * W] * c_ref_blur_w[i];
        }
        a[(size_t)x + (size_t)h * W] = r;
    }
}

struct BlurWeights {
    float w[9];
    int R;
};

// Separable Gaussian: same tap order as the reference loop, reading a copy.
template <bool HORIZ>
__global__ __launch_bounds__(256) void k_blur_sep(const float *__restrict__ src,
                                                  float *__restrict__ dst, int W, int H,
                                                  BlurWeights bw)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    float r = src[(size_t)x + (size_t)y * W] * bw.w[0];
    for (int i = 1; i <= bw.R; ++i) {
        if (HORIZ) {
            r += src[(size_t)clampi(x + i, 0, W - 1) + (size_t)y * W] * bw.w[i];
            r += src[(size_t)clampi(x - i, 0, W - 1) + (size_t)y * W] * bw.w[i];
        } else {
            r += src[(size_t)x + (size_t)clampi(y + i, 0, H - 1) * W] * bw.w[i];
            r += src[(size_t)x + (size_t)clampi(y - i, 0, H - 1) * W] * bw.w[i];
        }
    }
    dst[(size_t)x + (size_t)y * W] = r;
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

void launch_lighting(const FrameUniforms &u, const float *gbuf, const float *ao, int ao_is_one,
                     const float *shadow_map, const ShadowSamples &ss, float *fb, hipStream_t s)
{
    if (u.W <= 0 || u.H <= 0) return;
    hipLaunchKernelGGL(k_lighting, grid2(u.W, u.H), dim3(256), 0, s, u, gbuf, ao, ao_is_one,
                       shadow_map, ss, fb);
}

void launch_ssao_reference(const FrameUniforms &u, const float *samples, const float *pow3,
                           const float *depth, const float *gbuf, float *ao, hipStream_t s)
{
    if (u.W <= 0 || u.H <= 0) return;
    hipLaunchKernelGGL(k_ssao_ref, grid2(u.W, u.H), dim3(256), 0, s, u, samples, pow3, depth, gbuf, ao);
}

void launch_pcss_walk(const FrameUniforms &u, const float *gbuf, const float *shadow_map, const float *pairs,
                      uint32_t *offs, uint32_t *total, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcss_walk, dim3(1), dim3(64), 0, s, u, gbuf, shadow_map, pairs, offs, total);
}

void launch_ssao(const FrameUniforms &u, const SsaoArgs &sa, const float *depth,
                 const float *gbuf, float *ao, hipStream_t s)
{
    if (u.W <= 0 || u.H <= 0) return;
    hipLaunchKernelGGL(k_ssao, grid2(u.W, u.H), dim3(256), 0, s, u, sa, depth, gbuf, ao);
}

void launch_blur_recursive(float *ao, int W, int H, hipStream_t s)
{
    if (W <= 0 || H <= 0) return;
    hipLaunchKernelGGL(k_blur_rec_rows, dim3((H + 63) / 64), dim3(64), 0, s, ao, W, H);
    hipLaunchKernelGGL(k_blur_rec_cols, dim3((W + 63) / 64), dim3(64), 0, s, ao, W, H);
}

void launch_blur_separable(float *ao, float *tmp, int W, int H, const float *w, int R,
                           hipStream_t s)
{
    if (W <= 0 || H <= 0) return;
    BlurWeights bw;
    for (int i = 0; i < 9; ++i) bw.w[i] = i <= R ? w[i] : 0.f;
    bw.R = R;
    hipLaunchKernelGGL(k_blur_sep<true>, grid2(W, H), dim3(256), 0, s, ao, tmp, W, H, bw);
    hipLaunchKernelGGL(k_blur_sep<false>, grid2(W, H), dim3(256), 0, s, tmp, ao, W, H, bw);
}

void launch_resolve(const float *fb, int W, int H, int k, uint8_t *img, hipStream_t s)
{
    int ow = W / k, oh = H / k;
    if (ow <= 0 || oh <= 0) return;
    hipLaunchKernelGGL(k_resolve, grid2(ow, oh), dim3(256), 0, s, fb, W, H, k, img);
}

}  // namespace fgl
