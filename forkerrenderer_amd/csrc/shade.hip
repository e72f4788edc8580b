// shade.hip -- screen-space passes for gfx950: deferred lighting, kernel SSAO,
// AO blurs and the quantise + SSAA resolve.  All read/write SoA float planes
// in HBM with 64-wide rows per wavefront.
#include "fgl_kernels.hpp"

namespace fgl {

// ForkerGL::DrawScreenSpacePixels (src/forkergl.cpp:326-380): read the
// G-buffer (76 B/px) + AO (4 B/px), light, write the frame buffer (12 B/px).
__global__ __launch_bounds__(256) void k_lighting(FrameUniforms u, const float *__restrict__ g,
                                                  const float *__restrict__ ao, int ao_is_one,
                                                  const float *__restrict__ smap,
                                                  float *__restrict__ fb)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= u.W || y >= u.H) return;
    const size_t P = (size_t)u.W * u.H;
    const size_t i = (size_t)x + (size_t)y * u.W;
    Vec3 N = vec3(g[i], g[P + i], g[2 * P + i]);
    Vec3 Pw = vec3(g[3 * P + i], g[4 * P + i], g[5 * P + i]);
    Vec3 L = vec3(g[6 * P + i], g[7 * P + i], g[8 * P + i]);
    Vec3 al = vec3(g[9 * P + i], g[10 * P + i], g[11 * P + i]);
    Vec3 em = vec3(g[12 * P + i], g[13 * P + i], g[14 * P + i]);
    Vec3 pa = vec3(g[15 * P + i], g[16 * P + i], g[17 * P + i]);
    float type = g[18 * P + i];
    // AmbientOcclusionGBuffer starts at 1 (forkergl.cpp:80); x * 1.f == x.
    if (!ao_is_one) pa.x *= ao[i];
    Vec3 c = shade_point(u, smap, !(type < 0.5f), true, Pw, N, L, al, em, pa);
    fb[i] = c.x;
    fb[P + i] = c.y;
    fb[2 * P + i] = c.z;
}

// Kernel SSAO: the deterministic replacement for Render::DoSSAO's RNG
// hemisphere samples (src/render.cpp:214-289).  Tangent frame from the 4x4
// noise vector, n kernel offsets, projection to window space exactly as the
// reference does, depth gather, bias and range check; off-screen taps add no
// occlusion; ao = (1-occ)^3.
__global__ __launch_bounds__(256) void k_ssao(FrameUniforms u, SsaoArgs sa,
                                              const float *__restrict__ depth,
                                              const float *__restrict__ g,
                                              float *__restrict__ ao)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= u.W || y >= u.H) return;
    const size_t P = (size_t)u.W * u.H;
    const size_t i = (size_t)x + (size_t)y * u.W;
    Vec3 N = vec3(g[i], g[P + i], g[2 * P + i]);
    Vec3 Pw = vec3(g[3 * P + i], g[4 * P + i], g[5 * P + i]);
    float fd = depth[i];
    int ni = (x & 3) + 4 * (y & 3);
    Vec3 r = vec3(sa.noise[3 * ni], sa.noise[3 * ni + 1], sa.noise[3 * ni + 2]);
    Vec3 T = norm3(sub3(r, scale3(N, dot3(r, N))));
    Vec3 B = cross3(N, T);
    const float scale = 1.f / (float)sa.n_kernel;
    float occ = 0.f;
    for (int s = 0; s < sa.n_kernel; ++s) {
        Vec3 k = vec3(sa.kernel[3 * s], sa.kernel[3 * s + 1], sa.kernel[3 * s + 2]);
        Vec3 d = add3(add3(scale3(T, k.x), scale3(B, k.y)), scale3(N, k.z));
        Vec3 S = add3(Pw, scale3(d, sa.radius));
        Vec4 cs = mv4(u.vp, vec4(S.x, S.y, S.z, 1.f));
        Vec4 nd = divs4(cs, cs.w);
        Vec4 ss = mv4(u.viewport, nd);
        int sx = trunc_int(ss.x), sy = trunc_int(ss.y);
        if (sx < 0 || sy < 0 || sx >= u.W || sy >= u.H) continue;
        float cached = depth[(size_t)sx + (size_t)sy * u.W];
        if (ss.z >= cached + sa.bias) {
            float rc = fabsf(fd - cached) < sa.range ? 1.f : 0.f;
            occ += scale * rc;
        }
    }
    float a = 1.f - occ;
    ao[i] = (a * a) * a;
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

// Buffer3f::GenerateImage (src/buffer.cpp:113-126) and Render::DoSSAA
// (src/render.cpp:291-343): quantise each internal pixel with a truncating
// 32-bit conversion, sum the k x k bytes, divide as float, truncate.
__device__ __forceinline__ uint32_t quant(float c) { return (uint32_t)(trunc_int(c * 254.99f) & 255); }

__global__ __launch_bounds__(256) void k_resolve(const float *__restrict__ fb, int W, int H, int k,
                                                 uint8_t *__restrict__ img)
{
    const int ow = W / k, oh = H / k;
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= ow || y >= oh) return;
    const size_t P = (size_t)W * H;
    uint32_t s[3] = {0, 0, 0};
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) {
            size_t p = (size_t)(x * k + i) + (size_t)(y * k + j) * W;
            s[0] += quant(fb[2 * P + p]);  // B
            s[1] += quant(fb[P + p]);      // G
            s[2] += quant(fb[p]);          // R
        }
    const float k2 = (float)(k * k);
    uint8_t *o = img + ((size_t)x + (size_t)y * ow) * 3;
    o[0] = (uint8_t)(int)((float)s[0] / k2);
    o[1] = (uint8_t)(int)((float)s[1] / k2);
    o[2] = (uint8_t)(int)((float)s[2] / k2);
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
                     const float *shadow_map, float *fb, hipStream_t s)
{
    if (u.W <= 0 || u.H <= 0) return;
    hipLaunchKernelGGL(k_lighting, grid2(u.W, u.H), dim3(256), 0, s, u, gbuf, ao, ao_is_one,
                       shadow_map, fb);
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
