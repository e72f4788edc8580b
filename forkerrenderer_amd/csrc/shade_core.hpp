// shade_core.hpp -- per-pixel bodies of the screen-space passes (deferred
// lighting, kernel SSAO, quantise + SSAA resolve), shared by the single-view
// kernels (shade.hip) and the batched multi-view kernels (batch.hip).
#pragma once

#include "fgl_kernels.hpp"

namespace fgl {

// ForkerGL::DrawScreenSpacePixels (src/forkergl.cpp:326-380) at pixel (x,y):
// read the G-buffer (76 B/px) + AO (4 B/px), light, write the frame buffer
// (12 B/px).
__device__ __forceinline__ void lighting_px(const FrameUniforms &u, const float *__restrict__ g,
                                            const float *__restrict__ ao, int ao_is_one,
                                            const float *__restrict__ smap, float *__restrict__ fb,
                                            int x, int y)
{
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

// Kernel SSAO at pixel (x,y): the deterministic replacement for
// Render::DoSSAO's RNG hemisphere samples (src/render.cpp:214-289).  Tangent
// frame from the 4x4 noise vector, n kernel offsets, projection to window
// space exactly as the reference does, depth gather, bias and range check;
// off-screen taps add no occlusion; ao = (1-occ)^3.
__device__ __forceinline__ void ssao_px(const FrameUniforms &u, const SsaoArgs &sa,
                                        const float *__restrict__ depth,
                                        const float *__restrict__ g, float *__restrict__ ao,
                                        int x, int y)
{
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

// Buffer3f::GenerateImage (src/buffer.cpp:113-126) and Render::DoSSAA
// (src/render.cpp:291-343): quantise each internal pixel with a truncating
// 32-bit conversion, sum the k x k bytes, divide as float, truncate.
__device__ __forceinline__ uint32_t quant(float c) { return (uint32_t)(trunc_int(c * 254.99f) & 255); }

// Output pixel (x,y) of the W x H internal frame resolved by k.
__device__ __forceinline__ void resolve_px(const float *__restrict__ fb, int W, int H, int k,
                                           uint8_t *__restrict__ img, int x, int y)
{
    const int ow = W / k;
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

}  // namespace fgl
