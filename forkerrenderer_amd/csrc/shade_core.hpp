// shade_core.hpp -- per-pixel bodies of the screen-space passes (deferred
// lighting, kernel SSAO, quantise + SSAA resolve), shared by the single-view
// kernels (shade.hip) and the batched multi-view kernels (batch.hip).
#pragma once

#include <type_traits>

#include "fgl_kernels.hpp"

namespace fgl {

// ForkerGL::DrawScreenSpacePixels (src/forkergl.cpp:326-380) at pixel (x,y):
// read the G-buffer (76 B/px) + AO (4 B/px, or the network AO still on its
// coarse grid: 1 B of coverage and four cached grid values), light, write
// the frame buffer (12 B/px).
template <bool STREAM = false>
__device__ __forceinline__ void lighting_px(const FrameUniforms &u, const float *__restrict__ g,
                                            const float *__restrict__ ao, int ao_is_one,
                                            const float *__restrict__ smap, const ShadowSamples &ss,
                                            float *__restrict__ fb, int x, int y, const NetAo &net = NetAo{})
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
    if (net.lao) pa.x *= net.fmask[i] ? net_ao_interp(net.lao, net.lW, net.lH, net.f, x, y) : 1.f;
    else if (!ao_is_one) pa.x *= ao[i];
    const float *rdisk = nullptr;
    if (STREAM && ss.pairs) rdisk = ss.pairs + 2 * (ss.offs ? (size_t)ss.offs[i] : i * (size_t)ss.stride);
    Vec3 c = shade_point<STREAM>(u, smap, !(type < 0.5f), true, Pw, N, L, al, em, pa, rdisk);
    fb[i] = c.x;
    fb[P + i] = c.y;
    fb[2 * P + i] = c.z;
}

// Kernel SSAO at pixel (x,y): the deterministic replacement for
// Render::DoSSAO's RNG hemisphere samples (src/render.cpp:214-289).  Tangent
// frame from the 4x4 noise vector; the sample P + radius * (T kx + B ky +
// N kz) is projected with the view-projection and the viewport's diagonal
// and offset, linearised per pixel: with c = VP * (P, 1) and cT|cB|cN =
// VP * (dir, 0), each column is taken to window-space numerators
// (vp0 clip_x + vp3 clip_w, vp5 clip_y + vp7 clip_w, vp10 clip_z +
// vp11 clip_w, clip_w; window = numerator / clip_w), the direction columns
// times the radius, so a tap is e = ((ec + kx eT) + ky eB) + kz eN by fmaf
// and window x, y, z = e * (1 / e_w).  Depth gather, bias and range check;
// off-screen taps add no occlusion; ao = (1 - count / n)^3.  `tap(sx, sy)`
// reads the depth at an on-screen tap (from the depth plane, or from a
// workgroup's copy of the part around it: the same floats).
template <class Tap>
__device__ __forceinline__ void ssao_px_with(const FrameUniforms &u, const SsaoArgs &sa,
                                             const float *__restrict__ depth,
                                             const float *__restrict__ g, float *__restrict__ ao,
                                             int x, int y, const Tap &tap)
{
    const size_t P = (size_t)u.W * u.H;
    const size_t i = (size_t)x + (size_t)y * u.W;
    Vec3 N = vec3(g[i], g[P + i], g[2 * P + i]);
    Vec3 Pw = vec3(g[3 * P + i], g[4 * P + i], g[5 * P + i]);
    const float fd = depth[i];
    const int ni = (x & 3) + 4 * (y & 3);
    Vec3 r = vec3(sa.noise[3 * ni], sa.noise[3 * ni + 1], sa.noise[3 * ni + 2]);
    Vec3 T = norm3(sub3(r, scale3(N, dot3(r, N))));
    Vec3 B = cross3(N, T);
    const Vec4 c0 = mv4(u.vp, vec4(Pw.x, Pw.y, Pw.z, 1.f));
    float cT[4], cB[4], cN[4];
    const float cc[4] = {c0.x, c0.y, c0.z, c0.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const Vec3 row = vec3(u.vp[4 * j], u.vp[4 * j + 1], u.vp[4 * j + 2]);
        cT[j] = dot3(row, T);
        cB[j] = dot3(row, B);
        cN[j] = dot3(row, N);
    }
    const float *vp = u.viewport;
    const float radius = sa.radius;
    float ec[4], eT[4], eB[4], eN[4];
    ec[0] = fmaf(vp[0], cc[0], vp[3] * cc[3]);
    ec[1] = fmaf(vp[5], cc[1], vp[7] * cc[3]);
    ec[2] = fmaf(vp[10], cc[2], vp[11] * cc[3]);
    ec[3] = cc[3];
    const float *cols[3] = {cT, cB, cN};
    float *outs[3] = {eT, eB, eN};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const float *c = cols[q];
        float *e = outs[q];
        e[0] = radius * fmaf(vp[0], c[0], vp[3] * c[3]);
        e[1] = radius * fmaf(vp[5], c[1], vp[7] * c[3]);
        e[2] = radius * fmaf(vp[10], c[2], vp[11] * c[3]);
        e[3] = radius * c[3];
    }
    // The tap is on screen iff trunc_int(sxf) lies in [0, W): for sxf in
    // (-1, W) the truncation is in range, and outside it (NaN included) it
    // is not (trunc_int maps what does not fit an int to INT_MIN).  Testing
    // the float bounds is that test with fewer operations; W and H are
    // exact in float.
    const float Wf = (float)u.W, Hf = (float)u.H;
    // Every tap's e_w differs from ec[3] by at most kmax (|eT[3]| + |eB[3]|
    // + |eN[3]|) (widened by 0.1 % for that sum's roundings) plus three
    // roundings relative to ec[3], which the factor 2 between 2^-124 and
    // recip's 2^-125 (and 2^124, 2^125) covers.  When that interval lies in
    // [2^-124, 2^124] for every pixel of the wavefront, recip's range test
    // would pass for every tap and is left out (NaN or inf fails the test).
    const float dw = sa.kmax * (fabsf(eT[3]) + fabsf(eB[3]) + fabsf(eN[3])) * 1.001f;
    const bool w_safe = ec[3] - dw >= 0x1p-124f && ec[3] + dw <= 0x1p124f;
    auto taps = [&](auto in_range) {
        int count = 0;
        for (int s = 0; s < sa.n_kernel; ++s) {
            const float kx = sa.kernel[3 * s], ky = sa.kernel[3 * s + 1], kz = sa.kernel[3 * s + 2];
            float e[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) e[j] = fmaf(kz, eN[j], fmaf(ky, eB[j], fmaf(kx, eT[j], ec[j])));
            const float inv = decltype(in_range)::value ? recip_in_range(e[3]) : recip(e[3]);
            const float sxf = e[0] * inv;
            const float syf = e[1] * inv;
            if (!(sxf > -1.f && sxf < Wf && syf > -1.f && syf < Hf)) continue;
            const float sz = e[2] * inv;
            const float cached = tap((int)sxf, (int)syf);
            if (sz >= cached + sa.bias && fabsf(fd - cached) < sa.range) ++count;
        }
        return count;
    };
    const int count = __all(w_safe) ? taps(std::true_type{}) : taps(std::false_type{});
    const float a = 1.f - (float)count * (1.f / (float)sa.n_kernel);
    ao[i] = (a * a) * a;
}

// Render::DoSSAO (src/render.cpp:214-289) itself at pixel (x,y), with the
// pixel's 32 hemisphere samples taken from the random stream (samples: 32
// unit-sphere vectors per pixel, pixels in row order).  Each sample is
// flipped to the normal's side (RandomVectorInHemisphere), scaled by
// Lerp(0.1f, 1.0f, |d|^2) = (1 - 0.1f) * 1.0f + 0.1f * |d|^2 (Lerp(t, a, b),
// src/utility.h:26-29), offset by 0.075 from the position, projected with
// the view-projection and viewport matrices (v / w is v * (1/w)) and
// truncated; the depth read is data[x + y * W] in 32-bit int arithmetic, so
// an off-screen sample reads another pixel, or lies outside the buffer
// (undefined in the reference; here no occlusion).  Each occluding sample
// adds 1/32 (exact); ao = powf(1 - occlusion, 3) by table.  The samples are
// the pixel's own (384 bytes a pixel), so they are read 8 at a time (16-byte
// loads), and the 8 depths they project to are loaded together.
__device__ __forceinline__ void ssao_ref_px(const FrameUniforms &u, const float *__restrict__ samples,
                                            const float *__restrict__ pow3, const float *__restrict__ depth,
                                            const float *__restrict__ g, float *__restrict__ ao, int x, int y)
{
    const size_t P = (size_t)u.W * u.H;
    const size_t i = (size_t)x + (size_t)y * u.W;
    const Vec3 N = vec3(g[i], g[P + i], g[2 * P + i]);
    const Vec3 Pw = vec3(g[3 * P + i], g[4 * P + i], g[5 * P + i]);
    const float fd = depth[i];
    const float4 *smp = (const float4 *)(samples + i * 96);   // 16-byte aligned: 384 bytes a pixel
    int count = 0;
    for (int s0 = 0; s0 < 32; s0 += 8) {
        float v[24];
#pragma unroll
        for (int e = 0; e < 6; ++e) {
            const float4 q = smp[(s0 * 3) / 4 + e];
            v[4 * e] = q.x;
            v[4 * e + 1] = q.y;
            v[4 * e + 2] = q.z;
            v[4 * e + 3] = q.w;
        }
        size_t at[8];
        float sz[8];
        bool on[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            Vec3 d = vec3(v[3 * k], v[3 * k + 1], v[3 * k + 2]);
            if (!(dot3(d, N) > 0.f)) d = vec3(-d.x, -d.y, -d.z);
            float sc = len3(d);
            sc = (1 - 0.1f) * 1.0f + 0.1f * (sc * sc);
            d = scale3(d, sc);
            const Vec3 ws = add3(Pw, scale3(d, 0.075f));
            const Vec4 cs = mv4(u.vp, vec4(ws.x, ws.y, ws.z, 1.f));
            const Vec4 ndc = divs4(cs, cs.w);
            const Vec4 ss = mv4(u.viewport, ndc);
            const int sx = trunc_int(ss.x), sy = trunc_int(ss.y);
            const int idx = (int)((uint32_t)sx + (uint32_t)sy * (uint32_t)u.W);
            on[k] = !(idx < 0 || (size_t)idx >= P);
            at[k] = on[k] ? (size_t)idx : 0;
            sz[k] = ss.z;
        }
        float cached[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) cached[k] = depth[at[k]];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (on[k] && sz[k] >= cached[k] + 0.0005f && fabsf(fd - cached[k]) < 0.01f) ++count;
    }
    ao[i] = pow3[count];
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
