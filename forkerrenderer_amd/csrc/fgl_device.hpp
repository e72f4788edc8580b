// fgl_device.hpp -- device-resident scene data and the per-pixel programs of
// the reference's shaders, written for gfx950.
#pragma once

#include "fgl_math.hpp"

namespace fgl {

constexpr int kMaxLights = 4;
constexpr int kMaxDisk = 64;

// Texture in HBM: one packed texel per uint32 (r | g<<8 | b<<16), row y at y*w.
// TGAImage::Get's byte order (B,G,R; grey sets only b) is resolved at upload
// (src/tgaimage.h:149-181).
struct DevTexture {
    const uint32_t *texels;   // r | g << 8 | b << 16 per texel
    const uint8_t *zb;        // the b byte alone (what SampleFloat reads), 1 B per texel
    int w, h, wrap, filter;
};

struct DevMaterial {
    float ka[3], kd[3], ks[3], ke[3];
    int map_kd, map_ks, map_ke, map_bump;
    float pbr_ke[3], albedo[3], roughness, metalness;
    int map_base, map_rough, map_metal, map_ao, map_pbr_ke;
};

// One draw of a pass: a model's SoA attribute arrays, face index streams and
// uniforms.  vert_base/tri_base place it in the pass-wide vertex/triangle order.
struct DevDraw {
    const float *pos, *uv, *nrm, *tan;
    const int *vi, *ti, *ni, *face_mat;
    // per face corner (3 per face), the attributes the main vertex programs
    // read, interleaved as 3 float4: (pos.xyz, uv.x), (uv.y, nrm.xyz),
    // (tan.xyz, 0) -- one 48-byte record instead of four arrays behind three
    // index streams
    const float4 *corner;
    int n_pos, n_faces, vert_base, tri_base;
    int has_tangents, support_pbr;
    int view;      // batched views: index of the draw's view (0 otherwise)
    int pad0;
    float model[16];
    float normal[9];
    float lm[16];  // light_space * model (DepthShader, src/shaders/depthshader.h:23-25)
};

struct DevLight {
    int type;
    float vec[3];
    float color[3];
};

struct DevShadowCfg {
    int mode, grid, n_blocker, n_pcf;
    float pcf_filter, area_light;
    double blocker_filter, pcss_filter;
    float blocker_disk[2 * kMaxDisk];
    float pcf_disk[2 * kMaxDisk];
    // blocker_disk[i] * blocker_filter in double, narrowed to float (the
    // reference's per-sample offset; the same for every pixel)
    float blocker_off[2 * kMaxDisk];
};

struct FrameUniforms {
    float view[16], proj[16], light_space[16], viewport[16], shadow_viewport[16];
    float vp[16];  // proj * view (SetViewProjectionMatrix, src/render.cpp:115)
    float eye[3];
    int shadow_on;
    int n_lights, W, H, SW, SH;
    DevLight lights[kMaxLights];
    DevShadowCfg shadow;
};

// ---------------- textures (src/texture.h:41-145) ----------------

__device__ __forceinline__ Vec3 texel_rgb(const DevTexture &t, int x, int y)
{
    if (x < 0 || y < 0 || x >= t.w || y >= t.h) return vec3(0.f, 0.f, 0.f);
    uint32_t p = t.texels[(size_t)x + (size_t)y * t.w];
    return vec3((float)(p & 255u), (float)((p >> 8) & 255u), (float)((p >> 16) & 255u));
}

__device__ __forceinline__ Vec2 wrap_coord(const DevTexture &t, Vec2 c)
{
    if (t.wrap == 1) return vec2(c.x - floorf(c.x), c.y - floorf(c.y));
    if (t.wrap == 2) {
        int xi = trunc_int(floorf(c.x)), yi = trunc_int(floorf(c.y));
        float rx = c.x - (float)xi, ry = c.y - (float)yi;
        return vec2((xi % 2 == 0) ? rx : 1.f - rx, (yi % 2 == 0) ? ry : 1.f - ry);
    }
    if (t.wrap == 3) return vec2(clampf(c.x, 0.f, 1.f), clampf(c.y, 0.f, 1.f));
    return c;
}

__device__ __forceinline__ Vec3 filter_color(const DevTexture &t, Vec2 c)
{
    float w = (float)(t.w - 0.001), h = (float)(t.h - 0.001);
    if (t.filter == 1) {
        float px = c.x * w, py = c.y * h;
        float tlx = floorf(px - 0.5f), tly = floorf(py - 0.5f);
        float tx = px - (tlx + 0.5f);
        float ty = py - (tly + 0.5f);
        int s0x = trunc_int(tlx), s0y = trunc_int(tly);
        int s1x = trunc_int(tlx + 1.f), s2y = trunc_int(tly + 1.f);
        int s1y = s0y, s2x = s0x, s3x = s1x, s3y = s2y;
        if (t.wrap != 0) {
            s0x = clampi(s0x, 0, t.w - 1); s0y = clampi(s0y, 0, t.h - 1);
            s1x = clampi(s1x, 0, t.w - 1); s1y = clampi(s1y, 0, t.h - 1);
            s2x = clampi(s2x, 0, t.w - 1); s2y = clampi(s2y, 0, t.h - 1);
            s3x = clampi(s3x, 0, t.w - 1); s3y = clampi(s3y, 0, t.h - 1);
        }
        Vec3 c0 = texel_rgb(t, s0x, s0y), c1 = texel_rgb(t, s1x, s1y);
        Vec3 c2 = texel_rgb(t, s2x, s2y), c3 = texel_rgb(t, s3x, s3y);
        Vec3 cx1 = lerp3(tx, c0, c1), cx2 = lerp3(tx, c2, c3);
        return lerp3(ty, cx1, cx2);
    }
    return texel_rgb(t, trunc_int(floorf(c.x * w)), trunc_int(floorf(c.y * h)));
}

// filter_color's third channel alone, from the 1-byte plane: the same
// texel coordinates and the same per-channel arithmetic (lerp3 is
// component-wise), a quarter of the bytes.
__device__ __forceinline__ float texel_z(const DevTexture &t, int x, int y)
{
    if (x < 0 || y < 0 || x >= t.w || y >= t.h) return 0.f;
    return (float)t.zb[(size_t)x + (size_t)y * t.w];
}

__device__ __forceinline__ float lerp1(float t, float a, float b) { return a * (1 - t) + b * t; }

__device__ __forceinline__ float filter_z(const DevTexture &t, Vec2 c)
{
    float w = (float)(t.w - 0.001), h = (float)(t.h - 0.001);
    if (t.filter == 1) {
        float px = c.x * w, py = c.y * h;
        float tlx = floorf(px - 0.5f), tly = floorf(py - 0.5f);
        float tx = px - (tlx + 0.5f);
        float ty = py - (tly + 0.5f);
        int s0x = trunc_int(tlx), s0y = trunc_int(tly);
        int s1x = trunc_int(tlx + 1.f), s2y = trunc_int(tly + 1.f);
        int s1y = s0y, s2x = s0x, s3x = s1x, s3y = s2y;
        if (t.wrap != 0) {
            s0x = clampi(s0x, 0, t.w - 1); s0y = clampi(s0y, 0, t.h - 1);
            s1x = clampi(s1x, 0, t.w - 1); s1y = clampi(s1y, 0, t.h - 1);
            s2x = clampi(s2x, 0, t.w - 1); s2y = clampi(s2y, 0, t.h - 1);
            s3x = clampi(s3x, 0, t.w - 1); s3y = clampi(s3y, 0, t.h - 1);
        }
        float c0 = texel_z(t, s0x, s0y), c1 = texel_z(t, s1x, s1y);
        float c2 = texel_z(t, s2x, s2y), c3 = texel_z(t, s3x, s3y);
        float cx1 = lerp1(tx, c0, c1), cx2 = lerp1(tx, c2, c3);
        return lerp1(ty, cx1, cx2);
    }
    return texel_z(t, trunc_int(floorf(c.x * w)), trunc_int(floorf(c.y * h)));
}

// Texture::Sample / SampleFloat (src/texture.h:41-51)
__device__ __forceinline__ Vec3 tex_sample(const DevTexture *texs, int id, Vec2 uv)
{
    const DevTexture &t = texs[id];
    return divs3(filter_color(t, wrap_coord(t, uv)), 255.f);
}
__device__ __forceinline__ float tex_samplef(const DevTexture *texs, int id, Vec2 uv)
{
    const DevTexture &t = texs[id];
    return filter_z(t, wrap_coord(t, uv)) / 255.f;
}

// ---------------- vertex programs ----------------

// Varyings of one vertex after ProcessVertex with perspective-correct
// interpolation enabled (src/shaders/gshader.h:41-92).
struct Varying {
    Vec3 ws, nws, tws, lndc;
    Vec2 uv;
    float oow;
};

// BlinnPhong/PBR/G ProcessVertex (phongshader.h:35-85, pbrshader.h:35-85,
// gshader.h:41-92) of corner k of face f, with the corner's attributes as
// Mesh::Vert / TexCoord / Normal / Tangent fetch them (mesh.cpp:34-56).
__device__ __forceinline__ Varying main_vertex(const DevDraw &d, const FrameUniforms &u, int f, int k)
{
    Varying o;
    const float4 *r = d.corner + (size_t)(3 * f + k) * 3;
    const float4 r0 = r[0], r1 = r[1], r2 = r[2];
    Vec4 p = vec4(r0.x, r0.y, r0.z, 1.f);
    Vec4 ws = mv4(d.model, p);
    Vec4 vs = mv4(u.view, ws);
    Vec4 cs = mv4(u.proj, vs);
    Vec2 uv = vec2(r0.w, r1.x);
    Vec3 n = norm3(vec3(r1.y, r1.z, r1.w));
    Vec3 nws = mv3(d.normal, n);
    Vec3 tws = vec3(0.f, 0.f, 0.f);
    if (d.has_tangents) {
        Vec3 t = norm3(vec3(r2.x, r2.y, r2.z));
        tws = mv3(d.normal, t);
    }
    Vec4 l = vec4(0.f, 0.f, 0.f, 0.f);
    if (u.shadow_on) {
        l = mv4(u.light_space, ws);
        l = divs4(l, l.w);
    }
    float oow = recip(cs.w);
    ws = scale4(ws, oow);
    uv.x *= oow;
    uv.y *= oow;
    nws = scale3(nws, oow);
    if (d.has_tangents) tws = scale3(tws, oow);
    if (u.shadow_on) l = scale4(l, oow);
    o.ws = vec3(ws.x, ws.y, ws.z);
    o.uv = uv;
    o.nws = nws;
    o.tws = tws;
    o.lndc = vec3(l.x, l.y, l.z);
    o.oow = oow;
    return o;
}

// Clip-space -> NDC of the main shaders: positionCS / positionCS.w.
__device__ __forceinline__ Vec4 main_ndc(const DevDraw &d, const FrameUniforms &u, int vi)
{
    Vec4 p = vec4(d.pos[3 * vi], d.pos[3 * vi + 1], d.pos[3 * vi + 2], 1.f);
    Vec4 cs = mv4(u.proj, mv4(u.view, mv4(d.model, p)));
    return divs4(cs, cs.w);
}

// DepthShader::ProcessVertex (src/shaders/depthshader.h:21-28)
__device__ __forceinline__ Vec4 depth_ndc(const DevDraw &d, int vi)
{
    Vec4 p = vec4(d.pos[3 * vi], d.pos[3 * vi + 1], d.pos[3 * vi + 2], 1.f);
    Vec4 cs = mv4(d.lm, p);
    return divs4(cs, cs.w);
}

// ---------------- fragment geometry (gshader.h:95-141) ----------------

struct FragGeom {
    Vec3 pos, normal, lndc;
    Vec2 uv;
};

__device__ __forceinline__ FragGeom frag_geometry(const DevDraw &d, const FrameUniforms &u,
                                                  const DevTexture *texs, const Varying &v0,
                                                  const Varying &v1, const Varying &v2, Vec3 b,
                                                  int normal_map)
{
    FragGeom g;
    Vec3 pws = interp3(v0.ws, v1.ws, v2.ws, b);
    Vec2 uv = interp2(v0.uv, v1.uv, v2.uv, b);
    Vec3 nws = interp3(v0.nws, v1.nws, v2.nws, b);
    float w = recip(dot3(vec3(v0.oow, v1.oow, v2.oow), b));
    pws = scale3(pws, w);
    uv.x *= w;
    uv.y *= w;
    nws = scale3(nws, w);
    Vec3 N = norm3(nws);
    Vec3 normal = N;
    if (d.has_tangents && normal_map >= 0) {
        Vec3 tws = scale3(interp3(v0.tws, v1.tws, v2.tws, b), w);
        Vec3 T = norm3(add3(tws, vec3(0.001f, 0.001f, 0.001f)));
        T = norm3(sub3(T, scale3(N, dot3(T, N))));
        Vec3 B = norm3(cross3(N, T));
        Vec3 s = tex_sample(texs, normal_map, uv);
        s = norm3(sub3(scale3(s, 2.f), vec3(1.f, 1.f, 1.f)));
        normal = norm3(vec3(dot3(vec3(T.x, B.x, N.x), s), dot3(vec3(T.y, B.y, N.y), s),
                            dot3(vec3(T.z, B.z, N.z), s)));
    }
    g.pos = pws;
    g.normal = normal;
    g.uv = uv;
    g.lndc = vec3(0.f, 0.f, 0.f);
    if (u.shadow_on) g.lndc = scale3(interp3(v0.lndc, v1.lndc, v2.lndc, b), w);
    return g;
}

// ---------------- shadows (src/shaders/shadow.cpp:23-131) ----------------

__device__ __forceinline__ float sample_shadow(const float *map, int SW, int SH, float u, float v)
{
    // A NaN coordinate is treated as out of the map (the reference would
    // index with INT_MIN).
    if (!(u >= 0.f && u <= 1.f && v >= 0.f && v <= 1.f)) return __builtin_inff();
    int w = (int)((float)SW - 0.001f);
    int h = (int)((float)SH - 0.001f);
    int x = (int)((float)w * u);
    int y = (int)((float)h * v);
    float dd = map[(size_t)x + (size_t)y * SW];
    // (double)dd < 0.001: 0.001f is the float just above 0.001 (the one just
    // below is 0.00099999993), so for a float dd the two tests agree, NaN
    // included.
    return (dd < 0.001f) ? 1.f : dd;
}

__device__ __forceinline__ float pcf_disk(const float *map, int SW, int SH, Vec3 sc, float bias,
                                          float filter, const float *disk, int n)
{
    float vis = 0.f;
    float inv = 1.f / (float)n;
    for (int i = 0; i < n; ++i) {
        float s = sample_shadow(map, SW, SH, sc.x + disk[2 * i] * filter,
                                sc.y + disk[2 * i + 1] * filter);
        if (sc.z <= s + bias) vis += inv;
    }
    return vis;
}

// PCSS blocker term (FindAverageBlockDepth, shadow.cpp:65-91): the average
// depth of the samples that block, 0 when none does.  `disk` holds the
// offsets already scaled (blocker_off: the unit-disk samples times
// blocker_filter in double, as the reference scales them).
__device__ __forceinline__ float pcss_blocker_depth(const FrameUniforms &u, const float *map, Vec3 sc, float bias,
                                                    const float *disk, int n)
{
    float blocker = 0.f, nb = 0.f;
    for (int i = 0; i < n; ++i) {
        const float ou = disk[2 * i];
        const float ov = disk[2 * i + 1];
        float su = sc.x + ou;
        float sv = sc.y + ov;
        float s = sample_shadow(map, u.SW, u.SH, su, sv);
        if (sc.z > s + bias) {
            blocker += s;
            nb += 1.f;
        }
    }
    return (nb < 1.f) ? 0.f : blocker / nb;
}

// The random stream's samples (modes 4 and 5) are each pixel's own, 512
// bytes or more apart between neighbouring lanes: read one sample per load,
// every load would fetch a cache line per lane.  So they are taken
// kStreamGroup at a time: the group's pairs, then its texels, are loaded
// before any is used (a texel of a sample off the map is read at index 0 and
// replaced by +inf, as sample_shadow returns it), and a pixel's lines are
// used while they are cached.  The arithmetic and its order per sample are
// those of the reference's FindAverageBlockDepth and PCF (pcss_blocker_depth
// with the offsets scaled per sample, pcf_disk).
constexpr int kStreamGroup = 8;

// Shadow-map texel of (su, sv) as sample_shadow indexes it; in = on the map.
__device__ __forceinline__ size_t shadow_texel(int SW, int w, int h, float su, float sv, bool &in)
{
    in = su >= 0.f && su <= 1.f && sv >= 0.f && sv <= 1.f;
    const int x = in ? (int)((float)w * su) : 0, y = in ? (int)((float)h * sv) : 0;
    return (size_t)x + (size_t)y * SW;
}

__device__ __forceinline__ float shadow_value(bool in, float dd) { return !in ? __builtin_inff() : ((dd < 0.001f) ? 1.f : dd); }

// pcss_blocker_depth over stream samples (unit-disk samples scaled here in
// double, as the reference scales them).
__device__ __forceinline__ float pcss_blocker_stream(const FrameUniforms &u, const float *map, Vec3 sc, float bias,
                                                     const float *disk, int n)
{
    const DevShadowCfg &c = u.shadow;
    const float2 *d = (const float2 *)disk;
    const int w = (int)((float)u.SW - 0.001f), h = (int)((float)u.SH - 0.001f);
    float blocker = 0.f, nb = 0.f;
    int i = 0;
    for (; i + kStreamGroup <= n; i += kStreamGroup) {
        float2 ds[kStreamGroup];
#pragma unroll
        for (int e = 0; e < kStreamGroup; ++e) ds[e] = d[i + e];
        float dd[kStreamGroup];
        bool in[kStreamGroup];
#pragma unroll
        for (int e = 0; e < kStreamGroup; ++e) {
            const float su = sc.x + (float)((double)ds[e].x * c.blocker_filter);
            const float sv = sc.y + (float)((double)ds[e].y * c.blocker_filter);
            dd[e] = map[shadow_texel(u.SW, w, h, su, sv, in[e])];
        }
#pragma unroll
        for (int e = 0; e < kStreamGroup; ++e) {
            const float s = shadow_value(in[e], dd[e]);
            if (sc.z > s + bias) {
                blocker += s;
                nb += 1.f;
            }
        }
    }
    for (; i < n; ++i) {
        const float su = sc.x + (float)((double)d[i].x * c.blocker_filter);
        const float sv = sc.y + (float)((double)d[i].y * c.blocker_filter);
        const float s = sample_shadow(map, u.SW, u.SH, su, sv);
        if (sc.z > s + bias) {
            blocker += s;
            nb += 1.f;
        }
    }
    return (nb < 1.f) ? 0.f : blocker / nb;
}

// pcf_disk over stream samples.
__device__ __forceinline__ float pcf_stream(const float *map, int SW, int SH, Vec3 sc, float bias, float filter,
                                            const float *disk, int n)
{
    const float2 *d = (const float2 *)disk;
    const int w = (int)((float)SW - 0.001f), h = (int)((float)SH - 0.001f);
    float vis = 0.f;
    const float inv = 1.f / (float)n;
    int i = 0;
    for (; i + kStreamGroup <= n; i += kStreamGroup) {
        float2 ds[kStreamGroup];
#pragma unroll
        for (int e = 0; e < kStreamGroup; ++e) ds[e] = d[i + e];
        float dd[kStreamGroup];
        bool in[kStreamGroup];
#pragma unroll
        for (int e = 0; e < kStreamGroup; ++e)
            dd[e] = map[shadow_texel(SW, w, h, sc.x + ds[e].x * filter, sc.y + ds[e].y * filter, in[e])];
#pragma unroll
        for (int e = 0; e < kStreamGroup; ++e)
            if (sc.z <= shadow_value(in[e], dd[e]) + bias) vis += inv;
    }
    for (; i < n; ++i) {
        const float s = sample_shadow(map, SW, SH, sc.x + d[i].x * filter, sc.y + d[i].y * filter);
        if (sc.z <= s + bias) vis += inv;
    }
    return vis;
}

// Shadow coordinate and slope bias (CalculateShadowVisibility, shadow.cpp:109-117).
__device__ __forceinline__ Vec3 shadow_coord(Vec3 ndc) { return add3(scale3(ndc, 0.5f), vec3(0.5f, 0.5f, 0.5f)); }
__device__ __forceinline__ float shadow_bias(Vec3 n, Vec3 l) { return smax(0.009f * (1.f - dot3(n, l)), 0.007f); }

// STREAM: the kernel runs with the random stream's samples (modes 4 and 5),
// `rdisk` being this pixel's: the PCF samples, or the blocker samples
// followed by the PCF ones.  The kernels without them are compiled without
// the stream filters (their registers would cost the other modes occupancy);
// there modes 4 and 5 do not occur (the host launches the STREAM kernels for
// them) and would read as lit.
template <bool STREAM>
__device__ __forceinline__ float shadow_visibility(const FrameUniforms &u, const float *map,
                                                   Vec3 ndc, Vec3 n, Vec3 l, const float *rdisk)
{
    const DevShadowCfg &c = u.shadow;
    Vec3 sc = shadow_coord(ndc);
    float bias = shadow_bias(n, l);
    if (c.mode == 1) return pcf_disk(map, u.SW, u.SH, sc, bias, c.pcf_filter, c.pcf_disk, c.n_pcf);
    if (c.mode == 4) {
        if constexpr (STREAM) return pcf_stream(map, u.SW, u.SH, sc, bias, c.pcf_filter, rdisk, c.n_pcf);
        return 1.f;
    }
    if (c.mode == 5) {
        if constexpr (STREAM) {
            const float dB = pcss_blocker_stream(u, map, sc, bias, rdisk, c.n_blocker);
            if ((double)dB < 0.001) return 1.f;
            float pen = (sc.z - dB) * c.area_light / dB;
            float filter = (float)(c.pcss_filter * (double)pen);
            return pcf_stream(map, u.SW, u.SH, sc, bias, filter, rdisk + 2 * c.n_blocker, c.n_pcf);
        }
        return 1.f;
    }
    if (c.mode == 2) {
        const float dB = pcss_blocker_depth(u, map, sc, bias, c.blocker_off, c.n_blocker);
        if ((double)dB < 0.001) return 1.f;
        float pen = (sc.z - dB) * c.area_light / dB;
        float filter = (float)(c.pcss_filter * (double)pen);
        return pcf_disk(map, u.SW, u.SH, sc, bias, filter, c.pcf_disk, c.n_pcf);
    }
    if (c.mode == 3) {
        int r = c.grid / 2;
        float inv = 1.f / (float)(c.grid * c.grid);
        float tu = 1.f / (float)u.SW, tv = 1.f / (float)u.SH;
        float vis = 0.f;
        for (int dy = -r; dy <= r; ++dy)
            for (int dx = -r; dx <= r; ++dx) {
                float s = sample_shadow(map, u.SW, u.SH, sc.x + (float)dx * tu, sc.y + (float)dy * tv);
                if (sc.z <= s + bias) vis += inv;
            }
        return vis;
    }
    float s = sample_shadow(map, u.SW, u.SH, sc.x, sc.y);
    return (sc.z <= s + bias) ? 1.f : 0.f;
}

// ---------------- lighting (phongshader.h:171-215, pbrshader.h:182-288) ----------------

// powf(x, y) for the non-negative (or NaN) bases of the shading code, as
// exp2(y * log2(x)) with the hardware base-2 log and exp: a few ulps from the
// correctly rounded power, which the 8-bit output absorbs (a result moves by
// at most one 8-bit step, and only at a quantisation boundary), in ~8
// operations instead of the ~160 of the general powf.  pow(x, 0) = 1;
// denormal x and results below 2^-126 are rescaled (the hardware log and exp
// take and give normal floats only).
__device__ __forceinline__ float pow_pos(float x, float y)
{
    if (y == 0.f) return 1.f;
    const bool sub = x < 0x1p-126f;
    const float lx = __builtin_amdgcn_logf(sub ? x * 0x1p32f : x) - (sub ? 32.f : 0.f);
    const float e = y * lx;
    const bool tiny = e < -126.f;
    return __builtin_amdgcn_exp2f(tiny ? e + 32.f : e) * (tiny ? 0x1p-32f : 1.f);
}

__device__ __forceinline__ Vec3 pow3f(Vec3 a, float p)
{
    return vec3(pow_pos(a.x, p), pow_pos(a.y, p), pow_pos(a.z, p));
}

__device__ __forceinline__ Vec3 tonemap_gamma(Vec3 c)
{
    c = vec3(c.x / (c.x + 1.f), c.y / (c.y + 1.f), c.z / (c.z + 1.f));
    return clamp01_3(pow3f(c, 1.f / 2.2f));
}

__device__ __forceinline__ Vec3 bp_term(Vec3 L, Vec3 H, Vec3 N, float vis, Vec3 dl, Vec3 el,
                                        float ao, float ks, float shin, Vec3 lc, int shadow_on,
                                        bool with_emissive)
{
    float diff = smax(0.f, dot3(L, N));
    float spec = pow_pos(smax(0.f, dot3(H, N)), shin);
    Vec3 diffuse = scale3(scale3(dl, diff), ao);
    Vec3 specular = scale3(vec3(ks, ks, ks), spec);
    if (shadow_on) {
        float sh = (1 - vis) * 0.6f;
        float vv = 1 - sh;
        diffuse = scale3(diffuse, vv);
        specular = scale3(specular, vv);
    }
    Vec3 t = add3(diffuse, specular);
    if (with_emissive) t = add3(t, el);
    return mul3(t, lc);
}

__device__ __forceinline__ Vec3 pbr_lo(Vec3 L, Vec3 V, Vec3 H, Vec3 N, float vis, Vec3 al,
                                       float metal, float rough, Vec3 rad, int shadow_on)
{
    const float InvPi = (float)0.31830988618379067154;
    float NdotV = smax(dot3(N, V), 0.f);
    float NdotL = smax(dot3(N, L), 0.f);
    float NdotH = smax(dot3(N, H), 0.f);
    float HdotV = smax(dot3(H, V), 0.f);
    Vec3 F0 = lerp3(metal, vec3(0.04f, 0.04f, 0.04f), al);
    float a = rough * rough;
    float a2 = a * a;
    float NdotH2 = NdotH * NdotH;
    float den = (NdotH2 * (a2 - 1.f) + 1.f);
    float den2 = den * den;
    float NDF = a2 * InvPi / den2;
    float ka = rough + 1.f;
    float k = ka * ka / 8.f;
    float g1 = NdotV / (NdotV * (1 - k) + k);
    float g2 = NdotL / (NdotL * (1 - k) + k);
    float G = g1 * g2;
    float omh = smax(1.f - HdotV, 0.f);
    float p5 = pow_pos(omh, 5.f);
    Vec3 F = add3(F0, scale3(sub3(vec3(1.f, 1.f, 1.f), F0), p5));
    Vec3 DGF = scale3(F, NDF * G);
    float denom = 4 * NdotV * NdotL + 0.001f;
    Vec3 spec = divs3(DGF, denom);
    Vec3 kd = scale3(sub3(vec3(1.f, 1.f, 1.f), F), 1.f - metal);
    Vec3 brdf = add3(scale3(mul3(kd, al), InvPi), spec);
    Vec3 Lo = scale3(mul3(brdf, rad), NdotL);
    if (shadow_on) {
        float sh = (1 - vis) * 0.6f;
        float vv = 1 - sh;
        Lo = scale3(Lo, vv);
    }
    return Lo;
}

__device__ __forceinline__ Vec3 light_dir(const DevLight &l, Vec3 P)
{
    if (l.type == 0) return norm3(sub3(vec3(l.vec[0], l.vec[1], l.vec[2]), P));
    return norm3(vec3(-l.vec[0], -l.vec[1], -l.vec[2]));
}

// One surface point lit by every light.  With a single point light this is
// the reference's CalculateLight; `deferred_quirk` passes viewDir in the
// Blinn-Phong halfway slot as DrawScreenSpacePixels does (forkergl.cpp:369).
template <bool STREAM = false>
__device__ __forceinline__ Vec3 shade_point(const FrameUniforms &u, const float *smap, bool pbr,
                                            bool deferred_quirk, Vec3 P, Vec3 N, Vec3 lndc,
                                            Vec3 albedo, Vec3 emissive, Vec3 param,
                                            const float *rdisk = nullptr)
{
    const float gamma = (float)2.2;
    Vec3 Vd = norm3(sub3(vec3(u.eye[0], u.eye[1], u.eye[2]), P));
    if (!pbr) {
        Vec3 dl = pow3f(albedo, gamma);
        Vec3 el = pow3f(emissive, gamma);
        Vec3 c = scale3(mul3(vec3(0.3f, 0.3f, 0.3f), dl), param.x);
        for (int i = 0; i < u.n_lights; ++i) {
            const DevLight &l = u.lights[i];
            Vec3 L = light_dir(l, P);
            float vis = 1.f;
            if (!u.shadow_on) vis = 0.f;
            else if (i == 0) vis = shadow_visibility<STREAM>(u, smap, lndc, N, L, rdisk);
            Vec3 H = deferred_quirk ? Vd : norm3(add3(L, Vd));
            c = add3(c, bp_term(L, H, N, vis, dl, el, param.x, param.y, param.z,
                                vec3(l.color[0], l.color[1], l.color[2]), u.shadow_on, i == 0));
        }
        return tonemap_gamma(c);
    }
    Vec3 al = pow3f(albedo, gamma);
    Vec3 el = pow3f(emissive, gamma);
    Vec3 Lo = vec3(0.f, 0.f, 0.f);
    for (int i = 0; i < u.n_lights; ++i) {
        const DevLight &l = u.lights[i];
        Vec3 L = light_dir(l, P);
        float vis = 1.f;
        if (!u.shadow_on) vis = 0.f;
        else if (i == 0) vis = shadow_visibility<STREAM>(u, smap, lndc, N, L, rdisk);
        Vec3 H = norm3(add3(L, Vd));
        Vec3 lo = pbr_lo(L, Vd, H, N, vis, al, param.y, param.z,
                         vec3(l.color[0], l.color[1], l.color[2]), u.shadow_on);
        Lo = (i == 0) ? lo : add3(Lo, lo);
    }
    Vec3 color = Lo;
    color = add3(color, scale3(mul3(vec3(0.3f, 0.3f, 0.3f), al), param.x));
    color = add3(color, el);
    return tonemap_gamma(color);
}

}  // namespace fgl
