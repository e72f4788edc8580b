// raster_core.hpp -- per-vertex, per-triangle, per-tile and per-pixel bodies
// of the raster pipeline, shared by the single-view kernels (raster.hip) and
// the batched multi-view kernels (batch.hip), so both run the same
// arithmetic.
//
// Exactness: coverage is the reference's Barycentric (src/geometry.cpp:20-56).
// When every window coordinate satisfies |X|,|Y| <= 8191 the FP32 differences
// and FP64 cross products are exact integers, so they are computed in int32
// and only u = cx*inv, v = cy*inv, 1-(u+v) run in FP64 -- the same doubles.
// Otherwise the literal FP32/FP64 sequence runs.
#pragma once

#include "fgl_kernels.hpp"

namespace fgl {

constexpr int kExactLimit = 8191;

// Window coordinates of one vertex: Viewport * NDC with all four terms of
// each row, truncated like x86 cvttss2si (src/forkergl.cpp:244-249).
__device__ __forceinline__ int4 to_screen(const float *vp, Vec4 ndc)
{
    Vec4 c = mv4(vp, ndc);
    int X = trunc_int(c.x), Y = trunc_int(c.y);
    int exact = (X >= -kExactLimit && X <= kExactLimit && Y >= -kExactLimit && Y <= kExactLimit);
    return make_int4(X, Y, __float_as_int(c.z), exact);
}

// BoundBox::GenerateBoundBox (src/forkergl.cpp:145-163)
__device__ __forceinline__ void clamped_bbox(int4 a, int4 b, int4 c, int W, int H, int &minx,
                                             int &miny, int &maxx, int &maxy)
{
    minx = clampi(imin(a.x, imin(b.x, c.x)), 0, W - 1);
    miny = clampi(imin(a.y, imin(b.y, c.y)), 0, H - 1);
    maxx = clampi(imax(a.x, imax(b.x, c.x)), 0, W - 1);
    maxy = clampi(imax(a.y, imax(b.y, c.y)), 0, H - 1);
}

// Cross-product z of Barycentric, which does not depend on the pixel:
// |cz| > 1e-2 decides whether the triangle can cover anything.
__device__ __forceinline__ bool nondegenerate_general(int4 a, int4 b, int4 c)
{
    float Ax = (float)a.x, Ay = (float)a.y, Bx = (float)b.x, By = (float)b.y;
    float Cx = (float)c.x, Cy = (float)c.y;
    double s0x = (double)(Bx - Ax), s0y = (double)(Cx - Ax);
    double s1x = (double)(By - Ay), s1y = (double)(Cy - Ay);
    double uz = s0x * s1y - s0y * s1x;
    return fabs(uz) > 1e-2;
}

// Triangle setup of face f of draw d (pass-wide triangle id g): clamped
// bbox, degenerate / off-screen cull, screen-tile rectangle and count.
__device__ __forceinline__ void setup_tri(const DevDraw &dr, int d, int f,
                                          const int4 *__restrict__ screen, int W, int H, int TS,
                                          uint2 *__restrict__ rect, uint32_t *__restrict__ count,
                                          int4 *__restrict__ tri_v,
                                          uint32_t *__restrict__ big_list,
                                          uint32_t *__restrict__ big_count)
{
    const int g = dr.tri_base + f;
    const int v0 = dr.vert_base + dr.vi[3 * f], v1 = dr.vert_base + dr.vi[3 * f + 1];
    const int v2 = dr.vert_base + dr.vi[3 * f + 2];
    const int4 a = screen[v0], b = screen[v1], c = screen[v2];
    int minx, miny, maxx, maxy;
    clamped_bbox(a, b, c, W, H, minx, miny, maxx, maxy);
    const bool exact = a.w && b.w && c.w && W <= 8192 && H <= 8192;
    bool keep;
    if (exact) {
        int cz = (b.x - a.x) * (c.y - a.y) - (c.x - a.x) * (b.y - a.y);
        int ux0 = imin(a.x, imin(b.x, c.x)), ux1 = imax(a.x, imax(b.x, c.x));
        int uy0 = imin(a.y, imin(b.y, c.y)), uy1 = imax(a.y, imax(b.y, c.y));
        // Every clamped-bbox pixel of an off-screen triangle lies strictly
        // outside it, and in the exact regime one computed weight is then
        // strictly negative: nothing can be covered.
        keep = cz != 0 && !(ux1 < 0 || ux0 > W - 1 || uy1 < 0 || uy0 > H - 1);
    } else {
        keep = nondegenerate_general(a, b, c);
    }
    const uint32_t tx0 = minx / TS, tx1 = maxx / TS, ty0 = miny / TS, ty1 = maxy / TS;
    const uint32_t n = keep ? (tx1 - tx0 + 1) * (ty1 - ty0 + 1) : 0u;
    rect[g] = make_uint2(tx0 | (tx1 << 16), ty0 | (ty1 << 16));
    count[g] = n;
    tri_v[g] = make_int4(v0, v1, v2, d);
    // The big-triangle list only decides which workgroup writes a triangle's
    // pairs; their positions come from the scan, so its order is irrelevant.
    if (n > (uint32_t)kSmallTriTiles) big_list[atomicAdd(big_count, 1u)] = (uint32_t)g;
}

// Pair j of triangle t: key = tile_base + tile index, row-major over the
// triangle's tile rectangle.
__device__ __forceinline__ uint32_t pair_key(uint2 r, uint32_t j, int tiles_x, uint32_t tile_base)
{
    const uint32_t tx0 = r.x & 0xffff, tx1 = r.x >> 16, ty0 = r.y & 0xffff;
    const uint32_t wx = tx1 - tx0 + 1;
    return tile_base + (ty0 + j / wx) * (uint32_t)tiles_x + tx0 + j % wx;
}

// Per-triangle record of the rasteriser (registers; staged in LDS as TriSoA).
struct TriL {
    int X0, Y0;
    // exact: B-A, C-A as s0x = Bx-Ax, s0y = Cx-Ax, s1x = By-Ay, s1y = Cy-Ay;
    // otherwise the other two vertices X1, Y1, X2, Y2
    int q0, q1, q2, q3;
    float z0, z1, z2;
    int bb0, bb1;            // minx | miny << 16, maxx | maxy << 16
    int cz;                  // exact cross product z
    int exact;
    int g;
    double inv;              // 1.0 / cz (exact regime)
};

__device__ __forceinline__ void stage_tri(TriL &t, int g, int4 a, int4 b, int4 c, int W, int H)
{
    t.X0 = a.x; t.Y0 = a.y;
    t.z0 = __int_as_float(a.z); t.z1 = __int_as_float(b.z); t.z2 = __int_as_float(c.z);
    int minx, miny, maxx, maxy;
    clamped_bbox(a, b, c, W, H, minx, miny, maxx, maxy);
    t.bb0 = minx | (miny << 16);
    t.bb1 = maxx | (maxy << 16);
    t.exact = a.w && b.w && c.w && W <= 8192 && H <= 8192;
    t.g = g;
    if (t.exact) {
        t.q0 = b.x - a.x; t.q1 = c.x - a.x; t.q2 = b.y - a.y; t.q3 = c.y - a.y;
        t.cz = t.q0 * t.q3 - t.q1 * t.q2;
        t.inv = 1.0 / (double)t.cz;
    } else {
        t.q0 = b.x; t.q1 = b.y; t.q2 = c.x; t.q3 = c.y;
        t.cz = 0;
        t.inv = 0.0;
    }
}

// Coverage of pixel (px,py) by a staged triangle; on success the three
// barycentric weights exactly as Barycentric returns them.
__device__ __forceinline__ bool cover(const TriL &t, int px, int py, float &b0, float &b1, float &b2)
{
    if (t.exact) {
        int s0z = t.X0 - px, s1z = t.Y0 - py;
        int cx = t.q1 * s1z - s0z * t.q3;    // s0y * s1z - s0z * s1y
        int cy = s0z * t.q2 - t.q0 * s1z;    // s0z * s1x - s0x * s1z
        // exact sign pre-test: weights cx/cz, cy/cz and (cz-cx-cy)/cz
        if (t.cz > 0) {
            if (cx < 0 || cy < 0 || cx + cy > t.cz) return false;
        } else {
            if (cx > 0 || cy > 0 || cx + cy < t.cz) return false;
        }
        double u = (double)cx * t.inv, v = (double)cy * t.inv;
        float r0 = (float)(1.0 - (u + v)), r1 = (float)u, r2 = (float)v;
        if (r0 < 0.f || r1 < 0.f || r2 < 0.f) return false;
        b0 = r0; b1 = r1; b2 = r2;
        return true;
    }
    return barycentric_ref(t.X0, t.Y0, t.q0, t.q1, t.q2, t.q3, px, py, &b0, &b1, &b2);
}

// XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs, so
// give each XCD a contiguous run of tiles (neighbouring tiles share
// triangles and vertex data in that XCD's L2).
__device__ __forceinline__ int xcd_remap(int bid, int n)
{
    const int X = 8;
    int per = n / X, rem = n % X;
    int xcd = bid % X, local = bid / X;
    int start = xcd * per + imin(xcd, rem);
    return start + local;
}

// ---------------- order-independent tile rasteriser ----------------
//
// The submission-order depth fold of src/forkergl.cpp:180-199 -- replace when
// !(d >= best), fragments in increasing pass-wide id g -- is reproduced
// without any order:
//  * no NaN involved: the winner is the lowest g among fragments of minimum
//    depth (-0 == +0), or the initial value if it is not beaten: a minimum
//    over 64-bit keys (order(depth) << 32 | g + 1), initial value g + 1 = 0;
//  * a NaN fragment always replaces, and whatever follows a NaN replaces it:
//    with J the highest g of a NaN fragment at the pixel, the winner is the
//    key minimum over the fragments with g > J (the initial value excluded),
//    or J itself when there are none;
//  * a NaN initial depth with no NaN fragment: the key minimum over all
//    fragments (the initial value excluded), or the initial value.
// So any lane can fold any triangle into any pixel with LDS atomics, and the
// tile's triangle list needs no particular order.  Triangles whose bbox
// covers at most kSmallTriArea pixels of the tile are rasterised one per lane
// (the common case for dense meshes, where a wave-per-triangle walk would
// leave most lanes idle); larger ones one per wave, a lane per pixel.  Tiles
// with a NaN run the fragment walk a second time for the affected pixels.

constexpr int kSmallTriArea = 16;

// Monotone map of float depth to uint32 with -0 == +0.
__device__ __forceinline__ uint32_t depth_order(float d)
{
    const uint32_t u = __float_as_uint(d == 0.f ? 0.f : d);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float depth_from_order(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// The staged triangles of a chunk, one array per field: lanes reading
// consecutive triangles hit consecutive banks (an array of 64- or 80-byte
// records puts 8 to 32 lanes on one bank).
template <int BLOCK>
struct TriSoA {
    int X0[BLOCK], Y0[BLOCK], q0[BLOCK], q1[BLOCK], q2[BLOCK], q3[BLOCK];
    float z0[BLOCK], z1[BLOCK], z2[BLOCK];
    int bb0[BLOCK], bb1[BLOCK], cz[BLOCK], exact[BLOCK], g[BLOCK];
    double inv[BLOCK];
    __device__ __forceinline__ void put(int i, const TriL &t)
    {
        X0[i] = t.X0; Y0[i] = t.Y0; q0[i] = t.q0; q1[i] = t.q1; q2[i] = t.q2; q3[i] = t.q3;
        z0[i] = t.z0; z1[i] = t.z1; z2[i] = t.z2;
        bb0[i] = t.bb0; bb1[i] = t.bb1; cz[i] = t.cz; exact[i] = t.exact; g[i] = t.g;
        inv[i] = t.inv;
    }
    __device__ __forceinline__ TriL get(int i) const
    {
        TriL t;
        t.X0 = X0[i]; t.Y0 = Y0[i]; t.q0 = q0[i]; t.q1 = q1[i]; t.q2 = q2[i]; t.q3 = q3[i];
        t.z0 = z0[i]; t.z1 = z1[i]; t.z2 = z2[i];
        t.bb0 = bb0[i]; t.bb1 = bb1[i]; t.cz = cz[i]; t.exact = exact[i]; t.g = g[i];
        t.inv = inv[i];
        return t;
    }
};

template <int TS, int BLOCK>
struct RasterLds {
    TriSoA<BLOCK> tri;
    unsigned long long key[TS * TS];
    int nan_g[TS * TS];       // highest id of a NaN fragment, -1 if none
    uint8_t nan_init[TS * TS];
    int nan_seen;
};

// Depth of staged triangle t at a covered pixel: Dot(bary, depths) in the
// reference's order.
__device__ __forceinline__ float tri_depth(const TriL &t, float b0, float b1, float b2)
{
    float dd = 0.f;
    dd += b0 * t.z0;
    dd += b1 * t.z1;
    dd += b2 * t.z2;
    return dd;
}

// One walk over the tile's triangles.  PASS 0: every fragment; PASS 1 (NaN
// tiles): only pixels touched by a NaN, only fragments after their last NaN.
template <int TS, int BLOCK, int PASS>
__device__ __forceinline__ void raster_walk(const RasterArgs &a, int tile, RasterLds<TS, BLOCK> &L)
{
    constexpr int NW = BLOCK / 64;
    constexpr int QS = (NW == 1) ? TS : TS / 2;
    constexpr int NPX = QS * QS / 64;
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int x0 = tx * TS, y0 = ty * TS;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int qx0 = x0 + (NW == 1 ? 0 : (wave & 1) * QS);
    const int qy0 = y0 + (NW == 1 ? 0 : (wave >> 1) * QS);
    const int rx1 = qx0 + QS - 1, ry1 = qy0 + QS - 1;
    const uint32_t start = a.tile_start[tile], end = a.tile_end[tile];
    auto fold = [&](const TriL &t, int px, int py, float dd) {
        const int p = (py - y0) * TS + (px - x0);
        if (PASS == 0) {
            if (dd != dd) {
                atomicMax(&L.nan_g[p], t.g);
                L.nan_seen = 1;
                return;
            }
        } else {
            const int j = L.nan_g[p];
            if (!(j >= 0 || L.nan_init[p]) || t.g <= j || dd != dd) return;
        }
        atomicMin(&L.key[p], ((unsigned long long)depth_order(dd) << 32) | (unsigned long long)(uint32_t)(t.g + 1));
    };
    for (uint32_t base = start; base < end; base += BLOCK) {
        const int n = imin(BLOCK, (int)(end - base));
        __syncthreads();   // previous chunk's triangles fully consumed
        if (tid < n) {
            const int g = (int)a.vals[base + tid];
            const int4 tv = a.tri_v[g];
            TriL t;
            stage_tri(t, g, a.screen[tv.x], a.screen[tv.y], a.screen[tv.z], a.W, a.H);
            L.tri.put(tid, t);
        }
        __syncthreads();
        // small triangles: one per lane
        if (tid < n) {
            const TriL t = L.tri.get(tid);
            const int bx0 = imax(t.bb0 & 0xffff, x0), bx1 = imin(t.bb1 & 0xffff, x0 + TS - 1);
            const int by0 = imax(t.bb0 >> 16, y0), by1 = imin(t.bb1 >> 16, y0 + TS - 1);
            const int area = (bx1 >= bx0 && by1 >= by0) ? (bx1 - bx0 + 1) * (by1 - by0 + 1) : 0;
            if (area > 0 && area <= kSmallTriArea) {
                for (int py = by0; py <= by1; ++py)
                    for (int px = bx0; px <= bx1; ++px) {
                        float b0, b1, b2;
                        if (cover(t, px, py, b0, b1, b2)) fold(t, px, py, tri_depth(t, b0, b1, b2));
                    }
            }
        }
        // large triangles: one per wave over its quadrant, a lane per pixel
        unsigned long long mask[NW];
#pragma unroll
        for (int jj = 0; jj < NW; ++jj) {
            const int si = jj * 64 + lane;
            bool hit = false;
            if (si < n) {
                const int bb0 = L.tri.bb0[si], bb1 = L.tri.bb1[si];
                const int minx = bb0 & 0xffff, miny = bb0 >> 16;
                const int maxx = bb1 & 0xffff, maxy = bb1 >> 16;
                const int bx0 = imax(minx, x0), bx1 = imin(maxx, x0 + TS - 1);
                const int by0 = imax(miny, y0), by1 = imin(maxy, y0 + TS - 1);
                const int area = (bx1 >= bx0 && by1 >= by0) ? (bx1 - bx0 + 1) * (by1 - by0 + 1) : 0;
                hit = area > kSmallTriArea && !(maxx < qx0 || minx > rx1 || maxy < qy0 || miny > ry1);
            }
            mask[jj] = __ballot(hit);
        }
#pragma unroll
        for (int jj = 0; jj < NW; ++jj) {
            unsigned long long m = mask[jj];
            while (m) {
                const int i = jj * 64 + __builtin_ctzll(m);
                m &= m - 1;
                const TriL t = L.tri.get(i);
                const int minx = t.bb0 & 0xffff, miny = t.bb0 >> 16;
                const int maxx = t.bb1 & 0xffff, maxy = t.bb1 >> 16;
#pragma unroll
                for (int k = 0; k < NPX; ++k) {
                    const int lin = lane + 64 * k;
                    const int px = qx0 + lin % QS, py = qy0 + lin / QS;
                    if (px < a.W && py < a.H && px >= minx && px <= maxx && py >= miny && py <= maxy) {
                        float b0, b1, b2;
                        if (cover(t, px, py, b0, b1, b2)) fold(t, px, py, tri_depth(t, b0, b1, b2));
                    }
                }
            }
        }
    }
    __syncthreads();
}

template <int TS, int BLOCK, bool SHADOW>
__device__ __forceinline__ void raster_tile(const RasterArgs &a, int tile, RasterLds<TS, BLOCK> &L)
{
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int x0 = tx * TS, y0 = ty * TS;
    const int tid = threadIdx.x;
    if (tid == 0) L.nan_seen = 0;
    __syncthreads();
    for (int p = tid; p < TS * TS; p += BLOCK) {
        const int px = x0 + p % TS, py = y0 + p / TS;
        float init = 3.40282347e+38f;
        if (!a.fresh && px < a.W && py < a.H) init = a.depth[(size_t)px + (size_t)py * a.W];
        const bool ni = init != init;
        if (ni) L.nan_seen = 1;
        L.nan_init[p] = ni ? 1 : 0;
        L.nan_g[p] = -1;
        L.key[p] = ni ? ~0ull : (unsigned long long)depth_order(init) << 32;
    }
    raster_walk<TS, BLOCK, 0>(a, tile, L);   // ends with a barrier
    if (L.nan_seen) {
        for (int p = tid; p < TS * TS; p += BLOCK)
            if (L.nan_g[p] >= 0 || L.nan_init[p]) L.key[p] = ~0ull;
        raster_walk<TS, BLOCK, 1>(a, tile, L);
    }
    for (int p = tid; p < TS * TS; p += BLOCK) {
        const int px = x0 + p % TS, py = y0 + p / TS;
        if (px >= a.W || py >= a.H) continue;
        const size_t i = (size_t)px + (size_t)py * a.W;
        const unsigned long long key = L.key[p];
        int win = key == ~0ull ? -1 : (int)(uint32_t)(key & 0xffffffffu) - 1;
        bool from_key = win >= 0;
        if (!from_key && L.nan_g[p] >= 0) win = L.nan_g[p];   // nothing after the last NaN fragment
        if (win >= 0) {
            float dd = from_key ? depth_from_order((uint32_t)(key >> 32)) : 0.f;
            TriL t;
            float b0 = 0.f, b1 = 0.f, b2 = 0.f;
            const int4 tv = a.tri_v[win];
            if (!from_key || dd == 0.f || SHADOW) {
                stage_tri(t, win, a.screen[tv.x], a.screen[tv.y], a.screen[tv.z], a.W, a.H);
                cover(t, px, py, b0, b1, b2);
                if (!from_key || dd == 0.f) dd = tri_depth(t, b0, b1, b2);   // NaN, or a zero with its sign
            }
            a.depth[i] = dd;
            if (!SHADOW) {
                a.vis[i] = win;
            } else {
                // DepthShader::ProcessFragment (depthshader.h:30-36), with
                // the vertices' NDC z as the vertex pass computed them
                float z = dot3(vec3(a.ndc_z[tv.x], a.ndc_z[tv.y], a.ndc_z[tv.z]), vec3(b0, b1, b2));
                a.shadow_out[i] = z * 0.5f + 0.5f;
            }
        } else {
            if (a.fresh) a.depth[i] = 3.40282347e+38f;
            if (!SHADOW) a.vis[i] = -1;
            else if (a.fresh) a.shadow_out[i] = 0.f;
        }
    }
}

// Barycentric weights of the winning triangle at (px,py), recomputed with the
// same staged arithmetic as the coverage loop.
__device__ __forceinline__ Vec3 winner_bary(const int4 *__restrict__ screen,
                                            const int4 *__restrict__ tri_v, int g, int W, int H,
                                            int px, int py, int4 &tv)
{
    tv = tri_v[g];
    TriL t;
    stage_tri(t, g, screen[tv.x], screen[tv.y], screen[tv.z], W, H);
    float b0 = 0.f, b1 = 0.f, b2 = 0.f;
    cover(t, px, py, b0, b1, b2);
    return vec3(b0, b1, b2);
}

// GShader::ProcessFragment + the seven buffer writes (gshader.h:95-202,
// forkergl.cpp:211-223) for the winner (face f of draw dr, material m,
// barycentrics b) at pixel i: the 19 G-buffer planes.
__device__ __forceinline__ void gbuffer_shade(const ResolveArgs &a, const FrameUniforms &u, const DevDraw &dr,
                                              const DevMaterial &m, int f, const Vec3 &b, size_t i, size_t P)
{
    float *gb = a.gbuf;
    Varying v0 = main_vertex(dr, u, f, 0);
    Varying v1 = main_vertex(dr, u, f, 1);
    Varying v2 = main_vertex(dr, u, f, 2);
    FragGeom fg = frag_geometry(dr, u, a.textures, v0, v1, v2, b, m.map_bump);
    Vec3 albedo, emissive, param;
    float type;
    if (dr.support_pbr) {
        albedo = m.map_base >= 0 ? tex_sample(a.textures, m.map_base, fg.uv)
                                 : vec3(m.albedo[0], m.albedo[1], m.albedo[2]);
        emissive = m.map_pbr_ke >= 0 ? tex_sample(a.textures, m.map_pbr_ke, fg.uv)
                                     : vec3(m.ke[0], m.ke[1], m.ke[2]);
        float rough = m.map_rough >= 0 ? tex_samplef(a.textures, m.map_rough, fg.uv) : m.roughness;
        float metal = m.map_metal >= 0 ? tex_samplef(a.textures, m.map_metal, fg.uv) : m.metalness;
        float ao = m.map_ao >= 0 ? tex_samplef(a.textures, m.map_ao, fg.uv) : 1.f;
        param = vec3(ao, metal, rough);
        type = 1.f;
    } else {
        emissive = m.map_ke >= 0 ? tex_sample(a.textures, m.map_ke, fg.uv)
                                 : vec3(m.ke[0], m.ke[1], m.ke[2]);
        albedo = m.map_kd >= 0 ? tex_sample(a.textures, m.map_kd, fg.uv)
                               : vec3(m.kd[0], m.kd[1], m.kd[2]);
        float shin = m.map_ks >= 0 ? tex_samplef(a.textures, m.map_ks, fg.uv) + 5 : 1.f;
        param = vec3(1.f, m.ks[0], shin);
        type = 0.f;
    }
    gb[0 * P + i] = fg.normal.x; gb[1 * P + i] = fg.normal.y; gb[2 * P + i] = fg.normal.z;
    gb[3 * P + i] = fg.pos.x; gb[4 * P + i] = fg.pos.y; gb[5 * P + i] = fg.pos.z;
    if (u.shadow_on) {
        gb[6 * P + i] = fg.lndc.x; gb[7 * P + i] = fg.lndc.y; gb[8 * P + i] = fg.lndc.z;
    } else if (a.fresh) {
        gb[6 * P + i] = 0.f; gb[7 * P + i] = 0.f; gb[8 * P + i] = 0.f;
    }
    gb[9 * P + i] = albedo.x; gb[10 * P + i] = albedo.y; gb[11 * P + i] = albedo.z;
    gb[12 * P + i] = emissive.x; gb[13 * P + i] = emissive.y; gb[14 * P + i] = emissive.z;
    gb[15 * P + i] = param.x; gb[16 * P + i] = param.y; gb[17 * P + i] = param.z;
    gb[18 * P + i] = type;
}

// Geometry pass resolve of pixel (x,y); a pixel with no fragment gets the
// cleared values when `fresh`.  When every active lane of the wavefront
// resolves a face of the same draw with the same material (the common case:
// neighbouring pixels show the same mesh), the draw and material records are
// addressed through readfirstlane, so their fields (matrices, flags, texture
// descriptors) come in as scalar loads instead of one vector load per field.
__device__ __forceinline__ void gbuffer_px(const ResolveArgs &a, const FrameUniforms &u, int x, int y)
{
    const size_t i = (size_t)x + (size_t)y * a.W;
    const size_t P = (size_t)a.W * a.H;
    float *gb = a.gbuf;
    const int g = a.vis[i];
    if (g < 0) {
        if (a.fresh)
            for (int c = 0; c < kGbufPlanes; ++c) gb[c * P + i] = 0.f;
        return;
    }
    int4 tv;
    const Vec3 b = winner_bary(a.screen, a.tri_v, g, a.W, a.H, x, y, tv);
    const int d0 = __builtin_amdgcn_readfirstlane(tv.w);
    if (__ballot(tv.w != d0) == 0ull) {
        const DevDraw &dr = a.draws[d0];
        const int f = g - dr.tri_base;
        const int mi = dr.face_mat[f];
        const int m0 = __builtin_amdgcn_readfirstlane(mi);
        if (__ballot(mi != m0) == 0ull) {
            gbuffer_shade(a, u, dr, a.materials[m0], f, b, i, P);
            return;
        }
    }
    const DevDraw &dr = a.draws[tv.w];
    const int f = g - dr.tri_base;
    gbuffer_shade(a, u, dr, a.materials[dr.face_mat[f]], f, b, i, P);
}

}  // namespace fgl
