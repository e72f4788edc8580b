// raster.hip -- vertex transform, triangle setup/binning, the tile
// rasteriser and the per-pixel resolve passes for gfx950.
//
// Pipeline of one pass (all draws of the pass, in submission order):
//   k_vertex      one thread per (draw, position): window X, Y (int), z
//   k_setup       one thread per (draw, face): clamped bbox, degenerate and
//                 off-screen cull, screen-tile rectangle and count, and the
//                 triangle's pass-wide vertex ids
//   scan + emit   (tile, triangle) pairs in triangle order
//   stable radix sort by tile (sort.hip) -> per-tile lists in submission order
//   k_tile_ranges per-tile [start, end) in the sorted list
//   k_raster      one workgroup per tile: chunks of the tile's triangles are
//                 staged in LDS; each wavefront keeps, by ballot, only the
//                 triangles whose bbox meets its quadrant and walks them in
//                 list order; each lane folds the depth test for its pixels
//                 in submission order (no atomics) and records the winner
//   resolve       one thread per pixel: the pass's fragment program for the
//                 winning triangle (G-buffer or forward colour), coalesced
//                 SoA writes; the shadow pass resolves inside k_raster.
//
// Exactness: coverage is the reference's Barycentric (src/geometry.cpp:20-56).
// When every window coordinate satisfies |X|,|Y| <= 8191 the FP32 differences
// and FP64 cross products are exact integers, so they are computed in int32
// and only u = cx*inv, v = cy*inv, 1-(u+v) run in FP64 -- the same doubles.
// Otherwise the literal FP32/FP64 sequence runs.
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

template <bool SHADOW>
__global__ __launch_bounds__(256) void k_vertex(const DevDraw *__restrict__ draws, FrameUniforms u,
                                                int4 *__restrict__ out)
{
    const DevDraw &dr = draws[blockIdx.y];
    const int vi = blockIdx.x * blockDim.x + threadIdx.x;
    if (vi >= dr.n_pos) return;
    if (SHADOW) out[dr.vert_base + vi] = to_screen(u.shadow_viewport, depth_ndc(dr, vi));
    else out[dr.vert_base + vi] = to_screen(u.viewport, main_ndc(dr, u, vi));
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

__global__ __launch_bounds__(256) void k_setup(const DevDraw *__restrict__ draws,
                                               const int4 *__restrict__ screen, int W, int H,
                                               int TS, uint2 *__restrict__ rect,
                                               uint32_t *__restrict__ count,
                                               int4 *__restrict__ tri_v,
                                               uint32_t *__restrict__ big_list,
                                               uint32_t *__restrict__ big_count)
{
    const int d = blockIdx.y;
    const DevDraw &dr = draws[d];
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= dr.n_faces) return;
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

// Tiles of a triangle are emitted row-major over its tile rectangle at
// [off[t], off[t+1]).
__global__ __launch_bounds__(256) void k_emit_small(const uint32_t *__restrict__ off, int n_tri,
                                                    const uint2 *__restrict__ rect, int tiles_x,
                                                    uint32_t *__restrict__ keys,
                                                    uint32_t *__restrict__ vals)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tri) return;
    const uint32_t o = off[t], n = off[t + 1] - o;
    if (n == 0 || n > (uint32_t)kSmallTriTiles) return;
    const uint2 r = rect[t];
    const uint32_t tx0 = r.x & 0xffff, tx1 = r.x >> 16, ty0 = r.y & 0xffff;
    const uint32_t wx = tx1 - tx0 + 1;
    for (uint32_t j = 0; j < n; ++j) {
        keys[o + j] = (ty0 + j / wx) * tiles_x + tx0 + j % wx;
        vals[o + j] = (uint32_t)t;
    }
}

__global__ __launch_bounds__(256) void k_emit_big(const uint32_t *__restrict__ off,
                                                  const uint2 *__restrict__ rect, int tiles_x,
                                                  const uint32_t *__restrict__ big_list,
                                                  uint32_t *__restrict__ keys,
                                                  uint32_t *__restrict__ vals)
{
    const uint32_t t = big_list[blockIdx.x];
    const uint32_t o = off[t], n = off[t + 1] - o;
    const uint2 r = rect[t];
    const uint32_t tx0 = r.x & 0xffff, tx1 = r.x >> 16, ty0 = r.y & 0xffff;
    const uint32_t wx = tx1 - tx0 + 1;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        keys[o + j] = (ty0 + j / wx) * tiles_x + tx0 + j % wx;
        vals[o + j] = t;
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *__restrict__ p, uint32_t v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_fill_f32(float *__restrict__ p, float v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_tile_ranges(const uint32_t *__restrict__ keys, uint32_t n,
                                                     uint32_t *__restrict__ start,
                                                     uint32_t *__restrict__ end)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) start[k] = i;
    if (i == n - 1 || keys[i + 1] != k) end[k] = i + 1;
}

// Per-triangle record staged in LDS by the rasteriser.
struct TriL {
    int X0, Y0, X1, Y1, X2, Y2;
    int s0x, s0y, s1x, s1y;  // B-A, C-A (x then y); valid when exact
    float z0, z1, z2;
    int bb0, bb1;            // minx | miny << 16, maxx | maxy << 16
    int cz;                  // exact cross product z
    int exact;
    int g;
    double inv;              // 1.0 / cz (exact regime)
};

__device__ __forceinline__ void stage_tri(TriL &t, int g, int4 a, int4 b, int4 c, int W, int H)
{
    t.X0 = a.x; t.Y0 = a.y; t.X1 = b.x; t.Y1 = b.y; t.X2 = c.x; t.Y2 = c.y;
    t.z0 = __int_as_float(a.z); t.z1 = __int_as_float(b.z); t.z2 = __int_as_float(c.z);
    int minx, miny, maxx, maxy;
    clamped_bbox(a, b, c, W, H, minx, miny, maxx, maxy);
    t.bb0 = minx | (miny << 16);
    t.bb1 = maxx | (maxy << 16);
    t.exact = a.w && b.w && c.w && W <= 8192 && H <= 8192;
    t.g = g;
    if (t.exact) {
        t.s0x = b.x - a.x; t.s0y = c.x - a.x; t.s1x = b.y - a.y; t.s1y = c.y - a.y;
        t.cz = t.s0x * t.s1y - t.s0y * t.s1x;
        t.inv = 1.0 / (double)t.cz;
    } else {
        t.s0x = t.s0y = t.s1x = t.s1y = t.cz = 0;
        t.inv = 0.0;
    }
}

// Coverage of pixel (px,py) by a staged triangle; on success the three
// barycentric weights exactly as Barycentric returns them.
__device__ __forceinline__ bool cover(const TriL &t, int px, int py, float &b0, float &b1, float &b2)
{
    if (t.exact) {
        int s0z = t.X0 - px, s1z = t.Y0 - py;
        int cx = t.s0y * s1z - s0z * t.s1y;
        int cy = s0z * t.s1x - t.s0x * s1z;
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
    return barycentric_ref(t.X0, t.Y0, t.X1, t.Y1, t.X2, t.Y2, px, py, &b0, &b1, &b2);
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

// One workgroup per TS x TS tile.  With 4 wavefronts each owns a quadrant of
// QS = TS/2; with one wavefront it owns the tile.  Lane pixel k of a
// wavefront: lin = lane + 64k, (lin % QS, lin / QS) inside the quadrant.
template <int TS, int BLOCK, bool SHADOW>
__global__ __launch_bounds__(BLOCK) void k_raster(RasterArgs a)
{
    constexpr int NW = BLOCK / 64;
    constexpr int QS = (NW == 1) ? TS : TS / 2;
    constexpr int NPX = QS * QS / 64;
    __shared__ TriL lds[BLOCK];
    const int ntiles = a.tiles_x * a.tiles_y;
    const int tile = xcd_remap(blockIdx.x, ntiles);
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int qx0 = tx * TS + (NW == 1 ? 0 : (wave & 1) * QS);
    const int qy0 = ty * TS + (NW == 1 ? 0 : (wave >> 1) * QS);
    int px[NPX], py[NPX], win[NPX];
    float best[NPX];
    bool act[NPX];
#pragma unroll
    for (int k = 0; k < NPX; ++k) {
        const int lin = lane + 64 * k;
        px[k] = qx0 + lin % QS;
        py[k] = qy0 + lin / QS;
        act[k] = px[k] < a.W && py[k] < a.H;
        win[k] = -1;
        best[k] = (act[k] && !a.fresh) ? a.depth[(size_t)px[k] + (size_t)py[k] * a.W] : 3.40282347e+38f;
    }
    const int rx1 = qx0 + QS - 1, ry1 = qy0 + QS - 1;
    const uint32_t start = a.tile_start[tile], end = a.tile_end[tile];
    for (uint32_t base = start; base < end; base += BLOCK) {
        const int n = imin(BLOCK, (int)(end - base));
        if (tid < n) {
            const int g = (int)a.vals[base + tid];
            const int4 tv = a.tri_v[g];
            stage_tri(lds[tid], g, a.screen[tv.x], a.screen[tv.y], a.screen[tv.z], a.W, a.H);
        }
        __syncthreads();
        // Ballot per 64 staged triangles: does its bbox meet this quadrant?
        unsigned long long mask[NW];
#pragma unroll
        for (int j = 0; j < NW; ++j) {
            const int s = j * 64 + lane;
            bool hit = false;
            if (s < n) {
                const TriL &t = lds[s];
                const int minx = t.bb0 & 0xffff, miny = t.bb0 >> 16;
                const int maxx = t.bb1 & 0xffff, maxy = t.bb1 >> 16;
                hit = !(maxx < qx0 || minx > rx1 || maxy < qy0 || miny > ry1);
            }
            mask[j] = __ballot(hit);
        }
#pragma unroll
        for (int j = 0; j < NW; ++j) {
            unsigned long long m = mask[j];
            while (m) {
                const int i = j * 64 + __builtin_ctzll(m);
                m &= m - 1;
                const TriL &t = lds[i];
                const int minx = t.bb0 & 0xffff, miny = t.bb0 >> 16;
                const int maxx = t.bb1 & 0xffff, maxy = t.bb1 >> 16;
#pragma unroll
                for (int k = 0; k < NPX; ++k) {
                    if (act[k] && px[k] >= minx && px[k] <= maxx && py[k] >= miny && py[k] <= maxy) {
                        float b0, b1, b2;
                        if (cover(t, px[k], py[k], b0, b1, b2)) {
                            // Dot(bary, depths) and the test of forkergl.cpp:180-199
                            float dd = 0.f;
                            dd += b0 * t.z0;
                            dd += b1 * t.z1;
                            dd += b2 * t.z2;
                            if (!(dd >= best[k])) {
                                best[k] = dd;
                                win[k] = t.g;
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < NPX; ++k) {
        if (!act[k]) continue;
        const size_t i = (size_t)px[k] + (size_t)py[k] * a.W;
        if (win[k] >= 0 || a.fresh) a.depth[i] = best[k];
        if (!SHADOW) {
            a.vis[i] = win[k];
        } else if (win[k] >= 0) {
            // DepthShader::ProcessFragment (depthshader.h:30-36)
            const int4 tv = a.tri_v[win[k]];
            const DevDraw &dr = a.draws[tv.w];
            TriL t;
            stage_tri(t, win[k], a.screen[tv.x], a.screen[tv.y], a.screen[tv.z], a.W, a.H);
            float b0 = 0.f, b1 = 0.f, b2 = 0.f;
            cover(t, px[k], py[k], b0, b1, b2);
            const int f = win[k] - dr.tri_base;
            Vec4 n0 = depth_ndc(dr, dr.vi[3 * f]);
            Vec4 n1 = depth_ndc(dr, dr.vi[3 * f + 1]);
            Vec4 n2 = depth_ndc(dr, dr.vi[3 * f + 2]);
            float z = dot3(vec3(n0.z, n1.z, n2.z), vec3(b0, b1, b2));
            a.shadow_out[i] = z * 0.5f + 0.5f;
        } else if (a.fresh) {
            a.shadow_out[i] = 0.f;
        }
    }
}

// Barycentric weights of the winning triangle at (px,py), recomputed with the
// same staged arithmetic as the coverage loop.
__device__ __forceinline__ Vec3 winner_bary(const ResolveArgs &a, int g, int px, int py, int4 &tv)
{
    tv = a.tri_v[g];
    TriL t;
    stage_tri(t, g, a.screen[tv.x], a.screen[tv.y], a.screen[tv.z], a.W, a.H);
    float b0 = 0.f, b1 = 0.f, b2 = 0.f;
    cover(t, px, py, b0, b1, b2);
    return vec3(b0, b1, b2);
}

// Geometry pass resolve: GShader::ProcessFragment + the seven buffer writes
// (gshader.h:95-202, forkergl.cpp:211-223) for the visible fragment.
__global__ __launch_bounds__(256) void k_gbuffer_resolve(ResolveArgs a)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= a.W || y >= a.H) return;
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
    const Vec3 b = winner_bary(a, g, x, y, tv);
    const DevDraw &dr = a.draws[tv.w];
    const int f = g - dr.tri_base;
    const FrameUniforms &u = a.u;
    Varying v0 = main_vertex(dr, u, dr.vi[3 * f], dr.ti[3 * f], dr.ni[3 * f]);
    Varying v1 = main_vertex(dr, u, dr.vi[3 * f + 1], dr.ti[3 * f + 1], dr.ni[3 * f + 1]);
    Varying v2 = main_vertex(dr, u, dr.vi[3 * f + 2], dr.ti[3 * f + 2], dr.ni[3 * f + 2]);
    const DevMaterial &m = a.materials[dr.face_mat[f]];
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

// Forward pass resolve: BlinnPhongShader / PBRShader::ProcessFragment
// (phongshader.h:90-169, pbrshader.h:90-180) once for the visible fragment.
__global__ __launch_bounds__(256) void k_forward_resolve(ResolveArgs a)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= a.W || y >= a.H) return;
    const size_t i = (size_t)x + (size_t)y * a.W;
    const size_t P = (size_t)a.W * a.H;
    const int g = a.vis[i];
    if (g < 0) {
        if (a.fresh) {
            a.fb[i] = a.clear[0];
            a.fb[P + i] = a.clear[1];
            a.fb[2 * P + i] = a.clear[2];
        }
        return;
    }
    int4 tv;
    const Vec3 b = winner_bary(a, g, x, y, tv);
    const DevDraw &dr = a.draws[tv.w];
    const int f = g - dr.tri_base;
    const FrameUniforms &u = a.u;
    Varying v0 = main_vertex(dr, u, dr.vi[3 * f], dr.ti[3 * f], dr.ni[3 * f]);
    Varying v1 = main_vertex(dr, u, dr.vi[3 * f + 1], dr.ti[3 * f + 1], dr.ni[3 * f + 1]);
    Varying v2 = main_vertex(dr, u, dr.vi[3 * f + 2], dr.ti[3 * f + 2], dr.ni[3 * f + 2]);
    const DevMaterial &m = a.materials[dr.face_mat[f]];
    FragGeom fg = frag_geometry(dr, u, a.textures, v0, v1, v2, b, m.map_bump);
    Vec3 c;
    if (!dr.support_pbr) {
        Vec3 dc = m.map_kd >= 0 ? tex_sample(a.textures, m.map_kd, fg.uv)
                                : vec3(m.kd[0], m.kd[1], m.kd[2]);
        Vec3 em = m.map_ke >= 0 ? tex_sample(a.textures, m.map_ke, fg.uv)
                                : vec3(m.ke[0], m.ke[1], m.ke[2]);
        float shin = 1.f;
        if (m.map_ks >= 0) shin = tex_samplef(a.textures, m.map_ks, fg.uv) + 5;
        c = shade_point(u, a.shadow_map, false, false, fg.pos, fg.normal, fg.lndc, dc, em,
                        vec3(m.ka[0], m.ks[0], shin));
    } else {
        Vec3 al = m.map_base >= 0 ? tex_sample(a.textures, m.map_base, fg.uv)
                                  : vec3(m.albedo[0], m.albedo[1], m.albedo[2]);
        Vec3 em = m.map_pbr_ke >= 0 ? tex_sample(a.textures, m.map_pbr_ke, fg.uv)
                                    : vec3(m.pbr_ke[0], m.pbr_ke[1], m.pbr_ke[2]);
        float rough = m.map_rough >= 0 ? tex_samplef(a.textures, m.map_rough, fg.uv) : m.roughness;
        float metal = m.map_metal >= 0 ? tex_samplef(a.textures, m.map_metal, fg.uv) : m.metalness;
        float ao = m.map_ao >= 0 ? tex_samplef(a.textures, m.map_ao, fg.uv) : 1.f;
        c = shade_point(u, a.shadow_map, true, false, fg.pos, fg.normal, fg.lndc, al, em,
                        vec3(ao, metal, rough));
    }
    a.fb[i] = c.x;
    a.fb[P + i] = c.y;
    a.fb[2 * P + i] = c.z;
}

// ---------------- launch wrappers ----------------

static inline unsigned blocks_for(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

void launch_vertex(bool shadow, const DevDraw *draws, int n_draws, int max_pos,
                   const FrameUniforms &u, int4 *out, hipStream_t s)
{
    if (n_draws <= 0 || max_pos <= 0) return;
    dim3 grid(blocks_for(max_pos, 256), n_draws);
    if (shadow) hipLaunchKernelGGL(k_vertex<true>, grid, dim3(256), 0, s, draws, u, out);
    else hipLaunchKernelGGL(k_vertex<false>, grid, dim3(256), 0, s, draws, u, out);
}

void launch_setup(const DevDraw *draws, int n_draws, int max_faces, const int4 *screen, int W,
                  int H, int TS, uint2 *rect, uint32_t *count, int4 *tri_v, uint32_t *big_list,
                  uint32_t *big_count, hipStream_t s)
{
    if (n_draws <= 0 || max_faces <= 0) return;
    hipLaunchKernelGGL(k_setup, dim3(blocks_for(max_faces, 256), n_draws), dim3(256), 0, s, draws,
                       screen, W, H, TS, rect, count, tri_v, big_list, big_count);
}

void launch_emit(const uint32_t *off, int n_tri, const uint2 *rect, int tiles_x,
                 const uint32_t *big_list, uint32_t n_big, uint32_t *keys, uint32_t *vals,
                 hipStream_t s)
{
    if (n_tri <= 0) return;
    hipLaunchKernelGGL(k_emit_small, dim3(blocks_for(n_tri, 256)), dim3(256), 0, s, off, n_tri,
                       rect, tiles_x, keys, vals);
    if (n_big)
        hipLaunchKernelGGL(k_emit_big, dim3(n_big), dim3(256), 0, s, off, rect, tiles_x, big_list,
                           keys, vals);
}

void launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
}

void launch_fill_f32(float *p, float v, size_t n, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_f32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
}

void launch_tile_ranges(const uint32_t *keys, uint32_t n, uint32_t *start, uint32_t *end,
                        hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_tile_ranges, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, start,
                       end);
}

template <bool SHADOW>
static void launch_raster_t(int TS, const RasterArgs &a, hipStream_t s)
{
    const unsigned nt = (unsigned)(a.tiles_x * a.tiles_y);
    if (TS == 8) hipLaunchKernelGGL((k_raster<8, 64, SHADOW>), dim3(nt), dim3(64), 0, s, a);
    else if (TS == 16) hipLaunchKernelGGL((k_raster<16, 256, SHADOW>), dim3(nt), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_raster<32, 256, SHADOW>), dim3(nt), dim3(256), 0, s, a);
}

void launch_raster(bool shadow, int TS, const RasterArgs &a, hipStream_t s)
{
    if (a.tiles_x * a.tiles_y == 0) return;
    if (shadow) launch_raster_t<true>(TS, a, s);
    else launch_raster_t<false>(TS, a, s);
}

static inline dim3 grid2(int W, int H) { return dim3((W + 63) / 64, (H + 3) / 4); }

void launch_gbuffer_resolve(const ResolveArgs &a, hipStream_t s)
{
    if (a.W <= 0 || a.H <= 0) return;
    hipLaunchKernelGGL(k_gbuffer_resolve, grid2(a.W, a.H), dim3(256), 0, s, a);
}

void launch_forward_resolve(const ResolveArgs &a, hipStream_t s)
{
    if (a.W <= 0 || a.H <= 0) return;
    hipLaunchKernelGGL(k_forward_resolve, grid2(a.W, a.H), dim3(256), 0, s, a);
}

}  // namespace fgl
