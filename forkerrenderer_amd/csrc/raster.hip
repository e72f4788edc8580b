// raster.hip -- single-view kernels of the raster pipeline for gfx950:
// vertex transform, triangle setup/binning, the tile rasteriser and the
// per-pixel resolve passes.  Their bodies live in raster_core.hpp, shared
// with the batched multi-view kernels.
//
// Pipeline of one pass (all draws of the pass, in submission order):
//   k_vertex      one thread per (draw, position): window X, Y (int), z
//   k_setup       one thread per (draw, face): clamped bbox, degenerate and
//                 off-screen cull, screen-tile rectangle and count, and the
//                 triangle's pass-wide vertex ids
//   scan + emit   (tile, triangle) pairs in triangle order
//   radix sort by tile (sort.hip) -> per-tile triangle lists
//   k_tile_ranges per-tile [start, end) in the sorted list
//   k_raster      one workgroup per tile (raster_tile)
//   resolve       one thread per pixel: the pass's fragment program for the
//                 winning triangle (G-buffer or forward colour), coalesced
//                 SoA writes; the shadow pass resolves inside k_raster.
#include "raster_core.hpp"

namespace fgl {

// One thread per (draw, position); the first thread also zeroes *zero (the
// pass's big-triangle counter, which setup counts up next).
template <bool SHADOW>
__global__ __launch_bounds__(256) void k_vertex(const DevDraw *__restrict__ draws, FrameUniforms u,
                                                int4 *__restrict__ out, float *__restrict__ ndc_z,
                                                uint32_t *__restrict__ zero)
{
    const DevDraw &dr = draws[blockIdx.y];
    const int vi = blockIdx.x * blockDim.x + threadIdx.x;
    if (zero && (vi | blockIdx.y) == 0) *zero = 0u;
    if (vi >= dr.n_pos) return;
    if (SHADOW) {
        const Vec4 n = depth_ndc(dr, vi);
        out[dr.vert_base + vi] = to_screen(u.shadow_viewport, n);
        ndc_z[dr.vert_base + vi] = n.z;
    } else {
        out[dr.vert_base + vi] = to_screen(u.viewport, main_ndc(dr, u, vi));
    }
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
    setup_tri(dr, d, f, screen, W, H, TS, rect, count, tri_v, big_list, big_count);
}

// Pairs of triangle t at [off[t], off[t+1]), row-major over its tile
// rectangle; triangles with more than kSmallTriTiles tiles go to k_emit_big.
// The first n_clear threads also zero the per-tile ranges (range_a/range_b)
// that k_tile_ranges fills after the sort.
__global__ __launch_bounds__(256) void k_emit_small(const uint32_t *__restrict__ off, int n_tri,
                                                    const uint2 *__restrict__ rect, int tiles_x,
                                                    uint32_t *__restrict__ keys,
                                                    uint32_t *__restrict__ vals,
                                                    uint32_t *__restrict__ range_a,
                                                    uint32_t *__restrict__ range_b, int n_clear)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_clear) {
        range_a[t] = 0u;
        range_b[t] = 0u;
    }
    if (t >= n_tri) return;
    const uint32_t o = off[t], n = off[t + 1] - o;
    if (n == 0 || n > (uint32_t)kSmallTriTiles) return;
    const uint2 r = rect[t];
    for (uint32_t j = 0; j < n; ++j) {
        keys[o + j] = pair_key(r, j, tiles_x, 0u);
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
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        keys[o + j] = pair_key(r, j, tiles_x, 0u);
        vals[o + j] = t;
    }
}

// Publishes two device counters to host memory, then the sequence number
// that tells the host they are there.
__global__ __launch_bounds__(64) void k_publish2(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                                                 uint32_t seq, uint32_t *host)
{
    if (threadIdx.x != 0) return;
    const uint32_t va = *a, vb = *b;
    host[0] = va;
    host[1] = vb;
    __threadfence_system();
    host[2] = seq;
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

// Per-tile [start, end) of the sorted pair list.  The pair count is n, or
// min(n, *n_dev) when n_dev is given (capacity-bounded batched passes).
__global__ __launch_bounds__(256) void k_tile_ranges(const uint32_t *__restrict__ keys, uint32_t n,
                                                     const uint32_t *__restrict__ n_dev,
                                                     uint32_t *__restrict__ start,
                                                     uint32_t *__restrict__ end)
{
    if (n_dev) n = min(n, *n_dev);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) start[k] = i;
    if (i == n - 1 || keys[i + 1] != k) end[k] = i + 1;
}

// Six wavefronts per SIMD: the 16-pixel tiles' kernel fits 80 registers
// (from 88, five wavefronts) with an 8-byte spill, and the extra wavefront
// hides more of the LDS atomics' latency (main pass 0.325 -> 0.294 ms at
// 3840x2160); the 8-pixel tiles already fit seven, the 32-pixel ones stay
// at four (their LDS allows no more).
template <int TS, int BLOCK, bool SHADOW>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(TS == 32 ? 4 : 6))) void k_raster(RasterArgs a)
{
    __shared__ RasterLds<TS, BLOCK> lds;
    const int tile = xcd_remap(blockIdx.x, a.tiles_x * a.tiles_y);
    raster_tile<TS, BLOCK, SHADOW>(a, tile, lds);
}

__global__ __launch_bounds__(256) void k_gbuffer_resolve(ResolveArgs a, FrameUniforms u)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= a.W || y >= a.H) return;
    gbuffer_px(a, u, x, y);
}

// Forward pass resolve: BlinnPhongShader / PBRShader::ProcessFragment
// (phongshader.h:90-169, pbrshader.h:90-180) once for the visible fragment.
// The forward fragment program for the winner (face f of draw dr, material
// m, barycentrics b) at pixel i.
template <bool STREAM>
__device__ __forceinline__ void forward_shade(const ResolveArgs &a, const FrameUniforms &u, const DevDraw &dr,
                                              const DevMaterial &m, int f, const Vec3 &b, size_t i, size_t P)
{
    Varying v0 = main_vertex(dr, u, f, 0);
    Varying v1 = main_vertex(dr, u, f, 1);
    Varying v2 = main_vertex(dr, u, f, 2);
    FragGeom fg = frag_geometry(dr, u, a.textures, v0, v1, v2, b, m.map_bump);
    // the stochastic shadow filters: this fragment's samples in the stream
    const float *rdisk = nullptr;
    if (STREAM && a.ss.pairs) rdisk = a.ss.pairs + 2 * (a.ss.offs ? (size_t)a.ss.offs[i] : i * (size_t)a.ss.stride);
    Vec3 c;
    if (!dr.support_pbr) {
        Vec3 dc = m.map_kd >= 0 ? tex_sample(a.textures, m.map_kd, fg.uv)
                                : vec3(m.kd[0], m.kd[1], m.kd[2]);
        Vec3 em = m.map_ke >= 0 ? tex_sample(a.textures, m.map_ke, fg.uv)
                                : vec3(m.ke[0], m.ke[1], m.ke[2]);
        float shin = 1.f;
        if (m.map_ks >= 0) shin = tex_samplef(a.textures, m.map_ks, fg.uv) + 5;
        c = shade_point<STREAM>(u, a.shadow_map, false, false, fg.pos, fg.normal, fg.lndc, dc, em,
                        vec3(m.ka[0], m.ks[0], shin), rdisk);
    } else {
        Vec3 al = m.map_base >= 0 ? tex_sample(a.textures, m.map_base, fg.uv)
                                  : vec3(m.albedo[0], m.albedo[1], m.albedo[2]);
        Vec3 em = m.map_pbr_ke >= 0 ? tex_sample(a.textures, m.map_pbr_ke, fg.uv)
                                    : vec3(m.pbr_ke[0], m.pbr_ke[1], m.pbr_ke[2]);
        float rough = m.map_rough >= 0 ? tex_samplef(a.textures, m.map_rough, fg.uv) : m.roughness;
        float metal = m.map_metal >= 0 ? tex_samplef(a.textures, m.map_metal, fg.uv) : m.metalness;
        float ao = m.map_ao >= 0 ? tex_samplef(a.textures, m.map_ao, fg.uv) : 1.f;
        c = shade_point<STREAM>(u, a.shadow_map, true, false, fg.pos, fg.normal, fg.lndc, al, em,
                        vec3(ao, metal, rough), rdisk);
    }
    a.fb[i] = c.x;
    a.fb[P + i] = c.y;
    a.fb[2 * P + i] = c.z;
}

// Forward resolve, one thread per pixel; draw and material records through
// readfirstlane when the wavefront shares them (as in gbuffer_px).
template <bool STREAM>
__global__ __launch_bounds__(256) void k_forward_resolve(ResolveArgs a, FrameUniforms u)
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
    const Vec3 b = winner_bary(a.screen, a.tri_v, g, a.W, a.H, x, y, tv);
    const int d0 = __builtin_amdgcn_readfirstlane(tv.w);
    if (__ballot(tv.w != d0) == 0ull) {
        const DevDraw &dr = a.draws[d0];
        const int f = g - dr.tri_base;
        const int mi = dr.face_mat[f];
        const int m0 = __builtin_amdgcn_readfirstlane(mi);
        if (__ballot(mi != m0) == 0ull) {
            forward_shade<STREAM>(a, u, dr, a.materials[m0], f, b, i, P);
            return;
        }
    }
    const DevDraw &dr = a.draws[tv.w];
    const int f = g - dr.tri_base;
    forward_shade<STREAM>(a, u, dr, a.materials[dr.face_mat[f]], f, b, i, P);
}

// ---------------- the forward pass's fragment sequence ----------------
//
// The forward shaders run for every fragment that passes the depth test at
// the moment it is drawn (src/forkergl.cpp:169-204): triangles in
// submission order, each over its bounding box with px outer and py inner,
// a fragment passing when !(depth >= the depth buffer's value then).  With
// the stochastic shadow filters every such fragment draws from the random
// stream, so the samples a pixel's final colour uses depend on every passing
// fragment drawn before it, at any pixel.  k_frag_walk replays the depth
// test in that order inside a tile (the tile's triangles one after another
// in id order, a thread per pixel, the running depth in a register): the
// count walk counts each pixel's passing fragments; the emit walk writes
// them as (triangle id, q) from the pixel's offset in the exclusive scan of
// the counts over the pixels in column-major order (q = px * H + py).  A
// stable sort by triangle id then leaves them in the reference's order: by
// triangle, then px, then py.  Coverage and depth are the rasteriser's own
// (stage_tri, cover, tri_depth), so the last passing fragment of a pixel is
// the winner k_raster finds.
template <int TS, int BLOCK, bool EMIT>
__global__ __launch_bounds__(BLOCK) void k_frag_walk(FragWalkArgs a)
{
    constexpr int PPT = TS * TS / BLOCK;
    static_assert(PPT * BLOCK == TS * TS, "whole pixels per thread");
    __shared__ TriSoA<BLOCK> tri;
    const int tile = blockIdx.x;
    const int x0 = (tile % a.tiles_x) * TS, y0 = (tile / a.tiles_x) * TS;
    const int tid = threadIdx.x;
    int px[PPT], py[PPT];
    float run[PPT];
    uint32_t pos[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int p = tid + k * BLOCK;
        px[k] = x0 + p % TS;
        py[k] = y0 + p / TS;
        const bool in = px[k] < a.W && py[k] < a.H;
        run[k] = 3.40282347e+38f;
        if (in && !a.fresh) run[k] = a.depth[(size_t)px[k] + (size_t)py[k] * a.W];
        pos[k] = (EMIT && in) ? a.base[(size_t)px[k] * a.H + py[k]] : 0u;
        if (!in) px[k] = -1;   // inside no bounding box
    }
    const uint32_t start = a.tile_start[tile], end = a.tile_end[tile];
    for (uint32_t base = start; base < end; base += BLOCK) {
        const int n = imin(BLOCK, (int)(end - base));
        __syncthreads();   // the previous chunk fully walked
        if (tid < n) {
            const int g = (int)a.vals[base + tid];
            const int4 tv = a.tri_v[g];
            TriL t;
            stage_tri(t, g, a.screen[tv.x], a.screen[tv.y], a.screen[tv.z], a.W, a.H);
            tri.put(tid, t);
        }
        __syncthreads();
        for (int i = 0; i < n; ++i) {
            const int bb0 = tri.bb0[i], bb1 = tri.bb1[i];
            const int minx = bb0 & 0xffff, miny = bb0 >> 16, maxx = bb1 & 0xffff, maxy = bb1 >> 16;
            bool hit[PPT];
            bool any = false;
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                hit[k] = px[k] >= minx && px[k] <= maxx && py[k] >= miny && py[k] <= maxy;
                any |= hit[k];
            }
            if (!any) continue;
            const TriL t = tri.get(i);
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                float b0, b1, b2;
                if (!hit[k] || !cover(t, px[k], py[k], b0, b1, b2)) continue;
                const float dd = tri_depth(t, b0, b1, b2);
                if (dd >= run[k]) continue;
                run[k] = dd;
                if (EMIT) {
                    a.tri_out[pos[k]] = (uint32_t)t.g;
                    a.q_out[pos[k]] = (uint32_t)px[k] * (uint32_t)a.H + (uint32_t)py[k];
                }
                ++pos[k];
            }
        }
    }
    if (!EMIT) {
#pragma unroll
        for (int k = 0; k < PPT; ++k)
            if (px[k] >= 0) a.count[(size_t)px[k] * a.H + py[k]] = pos[k];
    }
}

// Light 0's shadow inputs of passing fragment k, as forward_shade computes
// them: shading normal, world position, light NDC (planes of n floats).
__global__ __launch_bounds__(256) void k_frag_shadow_inputs(ResolveArgs a, FrameUniforms u,
                                                            const uint32_t *__restrict__ tri,
                                                            const uint32_t *__restrict__ q, uint32_t n,
                                                            float *__restrict__ out)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int g = (int)tri[k];
    const uint32_t qq = q[k];
    const int px = (int)(qq / (uint32_t)a.H), py = (int)(qq - (uint32_t)px * (uint32_t)a.H);
    int4 tv;
    const Vec3 b = winner_bary(a.screen, a.tri_v, g, a.W, a.H, px, py, tv);
    const DevDraw &dr = a.draws[tv.w];
    const int f = g - dr.tri_base;
    const DevMaterial &m = a.materials[dr.face_mat[f]];
    Varying v0 = main_vertex(dr, u, f, 0);
    Varying v1 = main_vertex(dr, u, f, 1);
    Varying v2 = main_vertex(dr, u, f, 2);
    const FragGeom fg = frag_geometry(dr, u, a.textures, v0, v1, v2, b, m.map_bump);
    const size_t N = n;
    out[k] = fg.normal.x;
    out[N + k] = fg.normal.y;
    out[2 * N + k] = fg.normal.z;
    out[3 * N + k] = fg.pos.x;
    out[4 * N + k] = fg.pos.y;
    out[5 * N + k] = fg.pos.z;
    out[6 * N + k] = fg.lndc.x;
    out[7 * N + k] = fg.lndc.y;
    out[8 * N + k] = fg.lndc.z;
}

// A pixel's triangles pass at most once each, so exactly one passing
// fragment of a pixel with a winner carries the winner's id.
__global__ __launch_bounds__(256) void k_frag_pixel_offsets(const uint32_t *__restrict__ tri,
                                                            const uint32_t *__restrict__ q, uint32_t n, int W,
                                                            int H, const int *__restrict__ vis,
                                                            const uint32_t *__restrict__ offs, uint32_t stride,
                                                            uint32_t *__restrict__ offs_px)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint32_t qq = q[k];
    const uint32_t px = qq / (uint32_t)H, py = qq - px * (uint32_t)H;
    const size_t p = (size_t)px + (size_t)py * W;
    if (vis[p] == (int)tri[k]) offs_px[p] = offs ? offs[k] : k * stride;
}

// ---------------- launch wrappers ----------------

static inline unsigned blocks_for(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

void launch_vertex(bool shadow, const DevDraw *draws, int n_draws, int max_pos,
                   const FrameUniforms &u, int4 *out, float *ndc_z, uint32_t *zero, hipStream_t s)
{
    if (n_draws <= 0 || max_pos <= 0) {
        if (zero) launch_fill_u32(zero, 0u, 1, s);
        return;
    }
    dim3 grid(blocks_for(max_pos, 256), n_draws);
    if (shadow) hipLaunchKernelGGL(k_vertex<true>, grid, dim3(256), 0, s, draws, u, out, ndc_z, zero);
    else hipLaunchKernelGGL(k_vertex<false>, grid, dim3(256), 0, s, draws, u, out, ndc_z, zero);
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
                 uint32_t *range_a, uint32_t *range_b, int n_clear, hipStream_t s)
{
    const int n = n_tri > n_clear ? n_tri : n_clear;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_emit_small, dim3(blocks_for(n, 256)), dim3(256), 0, s, off, n_tri,
                       rect, tiles_x, keys, vals, range_a, range_b, n_clear);
    if (n_big)
        hipLaunchKernelGGL(k_emit_big, dim3(n_big), dim3(256), 0, s, off, rect, tiles_x, big_list,
                           keys, vals);
}

void launch_publish2(const uint32_t *a, const uint32_t *b, uint32_t seq, uint32_t *host, hipStream_t s)
{
    hipLaunchKernelGGL(k_publish2, dim3(1), dim3(64), 0, s, a, b, seq, host);
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

void launch_tile_ranges(const uint32_t *keys, uint32_t n, const uint32_t *n_dev, uint32_t *start,
                        uint32_t *end, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_tile_ranges, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, n_dev, start,
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

void launch_gbuffer_resolve(const ResolveArgs &a, const FrameUniforms &u, hipStream_t s)
{
    if (a.W <= 0 || a.H <= 0) return;
    hipLaunchKernelGGL(k_gbuffer_resolve, grid2(a.W, a.H), dim3(256), 0, s, a, u);
}

void launch_forward_resolve(const ResolveArgs &a, const FrameUniforms &u, hipStream_t s)
{
    if (a.W <= 0 || a.H <= 0) return;
    if (a.ss.pairs) hipLaunchKernelGGL(k_forward_resolve<true>, grid2(a.W, a.H), dim3(256), 0, s, a, u);
    else hipLaunchKernelGGL(k_forward_resolve<false>, grid2(a.W, a.H), dim3(256), 0, s, a, u);
}

template <bool EMIT>
static void launch_frag_walk_t(int TS, const FragWalkArgs &a, hipStream_t s)
{
    const unsigned nt = (unsigned)(a.tiles_x * a.tiles_y);
    if (TS == 8) hipLaunchKernelGGL((k_frag_walk<8, 64, EMIT>), dim3(nt), dim3(64), 0, s, a);
    else if (TS == 16) hipLaunchKernelGGL((k_frag_walk<16, 256, EMIT>), dim3(nt), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_frag_walk<32, 256, EMIT>), dim3(nt), dim3(256), 0, s, a);
}

void launch_frag_walk(bool emit, int TS, const FragWalkArgs &a, hipStream_t s)
{
    if (a.tiles_x * a.tiles_y == 0) return;
    if (emit) launch_frag_walk_t<true>(TS, a, s);
    else launch_frag_walk_t<false>(TS, a, s);
}

void launch_frag_shadow_inputs(const ResolveArgs &a, const FrameUniforms &u, const uint32_t *tri,
                               const uint32_t *q, uint32_t n, float *planes, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_frag_shadow_inputs, dim3(blocks_for(n, 256)), dim3(256), 0, s, a, u, tri, q, n, planes);
}

void launch_frag_pixel_offsets(const uint32_t *tri, const uint32_t *q, uint32_t n, int W, int H, const int *vis,
                               const uint32_t *offs, uint32_t stride, uint32_t *offs_px, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_frag_pixel_offsets, dim3(blocks_for(n, 256)), dim3(256), 0, s, tri, q, n, W, H, vis,
                       offs, stride, offs_px);
}

}  // namespace fgl
