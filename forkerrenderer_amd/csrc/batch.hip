// batch.hip -- batched-view kernels for gfx950.  Each pass covers every view
// of a batch in one launch; per-view parameters come from the DevView table
// (indexed by the draw's view, by a (view, tile) queue item, or by the view
// of a 64 x 4 pixel block).  The per-vertex, per-triangle, per-tile and
// per-pixel bodies are the single-view ones (raster_core.hpp,
// shade_core.hpp), so a view's results are those of its single-view render.
//
// The pair count of a pass is known only on the device: the emit kernels
// drop pairs beyond the capacity, the sort and tile-range kernels read the
// count from device memory, and an overflow flag tells the host to re-render
// with more room.  Nothing here needs the host mid-pass, so the whole
// sequence can be captured in a HIP graph.
#include "batch.hpp"
#include "raster_core.hpp"
#include "shade_core.hpp"

namespace fgl {

// Vertex and setup grids are flat lists of 256-item blocks: block b covers
// items blocks[b].y .. + 255 of draw blocks[b].x (draws differ in size by
// orders of magnitude in a batch, so a draws x largest-draw grid would be
// mostly empty).
template <int PASS>
__global__ __launch_bounds__(256) void k_bv_vertex(const DevDraw *__restrict__ draws,
                                                   const DevView *__restrict__ views,
                                                   const uint2 *__restrict__ blocks,
                                                   int4 *__restrict__ out, float *__restrict__ ndc_z)
{
    const uint2 bk = blocks[blockIdx.x];
    const DevDraw &dr = draws[bk.x];
    const int vi = (int)bk.y + threadIdx.x;
    if (vi >= dr.n_pos) return;
    const FrameUniforms &u = views[dr.view].u;
    if (PASS == kBvShadow) {
        const Vec4 n = depth_ndc(dr, vi);
        out[dr.vert_base + vi] = to_screen(u.shadow_viewport, n);
        ndc_z[dr.vert_base + vi] = n.z;
    } else {
        out[dr.vert_base + vi] = to_screen(u.viewport, main_ndc(dr, u, vi));
    }
}

template <int PASS>
__global__ __launch_bounds__(256) void k_bv_setup(const DevDraw *__restrict__ draws,
                                                  const DevView *__restrict__ views,
                                                  const uint2 *__restrict__ blocks,
                                                  const int4 *__restrict__ screen, int TS,
                                                  uint2 *__restrict__ rect,
                                                  uint32_t *__restrict__ count,
                                                  int4 *__restrict__ tri_v,
                                                  uint32_t *__restrict__ big_list,
                                                  uint32_t *__restrict__ big_count)
{
    const uint2 bk = blocks[blockIdx.x];
    const int d = (int)bk.x;
    const DevDraw &dr = draws[d];
    const int f = (int)bk.y + threadIdx.x;
    if (f >= dr.n_faces) return;
    const DevView &v = views[dr.view];
    const int W = PASS == kBvShadow ? v.S : v.W, H = PASS == kBvShadow ? v.S : v.H;
    setup_tri(dr, d, f, screen, W, H, TS, rect, count, tri_v, big_list, big_count);
}

template <int PASS>
__global__ __launch_bounds__(256) void k_bv_emit_small(BvBinArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_tri) return;
    const uint32_t o = a.off[t], n = a.off[t + 1] - o;
    if (n == 0 || n > (uint32_t)kSmallTriTiles) return;
    const DevView &v = a.views[a.draws[a.tri_v[t].w].view];
    const uint2 r = a.rect[t];
    for (uint32_t j = 0; j < n; ++j) {
        if (o + j >= a.cap) break;
        a.keys[o + j] = pair_key(r, j, v.tiles_x[PASS], v.tile_base[PASS]);
        a.vals[o + j] = (uint32_t)t;
    }
}

// Persistent: workgroups stride over the device-side big-triangle list.
template <int PASS>
__global__ __launch_bounds__(256) void k_bv_emit_big(BvBinArgs a)
{
    const uint32_t nbig = *a.big_count;
    for (uint32_t b = blockIdx.x; b < nbig; b += gridDim.x) {
        const uint32_t t = a.big_list[b];
        const uint32_t o = a.off[t], n = a.off[t + 1] - o;
        const DevView &v = a.views[a.draws[a.tri_v[t].w].view];
        const uint2 r = a.rect[t];
        for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
            if (o + j >= a.cap) break;
            a.keys[o + j] = pair_key(r, j, v.tiles_x[PASS], v.tile_base[PASS]);
            a.vals[o + j] = t;
        }
    }
}

__global__ void k_bv_overflow(const uint32_t *__restrict__ total, uint32_t cap, uint32_t *__restrict__ flag)
{
    if (threadIdx.x == 0) flag[0] = *total > cap ? 1u : 0u;
}

// Persistent rasteriser: each workgroup pulls (view, tile) items from the
// queue until it runs dry, and rasterises each as the single-view kernel
// rasterises that view's tile.
template <int TS, int BLOCK, bool SHADOW>
__global__ __launch_bounds__(BLOCK) void k_bv_raster(BvRasterArgs b)
{
    __shared__ RasterLds<TS, BLOCK> lds;
    __shared__ uint32_t s_item;
    constexpr int P = SHADOW ? kBvShadow : kBvMain;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(b.queue, 1u);
        __syncthreads();
        const uint32_t it = s_item;
        __syncthreads();
        if (it >= b.n_items) return;
        const uint32_t code = b.items[it];
        const DevView &v = b.views[code >> 24];
        const int tile = (int)(code & 0xffffffu);
        RasterArgs a;
        a.draws = b.draws;
        a.n_draws = 0;
        a.screen = b.screen;
        a.tri_v = b.tri_v;
        a.vals = b.vals;
        a.tile_start = b.tile_start + v.tile_base[P];
        a.tile_end = b.tile_end + v.tile_base[P];
        a.W = SHADOW ? v.S : v.W;
        a.H = SHADOW ? v.S : v.H;
        a.tiles_x = v.tiles_x[P];
        a.tiles_y = v.tiles_y[P];
        a.fresh = 1;
        a.depth = SHADOW ? b.shadow_depth + v.spx_off : b.depth + v.px_off;
        a.vis = b.vis + v.px_off;
        a.shadow_out = b.shadow + v.spx_off;
        a.ndc_z = b.ndc_z;
        raster_tile<TS, BLOCK, SHADOW>(a, tile, lds);
    }
}

// Pixel (x, y) of the view owning 64 x 4 block blockIdx.x; false when it is
// outside the view.
__device__ __forceinline__ bool bv_pixel(const BvPixelArgs &a, int &vi, int &x, int &y)
{
    vi = a.blk_view[blockIdx.x];
    const DevView &v = a.views[vi];
    const int lb = (int)(blockIdx.x - v.blk_base), nbx = (v.W + 63) / 64;
    x = (lb % nbx) * 64 + (threadIdx.x & 63);
    y = (lb / nbx) * 4 + (threadIdx.x >> 6);
    return x < v.W && y < v.H;
}

__global__ __launch_bounds__(256) void k_bv_gbuffer(BvPixelArgs b)
{
    int vi, x, y;
    if (!bv_pixel(b, vi, x, y)) return;
    const DevView &v = b.views[vi];
    ResolveArgs a;
    a.draws = b.draws;
    a.screen = b.screen;
    a.tri_v = b.tri_v;
    a.vis = b.vis + v.px_off;
    a.W = v.W;
    a.H = v.H;
    a.fresh = 1;
    a.clear[0] = a.clear[1] = a.clear[2] = 0.f;
    a.gbuf = b.gbuf + kGbufPlanes * v.px_off;
    a.fb = nullptr;
    a.materials = b.materials;
    a.textures = b.textures;
    a.shadow_map = nullptr;
    a.ss = ShadowSamples{nullptr, nullptr, 0u};
    gbuffer_px(a, v.u, x, y);
}

__global__ __launch_bounds__(256) void k_bv_lighting(BvPixelArgs b)
{
    int vi, x, y;
    if (!bv_pixel(b, vi, x, y)) return;
    const DevView &v = b.views[vi];
    const ShadowSamples none = {nullptr, nullptr, 0};
    lighting_px(v.u, b.gbuf + kGbufPlanes * v.px_off, b.ao + v.px_off, b.ao_is_one, b.shadow + v.spx_off, none,
                b.fb + 3 * v.px_off, x, y);
}

__global__ __launch_bounds__(256) void k_bv_resolve(BvPixelArgs b)
{
    int vi, x, y;
    if (!bv_pixel(b, vi, x, y)) return;
    const DevView &v = b.views[vi];
    resolve_px(b.fb + 3 * v.px_off, v.W, v.H, 1, b.img + 3 * v.px_off, x, y);
}

// ---------------- launch wrappers ----------------

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

void launch_bv_vertex(int pass, const DevDraw *draws, const DevView *views, const uint2 *blocks, int n_blocks,
                      int4 *screen, float *ndc_z, hipStream_t s)
{
    if (n_blocks <= 0) return;
    if (pass == kBvShadow)
        hipLaunchKernelGGL(k_bv_vertex<kBvShadow>, dim3(n_blocks), dim3(256), 0, s, draws, views, blocks, screen,
                           ndc_z);
    else
        hipLaunchKernelGGL(k_bv_vertex<kBvMain>, dim3(n_blocks), dim3(256), 0, s, draws, views, blocks, screen,
                           ndc_z);
}

void launch_bv_setup(int pass, const DevDraw *draws, const DevView *views, const uint2 *blocks, int n_blocks,
                     const int4 *screen, int TS, uint2 *rect, uint32_t *count, int4 *tri_v,
                     uint32_t *big_list, uint32_t *big_count, hipStream_t s)
{
    if (n_blocks <= 0) return;
    if (pass == kBvShadow)
        hipLaunchKernelGGL(k_bv_setup<kBvShadow>, dim3(n_blocks), dim3(256), 0, s, draws, views, blocks, screen, TS,
                           rect, count, tri_v, big_list, big_count);
    else
        hipLaunchKernelGGL(k_bv_setup<kBvMain>, dim3(n_blocks), dim3(256), 0, s, draws, views, blocks, screen, TS,
                           rect, count, tri_v, big_list, big_count);
}

void launch_bv_emit(int pass, const BvBinArgs &a, hipStream_t s)
{
    if (a.n_tri <= 0) return;
    // big triangles: a fixed grid strides over the device-side list
    const unsigned big_grid = 1024;
    if (pass == kBvShadow) {
        hipLaunchKernelGGL(k_bv_emit_small<kBvShadow>, dim3(nblk(a.n_tri, 256)), dim3(256), 0, s, a);
        hipLaunchKernelGGL(k_bv_emit_big<kBvShadow>, dim3(big_grid), dim3(256), 0, s, a);
    } else {
        hipLaunchKernelGGL(k_bv_emit_small<kBvMain>, dim3(nblk(a.n_tri, 256)), dim3(256), 0, s, a);
        hipLaunchKernelGGL(k_bv_emit_big<kBvMain>, dim3(big_grid), dim3(256), 0, s, a);
    }
}

void launch_bv_overflow(const uint32_t *total, uint32_t cap, uint32_t *flag, hipStream_t s)
{
    hipLaunchKernelGGL(k_bv_overflow, dim3(1), dim3(64), 0, s, total, cap, flag);
}

template <bool SHADOW>
static void bv_raster_t(int TS, unsigned n_wg, const BvRasterArgs &a, hipStream_t s)
{
    if (TS == 8) hipLaunchKernelGGL((k_bv_raster<8, 64, SHADOW>), dim3(n_wg), dim3(64), 0, s, a);
    else if (TS == 16) hipLaunchKernelGGL((k_bv_raster<16, 256, SHADOW>), dim3(n_wg), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_bv_raster<32, 256, SHADOW>), dim3(n_wg), dim3(256), 0, s, a);
}

void launch_bv_raster(int pass, int TS, unsigned n_wg, const BvRasterArgs &a, hipStream_t s)
{
    if (a.n_items == 0 || n_wg == 0) return;
    if (pass == kBvShadow) bv_raster_t<true>(TS, n_wg, a, s);
    else bv_raster_t<false>(TS, n_wg, a, s);
}

void launch_bv_gbuffer(const BvPixelArgs &a, hipStream_t s)
{
    if (a.n_blocks) hipLaunchKernelGGL(k_bv_gbuffer, dim3(a.n_blocks), dim3(256), 0, s, a);
}

void launch_bv_lighting(const BvPixelArgs &a, hipStream_t s)
{
    if (a.n_blocks) hipLaunchKernelGGL(k_bv_lighting, dim3(a.n_blocks), dim3(256), 0, s, a);
}

void launch_bv_resolve(const BvPixelArgs &a, hipStream_t s)
{
    if (a.n_blocks) hipLaunchKernelGGL(k_bv_resolve, dim3(a.n_blocks), dim3(256), 0, s, a);
}

}  // namespace fgl
