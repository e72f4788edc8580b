// batch.hpp -- device layout and launch interface of batched views: many
// deferred frames of different sizes and meshes rendered by one launch per
// pass (batch.hip kernels, batch.cpp host engine).
#pragma once

#include "fgl_kernels.hpp"

namespace fgl {

// Pass kinds of a batch: index into the per-pass fields below.
constexpr int kBvShadow = 0;
constexpr int kBvMain = 1;

// One view of a batch, in device memory.  Its targets are slices of the
// batch buffers: main targets start at pixel px_off (depth, vis, AO at
// px_off, G-buffer planes at 19 * px_off, frame planes at 3 * px_off, image
// bytes at 3 * px_off), shadow targets at spx_off.  Its tiles of pass p are
// global tiles tile_base[p] .. tile_base[p] + tiles_x[p] * tiles_y[p] - 1;
// its 64 x 4 pixel blocks are batch blocks blk_base ...
struct DevView {
    FrameUniforms u;   // ForkerGL state of the view (both passes)
    int W, H, S, shadow_on;
    int tiles_x[2], tiles_y[2];
    uint32_t tile_base[2];
    uint32_t blk_base, pad;
    size_t px_off, spx_off;
};

struct BvBinArgs {
    const DevDraw *draws;
    const DevView *views;
    const uint32_t *off;      // exclusive scan of per-triangle pair counts (total at off[n_tri])
    const uint2 *rect;
    const int4 *tri_v;
    const uint32_t *big_list;
    const uint32_t *big_count;
    int n_tri;
    uint32_t cap;             // pair capacity of keys / vals
    uint32_t *keys, *vals;
};

struct BvRasterArgs {
    const DevDraw *draws;
    const DevView *views;
    const int4 *screen;
    const int4 *tri_v;
    const uint32_t *vals;     // sorted pair values
    const uint32_t *tile_start, *tile_end;   // per global tile
    const uint32_t *items;    // (view << 24) | view-local tile
    uint32_t n_items;
    uint32_t *queue;          // next item (reset to 0 before the launch)
    float *depth, *shadow_depth, *shadow;
    int *vis;
    const float *ndc_z;       // shadow pass: per vertex NDC z
};

struct BvPixelArgs {
    const DevDraw *draws;
    const DevView *views;
    const uint8_t *blk_view;  // view of each 64 x 4 pixel block
    uint32_t n_blocks;
    const int4 *screen;
    const int4 *tri_v;
    const int *vis;
    const DevMaterial *materials;
    const DevTexture *textures;
    float *gbuf, *ao, *fb;
    const float *shadow;
    int ao_is_one;
    uint8_t *img;
};

// blocks: (draw, first vertex / face) of each 256-item block
void launch_bv_vertex(int pass, const DevDraw *draws, const DevView *views, const uint2 *blocks, int n_blocks,
                      int4 *screen, float *ndc_z, hipStream_t s);
void launch_bv_setup(int pass, const DevDraw *draws, const DevView *views, const uint2 *blocks, int n_blocks,
                     const int4 *screen, int TS, uint2 *rect, uint32_t *count, int4 *tri_v,
                     uint32_t *big_list, uint32_t *big_count, hipStream_t s);
void launch_bv_emit(int pass, const BvBinArgs &a, hipStream_t s);
// flag = (*total > cap): the pass's pairs did not fit
void launch_bv_overflow(const uint32_t *total, uint32_t cap, uint32_t *flag, hipStream_t s);
void launch_bv_raster(int pass, int TS, unsigned n_wg, const BvRasterArgs &a, hipStream_t s);
void launch_bv_gbuffer(const BvPixelArgs &a, hipStream_t s);
void launch_bv_lighting(const BvPixelArgs &a, hipStream_t s);
void launch_bv_resolve(const BvPixelArgs &a, hipStream_t s);

}  // namespace fgl
