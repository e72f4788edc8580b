// ctx_internal.hpp -- library-internal view of a context, for subsystems
// built on top of the raster/shading path (neural AO, batched views).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/fgl.h"
#include "fgl_kernels.hpp"

namespace fgl {

struct CtxView {
    hipStream_t stream;
    int W, H;
    float *depth;   // W*H
    float *gbuf;    // 19 planes
    float *ao;      // W*H
    float *fb;      // 3 planes
    float view[16];
    float proj[16];
    float viewport[16];
};

// Fills v from the context; flushes pending draws and materialises pending
// clears of depth, G-buffer and AO first.
int ctx_view(fgl_ctx *c, CtxView *v, bool need_ao = true);
// Marks the AO target as written in full (no pending clear).
void ctx_ao_written(fgl_ctx *c);
int ctx_fail(fgl_ctx *c, int code, const char *msg);

// FrameUniforms from the context's current ForkerGL state (what the next
// pass would use).
void ctx_uniforms(fgl_ctx *c, FrameUniforms *u);
// Draw record of `model` under model matrix M with the context's current
// light-space matrix: geometry pointers, normal matrix, light * model.  The
// pass-wide bases (vert_base, tri_base) and the view index are left zero.
int ctx_draw_record(fgl_ctx *c, int model, const float M[16], DevDraw *out);
// Uploads pending texture / material tables; returns their device copies.
int ctx_tables(fgl_ctx *c, const DevMaterial **materials, const DevTexture **textures);
hipStream_t ctx_stream(fgl_ctx *c);
bool ctx_timing(fgl_ctx *c);

// Per-view pixel layout of a batch of views, for the batched per-pixel
// kernels outside the batch engine (neural AO): the view's frame targets
// start at pixel px_off of the batch buffers and its 64 x 4 pixel blocks are
// blocks blk_base.. of the batch-wide frame block table; its lW x lH network
// grid (the frame divided by the network's scale, rounded up) starts at
// lo_off of the network buffers, and its 64 x 4 network-pixel blocks are
// blocks lblk_base.. of the network block table.
struct BatchPixView {
    int W, H, lW, lH;
    uint32_t blk_base, lblk_base;
    size_t px_off, lo_off;
    float view[16];
};

// Neural AO of every view of a batch (nao.hip): down pass from the batch
// depth / G-buffer, two fused kernels for the four convolutions and the head, up pass
// into the batch AO target.  nao_reserve allocates the batch-sized buffers
// beforehand (the infer call then allocates nothing, so it can be captured
// in a graph).
struct NaoBatchTables {
    const BatchPixView *views;
    const uint8_t *blk_view;     // view of each 64 x 4 frame-pixel block
    uint32_t n_blocks;
    const uint8_t *lblk_view;    // view of each 64 x 4 network-pixel block
    uint32_t n_lblocks;
    size_t total_px, total_lpx;  // frame and network pixels of all views
    const uint2 *items[2];       // tiles of the fused layers 1-2 and 3-4 (nao_conv_items)
    uint32_t n_items[2];
};
int nao_reserve(fgl_nao *n, size_t frame_px, size_t net_px);
// The up pass of a network AO still on its grid into a W x H AO plane.
void nao_launch_up(const NetAo &net, int W, int H, float *ao, hipStream_t s);
// A network whose context is being destroyed: it stops using the context.
void nao_detach(void *n);
// The context's AO target is `net` (the lighting pass reads it from its grid;
// reading or replacing the target first runs the up pass).  owner: the
// network whose buffers hold it.
void ctx_ao_from_network(fgl_ctx *c, const NetAo &net, const void *owner);
// Forget / write out a pending network AO held by `owner`'s buffers (before
// they are overwritten / moved or freed).
void ctx_drop_network_ao(fgl_ctx *c, const void *owner);
int ctx_settle_network_ao(fgl_ctx *c, const void *owner);
void ctx_register_nao(fgl_ctx *c, void *n);
void ctx_unregister_nao(fgl_ctx *c, void *n);
// Frame pixels per network pixel, each way (1, 2 or 4).
int nao_scale(const fgl_nao *n);
// Changes whenever nao_reserve moves the inference buffers (a graph that
// captured their addresses must then be captured again).
uint64_t nao_buffers_generation(const fgl_nao *n);
// Appends the tiles of a view's lW x lH network grid to the lists of the two
// fused inference kernels (layers 1-2, layers 3-4 with the head).
void nao_conv_items(int view, int lW, int lH, std::vector<uint2> items[2]);
int nao_infer_batch(fgl_nao *n, hipStream_t s, const NaoBatchTables &t, const float *depth, const float *gbuf,
                    float *ao);

// Scoped HIP-event timing of the kernels enqueued on the context stream
// while it lives, reported by fgl_read_timings under `name` (a string
// literal) when timing is enabled; a null name times nothing.
class CtxTimer {
public:
    CtxTimer(fgl_ctx *c, const char *name);
    ~CtxTimer();
    CtxTimer(const CtxTimer &) = delete;
    CtxTimer &operator=(const CtxTimer &) = delete;

private:
    void *impl_;
};

}  // namespace fgl
