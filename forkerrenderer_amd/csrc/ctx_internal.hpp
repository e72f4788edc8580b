// ctx_internal.hpp -- library-internal view of a context, for subsystems
// built on top of the raster/shading path (neural AO, batched views).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/fgl.h"

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
int ctx_view(fgl_ctx *c, CtxView *v);
// Marks the AO target as written in full (no pending clear).
void ctx_ao_written(fgl_ctx *c);
int ctx_fail(fgl_ctx *c, int code, const char *msg);

// Scoped HIP-event timing of the kernels enqueued on the context stream
// while it lives, reported by fgl_read_timings under `name` (a string
// literal) when timing is enabled.
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
