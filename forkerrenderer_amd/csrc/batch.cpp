// batch.cpp -- views (one deferred frame of Render::Render without SSAA over
// uploaded model instances) and the batched-view engine: every view of a
// batch rendered by one launch per pass, the pass sequence captured in a
// HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fgl.h"
#include "batch.hpp"
#include "ctx_internal.hpp"
#include "host_io.hpp"

using namespace fgl;

namespace {

struct Buf {
    void *p = nullptr;
    size_t cap = 0;
};

int grow(Buf &b, size_t bytes)
{
    if (b.p && bytes <= b.cap) return FGL_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    size_t want = bytes < 256 ? 256 : bytes;
    if (hipMalloc(&b.p, want) != hipSuccess) {
        b.p = nullptr;
        return FGL_ERR_NOMEM;
    }
    b.cap = want;
    return FGL_OK;
}

void release(Buf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

int bits_for(uint32_t n)
{
    int b = 0;
    while (b < 32 && (n >> b) != 0) ++b;
    return b;
}

struct ViewCopy {
    fgl_view_desc d;
    std::vector<int32_t> models;
    std::vector<float> mats;
};

// The view's ForkerGL state as Render::Render sets it (render.cpp:33-38,
// 60-70, 106-115): viewports, lights, eye, shadow filter; light-space matrix
// when the view casts shadows; camera.
int apply_view_state(fgl_ctx *c, const fgl_view_desc &v, bool shadow_phase_only)
{
    int rc;
    const float ratio = (float)v.width / v.height;
    if ((rc = fgl_set_viewport(c, 0, 0, 0, v.width, v.height))) return rc;
    const int S = v.shadow_on ? v.shadow_size : 0;
    if ((rc = fgl_set_viewport(c, 1, 0, 0, S ? S : v.width, S ? S : v.height))) return rc;
    if ((rc = fgl_set_lights(c, v.n_lights, v.lights))) return rc;
    if ((rc = fgl_set_eye(c, v.eye))) return rc;
    if ((rc = fgl_set_shadow(c, v.shadow_on, &v.shadow))) return rc;
    const float up[3] = {0.f, 1.f, 0.f};
    if (v.shadow_on) {
        float V[16], P[16], L[16];
        const float origin[3] = {0.f, 0.f, 0.f};
        make_lookat(v.lights[0].vec, origin, up, V);
        make_ortho(-3.0f * ratio, 3.0f * ratio, -3.0f, 3.0f, 0.1f, 20.f, P);
        mat_mul(P, V, L);
        if ((rc = fgl_set_matrix(c, FGL_MAT_LIGHT_SPACE, L))) return rc;
    }
    if (shadow_phase_only) return FGL_OK;
    float V[16], P[16];
    make_lookat(v.eye, v.look_at, up, V);
    make_perspective(45.f, ratio, 0.01f, 20.f, P);
    if ((rc = fgl_set_matrix(c, FGL_MAT_VIEW, V))) return rc;
    return fgl_set_matrix(c, FGL_MAT_PROJ, P);
}

int check_view(fgl_ctx *c, const fgl_view_desc &v)
{
    if (v.width < 1 || v.height < 1 || v.width > 8192 || v.height > 8192)
        return ctx_fail(c, FGL_ERR_INVALID, "view size must be 1..8192");
    if (v.shadow_on && (v.shadow_size < 1 || v.shadow_size > 8192 || v.n_lights < 1))
        return ctx_fail(c, FGL_ERR_INVALID, "a shadowed view needs shadow_size 1..8192 and a light");
    if (v.shadow_on && v.shadow.mode >= FGL_SHADOW_REF_PCF)
        return ctx_fail(c, FGL_ERR_INVALID, "views take the deterministic shadow filters (modes 0-3)");
    if (v.n_draws < 0 || (v.n_draws && (!v.models || !v.matrices)))
        return ctx_fail(c, FGL_ERR_INVALID, "bad draw list");
    return FGL_OK;
}

}  // namespace

struct fgl_batch {
    fgl_ctx *ctx = nullptr;
    int n = 0;
    int ao_mode = FGL_AO_NONE;
    fgl_nao *nao = nullptr;
    std::vector<ViewCopy> views;
    int TS[2] = {16, 16};
    int occupancy = 4;
    int n_cu = 256;
    uint32_t cap_req[2] = {0, 0};   // requested capacities (0 = auto)
    uint32_t cap[2] = {0, 0};       // allocated capacities
    uint64_t last_pairs[2] = {0, 0};
    bool measured = false;
    bool use_graph = true;
    hipGraphExec_t exec = nullptr;

    // layout
    std::vector<DevView> hv;
    size_t total_px = 0, total_spx = 0;
    uint32_t n_blocks = 0;
    int n_draws[2] = {0, 0}, max_pos[2] = {0, 0}, max_faces[2] = {0, 0};
    long long nt[2] = {0, 0}, nv[2] = {0, 0};
    uint32_t tiles_total[2] = {0, 0};
    uint32_t n_items[2] = {0, 0};
    bool have_shadow = false;

    // device
    Buf d_views, d_draws[2], d_items[2], d_blk_view, d_lblk_view, d_pix_views, d_conv_items[2], d_vblocks[2],
        d_fblocks[2];
    size_t total_lpx = 0;      // network pixels of all views (neural AO)
    uint32_t n_lblocks = 0;    // 64 x 4 network-pixel blocks of all views
    int n_vblocks[2] = {0, 0}, n_fblocks[2] = {0, 0};
    uint32_t n_conv_items[2] = {0, 0};
    Buf depth, vis, gbuf, ao, fb, img, shadow, shadow_depth;
    Buf screen, ndc_z, rect, counts, offs, scan_tmp, tri_v, big_list, tile_start, tile_end;
    Buf keys[2], vals[2], keys_alt[2], vals_alt[2], sort_tmp[2];
    Buf counters;   // [0] big count, [1] queue, [2], [3] overflow flags (shadow, main)
    uint32_t *h_flags = nullptr;
    uint32_t *h_tot = nullptr;
    const DevMaterial *materials = nullptr;
    const DevTexture *textures = nullptr;
    uint64_t nao_gen = 0;   // generation of the neural-AO buffers the graph captured
};

static int bfail(fgl_batch *b, int code, const char *msg) { return ctx_fail(b->ctx, code, msg); }

static void drop_graph(fgl_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
}

// Per-view uniforms, draw tables, target layout, tile items, pixel blocks.
static int plan(fgl_batch *b)
{
    fgl_ctx *c = b->ctx;
    int rc;
    drop_graph(b);
    b->hv.assign(b->n, DevView());
    std::vector<DevDraw> dd[2];
    std::vector<uint32_t> items[2];
    std::vector<uint8_t> blk_view, lblk_view;
    std::vector<BatchPixView> pix(b->n);
    const int nscale = b->ao_mode == FGL_AO_NEURAL ? nao_scale(b->nao) : 1;
    size_t lpx = 0;
    uint32_t lblocks = 0;
    std::vector<uint2> conv_items[2];
    size_t px = 0, spx = 0;
    uint32_t tiles[2] = {0, 0}, blocks = 0;
    b->have_shadow = false;
    for (int p = 0; p < 2; ++p) b->nt[p] = b->nv[p] = b->max_pos[p] = b->max_faces[p] = 0;
    for (int i = 0; i < b->n; ++i) {
        const fgl_view_desc &v = b->views[i].d;
        DevView &dv = b->hv[i];
        std::memset(&dv, 0, sizeof(dv));
        if ((rc = apply_view_state(c, v, false))) return rc;
        ctx_uniforms(c, &dv.u);
        const int S = v.shadow_on ? v.shadow_size : 0;
        dv.u.W = v.width;
        dv.u.H = v.height;
        dv.u.SW = S;
        dv.u.SH = S;
        dv.W = v.width;
        dv.H = v.height;
        dv.S = S;
        dv.shadow_on = v.shadow_on ? 1 : 0;
        dv.px_off = px;
        dv.spx_off = spx;
        px += (size_t)v.width * v.height;
        spx += (size_t)S * S;
        for (int p = 0; p < 2; ++p) {
            const int W = p == kBvShadow ? S : v.width, H = p == kBvShadow ? S : v.height;
            if (p == kBvShadow && !v.shadow_on) continue;
            dv.tiles_x[p] = (W + b->TS[p] - 1) / b->TS[p];
            dv.tiles_y[p] = (H + b->TS[p] - 1) / b->TS[p];
            dv.tile_base[p] = tiles[p];
            const uint32_t nt = (uint32_t)(dv.tiles_x[p] * dv.tiles_y[p]);
            for (uint32_t t = 0; t < nt; ++t) items[p].push_back(((uint32_t)i << 24) | t);
            tiles[p] += nt;
        }
        dv.blk_base = blocks;
        const uint32_t nb = (uint32_t)(((v.width + 63) / 64) * ((v.height + 3) / 4));
        blk_view.insert(blk_view.end(), nb, (uint8_t)i);
        blocks += nb;
        pix[i].W = v.width;
        pix[i].H = v.height;
        pix[i].lW = (v.width + nscale - 1) / nscale;
        pix[i].lH = (v.height + nscale - 1) / nscale;
        pix[i].blk_base = dv.blk_base;
        pix[i].lblk_base = lblocks;
        pix[i].px_off = dv.px_off;
        pix[i].lo_off = lpx;
        std::memcpy(pix[i].view, dv.u.view, sizeof(pix[i].view));
        {
            const uint32_t nl = (uint32_t)(((pix[i].lW + 63) / 64) * ((pix[i].lH + 3) / 4));
            lblk_view.insert(lblk_view.end(), nl, (uint8_t)i);
            lblocks += nl;
            lpx += (size_t)pix[i].lW * pix[i].lH;
        }
        if (b->ao_mode == FGL_AO_NEURAL) nao_conv_items(i, pix[i].lW, pix[i].lH, conv_items);
        // draw records: light * model with this view's light space
        for (int k = 0; k < v.n_draws; ++k) {
            DevDraw d;
            if ((rc = ctx_draw_record(c, b->views[i].models[k], &b->views[i].mats[16 * k], &d))) return rc;
            d.view = i;
            for (int p = 0; p < 2; ++p) {
                if (p == kBvShadow && !v.shadow_on) continue;
                DevDraw e = d;
                e.vert_base = (int)b->nv[p];
                e.tri_base = (int)b->nt[p];
                b->nv[p] += d.n_pos;
                b->nt[p] += d.n_faces;
                b->max_pos[p] = std::max(b->max_pos[p], d.n_pos);
                b->max_faces[p] = std::max(b->max_faces[p], d.n_faces);
                dd[p].push_back(e);
            }
        }
        if (v.shadow_on) b->have_shadow = true;
    }
    for (int p = 0; p < 2; ++p) {
        if (b->nt[p] > 0x7fffffffLL || b->nv[p] > 0x7fffffffLL)
            return bfail(b, FGL_ERR_CAPACITY, "batch pass too large");
        b->n_draws[p] = (int)dd[p].size();
        b->tiles_total[p] = tiles[p];
        b->n_items[p] = (uint32_t)items[p].size();
        if (tiles[p] > (1u << 30)) return bfail(b, FGL_ERR_CAPACITY, "too many tiles");
    }
    b->total_px = px;
    b->total_spx = spx;
    b->n_blocks = blocks;
    b->total_lpx = lpx;
    b->n_lblocks = lblocks;
    if ((rc = ctx_tables(c, &b->materials, &b->textures))) return rc;
    const long long mt = std::max(b->nt[0], b->nt[1]), mv = std::max(b->nv[0], b->nv[1]);
    const uint32_t mtiles = std::max(tiles[0], tiles[1]);
    if (grow(b->d_views, b->n * sizeof(DevView)) || grow(b->d_blk_view, blk_view.size() + 1) ||
        grow(b->d_lblk_view, lblk_view.size() + 1) ||
        grow(b->d_pix_views, b->n * sizeof(BatchPixView)) || grow(b->depth, px * 4) || grow(b->vis, px * 4) ||
        grow(b->gbuf, (size_t)kGbufPlanes * px * 4) || grow(b->ao, px * 4) || grow(b->fb, 3 * px * 4) ||
        grow(b->img, 3 * px) || grow(b->shadow, spx * 4) || grow(b->shadow_depth, spx * 4) ||
        grow(b->screen, (size_t)mv * sizeof(int4)) || grow(b->ndc_z, (size_t)mv * 4) ||
        grow(b->rect, (size_t)mt * sizeof(uint2)) ||
        grow(b->counts, (size_t)mt * 4) || grow(b->offs, ((size_t)mt + 1) * 4) ||
        grow(b->scan_tmp, scan_tmp_words((size_t)mt) * 4) || grow(b->tri_v, (size_t)mt * sizeof(int4)) ||
        grow(b->big_list, (size_t)mt * 4) || grow(b->tile_start, (size_t)mtiles * 4) ||
        grow(b->tile_end, (size_t)mtiles * 4) || grow(b->counters, 256))
        return bfail(b, FGL_ERR_NOMEM, "batch allocation failed");
    for (int p = 0; p < 2; ++p) {
        std::vector<uint2> vb, fb;
        for (size_t k = 0; k < dd[p].size(); ++k) {
            for (int v0 = 0; v0 < dd[p][k].n_pos; v0 += 256) vb.push_back(make_uint2((uint32_t)k, (uint32_t)v0));
            for (int f0 = 0; f0 < dd[p][k].n_faces; f0 += 256) fb.push_back(make_uint2((uint32_t)k, (uint32_t)f0));
        }
        b->n_vblocks[p] = (int)vb.size();
        b->n_fblocks[p] = (int)fb.size();
        if (grow(b->d_vblocks[p], (vb.size() + 1) * sizeof(uint2)) || grow(b->d_fblocks[p], (fb.size() + 1) * sizeof(uint2)))
            return bfail(b, FGL_ERR_NOMEM, "batch allocation failed");
        if ((!vb.empty() && hipMemcpy(b->d_vblocks[p].p, vb.data(), vb.size() * sizeof(uint2), hipMemcpyHostToDevice)) ||
            (!fb.empty() && hipMemcpy(b->d_fblocks[p].p, fb.data(), fb.size() * sizeof(uint2), hipMemcpyHostToDevice)))
            return bfail(b, FGL_ERR_DEVICE, "upload failed");
        if (grow(b->d_draws[p], (dd[p].size() + 1) * sizeof(DevDraw)) ||
            grow(b->d_items[p], (items[p].size() + 1) * 4))
            return bfail(b, FGL_ERR_NOMEM, "batch allocation failed");
        if (!dd[p].empty() &&
            hipMemcpy(b->d_draws[p].p, dd[p].data(), dd[p].size() * sizeof(DevDraw), hipMemcpyHostToDevice))
            return bfail(b, FGL_ERR_DEVICE, "upload failed");
        if (!items[p].empty() &&
            hipMemcpy(b->d_items[p].p, items[p].data(), items[p].size() * 4, hipMemcpyHostToDevice))
            return bfail(b, FGL_ERR_DEVICE, "upload failed");
    }
    if (hipMemcpy(b->d_views.p, b->hv.data(), b->n * sizeof(DevView), hipMemcpyHostToDevice) ||
        (!blk_view.empty() && hipMemcpy(b->d_blk_view.p, blk_view.data(), blk_view.size(), hipMemcpyHostToDevice)) ||
        (!lblk_view.empty() &&
         hipMemcpy(b->d_lblk_view.p, lblk_view.data(), lblk_view.size(), hipMemcpyHostToDevice)) ||
        hipMemcpy(b->d_pix_views.p, pix.data(), b->n * sizeof(BatchPixView), hipMemcpyHostToDevice))
        return bfail(b, FGL_ERR_DEVICE, "upload failed");
    if (b->ao_mode == FGL_AO_NEURAL) {
        if ((rc = nao_reserve(b->nao, px, lpx))) return bfail(b, rc, "neural AO buffers");
        for (int l = 0; l < 2; ++l) {
            b->n_conv_items[l] = (uint32_t)conv_items[l].size();
            if (grow(b->d_conv_items[l], (conv_items[l].size() + 1) * sizeof(uint2)))
                return bfail(b, FGL_ERR_NOMEM, "batch allocation failed");
            if (!conv_items[l].empty() && hipMemcpy(b->d_conv_items[l].p, conv_items[l].data(),
                                                    conv_items[l].size() * sizeof(uint2), hipMemcpyHostToDevice))
                return bfail(b, FGL_ERR_DEVICE, "upload failed");
        }
    }
    b->measured = false;
    return FGL_OK;
}

// Pair capacity of pass p: emits beyond `need` pairs are dropped (and
// flagged).  Buffers only grow.
static int ensure_capacity(fgl_batch *b, int p, uint32_t need)
{
    if (need != b->cap[p]) drop_graph(b);
    const uint32_t alloc = need < 1024 ? 1024 : need;
    if (grow(b->keys[p], (size_t)alloc * 4) || grow(b->vals[p], (size_t)alloc * 4) ||
        grow(b->keys_alt[p], (size_t)alloc * 4) || grow(b->vals_alt[p], (size_t)alloc * 4) ||
        grow(b->sort_tmp[p], radix_tmp_words(alloc) * 4))
        return bfail(b, FGL_ERR_NOMEM, "batch pair buffers");
    b->cap[p] = need;
    return FGL_OK;
}

// Enqueues one render of every view.  measure: read each pass's pair count
// back (host sync) and size the pair buffers before emitting -- not
// capturable; otherwise the pass runs with the allocated capacity and raises
// its overflow flag when the pairs do not fit.
static int enqueue(fgl_batch *b, bool measure, bool timed)
{
    hipStream_t s = ctx_stream(b->ctx);
    uint32_t *ctr = (uint32_t *)b->counters.p;
    const DevView *views = (const DevView *)b->d_views.p;
    int rc;
    for (int p = 0; p < 2; ++p) {
        if (p == kBvShadow && !b->have_shadow) {
            launch_fill_u32(ctr + 2, 0, 1, s);
            continue;
        }
        const DevDraw *draws = (const DevDraw *)b->d_draws[p].p;
        const int nt = (int)b->nt[p];
        uint32_t *offs = (uint32_t *)b->offs.p;
        {
            CtxTimer t(b->ctx, timed ? (p == kBvShadow ? "batch_shadow_binning" : "batch_main_binning") : nullptr);
            if (nt > 0) {
                launch_bv_vertex(p, draws, views, (const uint2 *)b->d_vblocks[p].p, b->n_vblocks[p], (int4 *)b->screen.p,
                                 (float *)b->ndc_z.p, s);
                launch_fill_u32(ctr, 0, 1, s);
                launch_bv_setup(p, draws, views, (const uint2 *)b->d_fblocks[p].p, b->n_fblocks[p],
                                (const int4 *)b->screen.p,
                                b->TS[p], (uint2 *)b->rect.p, (uint32_t *)b->counts.p, (int4 *)b->tri_v.p,
                                (uint32_t *)b->big_list.p, ctr, s);
                launch_exclusive_scan((const uint32_t *)b->counts.p, offs, (size_t)nt, (uint32_t *)b->scan_tmp.p, s);
            } else {
                launch_fill_u32(offs, 0, 1, s);
            }
            if (measure) {
                if (hipMemcpyAsync(b->h_tot, offs + nt, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess)
                    return bfail(b, FGL_ERR_DEVICE, "pair count readback failed");
                b->last_pairs[p] = *b->h_tot;
                const uint64_t want = b->cap_req[p] ? (uint64_t)b->cap_req[p]
                                                    : (uint64_t)*b->h_tot + *b->h_tot / 8 + 1024;
                if (want > 0xffffffffull) return bfail(b, FGL_ERR_CAPACITY, "too many (tile, triangle) pairs");
                if ((rc = ensure_capacity(b, p, (uint32_t)want))) return rc;
            }
            if (!b->keys[p].p) return bfail(b, FGL_ERR_INVALID, "batch capacity not sized");
            BvBinArgs ba;
            ba.draws = draws;
            ba.views = views;
            ba.off = offs;
            ba.rect = (const uint2 *)b->rect.p;
            ba.tri_v = (const int4 *)b->tri_v.p;
            ba.big_list = (const uint32_t *)b->big_list.p;
            ba.big_count = ctr;
            ba.n_tri = nt;
            ba.cap = b->cap[p];
            ba.keys = (uint32_t *)b->keys[p].p;
            ba.vals = (uint32_t *)b->vals[p].p;
            launch_bv_emit(p, ba, s);
            launch_bv_overflow(offs + nt, b->cap[p], ctr + 2 + p, s);
        }
        const uint32_t *skeys = (const uint32_t *)b->keys[p].p, *svals = (const uint32_t *)b->vals[p].p;
        {
            CtxTimer t(b->ctx, timed ? (p == kBvShadow ? "batch_shadow_sort" : "batch_main_sort") : nullptr);
            if (nt > 0) {
                const int alt = launch_radix_sort_impl((uint32_t *)b->keys[p].p, (uint32_t *)b->vals[p].p,
                                                       (uint32_t *)b->keys_alt[p].p, (uint32_t *)b->vals_alt[p].p,
                                                       b->cap[p], offs + nt, bits_for(b->tiles_total[p] - 1),
                                                       (uint32_t *)b->sort_tmp[p].p, s);
                if (alt) {
                    skeys = (const uint32_t *)b->keys_alt[p].p;
                    svals = (const uint32_t *)b->vals_alt[p].p;
                }
            }
            launch_fill_u32((uint32_t *)b->tile_start.p, 0, b->tiles_total[p], s);
            launch_fill_u32((uint32_t *)b->tile_end.p, 0, b->tiles_total[p], s);
            if (nt > 0)
                launch_tile_ranges(skeys, b->cap[p], offs + nt, (uint32_t *)b->tile_start.p,
                                   (uint32_t *)b->tile_end.p, s);
        }
        {
            CtxTimer t(b->ctx, timed ? (p == kBvShadow ? "batch_shadow_raster" : "batch_main_raster") : nullptr);
            launch_fill_u32(ctr + 1, 0, 1, s);
            BvRasterArgs ra;
            ra.draws = draws;
            ra.views = views;
            ra.screen = (const int4 *)b->screen.p;
            ra.ndc_z = (const float *)b->ndc_z.p;
            ra.tri_v = (const int4 *)b->tri_v.p;
            ra.vals = svals;
            ra.tile_start = (const uint32_t *)b->tile_start.p;
            ra.tile_end = (const uint32_t *)b->tile_end.p;
            ra.items = (const uint32_t *)b->d_items[p].p;
            ra.n_items = b->n_items[p];
            ra.queue = ctr + 1;
            ra.depth = (float *)b->depth.p;
            ra.shadow_depth = (float *)b->shadow_depth.p;
            ra.shadow = (float *)b->shadow.p;
            ra.vis = (int *)b->vis.p;
            const unsigned wg = std::min<unsigned>(b->n_items[p], (unsigned)(b->n_cu * b->occupancy));
            launch_bv_raster(p, b->TS[p], wg, ra, s);
        }
    }
    BvPixelArgs pa;
    pa.draws = (const DevDraw *)b->d_draws[kBvMain].p;
    pa.views = views;
    pa.blk_view = (const uint8_t *)b->d_blk_view.p;
    pa.n_blocks = b->n_blocks;
    pa.screen = (const int4 *)b->screen.p;
    pa.tri_v = (const int4 *)b->tri_v.p;
    pa.vis = (const int *)b->vis.p;
    pa.materials = b->materials;
    pa.textures = b->textures;
    pa.gbuf = (float *)b->gbuf.p;
    pa.ao = (float *)b->ao.p;
    pa.fb = (float *)b->fb.p;
    pa.shadow = (const float *)b->shadow.p;
    pa.ao_is_one = b->ao_mode == FGL_AO_NEURAL ? 0 : 1;
    pa.img = (uint8_t *)b->img.p;
    {
        CtxTimer t(b->ctx, timed ? "batch_gbuffer" : nullptr);
        launch_bv_gbuffer(pa, s);
    }
    if (b->ao_mode == FGL_AO_NEURAL) {
        CtxTimer t(b->ctx, timed ? "batch_neural_ao" : nullptr);
        NaoBatchTables nt;
        nt.views = (const BatchPixView *)b->d_pix_views.p;
        nt.blk_view = pa.blk_view;
        nt.n_blocks = b->n_blocks;
        nt.lblk_view = (const uint8_t *)b->d_lblk_view.p;
        nt.n_lblocks = b->n_lblocks;
        nt.total_px = b->total_px;
        nt.total_lpx = b->total_lpx;
        for (int l = 0; l < 2; ++l) {
            nt.items[l] = (const uint2 *)b->d_conv_items[l].p;
            nt.n_items[l] = b->n_conv_items[l];
        }
        if ((rc = nao_infer_batch(b->nao, s, nt, (const float *)b->depth.p, pa.gbuf, pa.ao)))
            return bfail(b, rc, "batched neural AO failed");
    }
    {
        CtxTimer t(b->ctx, timed ? "batch_lighting" : nullptr);
        launch_bv_lighting(pa, s);
    }
    {
        CtxTimer t(b->ctx, timed ? "batch_resolve" : nullptr);
        launch_bv_resolve(pa, s);
    }
    if (hipMemcpyAsync(b->h_flags, ctr + 2, 8, hipMemcpyDeviceToHost, s) != hipSuccess)
        return bfail(b, FGL_ERR_DEVICE, "flag readback failed");
    if (hipGetLastError() != hipSuccess) return bfail(b, FGL_ERR_DEVICE, "batch launch failed");
    return FGL_OK;
}

static int capture(fgl_batch *b)
{
    hipStream_t s = ctx_stream(b->ctx);
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return bfail(b, FGL_ERR_DEVICE, "graph capture failed to start");
    int rc = enqueue(b, false, false);
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess || !g) return bfail(b, FGL_ERR_DEVICE, "graph capture failed");
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->exec = nullptr;
        return bfail(b, FGL_ERR_DEVICE, "graph instantiation failed");
    }
    return FGL_OK;
}

extern "C" {

int fgl_render_view(fgl_ctx *c, const fgl_view_desc *v, int ao_mode, fgl_nao *nao)
{
    if (!c || !v) return FGL_ERR_INVALID;
    if (ao_mode != FGL_AO_NONE && ao_mode != FGL_AO_NEURAL) return ctx_fail(c, FGL_ERR_INVALID, "bad AO mode");
    if (ao_mode == FGL_AO_NEURAL && !nao) return ctx_fail(c, FGL_ERR_INVALID, "neural AO needs a network");
    int rc;
    if ((rc = check_view(c, *v))) return rc;
    // DoShadowPass (render.cpp:60-97)
    if ((rc = apply_view_state(c, *v, true))) return rc;
    if (v->shadow_on) {
        if ((rc = fgl_init_shadow_buffer(c, v->shadow_size, v->shadow_size))) return rc;
        if ((rc = fgl_begin_pass(c, FGL_PASS_SHADOW))) return rc;
        for (int i = 0; i < v->n_draws; ++i)
            if ((rc = fgl_draw_model(c, v->models[i], v->matrices + 16 * i))) return rc;
        if ((rc = fgl_end_pass(c))) return rc;
    }
    // DoGeometryPass (render.cpp:159-180)
    if ((rc = apply_view_state(c, *v, false))) return rc;
    if ((rc = fgl_init_targets(c, v->width, v->height))) return rc;
    if ((rc = fgl_begin_pass(c, FGL_PASS_GEOMETRY))) return rc;
    for (int i = 0; i < v->n_draws; ++i)
        if ((rc = fgl_draw_model(c, v->models[i], v->matrices + 16 * i))) return rc;
    if ((rc = fgl_end_pass(c))) return rc;
    // DoLightingPass (render.cpp:182-212), AO from the network
    if ((rc = fgl_begin_pass(c, FGL_PASS_LIGHTING))) return rc;
    if (ao_mode == FGL_AO_NEURAL && (rc = fgl_nao_infer(nao))) return rc;
    if ((rc = fgl_draw_screen_space_pixels(c))) return rc;
    return fgl_resolve(c, 1);
}

int fgl_batch_create(fgl_ctx *c, int n, const fgl_view_desc *views, int ao_mode, fgl_nao *nao, fgl_batch **out)
{
    if (!c || !out || n < 1 || !views) return FGL_ERR_INVALID;
    *out = nullptr;
    if (n > 255) return ctx_fail(c, FGL_ERR_INVALID, "a batch holds at most 255 views");
    if (ao_mode != FGL_AO_NONE && ao_mode != FGL_AO_NEURAL) return ctx_fail(c, FGL_ERR_INVALID, "bad AO mode");
    if (ao_mode == FGL_AO_NEURAL && !nao) return ctx_fail(c, FGL_ERR_INVALID, "neural AO needs a network");
    int rc;
    for (int i = 0; i < n; ++i)
        if ((rc = check_view(c, views[i]))) return rc;
    fgl_batch *b = new fgl_batch();
    b->ctx = c;
    b->n = n;
    b->ao_mode = ao_mode;
    b->nao = nao;
    b->views.resize(n);
    for (int i = 0; i < n; ++i) {
        ViewCopy &vc = b->views[i];
        vc.d = views[i];
        vc.models.assign(views[i].models, views[i].models + views[i].n_draws);
        vc.mats.assign(views[i].matrices, views[i].matrices + 16 * (size_t)views[i].n_draws);
        vc.d.models = vc.models.data();
        vc.d.matrices = vc.mats.data();
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || b->n_cu < 1)
        b->n_cu = 256;
    if (hipHostMalloc((void **)&b->h_flags, 16, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&b->h_tot, 16, hipHostMallocDefault) != hipSuccess) {
        fgl_batch_destroy(b);
        return ctx_fail(c, FGL_ERR_NOMEM, "pinned allocation failed");
    }
    if ((rc = plan(b))) {
        fgl_batch_destroy(b);
        return rc;
    }
    *out = b;
    return FGL_OK;
}

void fgl_batch_destroy(fgl_batch *b)
{
    if (!b) return;
    (void)hipStreamSynchronize(ctx_stream(b->ctx));   // teardown: nothing to report to
    drop_graph(b);
    for (int l = 0; l < 2; ++l) release(b->d_conv_items[l]);
    for (int p = 0; p < 2; ++p) {
        release(b->d_vblocks[p]);
        release(b->d_fblocks[p]);
    }
    Buf *bufs[] = {&b->d_views, &b->d_draws[0], &b->d_draws[1], &b->d_items[0], &b->d_items[1], &b->d_blk_view,
                   &b->d_lblk_view,
                   &b->d_pix_views, &b->depth, &b->vis, &b->gbuf, &b->ao, &b->fb, &b->img, &b->shadow,
                   &b->shadow_depth, &b->screen, &b->ndc_z, &b->rect, &b->counts, &b->offs, &b->scan_tmp, &b->tri_v,
                   &b->big_list, &b->tile_start, &b->tile_end, &b->counters};
    for (Buf *x : bufs) release(*x);
    for (int p = 0; p < 2; ++p) {
        release(b->keys[p]); release(b->vals[p]); release(b->keys_alt[p]); release(b->vals_alt[p]);
        release(b->sort_tmp[p]);
    }
    if (b->h_flags) (void)hipHostFree(b->h_flags);
    if (b->h_tot) (void)hipHostFree(b->h_tot);
    delete b;
}

int fgl_batch_set_tiles(fgl_batch *b, int shadow_tile, int main_tile)
{
    if (!b) return FGL_ERR_INVALID;
    for (int t : {shadow_tile, main_tile})
        if (t != 8 && t != 16 && t != 32) return bfail(b, FGL_ERR_INVALID, "tile must be 8, 16 or 32");
    if (b->TS[0] == shadow_tile && b->TS[1] == main_tile) return FGL_OK;
    if (hipStreamSynchronize(ctx_stream(b->ctx)) != hipSuccess) return bfail(b, FGL_ERR_DEVICE, "stream failed");
    b->TS[0] = shadow_tile;
    b->TS[1] = main_tile;
    return plan(b);
}

int fgl_batch_set_occupancy(fgl_batch *b, int wg_per_cu)
{
    if (!b) return FGL_ERR_INVALID;
    if (wg_per_cu < 1 || wg_per_cu > 16) return bfail(b, FGL_ERR_INVALID, "occupancy must be 1..16");
    if (b->occupancy != wg_per_cu) drop_graph(b);
    b->occupancy = wg_per_cu;
    return FGL_OK;
}

int fgl_batch_set_capacity(fgl_batch *b, uint32_t shadow_pairs, uint32_t main_pairs)
{
    if (!b) return FGL_ERR_INVALID;
    if (hipStreamSynchronize(ctx_stream(b->ctx)) != hipSuccess) return bfail(b, FGL_ERR_DEVICE, "stream failed");
    b->cap_req[0] = shadow_pairs;
    b->cap_req[1] = main_pairs;
    int rc;
    for (int p = 0; p < 2; ++p)
        if (b->cap_req[p] && (rc = ensure_capacity(b, p, b->cap_req[p]))) return rc;
    // explicit capacities for both passes: render without measuring
    b->measured = b->cap_req[0] && b->cap_req[1];
    return FGL_OK;
}

int fgl_batch_use_graph(fgl_batch *b, int on)
{
    if (!b) return FGL_ERR_INVALID;
    b->use_graph = on != 0;
    if (!on) drop_graph(b);
    return FGL_OK;
}

// The context's material / texture tables and the neural-AO buffers can move
// after the batch was planned (an upload that outgrows a table, a larger
// single-frame inference): re-read them before every render and capture the
// graph again when any address changed.
static int refresh_resources(fgl_batch *b)
{
    const DevMaterial *m = nullptr;
    const DevTexture *t = nullptr;
    int rc = ctx_tables(b->ctx, &m, &t);
    if (rc) return rc;
    if (m != b->materials || t != b->textures) {
        b->materials = m;
        b->textures = t;
        drop_graph(b);
    }
    if (b->ao_mode == FGL_AO_NEURAL) {
        if ((rc = nao_reserve(b->nao, b->total_px, b->total_lpx))) return bfail(b, rc, "neural AO buffers");
        const uint64_t g = nao_buffers_generation(b->nao);
        if (g != b->nao_gen) {
            b->nao_gen = g;
            drop_graph(b);
        }
    }
    return FGL_OK;
}

int fgl_batch_render(fgl_batch *b)
{
    if (!b) return FGL_ERR_INVALID;
    const bool timed = ctx_timing(b->ctx);
    int rc;
    if ((rc = refresh_resources(b))) return rc;
    if (!b->measured) {
        // first render: measure each pass's pairs and size the buffers
        if ((rc = enqueue(b, true, timed))) return rc;
        b->measured = true;
        return FGL_OK;
    }
    if (!b->use_graph || timed) return enqueue(b, false, timed);
    if (!b->exec && (rc = capture(b))) return rc;
    if (hipGraphLaunch(b->exec, ctx_stream(b->ctx)) != hipSuccess) return bfail(b, FGL_ERR_DEVICE, "graph launch failed");
    return FGL_OK;
}

int fgl_batch_sync(fgl_batch *b, int *rerenders)
{
    if (!b) return FGL_ERR_INVALID;
    if (rerenders) *rerenders = 0;
    hipStream_t s = ctx_stream(b->ctx);
    if (hipStreamSynchronize(s) != hipSuccess) return bfail(b, FGL_ERR_DEVICE, "batch failed");
    for (int attempt = 0; attempt < 2 && (b->h_flags[0] || b->h_flags[1]); ++attempt) {
        // a pass overflowed its pair capacity: re-render, sizing from the counts
        b->cap_req[0] = b->cap_req[1] = 0;
        drop_graph(b);
        int rc = refresh_resources(b);
        if (rc) return rc;
        rc = enqueue(b, true, false);
        if (rc) return rc;
        if (hipStreamSynchronize(s) != hipSuccess) return bfail(b, FGL_ERR_DEVICE, "batch failed");
        if (rerenders) ++*rerenders;
    }
    if (b->h_flags[0] || b->h_flags[1]) return bfail(b, FGL_ERR_CAPACITY, "batch pairs still overflow");
    return FGL_OK;
}

int fgl_batch_read(fgl_batch *b, int view, int which, void *dst, size_t bytes)
{
    if (!b || !dst || view < 0 || view >= b->n) return FGL_ERR_INVALID;
    const DevView &v = b->hv[view];
    const size_t P = (size_t)v.W * v.H, SP = (size_t)v.S * v.S;
    const char *src = nullptr;
    size_t n = 0;
    switch (which) {
    case FGL_TARGET_FRAME: src = (const char *)b->fb.p + 3 * v.px_off * 4; n = 3 * P * 4; break;
    case FGL_TARGET_DEPTH: src = (const char *)b->depth.p + v.px_off * 4; n = P * 4; break;
    case FGL_TARGET_GBUFFER: src = (const char *)b->gbuf.p + kGbufPlanes * v.px_off * 4; n = kGbufPlanes * P * 4; break;
    case FGL_TARGET_AO: src = (const char *)b->ao.p + v.px_off * 4; n = P * 4; break;
    case FGL_TARGET_IMAGE: src = (const char *)b->img.p + 3 * v.px_off; n = 3 * P; break;
    case FGL_TARGET_SHADOW: src = (const char *)b->shadow.p + v.spx_off * 4; n = SP * 4; break;
    case FGL_TARGET_SHADOW_DEPTH: src = (const char *)b->shadow_depth.p + v.spx_off * 4; n = SP * 4; break;
    default: return bfail(b, FGL_ERR_INVALID, "unknown target");
    }
    if (which == FGL_TARGET_AO && b->ao_mode == FGL_AO_NONE) return bfail(b, FGL_ERR_INVALID, "batch has no AO");
    if ((which == FGL_TARGET_SHADOW || which == FGL_TARGET_SHADOW_DEPTH) && !v.shadow_on)
        return bfail(b, FGL_ERR_INVALID, "view has no shadow pass");
    if (bytes != n) return bfail(b, FGL_ERR_INVALID, "size mismatch");
    if (hipStreamSynchronize(ctx_stream(b->ctx)) != hipSuccess ||
        hipMemcpy(dst, src, n, hipMemcpyDeviceToHost) != hipSuccess)
        return bfail(b, FGL_ERR_DEVICE, "readback failed");
    return FGL_OK;
}

int fgl_batch_stats(fgl_batch *b, uint64_t stats[10])
{
    if (!b || !stats) return FGL_ERR_INVALID;
    stats[0] = b->last_pairs[0];
    stats[1] = b->last_pairs[1];
    stats[2] = b->cap[0];
    stats[3] = b->cap[1];
    stats[4] = (uint64_t)b->nt[0];
    stats[5] = (uint64_t)b->nt[1];
    stats[6] = b->total_px;
    stats[7] = (uint64_t)b->n;
    stats[8] = std::min<uint64_t>(b->n_items[1], (uint64_t)b->n_cu * b->occupancy);
    stats[9] = b->tiles_total[1];
    return FGL_OK;
}

}  // extern "C"
