// context.cpp -- the fgl_* C ABI: device resources, ForkerGL-style state, pass
// orchestration on one HIP stream, readback, scene rendering.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <chrono>
#include <string>
#include <utility>
#include <mutex>
#include <vector>

#include "../../include/fgl.h"
#include "ctx_internal.hpp"
#include "mt_jump.hpp"
#include "fgl_kernels.hpp"
#include "host_io.hpp"

using namespace fgl;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct ModelRes {
    float *pos = nullptr, *uv = nullptr, *nrm = nullptr, *tan = nullptr;
    int *vi = nullptr, *ti = nullptr, *ni = nullptr, *fmat = nullptr;
    float *corner = nullptr;   // per face corner: pos, uv, nrm, tan (12 floats)
    int n_pos = 0, n_faces = 0;
    bool has_tan = false, pbr = false;
};

struct TimedKernel {
    std::string name;
    hipEvent_t a, b;
};

}  // namespace

struct fgl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int TS = 16;

    std::vector<DevTexture> textures;
    std::vector<void *> texture_mem;
    DevBuf d_textures;
    bool textures_dirty = true;
    std::vector<DevMaterial> materials;
    DevBuf d_materials;
    bool materials_dirty = true;
    std::vector<ModelRes> models;

    // ForkerGL state
    int W = 0, H = 0, SW = 0, SH = 0;
    float viewport[16], shadow_viewport[16], view[16], proj[16], light_space[16];
    float eye[3] = {0.f, 0.f, 0.f};
    std::vector<fgl_light> lights;
    int shadow_on = 1;
    fgl_shadow_cfg shadow_cfg;

    // targets
    DevBuf depth, fb, gbuf, ao, blur_tmp, shadow, shadow_depth, img, vis;
    int img_w = 0, img_h = 0;
    // Clears requested by fgl_init_* / fgl_clear_color that have not been
    // written yet: the next pass that covers every pixel applies them.
    bool pend_depth = false, pend_gbuf = false, pend_fb = false, pend_ao = false, pend_shadow = false;
    float fb_clear[3] = {0.f, 0.f, 0.f};
    int TS_shadow = 16;

    // current pass
    int pass = -1;
    struct Rec {
        int model;
        float m[16];
        bool has_normal;   // the shader's uNormalMatrix was given
        float n[9];
    };
    std::vector<Rec> draws;

    // pass scratch
    DevBuf d_draws, screen, rect, counts, offs, scan_tmp, keys, vals, keys_alt, vals_alt, sort_tmp,
        tile_start, tile_end, tri_v, big_list, counters, ndc_z;
    // forward pass with a stochastic shadow filter: per-pixel passing counts
    // (+ scan scratch), the passing fragments (triangle ids then pixels) and
    // the sort's other copy, its scratch, their shadow inputs, and each
    // winner's sample offset
    DevBuf frag_cnt, frag_tri, frag_alt, frag_tmp, frag_planes, frag_offs_px;
    bool frag_sorted_alt = false;
    std::vector<DevDraw> last_draws;
    uint32_t *h_total = nullptr;
    // pass counts published by the device into coherent pinned memory
    // (pair count, big-triangle count, sequence number), polled by the host
    uint32_t *h_pub = nullptr, *d_pub = nullptr;
    uint32_t pub_seq = 0;
    // the reference's random stream (device MtState), its sample buffers,
    // the PCSS sample offsets, the DoSSAO powf table
    DevBuf rng, rng_samples, rng_offs, ssao_pow3;
    uint32_t rng_seed = 5489u;
    // the parallel generator: base sequence, jumped chunk arrays, raw draws,
    // the per-workgroup counts of kept tries and their scan, the jump
    // polynomials uploaded so far; the last fill's tries (T of D draws) for
    // its commit
    DevBuf mt_base, mt_states, mt_raw, mt_pos;
    DevBuf mt_polys[2];          // jump polynomials for the two chunk lengths
    int mt_polys_n[2] = {0, 0};
    uint32_t fill_T = 0, fill_n = 0;
    int fill_D = 0;

    // device address of the allocation-free entry point's scratch array
    float *noalloc_scratch = nullptr;
    // neural AO still on its network grid (lao != nullptr): the AO target's
    // content until something reads or replaces it; owner = the network
    // whose buffers hold it
    NetAo ao_net{};
    const void *ao_net_owner = nullptr;
    std::vector<void *> naos;   // networks created on this context

    // timing
    bool timing = false;
    std::string timing_only;   // empty: every timed kernel
    std::vector<TimedKernel> timed;
};

// Scratch of fgl_ssao_noalloc: the unblurred AO that the separable blur
// reads, FGL_SSAO_NOALLOC_MAX_PIXELS floats on every device the library runs
// on, held until the process ends.  On the host `scratch` names a stand-in,
// not device memory: host code only ever uses the address that
// hipGetSymbolAddress returns for the current device (scratch_address).
static __device__ __align__(256) float scratch[FGL_SSAO_NOALLOC_MAX_PIXELS];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float *g_scratch_ptr[kMaxDevices] = {};

// Address of `scratch` on `device`, which must be the current device; looked
// up once per device.
static hipError_t scratch_address(int device, float **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = (float *)p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

static int fail(fgl_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(ctx, FGL_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static int ensure(fgl_ctx *c, DevBuf &b, size_t bytes)
{
    if (bytes <= b.cap && b.p) return FGL_OK;
    if (b.p) {
        HIPCHK(c, hipStreamSynchronize(c->stream));   // no kernel may still use it
        (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    if (hipMalloc(&b.p, want) != hipSuccess) {
        b.p = nullptr;
        return fail(c, FGL_ERR_NOMEM, "hipMalloc of " + std::to_string(want) + " bytes failed");
    }
    b.cap = want;
    return FGL_OK;
}

static void free_buf(DevBuf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// ---------------- timing ----------------

struct KernelTimer {
    fgl_ctx *c;
    TimedKernel tk;
    bool on;
    KernelTimer(fgl_ctx *ctx, const char *name)
        : c(ctx), on(ctx->timing && (ctx->timing_only.empty() || ctx->timing_only == name))
    {
        if (!on) return;
        tk.name = name;
        // a diagnostic: when its events cannot be made, the span is not timed
        if (hipEventCreate(&tk.a) != hipSuccess) {
            on = false;
            return;
        }
        if (hipEventCreate(&tk.b) != hipSuccess || hipEventRecord(tk.a, c->stream) != hipSuccess) {
            (void)hipEventDestroy(tk.a);
            on = false;
        }
    }
    ~KernelTimer()
    {
        if (!on) return;
        (void)hipEventRecord(tk.b, c->stream);
        c->timed.push_back(tk);
    }
};

// Two device counters to the host without a stream synchronisation: a
// one-thread kernel writes them and a fresh sequence number into coherent
// pinned memory, and the host polls for the sequence number (the GPU then
// idles only for the poll and the next launch, not for a full
// synchronisation).  A poll that outlasts a second falls back to
// synchronising the stream, which reports a failed kernel.
static int read_counts(fgl_ctx *c, const uint32_t *a, const uint32_t *b, uint32_t &va, uint32_t &vb)
{
    const uint32_t seq = ++c->pub_seq;
    launch_publish2(a, b, seq, c->d_pub, c->stream);
    HIPCHK(c, hipGetLastError());
    volatile uint32_t *pub = c->h_pub;
    auto t0 = std::chrono::steady_clock::now();
    for (uint32_t spin = 0; pub[2] != seq; ++spin) {
        if ((spin & 1023u) == 1023u &&
            std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
            if (pub[2] != seq) return fail(c, FGL_ERR_DEVICE, "pass counts were not published");
            break;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    va = pub[0];
    vb = pub[1];
    return FGL_OK;
}

// Restarts the context's random stream at its seed (std::mt19937(seed)).
static int rng_restart(fgl_ctx *c)
{
    MtState st;
    mt_seed_state(c->rng_seed, st);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->rng.p, &st, sizeof(st), hipMemcpyHostToDevice));
    return FGL_OK;
}

template <typename T>
static int upload_array(fgl_ctx *c, T **dst, const T *src, size_t n)
{
    *dst = nullptr;
    HIPCHK(c, hipMalloc((void **)dst, (n ? n : 1) * sizeof(T)));
    if (n) HIPCHK(c, hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return FGL_OK;
}

extern "C" {

int fgl_ctx_create(int device, fgl_ctx **out)
{
    if (!out) return FGL_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return FGL_ERR_DEVICE;
    if (device < 0 || device >= n) return FGL_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return FGL_ERR_DEVICE;
    fgl_ctx *c = new fgl_ctx();
    c->device = device;
    if (scratch_address(device, &c->noalloc_scratch) != hipSuccess) {
        delete c;
        return FGL_ERR_DEVICE;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return FGL_ERR_DEVICE;
    }
    if (hipHostMalloc((void **)&c->h_total, sizeof(uint32_t) * 4, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return FGL_ERR_NOMEM;
    }
    if (hipHostMalloc((void **)&c->h_pub, 64, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void **)&c->d_pub, c->h_pub, 0) != hipSuccess) {
        if (c->h_pub) (void)hipHostFree(c->h_pub);
        (void)hipHostFree(c->h_total);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return FGL_ERR_NOMEM;
    }
    std::memset(c->h_pub, 0, 64);
    float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    std::memcpy(c->viewport, I, sizeof(I));
    std::memcpy(c->shadow_viewport, I, sizeof(I));
    std::memcpy(c->view, I, sizeof(I));
    std::memcpy(c->proj, I, sizeof(I));
    std::memcpy(c->light_space, I, sizeof(I));
    std::memset(&c->shadow_cfg, 0, sizeof(c->shadow_cfg));
    // Render::DoSSAO's 1 - occlusion is k/32 (k = 0..32, exact); its
    // Pow(x, 3) is the host libm's powf, as in the reference.
    float pow3[33];
    for (int k = 0; k <= 32; ++k) {
        float occ = 0.f;
        for (int j = 0; j < k; ++j) occ += (1.f / 32) * 1.f;
        pow3[k] = powf(1.f - occ, 3);
    }
    if (ensure(c, c->ssao_pow3, sizeof(pow3)) || ensure(c, c->rng, sizeof(MtState)) ||
        hipMemcpy(c->ssao_pow3.p, pow3, sizeof(pow3), hipMemcpyHostToDevice) != hipSuccess ||
        rng_restart(c) != FGL_OK) {
        fgl_ctx_destroy(c);
        return FGL_ERR_NOMEM;
    }
    *out = c;
    return FGL_OK;
}

void fgl_ctx_destroy(fgl_ctx *c)
{
    if (!c) return;
    for (void *n : c->naos) nao_detach(n);   // networks outliving their context stop using it
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (void *p : c->texture_mem) (void)hipFree(p);
    for (auto &m : c->models) {
        (void)hipFree(m.pos); (void)hipFree(m.uv); (void)hipFree(m.nrm);
        if (m.tan) (void)hipFree(m.tan);
        (void)hipFree(m.vi); (void)hipFree(m.ti); (void)hipFree(m.ni); (void)hipFree(m.fmat); (void)hipFree(m.corner);
    }
    DevBuf *bufs[] = {&c->d_textures, &c->d_materials, &c->depth, &c->fb, &c->gbuf, &c->ao,
                      &c->blur_tmp, &c->shadow, &c->shadow_depth, &c->img, &c->vis, &c->d_draws,
                      &c->rng, &c->rng_samples, &c->rng_offs, &c->ssao_pow3,
                      &c->mt_base, &c->mt_states, &c->mt_raw, &c->mt_pos, &c->mt_polys[0], &c->mt_polys[1],
                      &c->screen, &c->rect, &c->counts, &c->offs, &c->scan_tmp, &c->keys,
                      &c->vals, &c->keys_alt, &c->vals_alt, &c->sort_tmp, &c->tile_start,
                      &c->tile_end, &c->tri_v, &c->big_list, &c->counters, &c->ndc_z,
                      &c->frag_cnt, &c->frag_tri, &c->frag_alt, &c->frag_tmp, &c->frag_planes,
                      &c->frag_offs_px};
    for (DevBuf *b : bufs) free_buf(*b);
    for (auto &t : c->timed) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    if (c->h_total) (void)hipHostFree(c->h_total);
    if (c->h_pub) (void)hipHostFree(c->h_pub);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *fgl_last_error(const fgl_ctx *c) { return c ? c->err.c_str() : "null context"; }

int fgl_sync(fgl_ctx *c)
{
    if (!c) return FGL_ERR_INVALID;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FGL_OK;
}

int fgl_set_tile_size(fgl_ctx *c, int tile)
{
    if (!c) return FGL_ERR_INVALID;
    if (tile != 8 && tile != 16 && tile != 32) return fail(c, FGL_ERR_INVALID, "tile must be 8, 16 or 32");
    c->TS = tile;
    c->TS_shadow = tile;
    return FGL_OK;
}

int fgl_set_pass_tile_size(fgl_ctx *c, int pass, int tile)
{
    if (!c) return FGL_ERR_INVALID;
    if (tile != 8 && tile != 16 && tile != 32) return fail(c, FGL_ERR_INVALID, "tile must be 8, 16 or 32");
    if (pass == FGL_PASS_SHADOW) c->TS_shadow = tile;
    else if (pass == FGL_PASS_FORWARD || pass == FGL_PASS_GEOMETRY) c->TS = tile;
    else return fail(c, FGL_ERR_INVALID, "pass has no raster stage");
    return FGL_OK;
}

int fgl_get_pass_tile_size(fgl_ctx *c, int pass, int *tile)
{
    if (!c || !tile) return FGL_ERR_INVALID;
    *tile = (pass == FGL_PASS_SHADOW) ? c->TS_shadow : c->TS;
    return FGL_OK;
}

// ---------------- resources ----------------

int fgl_upload_texture(fgl_ctx *c, int w, int h, int bpp, const uint8_t *bytes, int wrap,
                       int filter, int *out_id)
{
    if (!c || !bytes || !out_id) return FGL_ERR_INVALID;
    if (w <= 0 || h <= 0 || (bpp != 1 && bpp != 3 && bpp != 4) || wrap < 0 || wrap > 3 ||
        filter < 0 || filter > 1)
        return fail(c, FGL_ERR_INVALID, "bad texture parameters");
    std::vector<uint32_t> packed((size_t)w * h);
    std::vector<uint8_t> zb((size_t)w * h);
    for (size_t i = 0; i < packed.size(); ++i) {
        const uint8_t *p = bytes + i * bpp;
        uint32_t b = p[0], g = bpp > 1 ? p[1] : 0, r = bpp > 2 ? p[2] : 0;
        packed[i] = r | (g << 8) | (b << 16);
        zb[i] = (uint8_t)b;
    }
    void *mem = nullptr, *zmem = nullptr;
    HIPCHK(c, hipMalloc(&mem, packed.size() * 4));
    c->texture_mem.push_back(mem);
    HIPCHK(c, hipMalloc(&zmem, zb.size()));
    c->texture_mem.push_back(zmem);
    HIPCHK(c, hipMemcpy(mem, packed.data(), packed.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(zmem, zb.data(), zb.size(), hipMemcpyHostToDevice));
    DevTexture t;
    t.texels = (const uint32_t *)mem;
    t.zb = (const uint8_t *)zmem;
    t.w = w; t.h = h; t.wrap = wrap; t.filter = filter;
    c->textures.push_back(t);
    c->textures_dirty = true;
    *out_id = (int)c->textures.size() - 1;
    return FGL_OK;
}

int fgl_upload_material(fgl_ctx *c, const fgl_material *m, int *out_id)
{
    if (!c || !m || !out_id) return FGL_ERR_INVALID;
    const int32_t ids[] = {m->map_kd, m->map_ks, m->map_ke, m->map_bump, m->map_base,
                           m->map_rough, m->map_metal, m->map_ao, m->map_pbr_ke};
    for (int id : ids)
        if (id < -1 || id >= (int)c->textures.size())
            return fail(c, FGL_ERR_INVALID, "material references an unknown texture id");
    DevMaterial d;
    static_assert(sizeof(DevMaterial) == sizeof(fgl_material), "material layout");
    std::memcpy(&d, m, sizeof(d));
    c->materials.push_back(d);
    c->materials_dirty = true;
    *out_id = (int)c->materials.size() - 1;
    return FGL_OK;
}

int fgl_upload_model(fgl_ctx *c, const fgl_model_desc *d, int *out_id)
{
    if (!c || !d || !out_id) return FGL_ERR_INVALID;
    if (d->n_pos <= 0 || d->n_uv <= 0 || d->n_nrm <= 0 || d->n_faces < 0 || !d->pos || !d->uv ||
        !d->nrm || (d->n_faces > 0 && (!d->vi || !d->ti || !d->ni || !d->face_material)))
        return fail(c, FGL_ERR_INVALID, "bad model description");
    size_t nf = (size_t)d->n_faces;
    for (size_t i = 0; i < 3 * nf; ++i)
        if (d->vi[i] < 0 || d->vi[i] >= d->n_pos || d->ti[i] < 0 || d->ti[i] >= d->n_uv ||
            d->ni[i] < 0 || d->ni[i] >= d->n_nrm)
            return fail(c, FGL_ERR_INVALID, "face index out of range");
    for (size_t f = 0; f < nf; ++f)
        if (d->face_material[f] < 0 || d->face_material[f] >= (int)c->materials.size())
            return fail(c, FGL_ERR_INVALID, "face material out of range");
    ModelRes m;
    int rc;
    if ((rc = upload_array(c, &m.pos, d->pos, 3 * (size_t)d->n_pos))) return rc;
    if ((rc = upload_array(c, &m.uv, d->uv, 2 * (size_t)d->n_uv))) return rc;
    if ((rc = upload_array(c, &m.nrm, d->nrm, 3 * (size_t)d->n_nrm))) return rc;
    if (d->tan && (rc = upload_array(c, &m.tan, d->tan, 3 * (size_t)d->n_pos))) return rc;
    if ((rc = upload_array(c, &m.vi, (const int *)d->vi, 3 * nf))) return rc;
    if ((rc = upload_array(c, &m.ti, (const int *)d->ti, 3 * nf))) return rc;
    if ((rc = upload_array(c, &m.ni, (const int *)d->ni, 3 * nf))) return rc;
    if ((rc = upload_array(c, &m.fmat, (const int *)d->face_material, nf))) return rc;
    {
        std::vector<float> corner(3 * nf * 12, 0.f);
        for (size_t k = 0; k < 3 * nf; ++k) {
            float *o = &corner[12 * k];
            const int vi = d->vi[k], ti = d->ti[k], ni = d->ni[k];
            o[0] = d->pos[3 * vi]; o[1] = d->pos[3 * vi + 1]; o[2] = d->pos[3 * vi + 2];
            o[3] = d->uv[2 * ti]; o[4] = d->uv[2 * ti + 1];
            o[5] = d->nrm[3 * ni]; o[6] = d->nrm[3 * ni + 1]; o[7] = d->nrm[3 * ni + 2];
            if (d->tan) { o[8] = d->tan[3 * vi]; o[9] = d->tan[3 * vi + 1]; o[10] = d->tan[3 * vi + 2]; }
        }
        if ((rc = upload_array(c, &m.corner, corner.data(), corner.size()))) return rc;
    }
    m.n_pos = d->n_pos;
    m.n_faces = d->n_faces;
    m.has_tan = d->tan != nullptr;
    m.pbr = d->support_pbr != 0;
    c->models.push_back(m);
    *out_id = (int)c->models.size() - 1;
    return FGL_OK;
}

// ---------------- state ----------------

static int flush_pass(fgl_ctx *c);

// Writes the clear values still pending on the targets named by the flags.
static int materialize(fgl_ctx *c, bool depth, bool gbuf, bool fb, bool ao, bool shadow)
{
    const size_t P = (size_t)c->W * c->H, SP = (size_t)c->SW * c->SH;
    if (depth && c->pend_depth) { launch_fill_f32((float *)c->depth.p, FLT_MAX, P, c->stream); c->pend_depth = false; }
    if (gbuf && c->pend_gbuf) { launch_fill_f32((float *)c->gbuf.p, 0.f, kGbufPlanes * P, c->stream); c->pend_gbuf = false; }
    if (fb && c->pend_fb) {
        for (int k = 0; k < 3; ++k) launch_fill_f32((float *)c->fb.p + k * P, c->fb_clear[k], P, c->stream);
        c->pend_fb = false;
    }
    if (ao && c->ao_net.lao) {
        nao_launch_up(c->ao_net, c->W, c->H, (float *)c->ao.p, c->stream);
        c->ao_net = NetAo{};
        c->ao_net_owner = nullptr;
    }
    if (ao && c->pend_ao) { launch_fill_f32((float *)c->ao.p, 1.f, P, c->stream); c->pend_ao = false; }
    if (shadow && c->pend_shadow) {
        launch_fill_f32((float *)c->shadow.p, 0.f, SP, c->stream);
        launch_fill_f32((float *)c->shadow_depth.p, FLT_MAX, SP, c->stream);
        c->pend_shadow = false;
    }
    HIPCHK(c, hipGetLastError());
    return FGL_OK;
}

// InitFrameBuffer (zero), InitDepthBuffer (FLT_MAX), InitGeometryBuffers
// (zero, AO one): recorded, then written by the first pass that covers the
// targets (or materialised before any partial use).
int fgl_init_targets(fgl_ctx *c, int w, int h)
{
    if (!c) return FGL_ERR_INVALID;
    if (w <= 0 || h <= 0 || w > 16384 || h > 16384) return fail(c, FGL_ERR_INVALID, "bad target size");
    int rc = flush_pass(c);
    if (rc) return rc;
    size_t P = (size_t)w * h;
    if ((rc = ensure(c, c->depth, P * 4)) || (rc = ensure(c, c->fb, 3 * P * 4)) ||
        (rc = ensure(c, c->gbuf, kGbufPlanes * P * 4)) || (rc = ensure(c, c->ao, P * 4)) ||
        (rc = ensure(c, c->blur_tmp, P * 4)) || (rc = ensure(c, c->vis, P * 4)))
        return rc;
    c->W = w;
    c->H = h;
    c->pend_depth = c->pend_gbuf = c->pend_fb = c->pend_ao = true;
    c->ao_net = NetAo{};
    c->ao_net_owner = nullptr;
    c->fb_clear[0] = c->fb_clear[1] = c->fb_clear[2] = 0.f;
    return FGL_OK;
}

// One of InitFrameBuffer / InitDepthBuffer / InitGeometryBuffers
// (forkergl.cpp:55-81).  The main targets share one size; initialising one
// at a new size re-initialises all of them.
int fgl_init_buffer(fgl_ctx *c, int which, int w, int h)
{
    if (!c) return FGL_ERR_INVALID;
    if (which == FGL_TARGET_SHADOW) return fgl_init_shadow_buffer(c, w, h);
    if (which != FGL_TARGET_FRAME && which != FGL_TARGET_DEPTH && which != FGL_TARGET_GBUFFER)
        return fail(c, FGL_ERR_INVALID, "fgl_init_buffer: FRAME, DEPTH, GBUFFER or SHADOW");
    if (!c->depth.p || w != c->W || h != c->H) return fgl_init_targets(c, w, h);
    int rc = flush_pass(c);
    if (rc) return rc;
    if (which == FGL_TARGET_FRAME) {
        c->pend_fb = true;
        c->fb_clear[0] = c->fb_clear[1] = c->fb_clear[2] = 0.f;
    } else if (which == FGL_TARGET_DEPTH) {
        c->pend_depth = true;
    } else {
        c->pend_gbuf = c->pend_ao = true;
        c->ao_net = NetAo{};
        c->ao_net_owner = nullptr;
    }
    return FGL_OK;
}

int fgl_init_shadow_buffer(fgl_ctx *c, int w, int h)
{
    if (!c) return FGL_ERR_INVALID;
    if (w <= 0 || h <= 0 || w > 16384 || h > 16384) return fail(c, FGL_ERR_INVALID, "bad shadow size");
    int rc = flush_pass(c);
    if (rc) return rc;
    size_t P = (size_t)w * h;
    if ((rc = ensure(c, c->shadow, P * 4)) || (rc = ensure(c, c->shadow_depth, P * 4))) return rc;
    c->SW = w;
    c->SH = h;
    c->pend_shadow = true;
    return FGL_OK;
}

int fgl_clear_color(fgl_ctx *c, float r, float g, float b)
{
    if (!c || !c->fb.p) return fail(c, FGL_ERR_INVALID, "targets not initialised");
    int rc = flush_pass(c);
    if (rc) return rc;
    c->fb_clear[0] = r;
    c->fb_clear[1] = g;
    c->fb_clear[2] = b;
    c->pend_fb = true;  // overwrites every pixel, so it replaces any pending clear
    return FGL_OK;
}

// The draws recorded so far were issued under the current ForkerGL state
// (in the reference each draw runs at once with its shader's uniforms,
// src/model.cpp:77-90).  A setter that changes state the pending draws read
// (matrices, viewport, eye, lights, shadow settings) rasterises them first,
// so every draw sees the values set before it; setting an equal value keeps
// the pass batched.
static int flush_if_changed(fgl_ctx *c, const void *cur, const void *next, size_t bytes)
{
    if (c->draws.empty() || std::memcmp(cur, next, bytes) == 0) return FGL_OK;
    return flush_pass(c);
}

int fgl_set_viewport(fgl_ctx *c, int shadow, int x, int y, int w, int h)
{
    if (!c) return FGL_ERR_INVALID;
    float vp[16];
    make_viewport(x, y, w, h, vp);
    float *dst = shadow ? c->shadow_viewport : c->viewport;
    int rc = flush_if_changed(c, dst, vp, sizeof(vp));
    if (rc) return rc;
    std::memcpy(dst, vp, sizeof(vp));
    return FGL_OK;
}

int fgl_set_matrix(fgl_ctx *c, int which, const float m[16])
{
    if (!c || !m) return FGL_ERR_INVALID;
    float *dst = which == FGL_MAT_VIEW ? c->view
               : which == FGL_MAT_PROJ ? c->proj
               : which == FGL_MAT_LIGHT_SPACE ? c->light_space : nullptr;
    if (!dst) return fail(c, FGL_ERR_INVALID, "unknown matrix");
    int rc = flush_if_changed(c, dst, m, 16 * sizeof(float));
    if (rc) return rc;
    std::memcpy(dst, m, 16 * sizeof(float));
    return FGL_OK;
}

int fgl_set_eye(fgl_ctx *c, const float eye[3])
{
    if (!c || !eye) return FGL_ERR_INVALID;
    int rc = flush_if_changed(c, c->eye, eye, 3 * sizeof(float));
    if (rc) return rc;
    std::memcpy(c->eye, eye, 3 * sizeof(float));
    return FGL_OK;
}

int fgl_set_lights(fgl_ctx *c, int n, const fgl_light *lights)
{
    if (!c || n < 0 || n > FGL_MAX_LIGHTS || (n && !lights)) return fail(c, FGL_ERR_INVALID, "bad lights");
    if (!c->draws.empty() &&
        (n != (int)c->lights.size() || (n && std::memcmp(c->lights.data(), lights, n * sizeof(fgl_light)) != 0))) {
        int rc = flush_pass(c);
        if (rc) return rc;
    }
    c->lights.assign(lights, lights + n);
    return FGL_OK;
}

int fgl_set_shadow(fgl_ctx *c, int on, const fgl_shadow_cfg *cfg)
{
    if (!c) return FGL_ERR_INVALID;
    if (cfg) {
        if (cfg->mode < 0 || cfg->mode > 5 || cfg->n_blocker < 0 || cfg->n_blocker > FGL_MAX_DISK ||
            cfg->n_pcf < 0 || cfg->n_pcf > FGL_MAX_DISK || (cfg->mode == 3 && (cfg->grid < 1 || cfg->grid > 9)) ||
            ((cfg->mode == 1 || cfg->mode == 2 || cfg->mode >= 4) && cfg->n_pcf < 1) ||
            (cfg->mode == 5 && cfg->n_blocker < 1))
            return fail(c, FGL_ERR_INVALID, "bad shadow filter configuration");
    }
    if (!c->draws.empty() && ((on ? 1 : 0) != c->shadow_on ||
                              (cfg && std::memcmp(cfg, &c->shadow_cfg, sizeof(fgl_shadow_cfg)) != 0))) {
        int rc = flush_pass(c);
        if (rc) return rc;
    }
    if (cfg) c->shadow_cfg = *cfg;
    c->shadow_on = on ? 1 : 0;
    return FGL_OK;
}

static void build_uniforms(fgl_ctx *c, FrameUniforms &u)
{
    std::memset(&u, 0, sizeof(u));
    std::memcpy(u.view, c->view, sizeof(u.view));
    std::memcpy(u.proj, c->proj, sizeof(u.proj));
    std::memcpy(u.light_space, c->light_space, sizeof(u.light_space));
    std::memcpy(u.viewport, c->viewport, sizeof(u.viewport));
    std::memcpy(u.shadow_viewport, c->shadow_viewport, sizeof(u.shadow_viewport));
    mat_mul(c->proj, c->view, u.vp);
    std::memcpy(u.eye, c->eye, sizeof(u.eye));
    u.shadow_on = c->shadow_on;
    u.n_lights = (int)c->lights.size();
    for (int i = 0; i < u.n_lights; ++i) {
        u.lights[i].type = c->lights[i].type;
        std::memcpy(u.lights[i].vec, c->lights[i].vec, sizeof(float) * 3);
        std::memcpy(u.lights[i].color, c->lights[i].color, sizeof(float) * 3);
    }
    u.W = c->W;
    u.H = c->H;
    u.SW = c->SW;
    u.SH = c->SH;
    const fgl_shadow_cfg &s = c->shadow_cfg;
    u.shadow.mode = s.mode;
    u.shadow.grid = s.grid;
    u.shadow.n_blocker = s.n_blocker;
    u.shadow.n_pcf = s.n_pcf;
    u.shadow.pcf_filter = s.pcf_filter;
    u.shadow.area_light = s.area_light;
    u.shadow.blocker_filter = s.blocker_filter;
    u.shadow.pcss_filter = s.pcss_filter;
    std::memcpy(u.shadow.blocker_disk, s.blocker_disk, sizeof(u.shadow.blocker_disk));
    std::memcpy(u.shadow.pcf_disk, s.pcf_disk, sizeof(u.shadow.pcf_disk));
    for (int i = 0; i < 2 * kMaxDisk; ++i)
        u.shadow.blocker_off[i] = (float)((double)s.blocker_disk[i] * s.blocker_filter);
}

static int upload_tables(fgl_ctx *c)
{
    int rc;
    if (c->textures_dirty) {
        if ((rc = ensure(c, c->d_textures, (c->textures.size() + 1) * sizeof(DevTexture)))) return rc;
        if (!c->textures.empty())
            HIPCHK(c, hipMemcpyAsync(c->d_textures.p, c->textures.data(),
                                     c->textures.size() * sizeof(DevTexture),
                                     hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->textures_dirty = false;
    }
    if (c->materials_dirty) {
        if ((rc = ensure(c, c->d_materials, (c->materials.size() + 1) * sizeof(DevMaterial)))) return rc;
        if (!c->materials.empty())
            HIPCHK(c, hipMemcpyAsync(c->d_materials.p, c->materials.data(),
                                     c->materials.size() * sizeof(DevMaterial),
                                     hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->materials_dirty = false;
    }
    return FGL_OK;
}

// ---------------- passes ----------------

static int flush_pass(fgl_ctx *c);

int fgl_begin_pass(fgl_ctx *c, int pass)
{
    if (!c) return FGL_ERR_INVALID;
    if (pass != FGL_PASS_FORWARD && pass != FGL_PASS_GEOMETRY && pass != FGL_PASS_LIGHTING &&
        pass != FGL_PASS_SHADOW)
        return fail(c, FGL_ERR_INVALID, "unknown pass");
    int rc = flush_pass(c);
    if (rc) return rc;
    c->pass = pass;
    return FGL_OK;
}

int fgl_draw_model(fgl_ctx *c, int model, const float model_matrix[16])
{
    if (!c || !model_matrix) return FGL_ERR_INVALID;
    if (model < 0 || model >= (int)c->models.size()) return fail(c, FGL_ERR_INVALID, "unknown model");
    if (c->pass != FGL_PASS_FORWARD && c->pass != FGL_PASS_GEOMETRY && c->pass != FGL_PASS_SHADOW)
        return fail(c, FGL_ERR_INVALID, "no raster pass begun");
    fgl_ctx::Rec r;
    r.model = model;
    std::memcpy(r.m, model_matrix, sizeof(r.m));
    r.has_normal = false;
    c->draws.push_back(r);
    return FGL_OK;
}

int fgl_draw_model_normal(fgl_ctx *c, int model, const float model_matrix[16], const float normal_matrix[9])
{
    if (!c || !normal_matrix) return FGL_ERR_INVALID;
    int rc = fgl_draw_model(c, model, model_matrix);
    if (rc) return rc;
    c->draws.back().has_normal = true;
    std::memcpy(c->draws.back().n, normal_matrix, sizeof(c->draws.back().n));
    return FGL_OK;
}

static int bits_for(uint32_t n)
{
    int b = 0;
    while (b < 32 && (n >> b) != 0) ++b;
    return b;
}

// Draw record of one model instance (bases left zero): Model::Render's
// per-draw state -- model matrix, its normal matrix, light * model.
static void fill_draw(fgl_ctx *c, const ModelRes &m, const float M[16], DevDraw &d)
{
    std::memset(&d, 0, sizeof(d));
    d.pos = m.pos; d.uv = m.uv; d.nrm = m.nrm; d.tan = m.tan;
    d.vi = m.vi; d.ti = m.ti; d.ni = m.ni; d.face_mat = m.fmat;
    d.corner = (const float4 *)m.corner;
    d.n_pos = m.n_pos;
    d.n_faces = m.n_faces;
    d.has_tangents = m.has_tan;
    d.support_pbr = m.pbr;
    std::memcpy(d.model, M, sizeof(d.model));
    make_normal_matrix(d.model, d.normal);
    mat_mul(c->light_space, d.model, d.lm);
}

// The stream's next n vectors (RandomVectorInUnitDisk for D = 2,
// RandomVectorInUnitSphere for D = 3) into out, in stream order, without
// consuming them: rng_take then moves the stream past the first m <= n.  The
// draws are generated in parallel chunks of K blocks, each chunk's array
// reached by a jump from the context's (rng.hip, mt_jump.hpp); the draws
// needed are estimated from the acceptance rate (pi/4 for the disk, pi/6 for
// the sphere) with a margin, and doubled in the rare case that too few tries
// are kept.  A chunk's generation takes time in proportion to K, and the
// jumps (a workgroup each, one per CU) take a fixed time per 256 chunks:
// K = 256 while 256 chunks of it cover the request, else 512.
static constexpr int kMtChunkShort = 256, kMtChunkLong = 512;

static int rng_fill(fgl_ctx *c, int D, float *out, uint32_t n)
{
    int rc;
    c->fill_T = 0;
    c->fill_n = 0;
    c->fill_D = D;
    if (n == 0) return FGL_OK;
    const double keep = D == 2 ? 0.78539816339744831 : 0.52359877559829887;
    uint64_t R = (uint64_t)((double)n * D / keep * 1.01) + 4096;
    for (;;) {
        const int kind = R <= 256ull * 624 * kMtChunkShort ? 0 : 1;
        const uint64_t C = 624ull * (kind ? kMtChunkLong : kMtChunkShort);
        uint64_t Nc = (R + C - 1) / C;
        int K = kind ? kMtChunkLong : kMtChunkShort;
        if (Nc == 1) K = (int)((R + 623) / 624);   // one chunk: only the blocks needed
        const uint64_t len = Nc * (uint64_t)K * 624;
        if (len + 624 >= 0xffffffffull || Nc > 0x7fffffffull)
            return fail(c, FGL_ERR_CAPACITY, "random stream: too many draws in one request");
        const uint32_t T = (uint32_t)(len / (uint64_t)D);
        const uint32_t nb = mt_try_blocks(T);
        if ((rc = ensure(c, c->mt_raw, (len + 624) * 4)) ||
            (rc = ensure(c, c->mt_pos, ((size_t)nb + 1 + scan_tmp_words(nb)) * 4)))
            return rc;
        MtState *st = (MtState *)c->rng.p;
        uint32_t *raw = (uint32_t *)c->mt_raw.p, *pos = (uint32_t *)c->mt_pos.p;
        const uint32_t *states = nullptr;
        if (Nc > 1) {
            const int nj = (int)Nc - 1;
            if (nj > c->mt_polys_n[kind]) {
                std::vector<uint64_t> polys;
                if (!mt_jump_polys(C, nj, polys))
                    return fail(c, FGL_ERR_DEVICE, "random stream: no characteristic polynomial");
                if ((rc = ensure(c, c->mt_polys[kind], polys.size() * sizeof(uint64_t)))) return rc;
                HIPCHK(c, hipMemcpyAsync(c->mt_polys[kind].p, polys.data(), polys.size() * sizeof(uint64_t),
                                         hipMemcpyHostToDevice, c->stream));
                HIPCHK(c, hipStreamSynchronize(c->stream));
                c->mt_polys_n[kind] = nj;
            }
            if ((rc = ensure(c, c->mt_base, (size_t)kMtBaseWords * 4)) ||
                (rc = ensure(c, c->mt_states, (size_t)Nc * 624 * 4)))
                return rc;
            launch_mt_base(st, (uint32_t *)c->mt_base.p, c->stream);
            launch_mt_jump((const uint32_t *)c->mt_base.p, (const uint64_t *)c->mt_polys[kind].p, nj,
                           (uint32_t *)c->mt_states.p, c->stream);
            states = (const uint32_t *)c->mt_states.p;
        }
        launch_mt_gen(st, states, (int)Nc, K, raw, c->stream);
        launch_mt_count(D, raw, T, pos, c->stream);
        launch_exclusive_scan(pos, pos, nb, pos + nb + 1, c->stream);
        HIPCHK(c, hipMemcpyAsync(c->h_total, pos + nb, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->h_total[0] >= n) {
            launch_mt_take(D, raw, T, pos, n, out, c->stream);
            c->fill_T = T;
            c->fill_n = n;
            HIPCHK(c, hipGetLastError());
            return FGL_OK;
        }
        R *= 2;
    }
}

static int rng_take(fgl_ctx *c, uint32_t m)
{
    if (m == 0) return FGL_OK;
    if (m > c->fill_n) return fail(c, FGL_ERR_INVALID, "random stream: taking more vectors than were drawn");
    launch_mt_commit(c->fill_D, (MtState *)c->rng.p, (const uint32_t *)c->mt_raw.p, (const uint32_t *)c->mt_pos.p,
                     c->fill_T, m, c->stream);
    HIPCHK(c, hipGetLastError());
    return FGL_OK;
}

// The stochastic shadow filters' unit-disk samples for n shading points in
// the order they shade -- the lighting pass's pixels in row order, or the
// forward pass's passing fragments in raster order -- drawn from the
// context's stream, which ends past exactly the samples the points use.  g:
// the points' normal, world position and light NDC as planes of stride n
// (read for PCSS only).  ss: point i's samples start at pair ss.offs[i]
// (PCSS; offs[n] = the total) or at i * ss.stride (PCF).
static constexpr uint32_t kSegmentMax = kPcssSegmentMax;   // undecided PCSS points per walk segment, at most

static int stochastic_shadow_samples(fgl_ctx *c, const FrameUniforms &u, const float *g, size_t n, ShadowSamples &ss)
{
    int rc;
    const int mode = c->shadow_cfg.mode;
    const uint32_t nb = (uint32_t)c->shadow_cfg.n_blocker, np = (uint32_t)c->shadow_cfg.n_pcf;
    if (mode == FGL_SHADOW_REF_PCF) {
        const size_t total = n * np;
        if (total > 0xffffffffull) return fail(c, FGL_ERR_CAPACITY, "stochastic shadows: too many samples");
        if ((rc = ensure(c, c->rng_samples, (total > np ? total : np) * 2 * sizeof(float)))) return rc;
        float *pairs = (float *)c->rng_samples.p;
        if ((rc = rng_fill(c, 2, pairs, (uint32_t)total)) || (rc = rng_take(c, (uint32_t)total))) return rc;
        ss = {pairs, nullptr, np};
        return FGL_OK;
    }
    // PCSS takes n_pcf more samples only where its blocker search finds a
    // blocker.  Points the shadow map decides alone are counted in parallel;
    // the most samples the points can need (n_pcf more for every undecided
    // one) are drawn without consuming them; the undecided points are walked
    // in order; then the stream moves past exactly the samples used.
    if (n * (nb + np) > 0xffffffffull) return fail(c, FGL_ERR_CAPACITY, "stochastic shadows: too many samples");
    const size_t words = scan_tmp_words(n + 1), n4 = (n + 4) & ~(size_t)3;
    const size_t seg_words = 2 * (size_t)kSegmentMax + 2 * kPcssChainFbits + kPcssChainRes + 2 * ((size_t)kSegmentMax + 2) + 4;
    if ((rc = ensure(c, c->rng_offs, (6 * n4 + 4 * n4 + seg_words + words) * sizeof(uint32_t)))) return rc;
    uint32_t *base = (uint32_t *)c->rng_offs.p, *upos = base + n4, *list = upos + n4;
    uint32_t *found = list + n4, *offs = found + n4, *pbase = offs + n4;
    float4 *pt = (float4 *)(pbase + n4);
    unsigned long long *masks = (unsigned long long *)(pt + n4);
    unsigned long long *fbits = masks + kSegmentMax;
    uint32_t *res = (uint32_t *)(fbits + kPcssChainFbits);
    uint32_t *pred = res + kPcssChainRes;
    uint32_t *seg_out = pred + 2 * ((size_t)kSegmentMax + 2);
    uint32_t *tmp = seg_out + 4;
    launch_pcss_classify(u, g, n, (const float *)c->shadow.p, base, upos, c->stream);
    launch_exclusive_scan(base, base, n, tmp, c->stream);
    launch_exclusive_scan(upos, upos, n, tmp, c->stream);
    launch_pcss_compact(upos, n, list, c->stream);
    HIPCHK(c, hipMemcpyAsync(c->h_total, base + n, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->h_total + 1, upos + n, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const uint32_t n_unc = c->h_total[1];
    // the decided points' samples and n_blocker per undecided point, plus
    // n_pcf per undecided point: every offset the walk evaluates reads
    // inside it (point k of the undecided ones at most n_pcf k past its base)
    const size_t most = (size_t)c->h_total[0] + (size_t)np * n_unc;
    if ((rc = ensure(c, c->rng_samples, (most > nb + np ? most : nb + np) * 2 * sizeof(float)))) return rc;
    float *pairs = (float *)c->rng_samples.p;
    if ((rc = rng_fill(c, 2, pairs, (uint32_t)most))) return rc;
    // the undecided points in order, a segment at a time (shade.hip): the
    // band of miss counts each point is evaluated at follows the miss rate
    // seen so far
    launch_pcss_points(u, g, n, base, list, n_unc, pt, pbase, c->stream);
    uint32_t s0 = 0, M0 = 0, seg = kSegmentMax / 4;
    float rate = 0.25f;
    while (s0 < n_unc) {
        const uint32_t L = n_unc - s0 < seg ? n_unc - s0 : seg;
        launch_pcss_segment(u, (const float *)c->shadow.p, pairs, pt, pbase, s0, L, M0, rate, pred, tmp, masks, fbits,
                            res, found, seg_out, c->stream);
        HIPCHK(c, hipMemcpyAsync(c->h_total, seg_out, 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const uint32_t s1 = c->h_total[0], M1 = c->h_total[1];
        if (s1 <= s0 || s1 > n_unc) return fail(c, FGL_ERR_DEVICE, "stochastic shadows: the PCSS walk did not advance");
        if (s1 - s0 >= 64) rate = (float)(M1 - M0) / (float)(s1 - s0);
        // the next segment: longer while whole segments settle, shorter when few of it did
        if (s1 - s0 == L && seg < kSegmentMax) seg *= 2;
        else if (4 * (s1 - s0) < L && seg > 1024) seg /= 2;
        s0 = s1;
        M0 = M1;
    }
    launch_exclusive_scan(found, found, n_unc, tmp, c->stream);
    launch_pcss_offsets(base, upos, found, np, n, offs, c->stream);
    HIPCHK(c, hipMemcpyAsync(c->h_total, offs + n, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = rng_take(c, c->h_total[0]))) return rc;
    ss = {pairs, offs, 0};
    return FGL_OK;
}

static int flush_pass(fgl_ctx *c)
{
    if (c->draws.empty()) return FGL_OK;
    const int pass = c->pass == FGL_PASS_SHADOW ? kPassShadow
                   : c->pass == FGL_PASS_GEOMETRY ? kPassGeometry : kPassForward;
    const bool shadow = pass == kPassShadow;
    const int W = shadow ? c->SW : c->W, H = shadow ? c->SH : c->H;
    if (W <= 0 || H <= 0 || (shadow && !c->shadow.p) || (!shadow && !c->depth.p)) {
        c->draws.clear();
        return fail(c, FGL_ERR_INVALID, "render targets not initialised for this pass");
    }
    if (pass == kPassForward && c->shadow_on && !c->shadow.p) {
        c->draws.clear();
        return fail(c, FGL_ERR_INVALID, "forward pass with shadows needs a shadow buffer");
    }
    // The forward shaders with a stochastic shadow filter draw samples for
    // every fragment that passes the depth test, in raster order: the pass
    // replays that sequence (k_frag_walk) to find each winner's samples.
    const bool stoch = pass == kPassForward && c->shadow_on && !c->lights.empty() &&
                       (c->shadow_cfg.mode == FGL_SHADOW_REF_PCF || c->shadow_cfg.mode == FGL_SHADOW_REF_PCSS);
    if (stoch && (size_t)c->W * c->H > 0xffffffffull) {
        c->draws.clear();
        return fail(c, FGL_ERR_CAPACITY, "stochastic shadows in the forward pass: too many pixels");
    }
    int rc = upload_tables(c);
    if (rc) return rc;

    // draw table in submission order
    std::vector<DevDraw> dd(c->draws.size());
    long long nt = 0, nv = 0;
    int max_pos = 0, max_faces = 0;
    for (size_t i = 0; i < c->draws.size(); ++i) {
        const ModelRes &m = c->models[c->draws[i].model];
        DevDraw &d = dd[i];
        fill_draw(c, m, c->draws[i].m, d);
        if (c->draws[i].has_normal) std::memcpy(d.normal, c->draws[i].n, sizeof(d.normal));
        d.vert_base = (int)nv;
        d.tri_base = (int)nt;
        nt += m.n_faces;
        nv += m.n_pos;
        max_pos = m.n_pos > max_pos ? m.n_pos : max_pos;
        max_faces = m.n_faces > max_faces ? m.n_faces : max_faces;
    }
    const int n_draws = (int)c->draws.size();
    c->draws.clear();
    if (nt > 0x7fffffffLL || nv > 0x7fffffffLL) return fail(c, FGL_ERR_CAPACITY, "pass too large");
    // Whether this pass starts from the cleared targets: then it writes
    // every pixel (winner or clear value) and no separate clear runs.
    bool fresh;
    if (shadow) {
        fresh = c->pend_shadow;
    } else if (pass == kPassGeometry) {
        if (c->pend_depth != c->pend_gbuf && (rc = materialize(c, true, true, false, false, false))) return rc;
        fresh = c->pend_depth;
    } else {
        if (c->pend_depth != c->pend_fb && (rc = materialize(c, true, false, true, false, false))) return rc;
        if (c->shadow_on && (rc = materialize(c, false, false, false, false, true))) return rc;
        fresh = c->pend_depth;
    }
    if (nt == 0) return FGL_OK;

    if (dd.size() != c->last_draws.size() ||
        std::memcmp(dd.data(), c->last_draws.data(), dd.size() * sizeof(DevDraw)) != 0) {
        if ((rc = ensure(c, c->d_draws, dd.size() * sizeof(DevDraw)))) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(c->d_draws.p, dd.data(), dd.size() * sizeof(DevDraw), hipMemcpyHostToDevice));
        c->last_draws = dd;
    }

    FrameUniforms u;
    build_uniforms(c, u);
    const int TS = shadow ? c->TS_shadow : c->TS;
    const int tiles_x = (W + TS - 1) / TS, tiles_y = (H + TS - 1) / TS;
    const uint32_t ntiles = (uint32_t)(tiles_x * tiles_y);
    if ((rc = ensure(c, c->screen, (size_t)nv * sizeof(int4))) ||
        (rc = ensure(c, c->rect, (size_t)nt * sizeof(uint2))) ||
        (rc = ensure(c, c->counts, (size_t)nt * 4)) || (rc = ensure(c, c->offs, ((size_t)nt + 1) * 4)) ||
        (rc = ensure(c, c->scan_tmp, scan_tmp_words((size_t)nt) * 4)) ||
        (rc = ensure(c, c->tile_start, (size_t)ntiles * 4)) ||
        (rc = ensure(c, c->tile_end, (size_t)ntiles * 4)) ||
        (rc = ensure(c, c->tri_v, (size_t)nt * sizeof(int4))) ||
        (rc = ensure(c, c->big_list, (size_t)nt * 4)) || (rc = ensure(c, c->counters, 64)) ||
        (shadow && (rc = ensure(c, c->ndc_z, (size_t)nv * 4))))
        return rc;
    const DevDraw *ddraws = (const DevDraw *)c->d_draws.p;
    uint32_t *big_count = (uint32_t *)c->counters.p;
    {
        KernelTimer t(c, "vertex");
        launch_vertex(shadow, ddraws, n_draws, max_pos, u, (int4 *)c->screen.p,
                      shadow ? (float *)c->ndc_z.p : nullptr, big_count, c->stream);
    }
    {
        KernelTimer t(c, "setup");
        launch_setup(ddraws, n_draws, max_faces, (const int4 *)c->screen.p, W, H, TS,
                     (uint2 *)c->rect.p, (uint32_t *)c->counts.p, (int4 *)c->tri_v.p,
                     (uint32_t *)c->big_list.p, big_count, c->stream);
    }
    {
        KernelTimer t(c, "scan");
        launch_exclusive_scan((const uint32_t *)c->counts.p, (uint32_t *)c->offs.p, (size_t)nt,
                              (uint32_t *)c->scan_tmp.p, c->stream);
    }
    uint32_t npairs = 0, n_big = 0;
    if ((rc = read_counts(c, (const uint32_t *)c->offs.p + nt, big_count, npairs, n_big))) return rc;
    if (npairs == 0xffffffffu)
        return fail(c, FGL_ERR_CAPACITY, "the pass has 2^32 - 1 or more (tile, triangle) pairs; use a larger tile");
    if ((rc = ensure(c, c->keys, (size_t)npairs * 4 + 4)) || (rc = ensure(c, c->vals, (size_t)npairs * 4 + 4)) ||
        (rc = ensure(c, c->keys_alt, (size_t)npairs * 4 + 4)) ||
        (rc = ensure(c, c->vals_alt, (size_t)npairs * 4 + 4)) ||
        (rc = ensure(c, c->sort_tmp, radix_tmp_words(npairs) * 4)))
        return rc;
    {
        KernelTimer t(c, "emit");
        launch_emit((const uint32_t *)c->offs.p, (int)nt, (const uint2 *)c->rect.p, tiles_x,
                    (const uint32_t *)c->big_list.p, n_big, (uint32_t *)c->keys.p,
                    (uint32_t *)c->vals.p, (uint32_t *)c->tile_start.p, (uint32_t *)c->tile_end.p,
                    (int)ntiles, c->stream);
    }
    {
        KernelTimer t(c, "sort");
        launch_radix_sort((uint32_t *)c->keys.p, (uint32_t *)c->vals.p, (uint32_t *)c->keys_alt.p,
                          (uint32_t *)c->vals_alt.p, npairs, bits_for(ntiles - 1),
                          (uint32_t *)c->sort_tmp.p, c->stream);
    }
    {
        KernelTimer t(c, "tile_ranges");
        launch_tile_ranges((const uint32_t *)c->keys.p, npairs, nullptr, (uint32_t *)c->tile_start.p,
                           (uint32_t *)c->tile_end.p, c->stream);
    }
    uint32_t n_frag = 0;
    if (stoch) {
        // the passing fragments in raster order: count, scan, emit, sort
        KernelTimer t(c, "frag_walk");
        const size_t P = (size_t)W * H;
        if ((rc = ensure(c, c->frag_cnt, (P + 1 + scan_tmp_words(P)) * 4))) return rc;
        uint32_t *cnt = (uint32_t *)c->frag_cnt.p;
        FragWalkArgs fw;
        std::memset(&fw, 0, sizeof(fw));
        fw.screen = (const int4 *)c->screen.p;
        fw.tri_v = (const int4 *)c->tri_v.p;
        fw.vals = (const uint32_t *)c->vals.p;
        fw.tile_start = (const uint32_t *)c->tile_start.p;
        fw.tile_end = (const uint32_t *)c->tile_end.p;
        fw.W = W;
        fw.H = H;
        fw.tiles_x = tiles_x;
        fw.tiles_y = tiles_y;
        fw.fresh = fresh;
        fw.depth = (const float *)c->depth.p;
        fw.count = cnt;
        launch_frag_walk(false, TS, fw, c->stream);
        launch_exclusive_scan(cnt, cnt, P, cnt + P + 1, c->stream);
        HIPCHK(c, hipMemcpyAsync(c->h_total, cnt + P, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        n_frag = c->h_total[0];
        if (n_frag == 0xffffffffu) return fail(c, FGL_ERR_CAPACITY, "stochastic shadows: too many passing fragments");
        if ((rc = ensure(c, c->frag_tri, (size_t)n_frag * 8 + 8)) ||
            (rc = ensure(c, c->frag_alt, (size_t)n_frag * 8 + 8)) ||
            (rc = ensure(c, c->frag_tmp, radix_tmp_words(n_frag) * 4)))
            return rc;
        uint32_t *ftri = (uint32_t *)c->frag_tri.p, *fq = ftri + n_frag;
        uint32_t *atri = (uint32_t *)c->frag_alt.p, *aq = atri + n_frag;
        fw.base = cnt;
        fw.tri_out = ftri;
        fw.q_out = fq;
        launch_frag_walk(true, TS, fw, c->stream);
        c->frag_sorted_alt = launch_radix_sort_impl(ftri, fq, atri, aq, n_frag, nullptr, bits_for((uint32_t)(nt - 1)),
                                                    (uint32_t *)c->frag_tmp.p, c->stream) != 0;
    }
    RasterArgs a;
    std::memset(&a, 0, sizeof(a));
    a.draws = ddraws;
    a.n_draws = n_draws;
    a.screen = (const int4 *)c->screen.p;
    a.tri_v = (const int4 *)c->tri_v.p;
    a.vals = (const uint32_t *)c->vals.p;
    a.tile_start = (const uint32_t *)c->tile_start.p;
    a.tile_end = (const uint32_t *)c->tile_end.p;
    a.W = W;
    a.H = H;
    a.tiles_x = tiles_x;
    a.tiles_y = tiles_y;
    a.fresh = fresh;
    a.depth = shadow ? (float *)c->shadow_depth.p : (float *)c->depth.p;
    a.vis = (int *)c->vis.p;
    a.shadow_out = (float *)c->shadow.p;
    a.ndc_z = shadow ? (const float *)c->ndc_z.p : nullptr;
    {
        KernelTimer t(c, shadow ? "raster_shadow" : "raster");
        launch_raster(shadow, TS, a, c->stream);
    }
    if (shadow) {
        c->pend_shadow = false;
    } else {
        ResolveArgs r;
        std::memset(&r, 0, sizeof(r));
        r.draws = ddraws;
        r.screen = a.screen;
        r.tri_v = a.tri_v;
        r.vis = a.vis;
        r.W = W;
        r.H = H;
        r.fresh = fresh;
        std::memcpy(r.clear, c->fb_clear, sizeof(r.clear));
        r.gbuf = (float *)c->gbuf.p;
        r.fb = (float *)c->fb.p;
        r.materials = (const DevMaterial *)c->d_materials.p;
        r.textures = (const DevTexture *)c->d_textures.p;
        r.shadow_map = (const float *)c->shadow.p;
        if (pass == kPassGeometry) {
            KernelTimer t(c, "gbuffer_resolve");
            launch_gbuffer_resolve(r, u, c->stream);
            c->pend_gbuf = false;
        } else {
            if (stoch) {
                // each winner's samples: the offset of its fragment in the sequence
                KernelTimer t(c, "shadow_samples");
                const size_t P = (size_t)W * H;
                if ((rc = ensure(c, c->frag_offs_px, P * 4))) return rc;
                uint32_t *offs_px = (uint32_t *)c->frag_offs_px.p;
                launch_fill_u32(offs_px, 0u, P, c->stream);
                uint32_t *ftri = (uint32_t *)(c->frag_sorted_alt ? c->frag_alt.p : c->frag_tri.p), *fq = ftri + n_frag;
                const float *planes = nullptr;
                if (c->shadow_cfg.mode == FGL_SHADOW_REF_PCSS) {
                    if ((rc = ensure(c, c->frag_planes, (size_t)9 * n_frag * sizeof(float) + 4))) return rc;
                    launch_frag_shadow_inputs(r, u, ftri, fq, n_frag, (float *)c->frag_planes.p, c->stream);
                    planes = (const float *)c->frag_planes.p;
                }
                ShadowSamples ss = {nullptr, nullptr, 0};
                if ((rc = stochastic_shadow_samples(c, u, planes, n_frag, ss))) return rc;
                launch_frag_pixel_offsets(ftri, fq, n_frag, W, H, (const int *)c->vis.p, ss.offs, ss.stride, offs_px,
                                          c->stream);
                r.ss = {ss.pairs, offs_px, 0};
            }
            KernelTimer t(c, "forward_resolve");
            launch_forward_resolve(r, u, c->stream);
            c->pend_fb = false;
        }
        c->pend_depth = false;
    }
    HIPCHK(c, hipGetLastError());
    return FGL_OK;
}

int fgl_end_pass(fgl_ctx *c)
{
    if (!c) return FGL_ERR_INVALID;
    return flush_pass(c);
}

int fgl_draw_screen_space_pixels(fgl_ctx *c)
{
    if (!c) return FGL_ERR_INVALID;
    int rc = flush_pass(c);
    if (rc) return rc;
    if (!c->gbuf.p) return fail(c, FGL_ERR_INVALID, "targets not initialised");
    if (c->shadow_on && !c->shadow.p) return fail(c, FGL_ERR_INVALID, "shadows on but no shadow buffer");
    if ((rc = materialize(c, false, true, false, false, c->shadow_on != 0))) return rc;
    FrameUniforms u;
    build_uniforms(c, u);
    ShadowSamples ss = {nullptr, nullptr, 0};
    const int mode = c->shadow_cfg.mode;
    if (c->shadow_on && !c->lights.empty() && (mode == FGL_SHADOW_REF_PCF || mode == FGL_SHADOW_REF_PCSS)) {
        // the filters' unit-disk samples from the stream, pixels in row order
        KernelTimer t(c, "shadow_samples");
        if ((rc = stochastic_shadow_samples(c, u, (const float *)c->gbuf.p, (size_t)c->W * c->H, ss))) return rc;
    }
    {
        KernelTimer t(c, "lighting");
        // a pending network AO is interpolated from its grid by the lighting
        // pass itself (the same values the up pass would write)
        launch_lighting(u, (const float *)c->gbuf.p, (const float *)c->ao.p, c->pend_ao ? 1 : 0, c->ao_net,
                        (const float *)c->shadow.p, ss, (float *)c->fb.p, c->stream);
    }
    c->pend_fb = false;  // every pixel written
    HIPCHK(c, hipGetLastError());
    return FGL_OK;
}

int fgl_ssao(fgl_ctx *c, const fgl_ssao_cfg *cfg)
{
    if (!c || !cfg) return FGL_ERR_INVALID;
    if ((!cfg->stochastic && (cfg->n_kernel < 1 || cfg->n_kernel > 64)) || cfg->blur < 0 || cfg->blur > 2 ||
        (cfg->blur == 2 && (cfg->blur_radius < 1 || cfg->blur_radius > 8)))
        return fail(c, FGL_ERR_INVALID, "bad SSAO configuration");
    int rc = flush_pass(c);
    if (rc) return rc;
    if (!c->gbuf.p) return fail(c, FGL_ERR_INVALID, "targets not initialised");
    if ((rc = materialize(c, true, true, false, false, false))) return rc;
    c->pend_ao = false;  // every pixel written
    c->ao_net = NetAo{};
    c->ao_net_owner = nullptr;
    FrameUniforms u;
    build_uniforms(c, u);
    SsaoArgs sa;
    sa.n_kernel = cfg->n_kernel;
    std::memcpy(sa.kernel, cfg->kernel, sizeof(sa.kernel));
    std::memcpy(sa.noise, cfg->noise, sizeof(sa.noise));
    sa.radius = cfg->radius;
    sa.bias = cfg->bias;
    sa.range = cfg->range;
    sa.kmax = 0.f;
    for (int k = 0; k < 3 * sa.n_kernel; ++k) {
        const float a = std::fabs(sa.kernel[k]);
        if (!(a <= sa.kmax)) sa.kmax = a;   // a NaN component makes kmax NaN
    }
    // the whole AO stage (SSAO and blur, with the gap between them) as one span
    KernelTimer stage(c, "ssao_blur");
    if (cfg->stochastic) {
        // 32 hemisphere samples per pixel, pixels in row order, from the stream
        const size_t n = (size_t)c->W * c->H * 32;
        if (n > 0xffffffffull) return fail(c, FGL_ERR_CAPACITY, "stochastic SSAO: too many samples");
        if ((rc = ensure(c, c->rng_samples, n * 3 * sizeof(float)))) return rc;
        KernelTimer t(c, "ssao_reference");
        if ((rc = rng_fill(c, 3, (float *)c->rng_samples.p, (uint32_t)n)) || (rc = rng_take(c, (uint32_t)n)))
            return rc;
        launch_ssao_reference(u, (const float *)c->rng_samples.p, (const float *)c->ssao_pow3.p,
                              (const float *)c->depth.p, (const float *)c->gbuf.p, (float *)c->ao.p, c->stream);
    } else {
        KernelTimer t(c, "ssao");
        launch_ssao(u, sa, (const float *)c->depth.p, (const float *)c->gbuf.p, (float *)c->ao.p, c->stream);
    }
    if (cfg->blur == 1) {
        KernelTimer t(c, "blur_reference");
        launch_blur_recursive((float *)c->ao.p, c->W, c->H, c->stream);
    } else if (cfg->blur == 2) {
        KernelTimer t(c, "blur_gaussian");
        // one pass into the spare plane, which then becomes the AO target
        launch_blur_separable((const float *)c->ao.p, (float *)c->blur_tmp.p, c->W, c->H,
                                 cfg->blur_weights, cfg->blur_radius, c->stream);
        std::swap(c->ao, c->blur_tmp);
    }
    HIPCHK(c, hipGetLastError());
    return FGL_OK;
}

int fgl_ssao_noalloc(fgl_ctx *c, int w, int h, const float view[16], const float proj[16],
                     const float *depth, const float *gbuf, float *ao, const fgl_ssao_cfg *cfg)
{
    if (!c) return FGL_ERR_INVALID;
    if (!cfg || !view || !proj || !depth || !gbuf || !ao) return fail(c, FGL_ERR_INVALID, "null argument");
    if (w <= 0 || h <= 0) return fail(c, FGL_ERR_INVALID, "bad frame size");
    if ((long long)w * h > FGL_SSAO_NOALLOC_MAX_PIXELS)
        return fail(c, FGL_ERR_CAPACITY, "frame larger than FGL_SSAO_NOALLOC_MAX_PIXELS (" +
                                             std::to_string(FGL_SSAO_NOALLOC_MAX_PIXELS) + " pixels)");
    if (cfg->stochastic)
        return fail(c, FGL_ERR_INVALID, "the stochastic SSAO draws from the context's stream: use fgl_ssao");
    if (cfg->n_kernel < 1 || cfg->n_kernel > 64 || cfg->blur < 0 || cfg->blur > 2 ||
        (cfg->blur == 2 && (cfg->blur_radius < 1 || cfg->blur_radius > 8)))
        return fail(c, FGL_ERR_INVALID, "bad SSAO configuration");
    if (!c->noalloc_scratch) return fail(c, FGL_ERR_DEVICE, "no scratch address for this device");
    FrameUniforms u;
    std::memset(&u, 0, sizeof(u));
    std::memcpy(u.view, view, sizeof(u.view));
    std::memcpy(u.proj, proj, sizeof(u.proj));
    mat_mul(proj, view, u.vp);
    make_viewport(0, 0, w, h, u.viewport);
    u.W = w;
    u.H = h;
    SsaoArgs sa;
    sa.n_kernel = cfg->n_kernel;
    std::memcpy(sa.kernel, cfg->kernel, sizeof(sa.kernel));
    std::memcpy(sa.noise, cfg->noise, sizeof(sa.noise));
    sa.radius = cfg->radius;
    sa.bias = cfg->bias;
    sa.range = cfg->range;
    sa.kmax = 0.f;
    for (int k = 0; k < 3 * sa.n_kernel; ++k) {
        const float a = std::fabs(sa.kernel[k]);
        if (!(a <= sa.kmax)) sa.kmax = a;
    }
    // The separable blur reads the unblurred AO from the scratch array; the
    // SSAO kernel writes every pixel of it first, so nothing is cleared.
    float *raw = cfg->blur == 2 ? c->noalloc_scratch : ao;
    launch_ssao(u, sa, depth, gbuf, raw, c->stream);
    if (cfg->blur == 1)
        launch_blur_recursive(ao, w, h, c->stream);
    else if (cfg->blur == 2)
        launch_blur_separable(raw, ao, w, h, cfg->blur_weights, cfg->blur_radius, c->stream);
    HIPCHK(c, hipGetLastError());
    return FGL_OK;
}

int fgl_resolve(fgl_ctx *c, int k)
{
    if (!c) return FGL_ERR_INVALID;
    int rc = flush_pass(c);
    if (rc) return rc;
    if (!c->fb.p) return fail(c, FGL_ERR_INVALID, "targets not initialised");
    if (k < 1 || k > 16 || c->W / k <= 0 || c->H / k <= 0) return fail(c, FGL_ERR_INVALID, "bad SSAA factor");
    int ow = c->W / k, oh = c->H / k;
    if ((rc = materialize(c, false, false, true, false, false))) return rc;
    if ((rc = ensure(c, c->img, (size_t)ow * oh * 3))) return rc;
    c->img_w = ow;
    c->img_h = oh;
    KernelTimer t(c, "resolve");
    launch_resolve((const float *)c->fb.p, c->W, c->H, k, (uint8_t *)c->img.p, c->stream);
    HIPCHK(c, hipGetLastError());
    return FGL_OK;
}

// ---------------- readback ----------------

static int target_ptr(fgl_ctx *c, int which, void **p, size_t *bytes)
{
    size_t P = (size_t)c->W * c->H, SP = (size_t)c->SW * c->SH;
    switch (which) {
    case FGL_TARGET_FRAME: *p = c->fb.p; *bytes = 3 * P * 4; break;
    case FGL_TARGET_DEPTH: *p = c->depth.p; *bytes = P * 4; break;
    case FGL_TARGET_SHADOW: *p = c->shadow.p; *bytes = SP * 4; break;
    case FGL_TARGET_SHADOW_DEPTH: *p = c->shadow_depth.p; *bytes = SP * 4; break;
    case FGL_TARGET_GBUFFER: *p = c->gbuf.p; *bytes = kGbufPlanes * P * 4; break;
    case FGL_TARGET_AO: *p = c->ao.p; *bytes = P * 4; break;
    case FGL_TARGET_IMAGE: *p = c->img.p; *bytes = (size_t)c->img_w * c->img_h * 3; break;
    default: return fail(c, FGL_ERR_INVALID, "unknown target");
    }
    if (!*p) return fail(c, FGL_ERR_INVALID, "target not initialised");
    return FGL_OK;
}

int fgl_target_bytes(fgl_ctx *c, int which, size_t *out)
{
    if (!c || !out) return FGL_ERR_INVALID;
    void *p;
    return target_ptr(c, which, &p, out);
}

int fgl_read_target(fgl_ctx *c, int which, void *dst, size_t bytes)
{
    if (!c || !dst) return FGL_ERR_INVALID;
    int rc = flush_pass(c);
    if (rc) return rc;
    void *p;
    size_t n;
    if ((rc = target_ptr(c, which, &p, &n))) return rc;
    if (bytes != n) return fail(c, FGL_ERR_INVALID, "size mismatch: target has " + std::to_string(n) + " bytes");
    if ((rc = materialize(c, which == FGL_TARGET_DEPTH, which == FGL_TARGET_GBUFFER,
                          which == FGL_TARGET_FRAME, which == FGL_TARGET_AO,
                          which == FGL_TARGET_SHADOW || which == FGL_TARGET_SHADOW_DEPTH)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(dst, p, n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FGL_OK;
}

int fgl_write_target(fgl_ctx *c, int which, const void *src, size_t bytes)
{
    if (!c || !src) return FGL_ERR_INVALID;
    if (which == FGL_TARGET_IMAGE) return fail(c, FGL_ERR_INVALID, "image target is read-only");
    int rc = flush_pass(c);
    if (rc) return rc;
    void *p;
    size_t n;
    if ((rc = target_ptr(c, which, &p, &n))) return rc;
    if (bytes != n) return fail(c, FGL_ERR_INVALID, "size mismatch: target has " + std::to_string(n) + " bytes");
    // the whole target is replaced; a shadow target's companion is materialised
    if (which == FGL_TARGET_SHADOW || which == FGL_TARGET_SHADOW_DEPTH) {
        if ((rc = materialize(c, false, false, false, false, true))) return rc;
    }
    if (which == FGL_TARGET_DEPTH) c->pend_depth = false;
    if (which == FGL_TARGET_GBUFFER) c->pend_gbuf = false;
    if (which == FGL_TARGET_FRAME) c->pend_fb = false;
    if (which == FGL_TARGET_AO) {
        c->pend_ao = false;
        c->ao_net = NetAo{};
        c->ao_net_owner = nullptr;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(p, src, n, hipMemcpyHostToDevice));
    return FGL_OK;
}

int fgl_enable_timing(fgl_ctx *c, int on)
{
    if (!c) return FGL_ERR_INVALID;
    c->timing = on != 0;
    return FGL_OK;
}

int fgl_time_only(fgl_ctx *c, const char *name)
{
    if (!c) return FGL_ERR_INVALID;
    c->timing_only = name ? name : "";
    return FGL_OK;
}

int fgl_rng_seed(fgl_ctx *c, uint32_t seed)
{
    if (!c) return FGL_ERR_INVALID;
    int rc = flush_pass(c);
    if (rc) return rc;
    c->rng_seed = seed;
    return rng_restart(c);
}

int fgl_rng_vectors(fgl_ctx *c, int kind, uint32_t n, float *out)
{
    if (!c || (kind != 2 && kind != 3) || (n && !out)) return FGL_ERR_INVALID;
    int rc = flush_pass(c);
    if (rc) return rc;
    if (n == 0) return FGL_OK;
    const size_t bytes = (size_t)n * kind * sizeof(float);
    if ((rc = ensure(c, c->rng_samples, bytes))) return rc;
    if ((rc = rng_fill(c, kind, (float *)c->rng_samples.p, n)) || (rc = rng_take(c, n))) return rc;
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, c->rng_samples.p, bytes, hipMemcpyDeviceToHost));
    return FGL_OK;
}

int fgl_rng_draws(fgl_ctx *c, uint64_t *out)
{
    if (!c || !out) return FGL_ERR_INVALID;
    int rc = flush_pass(c);
    if (rc) return rc;
    MtState st;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&st, c->rng.p, sizeof(st), hipMemcpyDeviceToHost));
    *out = st.draws;
    return FGL_OK;
}

int fgl_read_timings(fgl_ctx *c, char (*names)[32], float *ms, int cap, int *out_n)
{
    if (!c || !out_n) return FGL_ERR_INVALID;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int n = 0;
    for (auto &t : c->timed) {
        if (n < cap && names && ms) {
            float v = 0.f;
            (void)hipEventElapsedTime(&v, t.a, t.b);
            std::snprintf(names[n], 32, "%s", t.name.c_str());
            ms[n] = v;
        }
        ++n;
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    c->timed.clear();
    *out_n = n;
    return FGL_OK;
}

int fgl_copy_benchmark(fgl_ctx *c, size_t bytes, int iters, float *ms_per_iter)
{
    if (!c || bytes < 16 || iters < 1) return FGL_ERR_INVALID;
    void *a = nullptr, *b = nullptr;
    HIPCHK(c, hipMalloc(&a, bytes));
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return fail(c, FGL_ERR_NOMEM, "copy buffer");
    }
    launch_fill_f32((float *)a, 1.f, bytes / 4, c->stream);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    launch_copy(a, b, bytes, c->stream);
    (void)hipEventRecord(e0, c->stream);
    for (int i = 0; i < iters; ++i) launch_copy(a, b, bytes, c->stream);
    (void)hipEventRecord(e1, c->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms_per_iter) *ms_per_iter = ms / iters;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return FGL_OK;
}

int fgl_check_recip(fgl_ctx *c, uint64_t *compared, uint64_t *mismatches, uint64_t *fast_path)
{
    if (!c) return FGL_ERR_INVALID;
    int rc = flush_pass(c);
    if (rc) return rc;
    unsigned long long *d = nullptr, h[3] = {0, 0, 0};
    HIPCHK(c, hipMalloc(&d, sizeof(h)));
    hipError_t e = hipMemcpyAsync(d, h, sizeof(h), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        launch_check_recip(d, c->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(c, FGL_ERR_DEVICE, hipGetErrorString(e));
    if (compared) *compared = h[0];
    if (mismatches) *mismatches = h[1];
    if (fast_path) *fast_path = h[2];
    return FGL_OK;
}

int fgl_sort_pairs(fgl_ctx *c, uint32_t *keys, uint32_t *vals, size_t n, int bits)
{
    if (!c || (n && (!keys || !vals)) || n > 0x80000000ull || bits < 1 || bits > 32) return FGL_ERR_INVALID;
    if (n == 0) return FGL_OK;
    int rc = flush_pass(c);
    if (rc) return rc;
    const size_t bytes = n * 4;
    const size_t tmp_bytes = radix_tmp_words(n) * 4;
    void *buf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    const size_t sz[5] = {bytes, bytes, bytes, bytes, tmp_bytes};
    for (int i = 0; i < 5; ++i)
        if (hipMalloc(&buf[i], sz[i]) != hipSuccess) {
            for (int j = 0; j < i; ++j) (void)hipFree(buf[j]);
            return fail(c, FGL_ERR_NOMEM, "sort buffers");
        }
    hipError_t e = hipMemcpyAsync(buf[0], keys, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(buf[1], vals, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        launch_radix_sort((uint32_t *)buf[0], (uint32_t *)buf[1], (uint32_t *)buf[2], (uint32_t *)buf[3], n, bits,
                          (uint32_t *)buf[4], c->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(keys, buf[0], bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(vals, buf[1], bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    for (int i = 0; i < 5; ++i) (void)hipFree(buf[i]);
    if (e != hipSuccess) return fail(c, FGL_ERR_DEVICE, hipGetErrorString(e));
    return FGL_OK;
}

}  // extern "C"

namespace fgl {

int ctx_view(fgl_ctx *c, CtxView *v, bool need_ao)
{
    int rc = flush_pass(c);
    if (rc) return rc;
    if (!c->depth.p) return fail(c, FGL_ERR_INVALID, "targets not initialised");
    if ((rc = materialize(c, true, true, false, need_ao, false))) return rc;
    v->stream = c->stream;
    v->W = c->W;
    v->H = c->H;
    v->depth = (float *)c->depth.p;
    v->gbuf = (float *)c->gbuf.p;
    v->ao = (float *)c->ao.p;
    v->fb = (float *)c->fb.p;
    std::memcpy(v->view, c->view, sizeof(v->view));
    std::memcpy(v->proj, c->proj, sizeof(v->proj));
    std::memcpy(v->viewport, c->viewport, sizeof(v->viewport));
    return FGL_OK;
}

void ctx_ao_written(fgl_ctx *c)
{
    c->pend_ao = false;
    c->ao_net = NetAo{};
    c->ao_net_owner = nullptr;
}

void ctx_ao_from_network(fgl_ctx *c, const NetAo &net, const void *owner)
{
    c->pend_ao = false;
    c->ao_net = net;
    c->ao_net_owner = owner;
}

void ctx_drop_network_ao(fgl_ctx *c, const void *owner)
{
    if (c->ao_net_owner != owner) return;
    c->ao_net = NetAo{};
    c->ao_net_owner = nullptr;
}

int ctx_settle_network_ao(fgl_ctx *c, const void *owner)
{
    if (!c->ao_net.lao || c->ao_net_owner != owner) return FGL_OK;
    return materialize(c, false, false, false, true, false);
}

void ctx_register_nao(fgl_ctx *c, void *n) { c->naos.push_back(n); }

void ctx_unregister_nao(fgl_ctx *c, void *n)
{
    for (size_t k = 0; k < c->naos.size(); ++k)
        if (c->naos[k] == n) {
            c->naos.erase(c->naos.begin() + (long)k);
            break;
        }
}

int ctx_fail(fgl_ctx *c, int code, const char *msg) { return fail(c, code, msg); }

void ctx_uniforms(fgl_ctx *c, FrameUniforms *u) { build_uniforms(c, *u); }

int ctx_draw_record(fgl_ctx *c, int model, const float M[16], DevDraw *out)
{
    if (model < 0 || model >= (int)c->models.size()) return fail(c, FGL_ERR_INVALID, "unknown model");
    fill_draw(c, c->models[model], M, *out);
    return FGL_OK;
}

int ctx_tables(fgl_ctx *c, const DevMaterial **materials, const DevTexture **textures)
{
    int rc = upload_tables(c);
    if (rc) return rc;
    *materials = (const DevMaterial *)c->d_materials.p;
    *textures = (const DevTexture *)c->d_textures.p;
    return FGL_OK;
}

hipStream_t ctx_stream(fgl_ctx *c) { return c->stream; }
bool ctx_timing(fgl_ctx *c) { return c->timing && c->timing_only.empty(); }

CtxTimer::CtxTimer(fgl_ctx *c, const char *name) : impl_(name ? new KernelTimer(c, name) : nullptr) {}
CtxTimer::~CtxTimer() { delete static_cast<KernelTimer *>(impl_); }

}  // namespace fgl

extern "C" {

// ---------------- scenes ----------------

struct fgl_scene {
    HostScene s;
};

int fgl_scene_load(const char *path, fgl_scene **out, char *err, size_t err_len)
{
    return fgl_scene_load_with(path, FGL_WRAP_NONE, FGL_FILTER_NEAREST, out, err, err_len);
}

int fgl_scene_load_with(const char *path, int wrap, int filter, fgl_scene **out, char *err, size_t err_len)
{
    if (!path || !out || wrap < FGL_WRAP_NONE || wrap > FGL_WRAP_CLAMP || filter < FGL_FILTER_NEAREST ||
        filter > FGL_FILTER_LINEAR)
        return FGL_ERR_INVALID;
    *out = nullptr;
    fgl_scene *sc = new fgl_scene();
    std::string e;
    if (!load_scene(path, sc->s, e, wrap, filter)) {
        if (err && err_len) std::snprintf(err, err_len, "%s", e.c_str());
        delete sc;
        return FGL_ERR_IO;
    }
    *out = sc;
    return FGL_OK;
}

void fgl_scene_free(fgl_scene *s) { delete s; }

int fgl_scene_info_get(const fgl_scene *s, fgl_scene_info *out)
{
    if (!s || !out) return FGL_ERR_INVALID;
    *out = s->s.info;
    return FGL_OK;
}

int fgl_scene_model(const fgl_scene *s, int i, fgl_model_desc *d, float mm[16])
{
    if (!s || !d || i < 0 || i >= (int)s->s.models.size()) return FGL_ERR_INVALID;
    const HostModel &m = *s->s.models[i];
    d->n_pos = (int)m.pos.size() / 3;
    d->n_uv = (int)m.uv.size() / 2;
    d->n_nrm = (int)m.nrm.size() / 3;
    d->n_faces = m.n_faces();
    d->pos = m.pos.data();
    d->uv = m.uv.data();
    d->nrm = m.nrm.data();
    d->tan = m.has_tangents ? m.tan.data() : nullptr;
    d->vi = m.vi.data();
    d->ti = m.ti.data();
    d->ni = m.ni.data();
    d->face_material = m.face_mat.data();
    d->support_pbr = m.support_pbr;
    if (mm) std::memcpy(mm, s->s.model_mats[i].data(), 16 * sizeof(float));
    return FGL_OK;
}

int fgl_scene_material(const fgl_scene *s, int i, fgl_material *out)
{
    if (!s || !out || i < 0 || i >= (int)s->s.materials.size()) return FGL_ERR_INVALID;
    *out = s->s.materials[i];
    return FGL_OK;
}

int fgl_scene_texture(const fgl_scene *s, int i, int *w, int *h, int *bpp, int *wrap, int *filter,
                      const uint8_t **bytes)
{
    if (!s || i < 0 || i >= (int)s->s.textures.size()) return FGL_ERR_INVALID;
    const HostTexture &t = s->s.textures[i];
    if (w) *w = t.w;
    if (h) *h = t.h;
    if (bpp) *bpp = t.bpp;
    if (wrap) *wrap = t.wrap;
    if (filter) *filter = t.filter;
    if (bytes) *bytes = t.bytes.data();
    return FGL_OK;
}

static int upload_scene(fgl_ctx *c, fgl_scene *sc)
{
    HostScene &s = sc->s;
    if (s.uploaded_ctx == c) return FGL_OK;
    int rc;
    std::vector<int> tex_id(s.textures.size()), mat_id(s.materials.size());
    for (size_t i = 0; i < s.textures.size(); ++i) {
        const HostTexture &t = s.textures[i];
        if ((rc = fgl_upload_texture(c, t.w, t.h, t.bpp, t.bytes.data(), t.wrap, t.filter, &tex_id[i])))
            return rc;
    }
    for (size_t i = 0; i < s.materials.size(); ++i) {
        fgl_material m = s.materials[i];
        int32_t *maps[] = {&m.map_kd, &m.map_ks, &m.map_ke, &m.map_bump, &m.map_base,
                           &m.map_rough, &m.map_metal, &m.map_ao, &m.map_pbr_ke};
        for (int32_t *p : maps)
            if (*p >= 0) *p = tex_id[*p];
        if ((rc = fgl_upload_material(c, &m, &mat_id[i]))) return rc;
    }
    s.dev_models.clear();
    std::vector<std::pair<const HostModel *, int>> done;
    for (size_t i = 0; i < s.models.size(); ++i) {
        const HostModel *hm = s.models[i].get();
        int id = -1;
        for (auto &p : done)
            if (p.first == hm) id = p.second;
        if (id < 0) {
            std::vector<int32_t> fm(hm->face_mat.size());
            for (size_t f = 0; f < fm.size(); ++f) fm[f] = mat_id[hm->face_mat[f]];
            fgl_model_desc d;
            fgl_scene_model(sc, (int)i, &d, nullptr);
            d.face_material = fm.data();
            if ((rc = fgl_upload_model(c, &d, &id))) return rc;
            done.push_back({hm, id});
        }
        s.dev_models.push_back(id);
    }
    s.uploaded_ctx = c;
    return FGL_OK;
}

int fgl_scene_upload(fgl_ctx *c, fgl_scene *sc, int32_t *model_ids)
{
    if (!c || !sc) return FGL_ERR_INVALID;
    int rc = upload_scene(c, sc);
    if (rc) return rc;
    if (model_ids)
        for (size_t i = 0; i < sc->s.dev_models.size(); ++i) model_ids[i] = sc->s.dev_models[i];
    return FGL_OK;
}

// Render::Render (src/render.cpp:40-58) with the scene's extension settings.
int fgl_render_scene(fgl_ctx *c, fgl_scene *sc)
{
    if (!c || !sc) return FGL_ERR_INVALID;
    int rc;
    if ((rc = upload_scene(c, sc))) return rc;
    const fgl_scene_info &I = sc->s.info;
    // one Render::Render run: the reference's stream starts fresh
    if ((I.ssao_on && I.ssao.stochastic) || (I.shadow_on && I.shadow.mode >= FGL_SHADOW_REF_PCF))
        if ((rc = rng_restart(c))) return rc;
    const int k = I.ssaa_on ? I.ssaa_k : 1;
    if (k < 1) return fail(c, FGL_ERR_INVALID, "bad SSAA factor");
    const int W = I.width * k, H = I.height * k;
    const float ratio = (float)I.width / I.height;
    // Preconfigure (render.cpp:33-38)
    fgl_set_viewport(c, 0, 0, 0, W, H);
    const int S = I.shadow_map_size > 0 ? I.shadow_map_size : 0;
    fgl_set_viewport(c, 1, 0, 0, S ? S : W, S ? S : H);
    fgl_set_lights(c, I.n_lights, I.lights);
    fgl_set_eye(c, I.eye);
    if ((rc = fgl_set_shadow(c, I.shadow_on, &I.shadow))) return rc;
    // DoShadowPass (render.cpp:60-97)
    if (I.shadow_on) {
        if ((rc = fgl_init_shadow_buffer(c, S ? S : W, S ? S : H))) return rc;
        float V[16], P[16], L[16];
        const float origin[3] = {0.f, 0.f, 0.f}, up[3] = {0.f, 1.f, 0.f};
        make_lookat(I.lights[0].vec, origin, up, V);
        make_ortho(-3.0f * ratio, 3.0f * ratio, -3.0f, 3.0f, 0.1f, 20.f, P);
        mat_mul(P, V, L);
        fgl_set_matrix(c, FGL_MAT_LIGHT_SPACE, L);
        fgl_begin_pass(c, FGL_PASS_SHADOW);
        for (size_t i = 0; i < sc->s.models.size(); ++i)
            if ((rc = fgl_draw_model(c, sc->s.dev_models[i], sc->s.model_mats[i].data()))) return rc;
        if ((rc = fgl_end_pass(c))) return rc;
    }
    // camera (render.cpp:106-115)
    float V[16], P[16];
    const float up[3] = {0.f, 1.f, 0.f};
    make_lookat(I.eye, I.look_at, up, V);
    if (I.camera_ortho) make_ortho(-1.f * ratio, 1.f * ratio, -1.f, 1.f, 0.01f, 20.f, P);
    else make_perspective(45.f, ratio, 0.01f, 20.f, P);
    fgl_set_matrix(c, FGL_MAT_VIEW, V);
    fgl_set_matrix(c, FGL_MAT_PROJ, P);
    if ((rc = fgl_init_targets(c, W, H))) return rc;
    if (!I.deferred) {
        // DoForwardPass (render.cpp:99-157)
        if ((rc = fgl_clear_color(c, 0.12f, 0.12f, 0.12f))) return rc;
        fgl_begin_pass(c, FGL_PASS_FORWARD);
        for (size_t i = 0; i < sc->s.models.size(); ++i)
            if ((rc = fgl_draw_model(c, sc->s.dev_models[i], sc->s.model_mats[i].data()))) return rc;
        if ((rc = fgl_end_pass(c))) return rc;
    } else {
        // DoGeometryPass + DoLightingPass (render.cpp:159-212)
        fgl_begin_pass(c, FGL_PASS_GEOMETRY);
        for (size_t i = 0; i < sc->s.models.size(); ++i)
            if ((rc = fgl_draw_model(c, sc->s.dev_models[i], sc->s.model_mats[i].data()))) return rc;
        if ((rc = fgl_end_pass(c))) return rc;
        fgl_begin_pass(c, FGL_PASS_LIGHTING);
        if (I.ssao_on && (rc = fgl_ssao(c, &I.ssao))) return rc;
        if ((rc = fgl_draw_screen_space_pixels(c))) return rc;
    }
    // DoSSAA (render.cpp:291-343); k = 1 quantises only
    return fgl_resolve(c, k);
}

// ---------------- host helpers ----------------

void fgl_make_model_matrix(float tx, float ty, float tz, float r, float s, float out[16])
{
    make_model_matrix(tx, ty, tz, r, s, out);
}
void fgl_make_lookat(const float eye[3], const float center[3], const float up[3], float out[16])
{
    make_lookat(eye, center, up, out);
}
void fgl_make_perspective(float fov, float aspect, float n, float f, float out[16])
{
    make_perspective(fov, aspect, n, f, out);
}
void fgl_make_ortho(float l, float r, float b, float t, float n, float f, float out[16])
{
    make_ortho(l, r, b, t, n, f, out);
}
void fgl_mat_mul(const float a[16], const float b[16], float out[16]) { mat_mul(a, b, out); }
void fgl_make_normal_matrix(const float m[16], float out[9]) { make_normal_matrix(m, out); }

int fgl_tga_write(const char *path, int w, int h, int bpp, const uint8_t *bytes, int rle)
{
    if (!path || !bytes || w <= 0 || h <= 0 || (bpp != 1 && bpp != 3 && bpp != 4)) return FGL_ERR_INVALID;
    return tga_write(path, w, h, bpp, bytes, rle != 0) ? FGL_OK : FGL_ERR_IO;
}

}  // extern "C"
