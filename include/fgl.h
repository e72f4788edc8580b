/*
 * fgl.h -- C ABI of the MI355X raster/shading path ("forkergl on gfx950").
 *
 * Plain pointers and sizes only; no C++ or HIP types cross this boundary.
 * Every entry point returns an int status (FGL_OK = 0, negative = error
 * kind) and records a message readable through fgl_last_error().
 *
 * The ABI replaces the static state and draw calls of the reference's
 * `struct ForkerGL` (src/forkergl.h:15-81), the per-face submission loop of
 * `Mesh::Draw` (src/mesh.cpp:10-25) and the per-pixel loops of the frame
 * driver (src/render.cpp:214-343).  Each declaration names the reference
 * interface it replaces.
 *
 * Threading: one host thread per context.  Work is queued on the context's
 * HIP stream; fgl_read_target / fgl_sync synchronise.  Contexts are
 * independent.
 */
#ifndef FGL_H
#define FGL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define FGL_OK 0
#define FGL_ERR_INVALID (-1)   /* bad argument, wrong state */
#define FGL_ERR_DEVICE (-2)    /* HIP runtime failure / no device */
#define FGL_ERR_IO (-3)        /* file missing or malformed */
#define FGL_ERR_NOMEM (-4)     /* device or host allocation failed */
#define FGL_ERR_CAPACITY (-5)  /* a fixed capacity was exceeded */

/* ---- enums mirroring the reference ---- */
/* ForkerGL::PassType (src/forkergl.h:25-31) */
#define FGL_PASS_FORWARD 0
#define FGL_PASS_GEOMETRY 1
#define FGL_PASS_LIGHTING 2
#define FGL_PASS_SHADOW 3
/* Texture::WrapMode / FilterMode (src/texture.h:15-27) */
#define FGL_WRAP_NONE 0
#define FGL_WRAP_REPEAT 1
#define FGL_WRAP_MIRROR 2
#define FGL_WRAP_CLAMP 3
#define FGL_FILTER_NEAREST 0
#define FGL_FILTER_LINEAR 1
/* Matrices held by ForkerGL / passed as shader uniforms */
#define FGL_MAT_VIEW 0
#define FGL_MAT_PROJ 1
#define FGL_MAT_LIGHT_SPACE 2
/* Shadow filters (src/shaders/shadow.h:14-23; modes 1-3 take explicit sample sets) */
#define FGL_SHADOW_HARD 0
#define FGL_SHADOW_PCF_DISK 1
#define FGL_SHADOW_PCSS 2
#define FGL_SHADOW_PCF_GRID 3
/* The reference's own filters with their samples drawn from the reference's
 * random stream (SOFT_SHADOW_PCF / SOFT_SHADOW_PCSS, src/shaders/shadow.h:14-23,
 * shadow.cpp:47-106): n_pcf (and n_blocker) RandomVectorInUnitDisk samples per
 * shading point from the context's std::mt19937 stream; the disk arrays are
 * unused.  The deferred lighting pass draws them for every pixel in row
 * order; the forward pass for every fragment that passes the depth test when
 * drawn (overdrawn ones included), triangles in submission order, px outer,
 * py inner (src/forkergl.cpp:169-204).  Batched views refuse them. */
#define FGL_SHADOW_REF_PCF 4
#define FGL_SHADOW_REF_PCSS 5
/* Render targets (ForkerGL's public buffers, src/forkergl.h:41-54) */
#define FGL_TARGET_FRAME 0        /* FrameBuffer: 3 float planes            */
#define FGL_TARGET_DEPTH 1        /* DepthBuffer: 1 float plane             */
#define FGL_TARGET_SHADOW 2       /* ShadowBuffer: 1 float plane (SW x SH)  */
#define FGL_TARGET_GBUFFER 3      /* Normal, WorldPos, LightSpaceNDC, Albedo,
                                     Emissive, Param (3 planes each),
                                     ShadingType (1 plane): 19 planes       */
#define FGL_TARGET_AO 4           /* AmbientOcclusionGBuffer: 1 plane       */
#define FGL_TARGET_IMAGE 5        /* 8-bit B,G,R image after quantise/SSAA  */
#define FGL_TARGET_SHADOW_DEPTH 6 /* depth buffer of the shadow pass        */

#define FGL_MAX_LIGHTS 4
#define FGL_MAX_DISK 64

typedef struct fgl_ctx fgl_ctx;

/* Material + PBRMaterial pair from one "newmtl" (src/materials/material.h:13-41,
 * pbrmaterial.h:13-45).  Texture ids are from fgl_upload_texture, -1 if absent. */
typedef struct {
    float ka[3], kd[3], ks[3], ke[3];
    int32_t map_kd, map_ks, map_ke, map_bump;
    float pbr_ke[3], albedo[3], roughness, metalness;
    int32_t map_base, map_rough, map_metal, map_ao, map_pbr_ke;
} fgl_material;

/* A loaded Model (src/model.h:15-67): attribute arrays plus the faces of all
 * its meshes concatenated in draw order (std::map order of mesh names,
 * src/model.cpp:77-90), one material id per face. */
typedef struct {
    int32_t n_pos, n_uv, n_nrm, n_faces;
    const float *pos;       /* n_pos*3 */
    const float *uv;        /* n_uv*2 */
    const float *nrm;       /* n_nrm*3 */
    const float *tan;       /* n_pos*3, NULL when the model has no tangents */
    const int32_t *vi, *ti, *ni;   /* n_faces*3 each: face-vertex indices */
    const int32_t *face_material;  /* n_faces */
    int32_t support_pbr;           /* Model::SupportPBR (src/model.cpp:52-68) */
} fgl_model_desc;

typedef struct {
    int32_t type;       /* 0 point (casts the shadow map), 1 directional */
    float vec[3];       /* position, or direction of travel */
    float color[3];
} fgl_light;

typedef struct {
    int32_t mode;               /* FGL_SHADOW_* */
    int32_t grid;               /* PCF_GRID: n (odd, <= 9) */
    int32_t n_blocker, n_pcf;   /* <= FGL_MAX_DISK */
    float blocker_disk[2 * FGL_MAX_DISK];
    float pcf_disk[2 * FGL_MAX_DISK];
    float pcf_filter;           /* PCF_DISK filter size (float)   */
    double blocker_filter;      /* PCSS blocker search size (FP64 literal in the reference) */
    double pcss_filter;         /* PCF_FILTER_SIZE used by PCSS (FP64 literal) */
    float area_light;           /* AREA_LIGHT_SIZE */
} fgl_shadow_cfg;

typedef struct {
    int32_t n_kernel;           /* <= 64 */
    float kernel[3 * 64];
    float noise[3 * 16];        /* 4 x 4 tile */
    float radius, bias, range;
    int32_t blur;               /* 0 none, 1 reference recursive 9-tap, 2 separable */
    int32_t blur_radius;        /* separable: taps = 2*R+1, R <= 8 */
    float blur_weights[9];
    int32_t stochastic;         /* 1: Render::DoSSAO exactly (render.cpp:214-289): 32
                                   RandomVectorInHemisphere samples per pixel from the
                                   context's random stream, radius 0.075, bias 0.0005,
                                   range 0.01; the kernel fields are unused */
} fgl_ssao_cfg;

/* ---- context ---- */
int fgl_ctx_create(int device, fgl_ctx **out);
void fgl_ctx_destroy(fgl_ctx *ctx);
const char *fgl_last_error(const fgl_ctx *ctx);
int fgl_sync(fgl_ctx *ctx);
/* Raster tuning: tile edge in pixels (8, 16 or 32).  Results do not depend
 * on it (each pixel's depth resolve is a minimum over (depth, submission id)
 * keys, whatever the tile). */
int fgl_set_tile_size(fgl_ctx *ctx, int tile);
/* Per-pass tile edge (FGL_PASS_SHADOW, or FORWARD/GEOMETRY for the main pass);
 * the autotuner (forkerrenderer_amd/autotune.py) sets these. */
int fgl_set_pass_tile_size(fgl_ctx *ctx, int pass, int tile);
int fgl_get_pass_tile_size(fgl_ctx *ctx, int pass, int *tile);

/* ---- resources (replace Model::Load's in-memory products) ---- */
/* Texture (src/texture.h:29-39): bpp 1/3/4, bytes in B,G,R[,A] order after
 * the loader's flips, texel (x,y) at (x + y*w)*bpp. */
int fgl_upload_texture(fgl_ctx *ctx, int w, int h, int bpp, const uint8_t *bytes, int wrap,
                       int filter, int *out_id);
int fgl_upload_material(fgl_ctx *ctx, const fgl_material *m, int *out_id);
int fgl_upload_model(fgl_ctx *ctx, const fgl_model_desc *d, int *out_id);

/* ---- state (ForkerGL statics, src/forkergl.cpp:55-132) ---- */
/* InitFrameBuffer/InitDepthBuffer/InitGeometryBuffers (forkergl.cpp:55-81) */
int fgl_init_targets(fgl_ctx *ctx, int w, int h);
/* InitShadowBuffer (forkergl.cpp:65-68); the reference uses the frame size */
int fgl_init_shadow_buffer(fgl_ctx *ctx, int w, int h);
/* InitFrameBuffer / InitDepthBuffer / InitGeometryBuffers one at a time
 * (forkergl.cpp:55-81): which = FGL_TARGET_FRAME, _DEPTH, _GBUFFER (with AO) or _SHADOW */
int fgl_init_buffer(fgl_ctx *ctx, int which, int w, int h);
/* ClearColor (forkergl.cpp:84-87) */
int fgl_clear_color(fgl_ctx *ctx, float r, float g, float b);
/* SetViewportMatrix (forkergl.cpp:89-102); shadow != 0 sets the shadow pass viewport */
int fgl_set_viewport(fgl_ctx *ctx, int shadow, int x, int y, int w, int h);
/* SetViewProjectionMatrix / SetLightSpaceMatrix and the shaders' uView/uProjection
 * uniforms (forkergl.cpp:109-127; render.cpp:106-131); row-major 4x4 */
int fgl_set_matrix(fgl_ctx *ctx, int which, const float m[16]);
/* uEyePos / uPointLight (src/shaders/phongshader.h:26-27) generalised to a light list */
int fgl_set_eye(fgl_ctx *ctx, const float eye[3]);
int fgl_set_lights(fgl_ctx *ctx, int n, const fgl_light *lights);
/* Shadow::SetShadowStatus + compile-time filter selection (src/shaders/shadow.h:11-45) */
int fgl_set_shadow(fgl_ctx *ctx, int on, const fgl_shadow_cfg *cfg);

/* ---- drawing ---- */
/* ForkerGL::SetPassType (forkergl.cpp:129-132). */
int fgl_begin_pass(fgl_ctx *ctx, int pass);
/* Model::Render(shader) -> Mesh::Draw -> ProcessVertex x3 -> DrawTriangle per face
 * (src/model.cpp:77-90, src/mesh.cpp:10-25, src/forkergl.cpp:239-261).  The shader
 * kind follows the pass: DepthShader (shadow), GShader (geometry), Blinn-Phong or
 * PBR by Model::SupportPBR (forward).  Draws are recorded and rasterised in
 * submission order by fgl_end_pass (or by any read), all draws of a pass in one
 * batched launch sequence. */
int fgl_draw_model(fgl_ctx *ctx, int model, const float model_matrix[16]);
/* The same draw with the shader's uNormalMatrix given by the caller (row-major
 * 3x3; src/shaders/gshader.h:26, phongshader.h:26, pbrshader.h:26, applied to
 * normals and tangents in ProcessVertex, gshader.h:50-54).  fgl_draw_model
 * uses MakeNormalMatrix(model_matrix) (src/geometry.cpp:60-68), the value
 * Render::* sets (src/render.cpp:129,146,187). */
int fgl_draw_model_normal(fgl_ctx *ctx, int model, const float model_matrix[16], const float normal_matrix[9]);
int fgl_end_pass(fgl_ctx *ctx);
/* ForkerGL::DrawScreenSpacePixels (forkergl.cpp:326-380): deferred lighting */
int fgl_draw_screen_space_pixels(fgl_ctx *ctx);
/* Render::DoSSAO (render.cpp:214-289) with a deterministic sample kernel, then the
 * configured blur (Buffer1f::TwoPassGaussianBlurDenoised, buffer.cpp:59-98, or separable) */
int fgl_ssao(fgl_ctx *ctx, const fgl_ssao_cfg *cfg);
/* Allocation-free variant of fgl_ssao.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The inputs
 * and the output are the caller's device buffers of one w x h frame:
 *   depth  w*h floats (the geometry pass's depth, ForkerGL::DepthBuffer),
 *   gbuf   the G-buffer planes laid out as fgl_read_target returns them
 *          (plane p at gbuf + p*w*h); planes 0-5 (normal, world position)
 *          are read,
 *   ao     w*h floats, written.
 * view / proj are the frame's row-major camera matrices; the viewport is
 * (0, 0, w, h), as Render::Render sets it.  cfg as fgl_ssao; the stochastic
 * mode is refused (it draws from the context's random stream).  The
 * unblurred AO that the separable blur reads is kept in a static device
 * array of FGL_SSAO_NOALLOC_MAX_PIXELS floats (32 MiB on every device the
 * library runs on, held until the process ends); larger frames fail with
 * FGL_ERR_CAPACITY.  Every call on a device shares that array: calls that
 * overlap on one device (two contexts, two host threads) race.  Work is
 * enqueued on the context's stream; nothing is allocated or freed, and the
 * context's targets and pending draws are left as they are.  Results equal
 * fgl_ssao's on the same frame bit for bit. */
#define FGL_SSAO_NOALLOC_MAX_PIXELS (1 << 23)
int fgl_ssao_noalloc(fgl_ctx *ctx, int w, int h, const float view[16], const float proj[16],
                     const float *depth, const float *gbuf, float *ao, const fgl_ssao_cfg *cfg);
/* The reference's random stream (src/utility.h:90-103: one function-static
 * std::mt19937, default seed 5489, through uniform_real_distribution<float>;
 * random vectors src/geometry.h:947-990).  The context owns one such stream on
 * the device; stochastic SSAO and the FGL_SHADOW_REF_* filters consume it in
 * the reference's order, and fgl_render_scene of a scene that draws from it
 * restarts it at the seed first (one Render::Render run per process in the
 * reference).  fgl_rng_seed sets the
 * seed and restarts the stream; fgl_rng_draws reports the 32-bit draws
 * consumed since the restart. */
int fgl_rng_seed(fgl_ctx *ctx, uint32_t seed);
int fgl_rng_draws(fgl_ctx *ctx, uint64_t *out_draws);
/* n results of RandomVectorInUnitSphere (kind 3: x, y, z each) or
 * RandomVectorInUnitDisk (kind 2: x, y each; src/geometry.h:957-976), drawn
 * from the context's stream on the device, copied to host memory `out`. */
int fgl_rng_vectors(fgl_ctx *ctx, int kind, uint32_t n, float *out);
/* Buffer3f::GenerateImage (buffer.cpp:113-126) + Render::DoSSAA (render.cpp:291-343):
 * k = 1 quantises only; the image is (W/k) x (H/k) B,G,R bytes. */
int fgl_resolve(fgl_ctx *ctx, int k);

/* ---- readback / upload of targets ---- */
/* Copies a target to host memory.  `bytes` must equal the target's size:
 * planes * W * H * 4 for float targets, (W/k)*(H/k)*3 for FGL_TARGET_IMAGE. */
int fgl_read_target(fgl_ctx *ctx, int which, void *dst, size_t bytes);
/* Overwrites a float target (tests feed G-buffers, depth or AO directly). */
int fgl_write_target(fgl_ctx *ctx, int which, const void *src, size_t bytes);
/* Size in bytes of a target as it is now. */
int fgl_target_bytes(fgl_ctx *ctx, int which, size_t *out_bytes);

/* ---- per-pass kernel timing (HIP events on the context stream) ---- */
/* Returns the number of timed kernels recorded since the last call; names and
 * milliseconds are written to the arrays (up to cap entries). */
int fgl_enable_timing(fgl_ctx *ctx, int on);
int fgl_read_timings(fgl_ctx *ctx, char (*names)[32], float *ms, int cap, int *out_n);
/* With timing on, records events only around the kernels timed under `name`
 * (NULL or "" = all).  Every pair of events costs the stream a few
 * microseconds, so a throughput run times only the kernel it reports. */
int fgl_time_only(fgl_ctx *ctx, const char *name);

/* Diagnostic with no counterpart in the reference: a plain device-to-device
 * copy of `bytes` run `iters` times, used to calibrate HBM byte counters and
 * bandwidth figures.  Allocates and frees its own buffers. */
int fgl_copy_benchmark(fgl_ctx *ctx, size_t bytes, int iters, float *ms_per_iter);

/* Diagnostic with no counterpart in the reference: the binning's device
 * radix sort on caller arrays (n <= 2^31 key/value pairs, keys < 2^bits,
 * bits 1..32): a stable sort of the pairs by key, in place.  Allocates and
 * frees its own buffers. */
int fgl_sort_pairs(fgl_ctx *ctx, uint32_t *keys, uint32_t *vals, size_t n, int bits);

/* Diagnostic with no counterpart in the reference: the kernels' correctly
 * rounded reciprocal (hardware reciprocal + one fma Newton step for |w| in
 * [2^-125, 2^125], the IEEE division elsewhere) compared with the IEEE
 * quotient 1.f / w for every non-NaN float w on this device.  Writes the
 * number of floats compared, of mismatches, and of floats on the fast path. */
int fgl_check_recip(fgl_ctx *ctx, uint64_t *compared, uint64_t *mismatches, uint64_t *fast_path);

/* ---- scene files: .scene/.obj/.mtl/.tga (src/scene.cpp, src/model.cpp, src/tgaimage.cpp) ---- */
typedef struct fgl_scene fgl_scene;
typedef struct {
    int32_t width, height, ssaa_on, ssaa_k, ssao_on, shadow_on, deferred;
    int32_t camera_ortho;
    float eye[3], look_at[3];
    int32_t n_lights;
    fgl_light lights[FGL_MAX_LIGHTS];
    int32_t n_models, n_materials, n_textures;
    /* extensions (extra .scene keys the reference parser ignores) */
    int32_t shadow_map_size;      /* 0: frame size (reference) */
    fgl_shadow_cfg shadow;
    fgl_ssao_cfg ssao;
} fgl_scene_info;

int fgl_scene_load(const char *path, fgl_scene **out, char *err, size_t err_len);
/* Scene(filename) with ForkerGL::TextureWrapping / TextureFiltering set to wrap
 * / filter when the scene is loaded (src/forkergl.cpp:15-16, 44-52): every
 * texture Model::Load reads takes those modes (src/model.cpp:425-426) unless
 * a `texture` key of the scene sets others.  fgl_scene_load uses the
 * reference's defaults, FGL_WRAP_NONE and FGL_FILTER_NEAREST. */
int fgl_scene_load_with(const char *path, int wrap, int filter, fgl_scene **out, char *err, size_t err_len);
/* Uploads the scene's textures, materials and models into the context (once
 * per context; fgl_render_scene does it on first use); model_ids (n_models
 * entries, may be NULL) receives the context's id of each of the scene's
 * model lines. */
int fgl_scene_upload(fgl_ctx *ctx, fgl_scene *s, int32_t *model_ids);
void fgl_scene_free(fgl_scene *s);
int fgl_scene_info_get(const fgl_scene *s, fgl_scene_info *out);
/* Model i as uploaded (pointers stay valid until fgl_scene_free) and its model matrix */
int fgl_scene_model(const fgl_scene *s, int i, fgl_model_desc *out, float model_matrix[16]);
int fgl_scene_material(const fgl_scene *s, int i, fgl_material *out);
/* Texture i: size, bpp, wrap, filter and a pointer to its bytes */
int fgl_scene_texture(const fgl_scene *s, int i, int *w, int *h, int *bpp, int *wrap,
                      int *filter, const uint8_t **bytes);
/* Render::Render (render.cpp:40-58) on the GPU: shadow pass, forward or
 * geometry + [SSAO + blur] + lighting, then quantise/SSAA. Uploads the scene
 * on first use. */
int fgl_render_scene(fgl_ctx *ctx, fgl_scene *s);

/* ---- neural ambient occlusion (new; no counterpart in the reference) ----
 * A BF16 convolutional network that replaces Render::DoSSAO + blur
 * (src/render.cpp:214-289).  It runs on a grid `scale` times coarser than
 * the frame (1, 2 or 4): per network pixel, 8 inputs from its scale x scale
 * block (the mean view-space depth, normal and position of the covered
 * pixels, and the covered fraction) -> 4 x conv3x3 (32 ch, ReLU, dilated
 * 1/2/6/12) -> conv1x1 + sigmoid, interpolated bilinearly back to the covered
 * frame pixels (AO 1 elsewhere).  Training runs on the device (forward,
 * backward, Adam) on network-grid crops of frames the plain pipeline
 * rendered, against the block means of their AO. */
typedef struct fgl_nao fgl_nao;
int fgl_nao_param_count(void);
/* batch crops of crop x crop network pixels per training step */
int fgl_nao_create(fgl_ctx *ctx, int batch, int crop, int scale, fgl_nao **out);
void fgl_nao_destroy(fgl_nao *n);
/* state = [params | Adam m | Adam v] (3 * param_count floats); with_adam = 0
 * loads params only and resets Adam */
int fgl_nao_set_state(fgl_nao *n, const float *state, int with_adam, int step);
int fgl_nao_get_state(fgl_nao *n, float *state, int *step);
/* crops (origins xy[2*i], in network pixels) of the context's current frame
 * -- depth, G-buffer and AO target -- into batch slots slot0.. */
int fgl_nao_stage_crops(fgl_nao *n, int slot0, int count, const int *xy);
/* one Adam step on the staged batch; stats = [sum sq err, sum |err|, covered] */
int fgl_nao_train_step(fgl_nao *n, float lr, float *stats);
int fgl_nao_gradients(fgl_nao *n, float *grads, float *stats);
int fgl_nao_read_batch(fgl_nao *n, uint16_t *feat_bf16, float *target, uint8_t *mask);
/* neural AO of the context's current frame into its AO target */
int fgl_nao_infer(fgl_nao *n);
/* fgl_nao_infer's kernels: 1 (default) two fused kernels (layers 1-2, layers
 * 3-4 + head; activations of layers 1 and 3 stay on chip), 0 one kernel per
 * layer (the training forward's); both give the same bits */
int fgl_nao_set_fused(fgl_nao *n, int fused);

/* ---- views and batched views (new; no counterpart in the reference) ----
 * A view is one deferred frame of Render::Render (src/render.cpp:24-58)
 * without SSAA: optional shadow pass (light-space ortho +-3*aspect, near 0.1,
 * far 20, looking from lights[0] at the origin), geometry pass (perspective
 * 45 deg, 0.01..20), ambient occlusion (none or neural), lighting
 * (ForkerGL::DrawScreenSpacePixels) and quantisation, over a list of model
 * instances already uploaded into the context. */
#define FGL_AO_NONE 0
#define FGL_AO_NEURAL 2
typedef struct {
    int32_t width, height;     /* frame size (pixels) */
    int32_t shadow_size;       /* shadow map edge (square); used when shadow_on */
    int32_t shadow_on;
    float eye[3], look_at[3];
    int32_t n_lights;
    fgl_light lights[FGL_MAX_LIGHTS];
    fgl_shadow_cfg shadow;
    int32_t n_draws;
    const int32_t *models;     /* n_draws model ids */
    const float *matrices;     /* n_draws row-major 4x4 model matrices */
} fgl_view_desc;

/* One view through the context's ForkerGL calls (the single-view path);
 * nao is used when ao_mode == FGL_AO_NEURAL.  Results in the context targets. */
int fgl_render_view(fgl_ctx *ctx, const fgl_view_desc *view, int ao_mode, fgl_nao *nao);

/* A batch renders many views of different sizes and meshes together: every
 * pass is one launch over all views (persistent rasteriser workgroups pull
 * (view, tile) items from a device queue), the pass sequence is captured in
 * a HIP graph and replayed.  Results equal fgl_render_view's bit for bit. */
typedef struct fgl_batch fgl_batch;
int fgl_batch_create(fgl_ctx *ctx, int n_views, const fgl_view_desc *views, int ao_mode, fgl_nao *nao,
                     fgl_batch **out);
void fgl_batch_destroy(fgl_batch *b);
/* raster tile edges (8, 16, 32) of the shadow and geometry passes */
int fgl_batch_set_tiles(fgl_batch *b, int shadow_tile, int main_tile);
/* persistent rasteriser workgroups per compute unit (1..16) */
int fgl_batch_set_occupancy(fgl_batch *b, int wg_per_cu);
/* (tile, triangle) pair capacity per pass; 0 = measured + margin.  A pass
 * that overflows is detected by fgl_batch_sync and re-rendered. */
int fgl_batch_set_capacity(fgl_batch *b, uint32_t shadow_pairs, uint32_t main_pairs);
int fgl_batch_use_graph(fgl_batch *b, int on);
/* enqueues one render of every view (graph replay when enabled) */
int fgl_batch_render(fgl_batch *b);
/* waits; re-renders when a pass overflowed its capacity.  *rerenders (optional)
 * receives the number of such re-renders. */
int fgl_batch_sync(fgl_batch *b, int *rerenders);
/* target of one view, as fgl_read_target (FRAME, DEPTH, GBUFFER, AO, IMAGE,
 * SHADOW, SHADOW_DEPTH) */
int fgl_batch_read(fgl_batch *b, int view, int which, void *dst, size_t bytes);
/* stats[0..9]: (tile, triangle) pairs of the last measured render (shadow,
 * main), pair capacities (shadow, main), triangles (shadow, main), pixels,
 * views, persistent rasteriser workgroups, main-pass tiles */
int fgl_batch_stats(fgl_batch *b, uint64_t stats[10]);

/* ---- host matrix builders (src/geometry.cpp:60-179), shared by the CLI and tests ---- */
void fgl_make_model_matrix(float tx, float ty, float tz, float roty_deg, float scale, float out[16]);
void fgl_make_lookat(const float eye[3], const float center[3], const float up[3], float out[16]);
void fgl_make_perspective(float fov, float aspect, float n, float f, float out[16]);
void fgl_make_ortho(float l, float r, float b, float t, float n, float f, float out[16]);
void fgl_mat_mul(const float a[16], const float b[16], float out[16]);
/* MakeNormalMatrix (src/geometry.cpp:60-68): upper 3x3 of the inverse transpose */
void fgl_make_normal_matrix(const float m[16], float out[9]);

/* ---- TGA codec (src/tgaimage.cpp:43-300) ---- */
int fgl_tga_write(const char *path, int w, int h, int bpp, const uint8_t *bytes, int rle);

#ifdef __cplusplus
}
#endif
#endif
