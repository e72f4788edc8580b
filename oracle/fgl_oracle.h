/*
 * fgl_oracle.h -- CPU restatement of ForkerRenderer's rasterisation and
 * shading arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path
 * is compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (forkerrenderer_amd) never does.
 *
 * Every function restates the reference's own operation order so that float
 * results are bit-identical on x86-64 SSE with -ffp-contract=off:
 *   - vector/matrix products are index-ordered dot products starting from
 *     0.f (src/geometry.h:774-793, 881-888);
 *   - v / s is v * (1/s)  (src/geometry.h:180-185, 336-341, 519-525);
 *   - Barycentric mixes FP32 differences with an FP64 cross product
 *     (src/geometry.cpp:29-56);
 *   - Clamp/Max use std::min/std::max compare semantics (src/utility.h:31-49).
 *
 * Data layout: every render target is a float plane of W*H values indexed
 * x + y*W (row 0 = bottom of the written TGA).  Three-channel targets are
 * three planes (SoA); the reference's AoS Buffer3f holds the same values.
 */
#ifndef FGL_ORACLE_H
#define FGL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Texture as held by the reference after TGAImage::ReadTgaFile and the
 * loader's vertical flip (src/model.cpp:411-427): bpp bytes per texel in
 * B,G,R[,A] order, texel (x,y) at (x + y*w)*bpp. */
typedef struct {
    int w, h, bpp;
    const unsigned char *data;
    int wrap;   /* 0 NoWrap, 1 Repeat, 2 MirroredRepeat, 3 ClampToEdge (src/texture.h:15-21) */
    int filter; /* 0 Nearest, 1 Linear (src/texture.h:23-27) */
} orc_texture;

/* Material + PBRMaterial pair built from one .mtl "newmtl" entry
 * (src/materials/material.h:13-41, pbrmaterial.h:13-45).  Map fields are
 * indices into the texture table, -1 when absent. */
typedef struct {
    float ka[3], kd[3], ks[3], ke[3];
    int map_kd, map_ks, map_ke, map_bump;
    float pbr_ke[3], albedo[3], roughness, metalness;
    int map_base, map_rough, map_metal, map_ao, map_pbr_ke;
} orc_material;

/* One Mesh::Draw call: the mesh's face-vertex index streams
 * (src/mesh.h:53-61; tangent index == position index, src/model.cpp:430-476),
 * the owning model's attribute arrays and flags, and the model matrix. */
typedef struct {
    int n_faces;
    const int *vi, *ti, *ni;          /* 3*n_faces each */
    const float *pos, *uv, *nrm, *tan;/* tan may be NULL when !has_tangents */
    int has_tangents, support_pbr, material;
    float model[16];                  /* row-major */
    float normal[9];                  /* MakeNormalMatrix(model), row-major */
} orc_draw;

typedef struct {
    int type;         /* 0 point (shadow-casting), 1 directional */
    float vec[3];     /* position (point) or direction of travel (dir) */
    float color[3];
} orc_light;

/* Shadow visibility filter.  mode 0 is the reference's hard shadow
 * (src/shaders/shadow.cpp:38-45); mode 1 is PCF over a fixed disk sample set
 * (the reference's PCF with its RNG samples replaced by a given set,
 * shadow.cpp:47-63); mode 2 is PCSS with fixed blocker/PCF sets
 * (shadow.cpp:65-106); mode 3 is an n x n texel-grid PCF (new); modes 4 and
 * 5 are the reference's PCF and PCSS drawing their disk samples from the
 * reference's random stream (ref_rng.cpp; the disk arrays are unused). */
typedef struct {
    int mode;
    int grid;                     /* mode 3: n (odd) */
    int n_blocker, n_pcf;
    const float *blocker_disk;    /* 2*n_blocker unit-disk samples */
    const float *pcf_disk;        /* 2*n_pcf unit-disk samples */
    float pcf_filter;             /* mode 1 filter size (float, as passed to PCF) */
    double blocker_filter;        /* PCSS_BLOCKER_SEARCH_FILTER_SIZE (double literal) */
    double pcss_filter;           /* PCF_FILTER_SIZE as used by PCSS (double literal) */
    float area_light;             /* AREA_LIGHT_SIZE (float literal) */
} orc_shadow_cfg;

typedef struct {
    float view[16], proj[16], light_space[16];
    float viewport[16];          /* main pass ForkerGL::SetViewportMatrix */
    float shadow_viewport[16];   /* shadow pass viewport (reference: same as main) */
    float eye[3];
    int shadow_on;
    orc_shadow_cfg shadow;
    int n_lights;
    orc_light lights[4];
} orc_frame;

/* ---- the reference's random stream (ref_rng.cpp; src/utility.h:90-103,
 * src/geometry.h:947-990) ---- */
void orc_rng_seed(unsigned int seed);
unsigned long long orc_rng_draws(void);
void orc_rng_raw(int n, unsigned int *out);
void orc_rng_uniform(int n, float *out);
void orc_rng_sphere(int n, float *out);
void orc_rng_disk(int n, float *out);
void orc_rng_hemisphere(const float n[3], float out[3]);

/* ---- L0 math (src/geometry.h, src/geometry.cpp) ---- */
void orc_mat_mul(const float *a, const float *b, float *out);
void orc_make_model_matrix(float tx, float ty, float tz, float roty_deg, float scale, float *out);
void orc_make_lookat(const float *eye, const float *center, const float *up, float *out);
void orc_make_perspective(float fov, float aspect, float n, float f, float *out);
void orc_make_ortho(float l, float r, float b, float t, float n, float f, float *out);
void orc_make_normal_matrix(const float *m, float *out9);
void orc_make_viewport(int x, int y, int w, int h, float *out);
int  orc_trunc_int(float v);
int  orc_barycentric(int ax, int ay, int bx, int by, int cx, int cy, int px, int py, float *out3);

/* ---- texture (src/texture.h) ---- */
void  orc_texture_sample(const orc_texture *t, float u, float v, float *out3);
float orc_texture_sample_float(const orc_texture *t, float u, float v);

/* ---- raster passes (src/forkergl.cpp, src/render.cpp) ---- */
void orc_shadow_pass(const orc_draw *draws, int n_draws, const orc_frame *fr,
                     int W, int H, float *depth, float *shadow);
void orc_geometry_pass(const orc_draw *draws, int n_draws, const orc_material *mats,
                       const orc_texture *texs, const orc_frame *fr, int W, int H,
                       float *depth, float *const *gbuf /* 19 planes */);
void orc_forward_pass(const orc_draw *draws, int n_draws, const orc_material *mats,
                      const orc_texture *texs, const orc_frame *fr, int W, int H,
                      const float *shadow_map, int SW, int SH,
                      float *depth, float *const *fb /* 3 planes */);

/* ---- screen-space passes ---- */
void orc_lighting_pass(const orc_frame *fr, int W, int H, const float *const *gbuf,
                       const float *ao, const float *shadow_map, int SW, int SH,
                       float *const *fb);
float orc_shadow_visibility(const orc_frame *fr, const float *shadow_map, int SW, int SH,
                            const float *ndc3, const float *n3, const float *l3);
void orc_shade_blinn_phong(const float *L, const float *H, const float *N, float vis,
                           const float *diffuse, const float *emissive, const float *param,
                           const float *light_color, int shadow_on, float *out3);
void orc_shade_pbr(const float *L, const float *V, const float *Hv, const float *N, float vis,
                   const float *albedo, const float *emissive, const float *param,
                   const float *radiance, int shadow_on, float *out3);
/* Render::DoSSAO with the reference's random stream (consumes 32 hemisphere
 * samples per pixel from it) */
void orc_ssao_reference(const orc_frame *fr, int W, int H, const float *depth,
                        const float *const *nrm, const float *const *pos, float *ao);
void orc_ssao_kernel(const orc_frame *fr, int W, int H, const float *depth,
                     const float *const *normal3, const float *const *pos3,
                     const float *kernel, int n_kernel, const float *noise16,
                     float radius, float bias, float range, float *ao);
void orc_set_rows(int y0, int y1);
void orc_blur_recursive(float *a, int W, int H);
void orc_blur_separable(float *a, int W, int H, const float *weights, int R, float *tmp);
void orc_quantize_rgb(const float *const *fb, int W, int H, unsigned char *out_bgr);
void orc_quantize_grey(const float *plane, int W, int H, unsigned char *out);
void orc_ssaa_resolve(const unsigned char *in_bgr, int W, int H, int k, unsigned char *out_bgr);

#ifdef __cplusplus
}
#endif
#endif
