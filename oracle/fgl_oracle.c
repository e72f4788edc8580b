/*
 * fgl_oracle.c -- CPU restatement of ForkerRenderer's rasterisation and
 * shading path (TEST INFRASTRUCTURE: the checker, never the product).
 *
 * Build: gcc -std=c11 -O2 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 * No -march=native, no fast-math: every float op is a separately rounded
 * IEEE single (or double where the reference promotes) operation, matching
 * the reference compiled for x86-64 SSE.
 *
 * Each function cites the reference file:line whose arithmetic it restates.
 * Where the reference is undefined (out-of-bounds reads, NaN indices) the
 * oracle defines a behaviour and says so in a comment; the HIP kernels
 * implement the same definition.
 */
#include "fgl_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Row window for timing bounded samples: raster and screen passes only
 * produce pixels with y in [g_y0, g_y1).  Default: every row. */
static int g_y0 = 0, g_y1 = INT_MAX;
void orc_set_rows(int y0, int y1) { g_y0 = y0; g_y1 = y1; }

typedef struct { float x, y; } V2;
typedef struct { float x, y, z; } V3;
typedef struct { float x, y, z, w; } V4;

/* ---------------- scalar helpers (src/utility.h:26-103) ---------------- */

/* std::max(a,b) == (a < b) ? b : a ; std::min(a,b) == (b < a) ? b : a */
static inline float smax(float a, float b) { return (a < b) ? b : a; }
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline int imax(int a, int b) { return (a < b) ? b : a; }
static inline int imin(int a, int b) { return (b < a) ? b : a; }
/* Clamp(v,lo,hi) = std::min(hi, std::max(v, lo))  (utility.h:31-35) */
static inline float clampf(float v, float lo, float hi) { return smin(hi, smax(v, lo)); }
static inline int clampi(int v, int lo, int hi) { return imin(hi, imax(v, lo)); }

/* float -> int conversion as x86 cvttss2si: truncation, and INT_MIN for
 * NaN or values outside [-2^31, 2^31) (the reference's Point2i(coord.x,...)
 * at src/forkergl.cpp:246-248 compiles to it). */
int orc_trunc_int(float v)
{
    if (v >= -2147483648.0f && v < 2147483648.0f) return (int)v;
    return INT_MIN;
}

/* ---------------- vector primitives (src/geometry.h) ---------------- */

/* Dot: ret = 0.f; ret += a[i]*b[i] in index order (geometry.h:881-888) */
static inline float dot3(V3 a, V3 b)
{
    float r = 0.f;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
static inline float dot4(V4 a, V4 b)
{
    float r = 0.f;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    r += a.w * b.w;
    return r;
}
static inline V3 v3(float x, float y, float z) { V3 r = {x, y, z}; return r; }
static inline V3 add3(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub3(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 mul3(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V3 scale3(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
/* v / s == v * ((float)1 / s)  (geometry.h:336-341) */
static inline V3 divs3(V3 a, float s) { float inv = 1.f / s; return scale3(a, inv); }
/* Length = sqrt((x*x + y*y) + z*z) (geometry.h:370-371) */
static inline float len3(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
/* Normalize(v) = v * 1 / Length(v)  (geometry.h:906-910) */
static inline V3 norm3(V3 a) { return divs3(scale3(a, 1.f), len3(a)); }
static inline V3 cross3(V3 a, V3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline V3 pow3(V3 a, float p) { return v3(powf(a.x, p), powf(a.y, p), powf(a.z, p)); }
static inline V3 clamp01_3(V3 a)
{
    return v3(clampf(a.x, 0.f, 1.f), clampf(a.y, 0.f, 1.f), clampf(a.z, 0.f, 1.f));
}
/* vector Lerp: (1 - t) * v1 + t * v2 (geometry.h:912-916) */
static inline V3 lerp3(float t, V3 a, V3 b) { return add3(scale3(a, 1 - t), scale3(b, t)); }

static inline V4 row4(const float *m, int r) { V4 v = {m[4 * r], m[4 * r + 1], m[4 * r + 2], m[4 * r + 3]}; return v; }
static inline V3 row3(const float *m, int r) { return v3(m[3 * r], m[3 * r + 1], m[3 * r + 2]); }
/* Matrix x Vector: row dots (geometry.h:774-782) */
static inline V4 mv4(const float *m, V4 v)
{
    V4 o;
    o.x = dot4(row4(m, 0), v);
    o.y = dot4(row4(m, 1), v);
    o.z = dot4(row4(m, 2), v);
    o.w = dot4(row4(m, 3), v);
    return o;
}
static inline V3 mv3(const float *m9, V3 v)
{
    return v3(dot3(row3(m9, 0), v), dot3(row3(m9, 1), v), dot3(row3(m9, 2), v));
}
/* Vector4 *= s and /= s (geometry.h:512-534) */
static inline V4 scale4(V4 a, float s) { V4 r = {a.x * s, a.y * s, a.z * s, a.w * s}; return r; }
static inline V4 divs4(V4 a, float s) { float inv = 1.f / s; return scale4(a, inv); }

/* Matrix x Matrix: ret[i][j] = Dot(lhs.Row(i), rhs.Col(j)) (geometry.h:784-793) */
void orc_mat_mul(const float *a, const float *b, float *out)
{
    float t[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            V4 col = {b[j], b[4 + j], b[8 + j], b[12 + j]};
            t[4 * i + j] = dot4(row4(a, i), col);
        }
    memcpy(out, t, sizeof(t));
}

/* Cofactor-expansion determinant along row 0, recursive, with the int
 * sign factor applied to the minor's determinant (geometry.h:580-627). */
static float det_n(const float *m, int n)
{
    if (n == 1) return m[0];
    float ret = 0;
    float minor[16];
    for (int i = 0; i < n; ++i) {
        int k = 0;
        for (int r = 1; r < n; ++r)
            for (int c = 0; c < n; ++c)
                if (c != i) minor[k++] = m[r * n + c];
        float cof = det_n(minor, n - 1) * (float)((i % 2) ? -1 : 1);
        ret += m[i] * cof;
    }
    return ret;
}
static float cofactor_n(const float *m, int n, int row, int col)
{
    float minor[16];
    int k = 0;
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c)
            if (r != row && c != col) minor[k++] = m[r * n + c];
    return det_n(minor, n - 1) * (float)(((row + col) % 2) ? -1 : 1);
}

/* MakeNormalMatrix: upper 3x3 of Inverse(M)^T where Inverse = Adjugate / (Float)Det
 * and Adjugate = transpose of the cofactor matrix (geometry.cpp:60-68,
 * geometry.h:705-716).  Inverse(M)^T[i][j] = Cof[i][j] * (1/det). */
void orc_make_normal_matrix(const float *m, float *out9)
{
    float det = det_n(m, 4);
    float inv = (float)1 / det;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out9[3 * i + j] = cofactor_n(m, 4, i, j) * inv;
}

static void identity(float *m) { memset(m, 0, 16 * sizeof(float)); m[0] = m[5] = m[10] = m[15] = 1.f; }

/* MakeModelMatrix = T * R_y * S (geometry.cpp:92-108).  The unqualified
 * cos/sin(float) resolve to the C library's double versions under
 * libstdc++'s <cmath>; the results are narrowed to float on store. */
void orc_make_model_matrix(float tx, float ty, float tz, float roty_deg, float scale, float *out)
{
    const float Pi = (float)3.14159265358979323846;
    float S[16], R[16], T[16], TR[16];
    identity(S); S[0] = S[5] = S[10] = scale;
    identity(R);
    float rad = roty_deg * Pi / 180.f;
    R[0] = (float)cos((double)rad);
    R[2] = (float)sin((double)rad);
    R[8] = (float)(-sin((double)rad));
    R[10] = (float)cos((double)rad);
    identity(T); T[3] = tx; T[7] = ty; T[11] = tz; T[15] = 1.f;
    orc_mat_mul(T, R, TR);
    orc_mat_mul(TR, S, out);
}

/* MakeLookAtMatrix (geometry.cpp:110-128) */
void orc_make_lookat(const float *eye, const float *center, const float *up, float *out)
{
    V3 e = v3(eye[0], eye[1], eye[2]), c = v3(center[0], center[1], center[2]);
    V3 u = v3(up[0], up[1], up[2]);
    V3 front = norm3(sub3(c, e));
    V3 right = norm3(cross3(front, u));
    V3 upv = norm3(cross3(right, front));
    float R[16], T[16];
    identity(R);
    R[0] = right.x; R[1] = right.y; R[2] = right.z; R[3] = 0.f;
    R[4] = upv.x; R[5] = upv.y; R[6] = upv.z; R[7] = 0.f;
    R[8] = -front.x; R[9] = -front.y; R[10] = -front.z; R[11] = 0.f;
    identity(T);
    T[3] = -e.x; T[7] = -e.y; T[11] = -e.z; T[15] = 1.f;
    orc_mat_mul(R, T, out);
}

/* MakePerspectiveMatrix(fov, aspect, n, f) (geometry.cpp:130-145);
 * std::tan(float) is tanf. */
void orc_make_perspective(float fov, float aspect, float n, float f, float *out)
{
    const float Pi = (float)3.14159265358979323846;
    float t = tanf((fov / 2.f) * Pi / 180.f);
    identity(out);
    out[0] = 1.f / (aspect * t);
    out[5] = 1.f / t;
    out[10] = -(f + n) / (f - n);
    out[11] = -2 * f * n / (f - n);
    out[14] = -1;
}

/* MakeOrthographicMatrix (geometry.cpp:166-179) */
void orc_make_ortho(float l, float r, float b, float t, float n, float f, float *out)
{
    identity(out);
    out[0] = 2.f / (r - l);
    out[5] = 2.f / (t - b);
    out[10] = -2.f / (f - n);
    out[3] = -(r + l) / (r - l);
    out[7] = -(t + b) / (t - b);
    out[11] = -(f + n) / (f - n);
    out[15] = 1.f;
}

/* ForkerGL::SetViewportMatrix (forkergl.cpp:89-102) */
void orc_make_viewport(int x, int y, int w, int h, float *out)
{
    identity(out);
    out[0] = w / 2.f;
    out[5] = h / 2.f;
    out[3] = x + w / 2.f;
    out[7] = y + h / 2.f;
    out[10] = 1 / 2.f;
    out[11] = 1 / 2.f;
}

/* Barycentric (geometry.cpp:20-56): FP32 coordinate differences widened to
 * double, FP64 cross product, 1.0/c.z, narrowing to float.  Returns 1 and
 * the three weights if all are >= 0, else 0. */
int orc_barycentric(int ax, int ay, int bx, int by, int cx, int cy, int px, int py, float *out3)
{
    float A[2] = {(float)ax, (float)ay}, B[2] = {(float)bx, (float)by};
    float C[2] = {(float)cx, (float)cy}, P[2] = {(float)px, (float)py};
    double s[2][3];
    for (int i = 0; i < 2; ++i) {
        s[i][0] = (double)(B[i] - A[i]);
        s[i][1] = (double)(C[i] - A[i]);
        s[i][2] = (double)(A[i] - P[i]);
    }
    double bx_ = s[0][1] * s[1][2] - s[0][2] * s[1][1];
    double by_ = s[0][2] * s[1][0] - s[0][0] * s[1][2];
    double bz_ = s[0][0] * s[1][1] - s[0][1] * s[1][0];
    if (fabs(bz_) > 1e-2) {
        double inv = 1.0 / bz_;
        bx_ *= inv;
        by_ *= inv;
        float r0 = (float)(1.0 - (bx_ + by_)), r1 = (float)bx_, r2 = (float)by_;
        if (r0 < 0.f || r1 < 0.f || r2 < 0.f) return 0;
        out3[0] = r0; out3[1] = r1; out3[2] = r2;
        return 1;
    }
    return 0;
}

/* ---------------- textures (src/texture.h:41-145) ---------------- */

static V3 texel(const orc_texture *t, int x, int y)
{
    /* TGAImage::Get: black outside the image (tgaimage.cpp:304-311);
     * Color3(color.r, color.g, color.b) with bgra byte order; a grey
     * texel sets only b (tgaimage.h:173). */
    if (x < 0 || y < 0 || x >= t->w || y >= t->h) return v3(0.f, 0.f, 0.f);
    const unsigned char *p = t->data + ((size_t)x + (size_t)y * t->w) * t->bpp;
    unsigned char b = p[0], g = t->bpp > 1 ? p[1] : 0, r = t->bpp > 2 ? p[2] : 0;
    return v3((float)r, (float)g, (float)b);
}

static V2 wrap_coord(const orc_texture *t, float u, float v)
{
    V2 r;
    if (t->wrap == 1) {           /* Repeat */
        r.x = u - floorf(u);
        r.y = v - floorf(v);
    } else if (t->wrap == 2) {    /* MirroredRepeat */
        int xi = orc_trunc_int(floorf(u)), yi = orc_trunc_int(floorf(v));
        float rx = u - (float)xi, ry = v - (float)yi;
        r.x = (xi % 2 == 0) ? rx : 1.f - rx;
        r.y = (yi % 2 == 0) ? ry : 1.f - ry;
    } else if (t->wrap == 3) {    /* ClampToEdge */
        r.x = clampf(u, 0.f, 1.f);
        r.y = clampf(v, 0.f, 1.f);
    } else {
        r.x = u;
        r.y = v;
    }
    return r;
}

static V3 filter_color(const orc_texture *t, V2 c)
{
    float w = (float)(t->w - 0.001), h = (float)(t->h - 0.001);
    if (t->filter == 1) {
        float px = c.x * w, py = c.y * h;
        float tlx = floorf(px - 0.5f), tly = floorf(py - 0.5f);
        float tx = px - (tlx + 0.5f);
        float ty = py - (tly + 0.5f);
        int s0x = orc_trunc_int(tlx), s0y = orc_trunc_int(tly);
        int s1x = orc_trunc_int(tlx + 1.f), s1y = s0y;
        int s2x = s0x, s2y = orc_trunc_int(tly + 1.f);
        int s3x = s1x, s3y = s2y;
        if (t->wrap != 0) {
            s0x = clampi(s0x, 0, t->w - 1); s0y = clampi(s0y, 0, t->h - 1);
            s1x = clampi(s1x, 0, t->w - 1); s1y = clampi(s1y, 0, t->h - 1);
            s2x = clampi(s2x, 0, t->w - 1); s2y = clampi(s2y, 0, t->h - 1);
            s3x = clampi(s3x, 0, t->w - 1); s3y = clampi(s3y, 0, t->h - 1);
        }
        V3 c0 = texel(t, s0x, s0y), c1 = texel(t, s1x, s1y);
        V3 c2 = texel(t, s2x, s2y), c3 = texel(t, s3x, s3y);
        V3 cx1 = lerp3(tx, c0, c1), cx2 = lerp3(tx, c2, c3);
        return lerp3(ty, cx1, cx2);
    }
    return texel(t, orc_trunc_int(floorf(c.x * w)), orc_trunc_int(floorf(c.y * h)));
}

void orc_texture_sample(const orc_texture *t, float u, float v, float *out3)
{
    V3 c = divs3(filter_color(t, wrap_coord(t, u, v)), 255.f);
    out3[0] = c.x; out3[1] = c.y; out3[2] = c.z;
}

float orc_texture_sample_float(const orc_texture *t, float u, float v)
{
    return filter_color(t, wrap_coord(t, u, v)).z / 255.f;
}

static V3 tex_sample(const orc_texture *texs, int id, V2 uv)
{
    float o[3];
    orc_texture_sample(&texs[id], uv.x, uv.y, o);
    return v3(o[0], o[1], o[2]);
}
static float tex_samplef(const orc_texture *texs, int id, V2 uv)
{
    return orc_texture_sample_float(&texs[id], uv.x, uv.y);
}

/* ---------------- vertex stage ---------------- */

/* Per-triangle state that the reference keeps in the shader object between
 * ProcessVertex and ProcessFragment (e.g. phongshader.h:11-32). */
typedef struct {
    V4 ndc[3];
    V3 ws[3];       /* world position * 1/w */
    V2 uv[3];
    V3 nws[3];
    V3 tws[3];
    V3 lndc[3];
    float oow[3];
    V3 depth_ndc[3];/* DepthShader vPositionNDC columns */
} TriVaryings;

/* DepthShader::ProcessVertex (depthshader.h:21-28) */
static void depth_vertex(const orc_draw *d, const float *LM, int f, int k, TriVaryings *tv)
{
    int vi = d->vi[3 * f + k];
    V4 p = {d->pos[3 * vi], d->pos[3 * vi + 1], d->pos[3 * vi + 2], 1.f};
    V4 cs = mv4(LM, p);
    V4 ndc = divs4(cs, cs.w);
    tv->depth_ndc[k] = v3(ndc.x, ndc.y, ndc.z);
    tv->ndc[k] = ndc;
}

/* BlinnPhongShader / PBRShader / GShader::ProcessVertex, identical bodies
 * (phongshader.h:35-85, pbrshader.h:35-85, gshader.h:41-92), with the
 * normal/tangent normalised on fetch (mesh.cpp:45-56). */
static void main_vertex(const orc_draw *d, const orc_frame *fr, int f, int k, TriVaryings *tv)
{
    int vi = d->vi[3 * f + k], ti = d->ti[3 * f + k], ni = d->ni[3 * f + k];
    V4 p = {d->pos[3 * vi], d->pos[3 * vi + 1], d->pos[3 * vi + 2], 1.f};
    V4 ws = mv4(d->model, p);
    V4 vs = mv4(fr->view, ws);
    V4 cs = mv4(fr->proj, vs);
    V2 uv = {d->uv[2 * ti], d->uv[2 * ti + 1]};
    V3 n = norm3(v3(d->nrm[3 * ni], d->nrm[3 * ni + 1], d->nrm[3 * ni + 2]));
    V3 nws = mv3(d->normal, n);
    V3 tws = v3(0.f, 0.f, 0.f);
    if (d->has_tangents) {
        V3 t = norm3(v3(d->tan[3 * vi], d->tan[3 * vi + 1], d->tan[3 * vi + 2]));
        tws = mv3(d->normal, t);
    }
    V4 l = {0.f, 0.f, 0.f, 0.f};
    if (fr->shadow_on) {
        l = mv4(fr->light_space, ws);
        l = divs4(l, l.w);
    }
    float oow = 1.f / cs.w;
    tv->oow[k] = oow;
    ws = scale4(ws, oow);
    uv.x *= oow; uv.y *= oow;
    nws = scale3(nws, oow);
    if (d->has_tangents) tws = scale3(tws, oow);
    if (fr->shadow_on) l = scale4(l, oow);
    tv->uv[k] = uv;
    tv->ws[k] = v3(ws.x, ws.y, ws.z);
    tv->nws[k] = nws;
    if (d->has_tangents) tv->tws[k] = tws;
    if (fr->shadow_on) tv->lndc[k] = v3(l.x, l.y, l.z);
    tv->ndc[k] = divs4(cs, cs.w);
}

/* Matrix3x3 (columns = vertices) * bary: row dots (phongshader.h:94-97) */
static inline V3 interp3(const V3 *c, const float *b)
{
    V3 bb = v3(b[0], b[1], b[2]);
    return v3(dot3(v3(c[0].x, c[1].x, c[2].x), bb), dot3(v3(c[0].y, c[1].y, c[2].y), bb),
              dot3(v3(c[0].z, c[1].z, c[2].z), bb));
}
static inline V2 interp2(const V2 *c, const float *b)
{
    V3 bb = v3(b[0], b[1], b[2]);
    V2 r = {dot3(v3(c[0].x, c[1].x, c[2].x), bb), dot3(v3(c[0].y, c[1].y, c[2].y), bb)};
    return r;
}

/* ---------------- fragment helpers ---------------- */

typedef struct {
    V3 pos, normal, lndc;
    V2 uv;
} FragGeom;

/* Interpolation, PCI and TBN normal mapping shared by all three fragment
 * programs (gshader.h:95-141, phongshader.h:90-128, pbrshader.h:90-129). */
static FragGeom frag_geometry(const orc_draw *d, const orc_material *m, const orc_texture *texs,
                              const orc_frame *fr, const TriVaryings *tv, const float *b,
                              int normal_map)
{
    FragGeom g;
    V3 pws = interp3(tv->ws, b);
    V2 uv = interp2(tv->uv, b);
    V3 nws = interp3(tv->nws, b);
    V3 bb = v3(b[0], b[1], b[2]);
    float w = 1.f / dot3(v3(tv->oow[0], tv->oow[1], tv->oow[2]), bb);
    pws = scale3(pws, w);
    uv.x *= w; uv.y *= w;
    nws = scale3(nws, w);
    V3 N = norm3(nws);
    V3 normal;
    if (d->has_tangents && normal_map >= 0) {
        V3 tws = interp3(tv->tws, b);
        tws = scale3(tws, w);
        V3 T = norm3(add3(tws, v3(0.001f, 0.001f, 0.001f)));
        T = norm3(sub3(T, scale3(N, dot3(T, N))));
        V3 B = norm3(cross3(N, T));
        V3 s = tex_sample(texs, normal_map, uv);
        s = norm3(sub3(scale3(s, 2.f), v3(1.f, 1.f, 1.f)));
        /* TBN rows are (T_i, B_i, N_i) */
        V3 r0 = v3(T.x, B.x, N.x), r1 = v3(T.y, B.y, N.y), r2 = v3(T.z, B.z, N.z);
        normal = norm3(v3(dot3(r0, s), dot3(r1, s), dot3(r2, s)));
    } else {
        normal = N;
    }
    g.pos = pws;
    g.normal = normal;
    g.uv = uv;
    g.lndc = v3(0.f, 0.f, 0.f);
    if (fr->shadow_on) g.lndc = scale3(interp3(tv->lndc, b), w);
    (void)m;
    return g;
}

/* ---------------- shadows (src/shaders/shadow.cpp) ---------------- */

/* SampleShadowMap (shadow.cpp:23-35).  A NaN coordinate is treated like an
 * out-of-range one (the reference would index with INT_MIN: undefined). */
static float sample_shadow(const float *map, int SW, int SH, float u, float v)
{
    if (!(u >= 0.f && u <= 1.f && v >= 0.f && v <= 1.f)) return INFINITY;
    int w = (int)((float)SW - 0.001f);
    int h = (int)((float)SH - 0.001f);
    int x = (int)((float)w * u);
    int y = (int)((float)h * v);
    float d = map[(size_t)x + (size_t)y * SW];
    return ((double)d < 0.001) ? 1.f : d;
}

static float pcf_disk(const float *map, int SW, int SH, V3 sc, float bias, float filter,
                      const float *disk, int n)
{
    float vis = 0.f;
    float inv = 1.f / (float)n;
    for (int i = 0; i < n; ++i) {
        float u = sc.x + disk[2 * i] * filter;
        float v = sc.y + disk[2 * i + 1] * filter;
        float s = sample_shadow(map, SW, SH, u, v);
        if (sc.z <= s + bias) vis += inv;
    }
    return vis;
}

float orc_shadow_visibility(const orc_frame *fr, const float *map, int SW, int SH,
                            const float *ndc3, const float *n3, const float *l3)
{
    const orc_shadow_cfg *c = &fr->shadow;
    V3 ndc = v3(ndc3[0], ndc3[1], ndc3[2]);
    /* shadow.cpp:113-117 */
    V3 sc = add3(scale3(ndc, 0.5f), v3(0.5f, 0.5f, 0.5f));
    float bias = smax(0.009f * (1.f - dot3(v3(n3[0], n3[1], n3[2]), v3(l3[0], l3[1], l3[2]))), 0.007f);
    if (c->mode == 1) return pcf_disk(map, SW, SH, sc, bias, c->pcf_filter, c->pcf_disk, c->n_pcf);
    if (c->mode == 4) {
        /* SOFT_SHADOW_PCF with the reference's random stream: PCF
         * (shadow.cpp:47-63) draws n_pcf RandomVectorInUnitDisk samples, one
         * per iteration (nothing else draws in between, so drawing them
         * first consumes the stream identically). */
        float disk[2 * 64];
        orc_rng_disk(c->n_pcf, disk);
        return pcf_disk(map, SW, SH, sc, bias, c->pcf_filter, disk, c->n_pcf);
    }
    if (c->mode == 5) {
        /* SOFT_SHADOW_PCSS with the reference's random stream
         * (shadow.cpp:65-106): n_blocker blocker-search samples, and n_pcf
         * PCF samples only when a blocker was found. */
        float bd[2 * 64], pd[2 * 64];
        orc_rng_disk(c->n_blocker, bd);
        float blocker = 0.f, nb = 0.f;
        for (int i = 0; i < c->n_blocker; ++i) {
            float u = sc.x + (float)((double)bd[2 * i] * c->blocker_filter);
            float v = sc.y + (float)((double)bd[2 * i + 1] * c->blocker_filter);
            float s = sample_shadow(map, SW, SH, u, v);
            if (sc.z > s + bias) {
                blocker += s;
                nb += 1.f;
            }
        }
        float dB = (nb < 1.f) ? 0.f : blocker / nb;
        if ((double)dB < 0.001) return 1.f;
        float pen = (sc.z - dB) * c->area_light / dB;
        float filter = (float)(c->pcss_filter * (double)pen);
        orc_rng_disk(c->n_pcf, pd);
        return pcf_disk(map, SW, SH, sc, bias, filter, pd, c->n_pcf);
    }
    if (c->mode == 2) {
        /* FindAverageBlockDepth + PCSS (shadow.cpp:65-106) */
        float blocker = 0.f, nb = 0.f;
        for (int i = 0; i < c->n_blocker; ++i) {
            float u = sc.x + (float)((double)c->blocker_disk[2 * i] * c->blocker_filter);
            float v = sc.y + (float)((double)c->blocker_disk[2 * i + 1] * c->blocker_filter);
            float s = sample_shadow(map, SW, SH, u, v);
            if (sc.z > s + bias) {
                blocker += s;
                nb += 1.f;
            }
        }
        float dB = (nb < 1.f) ? 0.f : blocker / nb;
        if ((double)dB < 0.001) return 1.f;
        float pen = (sc.z - dB) * c->area_light / dB;
        float filter = (float)(c->pcss_filter * (double)pen);
        return pcf_disk(map, SW, SH, sc, bias, filter, c->pcf_disk, c->n_pcf);
    }
    if (c->mode == 3) {
        /* n x n texel grid PCF (new; config C2's 5x5 PCF) */
        int r = c->grid / 2;
        float inv = 1.f / (float)(c->grid * c->grid);
        float tu = 1.f / (float)SW, tv = 1.f / (float)SH;
        float vis = 0.f;
        for (int dy = -r; dy <= r; ++dy)
            for (int dx = -r; dx <= r; ++dx) {
                float s = sample_shadow(map, SW, SH, sc.x + (float)dx * tu, sc.y + (float)dy * tv);
                if (sc.z <= s + bias) vis += inv;
            }
        return vis;
    }
    /* HardShadow (shadow.cpp:38-45) */
    float s = sample_shadow(map, SW, SH, sc.x, sc.y);
    return (sc.z <= s + bias) ? 1.f : 0.f;
}

/* ---------------- lighting ---------------- */

/* BlinnPhongShader::CalculateLight (phongshader.h:171-215), generalised to
 * several lights: light 0 carries the emissive term exactly as the
 * reference's single light does. */
static V3 bp_light_term(V3 L, V3 H, V3 N, float vis, V3 dl, V3 el, float ao, float ks,
                        float shin, V3 lc, int shadow_on, int with_emissive)
{
    float diff = smax(0.f, dot3(L, N));
    float spec = powf(smax(0.f, dot3(H, N)), shin);
    V3 diffuse = scale3(scale3(dl, diff), ao);
    V3 specular = scale3(v3(ks, ks, ks), spec);
    if (shadow_on) {
        float sh = (1 - vis) * 0.6f;
        float vv = 1 - sh;
        diffuse = scale3(diffuse, vv);
        specular = scale3(specular, vv);
    }
    V3 t = add3(diffuse, specular);
    if (with_emissive) t = add3(t, el);
    return mul3(t, lc);
}

static V3 tonemap_gamma(V3 c)
{
    c = v3(c.x / (c.x + 1.f), c.y / (c.y + 1.f), c.z / (c.z + 1.f));
    c = pow3(c, 1.f / 2.2f);
    return clamp01_3(c);
}

void orc_shade_blinn_phong(const float *L, const float *H, const float *N, float vis,
                           const float *diffuse, const float *emissive, const float *param,
                           const float *light_color, int shadow_on, float *out3)
{
    const float gamma = (float)2.2;
    V3 dl = pow3(v3(diffuse[0], diffuse[1], diffuse[2]), gamma);
    V3 el = pow3(v3(emissive[0], emissive[1], emissive[2]), gamma);
    float ao = param[0], ks = param[1], shin = param[2];
    V3 ambient = scale3(mul3(v3(0.3f, 0.3f, 0.3f), dl), ao);
    V3 t = bp_light_term(v3(L[0], L[1], L[2]), v3(H[0], H[1], H[2]), v3(N[0], N[1], N[2]), vis,
                         dl, el, ao, ks, shin, v3(light_color[0], light_color[1], light_color[2]),
                         shadow_on, 1);
    V3 c = tonemap_gamma(add3(ambient, t));
    out3[0] = c.x; out3[1] = c.y; out3[2] = c.z;
}

/* PBRShader helpers (pbrshader.h:255-288) */
static float ggx_d(float NdotH, float r)
{
    float a = r * r;
    float a2 = a * a;
    float NdotH2 = NdotH * NdotH;
    float den = (NdotH2 * (a2 - 1.f) + 1.f);
    float den2 = den * den;
    const float InvPi = (float)0.31830988618379067154;
    return a2 * InvPi / den2;
}
static float schlick_ggx(float NdotV, float r)
{
    float a = r + 1.f;
    float k = a * a / 8.f;
    return NdotV / (NdotV * (1 - k) + k);
}

/* Outgoing radiance of one light (pbrshader.h:196-235) */
static V3 pbr_lo(V3 L, V3 V, V3 H, V3 N, float vis, V3 al, float metal, float rough, V3 rad,
                 int shadow_on)
{
    const float InvPi = (float)0.31830988618379067154;
    float NdotV = smax(dot3(N, V), 0.f);
    float NdotL = smax(dot3(N, L), 0.f);
    float NdotH = smax(dot3(N, H), 0.f);
    float HdotV = smax(dot3(H, V), 0.f);
    V3 F0 = lerp3(metal, v3(0.04f, 0.04f, 0.04f), al);
    float NDF = ggx_d(NdotH, rough);
    float G = schlick_ggx(NdotV, rough) * schlick_ggx(NdotL, rough);
    float omh = smax(1.f - HdotV, 0.f);
    float p5 = powf(omh, 5.f);
    V3 F = add3(F0, scale3(sub3(v3(1.f, 1.f, 1.f), F0), p5));
    V3 DGF = scale3(F, NDF * G);
    float den = 4 * NdotV * NdotL + 0.001f;
    V3 spec = divs3(DGF, den);
    V3 kd = sub3(v3(1.f, 1.f, 1.f), F);
    kd = scale3(kd, 1.f - metal);
    V3 brdf = add3(scale3(mul3(kd, al), InvPi), spec);
    V3 Lo = scale3(mul3(brdf, rad), NdotL);
    if (shadow_on) {
        float sh = (1 - vis) * 0.6f;
        float vv = 1 - sh;
        Lo = scale3(Lo, vv);
    }
    return Lo;
}

static V3 pbr_finish(V3 Lo, V3 al, float ao, V3 el)
{
    V3 color = Lo;
    V3 ambient = scale3(mul3(v3(0.3f, 0.3f, 0.3f), al), ao);
    color = add3(color, ambient);
    color = add3(color, el);
    return tonemap_gamma(color);
}

void orc_shade_pbr(const float *L, const float *V, const float *Hv, const float *N, float vis,
                   const float *albedo, const float *emissive, const float *param,
                   const float *radiance, int shadow_on, float *out3)
{
    const float gamma = (float)2.2;
    V3 al = pow3(v3(albedo[0], albedo[1], albedo[2]), gamma);
    V3 el = pow3(v3(emissive[0], emissive[1], emissive[2]), gamma);
    V3 Lo = pbr_lo(v3(L[0], L[1], L[2]), v3(V[0], V[1], V[2]), v3(Hv[0], Hv[1], Hv[2]),
                   v3(N[0], N[1], N[2]), vis, al, param[1], param[2],
                   v3(radiance[0], radiance[1], radiance[2]), shadow_on);
    V3 c = pbr_finish(Lo, al, param[0], el);
    out3[0] = c.x; out3[1] = c.y; out3[2] = c.z;
}

/* Direction towards light l from P (point: Normalize(pos - P); dir: the
 * normalised reverse of the travel direction). */
static V3 light_dir(const orc_light *l, V3 P)
{
    if (l->type == 0) return norm3(sub3(v3(l->vec[0], l->vec[1], l->vec[2]), P));
    return norm3(v3(-l->vec[0], -l->vec[1], -l->vec[2]));
}

/* Multi-light shading of one surface point.  With one point light this is
 * exactly BlinnPhongShader::CalculateLight / PBRShader::CalculateLight with
 * the halfway slot as each caller fills it: forward passes the halfway
 * vector (phongshader.h:165), the deferred pass passes viewDir
 * (forkergl.cpp:369).  Only light 0 is shadowed (the shadow map is rendered
 * from the point light, render.cpp:73-78). */
static V3 shade_point(const orc_frame *fr, const float *map, int SW, int SH, int pbr,
                      int deferred_halfway_quirk, V3 P, V3 N, V3 lndc, V3 albedo, V3 emissive,
                      V3 param)
{
    const float gamma = (float)2.2;
    V3 eye = v3(fr->eye[0], fr->eye[1], fr->eye[2]);
    V3 Vd = norm3(sub3(eye, P));
    if (!pbr) {
        V3 dl = pow3(albedo, gamma);
        V3 el = pow3(emissive, gamma);
        V3 c = scale3(mul3(v3(0.3f, 0.3f, 0.3f), dl), param.x);
        for (int i = 0; i < fr->n_lights; ++i) {
            const orc_light *l = &fr->lights[i];
            V3 L = light_dir(l, P);
            float vis = 1.f;
            if (i == 0 && fr->shadow_on) {
                float nd[3] = {lndc.x, lndc.y, lndc.z}, nn[3] = {N.x, N.y, N.z}, ll[3] = {L.x, L.y, L.z};
                vis = orc_shadow_visibility(fr, map, SW, SH, nd, nn, ll);
            } else if (!fr->shadow_on) {
                vis = 0.f;
            }
            V3 H = deferred_halfway_quirk ? Vd : norm3(add3(L, Vd));
            V3 t = bp_light_term(L, H, N, vis, dl, el, param.x, param.y, param.z,
                                 v3(l->color[0], l->color[1], l->color[2]), fr->shadow_on, i == 0);
            c = add3(c, t);
        }
        return tonemap_gamma(c);
    }
    V3 al = pow3(albedo, gamma);
    V3 el = pow3(emissive, gamma);
    V3 Lo = v3(0.f, 0.f, 0.f);
    for (int i = 0; i < fr->n_lights; ++i) {
        const orc_light *l = &fr->lights[i];
        V3 L = light_dir(l, P);
        float vis = 1.f;
        if (i == 0 && fr->shadow_on) {
            float nd[3] = {lndc.x, lndc.y, lndc.z}, nn[3] = {N.x, N.y, N.z}, ll[3] = {L.x, L.y, L.z};
            vis = orc_shadow_visibility(fr, map, SW, SH, nd, nn, ll);
        } else if (!fr->shadow_on) {
            vis = 0.f;
        }
        V3 H = norm3(add3(L, Vd));
        V3 lo = pbr_lo(L, Vd, H, N, vis, al, param.y, param.z,
                       v3(l->color[0], l->color[1], l->color[2]), fr->shadow_on);
        Lo = (i == 0) ? lo : add3(Lo, lo);
    }
    return pbr_finish(Lo, al, param.x, el);
}

/* ---------------- raster core (forkergl.cpp:145-261) ---------------- */

typedef struct {
    int X[3], Y[3];
    float z[3];
    int minx, miny, maxx, maxy;
} TriSetup;

static void tri_setup(const float *vp, const V4 *ndc, int W, int H, TriSetup *t)
{
    for (int i = 0; i < 3; ++i) {
        V4 c = mv4(vp, ndc[i]);
        t->X[i] = orc_trunc_int(c.x);
        t->Y[i] = orc_trunc_int(c.y);
        t->z[i] = c.z;
    }
    t->minx = clampi(imin(t->X[0], imin(t->X[1], t->X[2])), 0, W - 1);
    t->miny = clampi(imin(t->Y[0], imin(t->Y[1], t->Y[2])), 0, H - 1);
    t->maxx = clampi(imax(t->X[0], imax(t->X[1], t->X[2])), 0, W - 1);
    t->maxy = clampi(imax(t->Y[0], imax(t->Y[1], t->Y[2])), 0, H - 1);
}

void orc_shadow_pass(const orc_draw *draws, int n_draws, const orc_frame *fr, int W, int H,
                     float *depth, float *shadow)
{
    for (size_t i = 0; i < (size_t)W * H; ++i) { depth[i] = FLT_MAX; shadow[i] = 0.f; }
    for (int di = 0; di < n_draws; ++di) {
        const orc_draw *d = &draws[di];
        float LM[16];
        orc_mat_mul(fr->light_space, d->model, LM);
        for (int f = 0; f < d->n_faces; ++f) {
            TriVaryings tv;
            for (int k = 0; k < 3; ++k) depth_vertex(d, LM, f, k, &tv);
            TriSetup t;
            tri_setup(fr->shadow_viewport, tv.ndc, W, H, &t);
            V3 dz = v3(t.z[0], t.z[1], t.z[2]);
            int py0 = t.miny > g_y0 ? t.miny : g_y0, py1 = t.maxy < g_y1 - 1 ? t.maxy : g_y1 - 1;
            for (int px = t.minx; px <= t.maxx; ++px)
                for (int py = py0; py <= py1; ++py) {
                    float b[3];
                    if (!orc_barycentric(t.X[0], t.Y[0], t.X[1], t.Y[1], t.X[2], t.Y[2], px, py, b))
                        continue;
                    float cur = dot3(v3(b[0], b[1], b[2]), dz);
                    size_t idx = (size_t)px + (size_t)py * W;
                    if (cur >= depth[idx]) continue;
                    depth[idx] = cur;
                    /* DepthShader::ProcessFragment (depthshader.h:30-36) */
                    V3 p = interp3(tv.depth_ndc, b);
                    shadow[idx] = p.z * 0.5f + 0.5f;
                }
        }
    }
}

void orc_geometry_pass(const orc_draw *draws, int n_draws, const orc_material *mats,
                       const orc_texture *texs, const orc_frame *fr, int W, int H,
                       float *depth, float *const *g)
{
    size_t N = (size_t)W * H;
    for (size_t i = 0; i < N; ++i) {
        depth[i] = FLT_MAX;
        for (int c = 0; c < 19; ++c) g[c][i] = 0.f;
    }
    for (int di = 0; di < n_draws; ++di) {
        const orc_draw *d = &draws[di];
        const orc_material *m = &mats[d->material];
        for (int f = 0; f < d->n_faces; ++f) {
            TriVaryings tv;
            memset(&tv, 0, sizeof(tv));
            for (int k = 0; k < 3; ++k) main_vertex(d, fr, f, k, &tv);
            TriSetup t;
            tri_setup(fr->viewport, tv.ndc, W, H, &t);
            V3 dz = v3(t.z[0], t.z[1], t.z[2]);
            int py0 = t.miny > g_y0 ? t.miny : g_y0, py1 = t.maxy < g_y1 - 1 ? t.maxy : g_y1 - 1;
            for (int px = t.minx; px <= t.maxx; ++px)
                for (int py = py0; py <= py1; ++py) {
                    float b[3];
                    if (!orc_barycentric(t.X[0], t.Y[0], t.X[1], t.Y[1], t.X[2], t.Y[2], px, py, b))
                        continue;
                    float cur = dot3(v3(b[0], b[1], b[2]), dz);
                    size_t idx = (size_t)px + (size_t)py * W;
                    if (cur >= depth[idx]) continue;
                    depth[idx] = cur;
                    /* GShader::ProcessFragment (gshader.h:95-202) and the
                     * seven buffer writes (forkergl.cpp:211-223) */
                    FragGeom fg = frag_geometry(d, m, texs, fr, &tv, b, m->map_bump);
                    V3 albedo, emissive, param;
                    float type;
                    if (d->support_pbr) {
                        albedo = m->map_base >= 0 ? tex_sample(texs, m->map_base, fg.uv)
                                                  : v3(m->albedo[0], m->albedo[1], m->albedo[2]);
                        emissive = m->map_pbr_ke >= 0 ? tex_sample(texs, m->map_pbr_ke, fg.uv)
                                                      : v3(m->ke[0], m->ke[1], m->ke[2]);
                        float rough = m->map_rough >= 0 ? tex_samplef(texs, m->map_rough, fg.uv) : m->roughness;
                        float metal = m->map_metal >= 0 ? tex_samplef(texs, m->map_metal, fg.uv) : m->metalness;
                        float ao = m->map_ao >= 0 ? tex_samplef(texs, m->map_ao, fg.uv) : 1.f;
                        param = v3(ao, metal, rough);
                        type = 1.f;
                    } else {
                        emissive = m->map_ke >= 0 ? tex_sample(texs, m->map_ke, fg.uv)
                                                  : v3(m->ke[0], m->ke[1], m->ke[2]);
                        albedo = m->map_kd >= 0 ? tex_sample(texs, m->map_kd, fg.uv)
                                                : v3(m->kd[0], m->kd[1], m->kd[2]);
                        float shin = m->map_ks >= 0 ? tex_samplef(texs, m->map_ks, fg.uv) + 5 : 1.f;
                        param = v3(1.f, m->ks[0], shin);
                        type = 0.f;
                    }
                    V3 vals[6] = {fg.normal, fg.pos, fg.lndc, albedo, emissive, param};
                    for (int c = 0; c < 6; ++c) {
                        if (c == 2 && !fr->shadow_on) continue;
                        g[3 * c][idx] = vals[c].x;
                        g[3 * c + 1][idx] = vals[c].y;
                        g[3 * c + 2][idx] = vals[c].z;
                    }
                    g[18][idx] = type;
                }
        }
    }
}

void orc_forward_pass(const orc_draw *draws, int n_draws, const orc_material *mats,
                      const orc_texture *texs, const orc_frame *fr, int W, int H,
                      const float *smap, int SW, int SH, float *depth, float *const *fb)
{
    size_t N = (size_t)W * H;
    for (size_t i = 0; i < N; ++i) {
        depth[i] = FLT_MAX;
        fb[0][i] = fb[1][i] = fb[2][i] = 0.12f;  /* ClearColor (render.cpp:103) */
    }
    for (int di = 0; di < n_draws; ++di) {
        const orc_draw *d = &draws[di];
        const orc_material *m = &mats[d->material];
        for (int f = 0; f < d->n_faces; ++f) {
            TriVaryings tv;
            memset(&tv, 0, sizeof(tv));
            for (int k = 0; k < 3; ++k) main_vertex(d, fr, f, k, &tv);
            TriSetup t;
            tri_setup(fr->viewport, tv.ndc, W, H, &t);
            V3 dz = v3(t.z[0], t.z[1], t.z[2]);
            int py0 = t.miny > g_y0 ? t.miny : g_y0, py1 = t.maxy < g_y1 - 1 ? t.maxy : g_y1 - 1;
            for (int px = t.minx; px <= t.maxx; ++px)
                for (int py = py0; py <= py1; ++py) {
                    float b[3];
                    if (!orc_barycentric(t.X[0], t.Y[0], t.X[1], t.Y[1], t.X[2], t.Y[2], px, py, b))
                        continue;
                    float cur = dot3(v3(b[0], b[1], b[2]), dz);
                    size_t idx = (size_t)px + (size_t)py * W;
                    if (cur >= depth[idx]) continue;
                    depth[idx] = cur;
                    FragGeom fg = frag_geometry(d, m, texs, fr, &tv, b, m->map_bump);
                    V3 c;
                    if (!d->support_pbr) {
                        /* BlinnPhongShader::ProcessFragment (phongshader.h:152-168) */
                        V3 dc = m->map_kd >= 0 ? tex_sample(texs, m->map_kd, fg.uv)
                                               : v3(m->kd[0], m->kd[1], m->kd[2]);
                        V3 em = m->map_ke >= 0 ? tex_sample(texs, m->map_ke, fg.uv)
                                               : v3(m->ke[0], m->ke[1], m->ke[2]);
                        float shin = 1.f;
                        if (m->map_ks >= 0) shin = tex_samplef(texs, m->map_ks, fg.uv) + 5;
                        c = shade_point(fr, smap, SW, SH, 0, 0, fg.pos, fg.normal, fg.lndc, dc, em,
                                        v3(m->ka[0], m->ks[0], shin));
                    } else {
                        /* PBRShader::ProcessFragment (pbrshader.h:153-179) */
                        V3 al = m->map_base >= 0 ? tex_sample(texs, m->map_base, fg.uv)
                                                 : v3(m->albedo[0], m->albedo[1], m->albedo[2]);
                        V3 em = m->map_pbr_ke >= 0 ? tex_sample(texs, m->map_pbr_ke, fg.uv)
                                                   : v3(m->pbr_ke[0], m->pbr_ke[1], m->pbr_ke[2]);
                        float rough = m->map_rough >= 0 ? tex_samplef(texs, m->map_rough, fg.uv) : m->roughness;
                        float metal = m->map_metal >= 0 ? tex_samplef(texs, m->map_metal, fg.uv) : m->metalness;
                        float ao = m->map_ao >= 0 ? tex_samplef(texs, m->map_ao, fg.uv) : 1.f;
                        c = shade_point(fr, smap, SW, SH, 1, 0, fg.pos, fg.normal, fg.lndc, al, em,
                                        v3(ao, metal, rough));
                    }
                    fb[0][idx] = c.x; fb[1][idx] = c.y; fb[2][idx] = c.z;
                }
        }
    }
}

/* ForkerGL::DrawScreenSpacePixels (forkergl.cpp:326-380) */
void orc_lighting_pass(const orc_frame *fr, int W, int H, const float *const *g, const float *ao,
                       const float *smap, int SW, int SH, float *const *fb)
{
    for (int y = g_y0 > 0 ? g_y0 : 0; y < H && y < g_y1; ++y)
        for (int x = 0; x < W; ++x) {
            size_t i = (size_t)x + (size_t)y * W;
            V3 N = v3(g[0][i], g[1][i], g[2][i]);
            V3 P = v3(g[3][i], g[4][i], g[5][i]);
            V3 L = v3(g[6][i], g[7][i], g[8][i]);
            V3 al = v3(g[9][i], g[10][i], g[11][i]);
            V3 em = v3(g[12][i], g[13][i], g[14][i]);
            V3 pa = v3(g[15][i], g[16][i], g[17][i]);
            float type = g[18][i];
            pa.x *= ao[i];
            V3 c = shade_point(fr, smap, SW, SH, !(type < 0.5f), 1, P, N, L, al, em, pa);
            fb[0][i] = c.x; fb[1][i] = c.y; fb[2][i] = c.z;
        }
}

/* ---------------- SSAO, blur, resolve ---------------- */

/* Render::DoSSAO (render.cpp:214-289) with the reference's random stream
 * (ref_rng.cpp): every pixel in row order (y outer, x inner) draws 32
 * RandomVectorInHemisphere(normal) samples, scaled by
 * Lerp(0.1f, 1.0f, |d|^2) = (1 - 0.1f) * 1.0f + 0.1f * |d|^2
 * (Lerp(t, v1, v2), utility.h:26-29), radius 0.075f, projected with the
 * view-projection and viewport matrices (v / w is v * (1/w)), truncated to
 * Point2i; the depth read is DepthBuffer.GetValue(x, y) = data[x + y * W]
 * with 32-bit int arithmetic, which for an off-screen sample still lands
 * inside the buffer (another pixel) or outside it (undefined in the
 * reference: here no occlusion).  Occlusion adds 1/32 per sample past the
 * 0.0005 bias with |frag - cached| < 0.01; ao = powf(1 - occlusion, 3). */
void orc_ssao_reference(const orc_frame *fr, int W, int H, const float *depth,
                        const float *const *nrm, const float *const *pos, float *ao)
{
    float VP[16];
    orc_mat_mul(fr->proj, fr->view, VP);
    const float *vpm = fr->viewport;
    const int numSample = 32;
    const float kernelScale = 1.f / numSample;
    const float radius = 0.075f, rangeCheckRadius = 0.01f, bias = 0.0005f;
    const int64_t n_px = (int64_t)W * H;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t i = (size_t)x + (size_t)y * W;
            V3 P = v3(pos[0][i], pos[1][i], pos[2][i]);
            float nn[3] = {nrm[0][i], nrm[1][i], nrm[2][i]};
            float fd = depth[i];
            float occ = 0.f;
            for (int s = 0; s < numSample; ++s) {
                float d[3];
                orc_rng_hemisphere(nn, d);
                V3 dir = v3(d[0], d[1], d[2]);
                float scale = len3(dir);
                scale = (1 - 0.1f) * 1.0f + 0.1f * (scale * scale);
                dir = scale3(dir, scale);
                V3 ws = add3(P, scale3(dir, radius));
                V4 cs = mv4(VP, (V4){ws.x, ws.y, ws.z, 1.f});
                V4 ndc = divs4(cs, cs.w);
                V4 ss = mv4(vpm, ndc);
                int sx = orc_trunc_int(ss.x), sy = orc_trunc_int(ss.y);
                int32_t idx = (int32_t)((uint32_t)sx + (uint32_t)sy * (uint32_t)W);
                if (idx < 0 || (int64_t)idx >= n_px) continue;
                float cached = depth[idx];
                if (ss.z >= cached + bias) {
                    float rc = fabsf(fd - cached) < rangeCheckRadius ? 1.f : 0.f;
                    occ += kernelScale * rc;
                }
            }
            occ = 1.f - occ;
            ao[i] = powf(occ, 3);
        }
}

/* Kernel SSAO (new, deterministic; the configs' replacement for the
 * RNG-driven Render::DoSSAO, render.cpp:214-289).  Per pixel: a tangent
 * frame from the 4x4 noise vector, n kernel offsets in world space, the
 * reference's projection to window space, depth-buffer gather, bias and
 * range check.  Off-screen taps add no occlusion. */
void orc_ssao_kernel(const orc_frame *fr, int W, int H, const float *depth,
                     const float *const *nrm, const float *const *pos, const float *kernel,
                     int n_kernel, const float *noise, float radius, float bias, float range,
                     float *ao)
{
    /* Kernel SSAO with the projection linearised per pixel: c = VP * (P, 1)
     * and cT, cB, cN = VP * (T|B|N, 0); each column is taken to window-space
     * numerators (vp0 clip_x + vp3 clip_w, vp5 clip_y + vp7 clip_w,
     * vp10 clip_z + vp11 clip_w, clip_w) by fmaf, the direction columns
     * times the radius; a tap is e = ((ec + kx eT) + ky eB) + kz eN by fmaf
     * and window x, y, z = e * (1 / e_w); occlusion as Render::DoSSAO tests
     * it (render.cpp:250-275). */
    float VP[16];
    orc_mat_mul(fr->proj, fr->view, VP);
    const float *vp = fr->viewport;
    const float scale = 1.f / (float)n_kernel;
    for (int y = g_y0 > 0 ? g_y0 : 0; y < H && y < g_y1; ++y)
        for (int x = 0; x < W; ++x) {
            size_t i = (size_t)x + (size_t)y * W;
            V3 P = v3(pos[0][i], pos[1][i], pos[2][i]);
            V3 N = v3(nrm[0][i], nrm[1][i], nrm[2][i]);
            float fd = depth[i];
            int ni = (x & 3) + 4 * (y & 3);
            V3 r = v3(noise[3 * ni], noise[3 * ni + 1], noise[3 * ni + 2]);
            V3 T = norm3(sub3(r, scale3(N, dot3(r, N))));
            V3 B = cross3(N, T);
            V4 c0 = mv4(VP, (V4){P.x, P.y, P.z, 1.f});
            float cT[4], cB[4], cN[4], cc[4];
            for (int j = 0; j < 4; ++j) {
                V3 row = v3(VP[4 * j], VP[4 * j + 1], VP[4 * j + 2]);
                cT[j] = dot3(row, T);
                cB[j] = dot3(row, B);
                cN[j] = dot3(row, N);
            }
            cc[0] = c0.x; cc[1] = c0.y; cc[2] = c0.z; cc[3] = c0.w;
            float ec[4], e3[3][4];
            ec[0] = fmaf(vp[0], cc[0], vp[3] * cc[3]);
            ec[1] = fmaf(vp[5], cc[1], vp[7] * cc[3]);
            ec[2] = fmaf(vp[10], cc[2], vp[11] * cc[3]);
            ec[3] = cc[3];
            const float *cols[3] = {cT, cB, cN};
            for (int q = 0; q < 3; ++q) {
                const float *c = cols[q];
                e3[q][0] = radius * fmaf(vp[0], c[0], vp[3] * c[3]);
                e3[q][1] = radius * fmaf(vp[5], c[1], vp[7] * c[3]);
                e3[q][2] = radius * fmaf(vp[10], c[2], vp[11] * c[3]);
                e3[q][3] = radius * c[3];
            }
            int count = 0;
            for (int s = 0; s < n_kernel; ++s) {
                const float kx = kernel[3 * s], ky = kernel[3 * s + 1], kz = kernel[3 * s + 2];
                float e[4];
                for (int j = 0; j < 4; ++j)
                    e[j] = fmaf(kz, e3[2][j], fmaf(ky, e3[1][j], fmaf(kx, e3[0][j], ec[j])));
                const float inv = 1.f / e[3];
                const float sxf = e[0] * inv;
                const float syf = e[1] * inv;
                const float sz = e[2] * inv;
                int sx = orc_trunc_int(sxf), sy = orc_trunc_int(syf);
                if (sx < 0 || sy < 0 || sx >= W || sy >= H) continue;
                float cached = depth[(size_t)sx + (size_t)sy * W];
                if (sz >= cached + bias && fabsf(fd - cached) < range) ++count;
            }
            float a = 1.f - (float)count * scale;
            ao[i] = (a * a) * a;
        }
}

/* Buffer1f::TwoPassGaussianBlurDenoised (buffer.cpp:59-98): in-place, so a
 * tap to the left (or below) reads the already-blurred value. */
void orc_blur_recursive(float *a, int W, int H)
{
    const float w[5] = {(float)0.227027, (float)0.1945946, (float)0.1216216, (float)0.054054,
                        (float)0.016216};
    for (int h = 0; h < H; ++h)
        for (int x = 0; x < W; ++x) {
            float r = a[(size_t)x + (size_t)h * W] * w[0];
            for (int i = 1; i < 5; ++i) {
                int x1 = clampi(x + i, 0, W - 1), x2 = clampi(x - i, 0, W - 1);
                r += a[(size_t)x1 + (size_t)h * W] * w[i];
                r += a[(size_t)x2 + (size_t)h * W] * w[i];
            }
            a[(size_t)x + (size_t)h * W] = r;
        }
    for (int h = 0; h < H; ++h)
        for (int x = 0; x < W; ++x) {
            float r = a[(size_t)x + (size_t)h * W] * w[0];
            for (int i = 1; i < 5; ++i) {
                int h1 = clampi(h + i, 0, H - 1), h2 = clampi(h - i, 0, H - 1);
                r += a[(size_t)x + (size_t)h1 * W] * w[i];
                r += a[(size_t)x + (size_t)h2 * W] * w[i];
            }
            a[(size_t)x + (size_t)h * W] = r;
        }
}

/* Separable Gaussian (new; the configs' sigma=2 blur): same tap order as the
 * reference's loop, but reading a copy so the stencil is not recursive. */
void orc_blur_separable(float *a, int W, int H, const float *wt, int R, float *tmp)
{
    for (int h = 0; h < H; ++h)
        for (int x = 0; x < W; ++x) {
            float r = a[(size_t)x + (size_t)h * W] * wt[0];
            for (int i = 1; i <= R; ++i) {
                int x1 = clampi(x + i, 0, W - 1), x2 = clampi(x - i, 0, W - 1);
                r += a[(size_t)x1 + (size_t)h * W] * wt[i];
                r += a[(size_t)x2 + (size_t)h * W] * wt[i];
            }
            tmp[(size_t)x + (size_t)h * W] = r;
        }
    for (int h = 0; h < H; ++h)
        for (int x = 0; x < W; ++x) {
            float r = tmp[(size_t)x + (size_t)h * W] * wt[0];
            for (int i = 1; i <= R; ++i) {
                int h1 = clampi(h + i, 0, H - 1), h2 = clampi(h - i, 0, H - 1);
                r += tmp[(size_t)x + (size_t)h1 * W] * wt[i];
                r += tmp[(size_t)x + (size_t)h2 * W] * wt[i];
            }
            a[(size_t)x + (size_t)h * W] = r;
        }
}

/* Buffer3f::GenerateImage (buffer.cpp:113-126): TGAColor(c*254.99f, ...)
 * narrows float -> uint8 through a 32-bit truncating conversion. */
void orc_quantize_rgb(const float *const *fb, int W, int H, unsigned char *out)
{
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        out[3 * i + 0] = (unsigned char)orc_trunc_int(fb[2][i] * 254.99f);
        out[3 * i + 1] = (unsigned char)orc_trunc_int(fb[1][i] * 254.99f);
        out[3 * i + 2] = (unsigned char)orc_trunc_int(fb[0][i] * 254.99f);
    }
}

/* Buffer1f::GenerateImage (buffer.cpp:20-32) */
void orc_quantize_grey(const float *p, int W, int H, unsigned char *out)
{
    for (size_t i = 0; i < (size_t)W * H; ++i) out[i] = (unsigned char)orc_trunc_int(p[i] * 255);
}

/* Render::DoSSAA (render.cpp:291-343): integer k x k box sum of the 8-bit
 * image, divided as (int)((float)sum / (float)(k*k)). */
void orc_ssaa_resolve(const unsigned char *in, int W, int H, int k, unsigned char *out)
{
    int ow = W / k, oh = H / k;
    float k2 = (float)(k * k);
    for (int x = 0; x < ow; ++x)
        for (int y = 0; y < oh; ++y) {
            int s[3] = {0, 0, 0};
            for (int i = 0; i < k; ++i)
                for (int j = 0; j < k; ++j) {
                    size_t p = ((size_t)(x * k + i) + (size_t)(y * k + j) * W) * 3;
                    s[0] += in[p]; s[1] += in[p + 1]; s[2] += in[p + 2];
                }
            size_t o = ((size_t)x + (size_t)y * ow) * 3;
            for (int c = 0; c < 3; ++c) out[o + c] = (unsigned char)(int)((float)s[c] / k2);
        }
}
