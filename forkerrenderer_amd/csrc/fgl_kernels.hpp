// fgl_kernels.hpp -- launch interface between the host context and the
// gfx950 kernels (raster.hip, sort.hip, shade.hip).
#pragma once

#include "fgl_device.hpp"

namespace fgl {

constexpr int kPassForward = 0;
constexpr int kPassGeometry = 1;
constexpr int kPassShadow = 3;
constexpr int kGbufPlanes = 19;
// Triangles whose tile rectangle has more tiles than this are emitted by a
// whole workgroup each (k_emit_big) instead of by one thread.
constexpr int kSmallTriTiles = 16;

struct RasterArgs {
    const DevDraw *draws;
    int n_draws;
    const int4 *screen;        // per vertex: X, Y, z bits, exact flag
    const int4 *tri_v;         // per triangle: 3 pass-wide vertex ids, draw id
    const uint32_t *vals;      // sorted pair values (triangle ids)
    const uint32_t *tile_start, *tile_end;
    int W, H, tiles_x, tiles_y;
    int fresh;                 // targets still hold their clear values: start from them
    float *depth;
    int *vis;                  // winning triangle per pixel, -1 if none (this pass)
    float *shadow_out;         // shadow pass: ShadowBuffer plane
    const float *ndc_z;        // shadow pass: per vertex, z of DepthShader::ProcessVertex's NDC
};

// Per-pixel unit-disk samples of the stochastic shadow filters: pixel i's
// samples start at pair offs[i] of `pairs`, or at i * stride when offs is
// null (fixed count per pixel).
struct ShadowSamples {
    const float *pairs;
    const uint32_t *offs;
    uint32_t stride;
};

struct ResolveArgs {
    const DevDraw *draws;
    const int4 *screen;
    const int4 *tri_v;
    const int *vis;
    int W, H, fresh;
    float clear[3];            // forward pass background colour
    float *gbuf;               // 19 planes (geometry pass)
    float *fb;                 // 3 planes (forward pass)
    const DevMaterial *materials;
    const DevTexture *textures;
    const float *shadow_map;
    ShadowSamples ss;          // forward pass, stochastic shadow filters: the winners' samples
};

// The forward pass's depth test replayed in submission order (k_frag_walk):
// per tile, triangles in pass-wide id order, each pixel's running depth in a
// register.  Pixels are indexed column-major, q = px * H + py.
struct FragWalkArgs {
    const int4 *screen;
    const int4 *tri_v;
    const uint32_t *vals;      // per-tile triangle lists (sorted pairs)
    const uint32_t *tile_start, *tile_end;
    int W, H, tiles_x, tiles_y;
    int fresh;                 // the depth target still holds its clear value
    const float *depth;        // depth target before the pass
    uint32_t *count;           // count walk: passing fragments per pixel (q)
    const uint32_t *base;      // emit walk: exclusive scan of count
    uint32_t *tri_out, *q_out; // emit walk: the fragments' triangle ids and pixels
};

// Neural AO still on its network grid, the up pass not yet run: the AO of
// frame pixel (x, y) is 1 where the pixel is uncovered (fmask 0), else the
// network AO interpolated bilinearly between the centres of the f x f blocks
// around the pixel's centre (block (X, Y) centred at frame coordinate
// f X + f/2; edges clamp).  lao == nullptr: not in use.
struct NetAo {
    const float *lao;       // lW x lH network-grid AO
    const uint8_t *fmask;   // frame-pixel coverage
    int lW, lH, f;
};

__device__ __forceinline__ float net_ao_interp(const float *__restrict__ L, int lW, int lH, int f, int x, int y)
{
    const float inv = 1.f / (float)f;
    const float u = ((float)x + 0.5f) * inv - 0.5f, w = ((float)y + 0.5f) * inv - 0.5f;
    const float fu = floorf(u), fw = floorf(w);
    const float tx = u - fu, ty = w - fw;
    const int X0 = (int)fu, Y0 = (int)fw;
    const int xa = imax(0, imin(lW - 1, X0)), xb = imax(0, imin(lW - 1, X0 + 1));
    const int ya = imax(0, imin(lH - 1, Y0)), yb = imax(0, imin(lH - 1, Y0 + 1));
    const float top = L[(size_t)ya * lW + xa] * (1.f - tx) + L[(size_t)ya * lW + xb] * tx;
    const float bot = L[(size_t)yb * lW + xa] * (1.f - tx) + L[(size_t)yb * lW + xb] * tx;
    return top * (1.f - ty) + bot * ty;
}

struct SsaoArgs {
    int n_kernel;
    float kernel[3 * 64];
    float noise[3 * 16];
    float radius, bias, range;
    float kmax;   // largest |component| over the n_kernel kernel vectors
};

// raster.hip
// Window records of every (draw, position); shadow pass: also the NDC z of
// each vertex into ndc_z; zero (may be null): a word set to 0 on the way.
void launch_vertex(bool shadow, const DevDraw *draws, int n_draws, int max_pos,
                   const FrameUniforms &u, int4 *out, float *ndc_z, uint32_t *zero, hipStream_t s);
void launch_setup(const DevDraw *draws, int n_draws, int max_faces, const int4 *screen, int W,
                  int H, int TS, uint2 *rect, uint32_t *count, int4 *tri_v, uint32_t *big_list,
                  uint32_t *big_count, hipStream_t s);
// Pairs of the pass (offsets from the scan of the per-triangle counts); the
// first n_clear entries of range_a and range_b are zeroed on the way.
void launch_emit(const uint32_t *off, int n_tri, const uint2 *rect, int tiles_x,
                 const uint32_t *big_list, uint32_t n_big, uint32_t *keys, uint32_t *vals,
                 uint32_t *range_a, uint32_t *range_b, int n_clear, hipStream_t s);
// host: a device pointer to host memory; writes *a, *b, then seq.
void launch_publish2(const uint32_t *a, const uint32_t *b, uint32_t seq, uint32_t *host, hipStream_t s);
void launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
void launch_fill_f32(float *p, float v, size_t n, hipStream_t s);
// n_dev (optional): device-side pair count, used as min(n, *n_dev)
void launch_tile_ranges(const uint32_t *keys, uint32_t n, const uint32_t *n_dev, uint32_t *start,
                        uint32_t *end, hipStream_t s);
void launch_raster(bool shadow, int TS, const RasterArgs &a, hipStream_t s);
void launch_gbuffer_resolve(const ResolveArgs &a, const FrameUniforms &u, hipStream_t s);
void launch_forward_resolve(const ResolveArgs &a, const FrameUniforms &u, hipStream_t s);
// emit = false: the count walk; true: the emit walk.
void launch_frag_walk(bool emit, int TS, const FragWalkArgs &a, hipStream_t s);
// Light 0's shadow inputs of the n passing fragments (triangle tri[k] at
// pixel q[k], in raster order): shading normal, world position and light
// NDC as planes of n floats (9 n in all), as the forward shaders compute them.
void launch_frag_shadow_inputs(const ResolveArgs &a, const FrameUniforms &u, const uint32_t *tri,
                               const uint32_t *q, uint32_t n, float *planes, hipStream_t s);
// offs_px[p] for each pixel whose winner vis[p] is passing fragment k:
// offs[k], or k * stride when offs is null.
void launch_frag_pixel_offsets(const uint32_t *tri, const uint32_t *q, uint32_t n, int W, int H, const int *vis,
                               const uint32_t *offs, uint32_t stride, uint32_t *offs_px, hipStream_t s);

// sort.hip
// Exclusive scan of n uint32 values into out (out may alias in); the total
// is written to out[n].  tmp needs scan_tmp_words(n) words.
size_t scan_tmp_words(size_t n);
void launch_exclusive_scan(const uint32_t *in, uint32_t *out, size_t n, uint32_t *tmp,
                           hipStream_t s);
// Stable LSD radix sort of (key, value) pairs by the low `bits` key bits.
// Result ends in (keys, vals); alt buffers are scratch of the same size.
size_t radix_tmp_words(size_t n);
void launch_radix_sort(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                       size_t n, int bits, uint32_t *tmp, hipStream_t s);
// The passes only: returns 1 when the result is in the alt buffers.  With
// n_dev, n is the capacity and min(n, *n_dev) items are sorted.
int launch_radix_sort_impl(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                           size_t n, const uint32_t *n_dev, int bits, uint32_t *tmp, hipStream_t s);

// rng.hip: the reference's std::mt19937 stream (624 words, the next output
// index, 32-bit draws consumed).
struct MtState {
    uint32_t mt[624];
    uint32_t idx, pad;
    unsigned long long draws;
};
void mt_seed_state(uint32_t seed, MtState &st);
// The stream's next draws in parallel chunks of K blocks (624 K draws each):
// base = the context's array and the 32 blocks after it (kMtBaseWords
// words); states[c] (c = 1..n_jumps, 624 words each) = chunk c's starting
// array from jump polynomial c (mt_jump.hpp; kMtPolyWords words each);
// raw = the tempered draws from the context's next one on,
// (624 - idx) + n_chunks 624 K of them.
constexpr int kMtBaseWords = 33 * 624;
void launch_mt_base(const MtState *st, uint32_t *base, hipStream_t s);
void launch_mt_jump(const uint32_t *base, const uint64_t *polys, int n_jumps, uint32_t *states, hipStream_t s);
void launch_mt_gen(const MtState *st, const uint32_t *states, int n_chunks, int K, uint32_t *raw, hipStream_t s);
// Rejection-sampled unit vectors from raw's T tries of D draws (D = 3:
// RandomVectorInUnitSphere (x, y, z); 2: RandomVectorInUnitDisk (x, y)),
// kMtTryTile tries to a workgroup: part[b] = workgroup b's count of kept
// tries (mt_try_blocks(T) of them); after an exclusive scan of part (the
// total at part[nb]) the first n kept vectors go to out in stream order, and
// a commit moves the stream past the try that kept the m-th.
constexpr uint32_t kMtTryTile = 2048;
uint32_t mt_try_blocks(uint32_t T);
void launch_mt_count(int D, const uint32_t *raw, uint32_t T, uint32_t *part, hipStream_t s);
void launch_mt_take(int D, const uint32_t *raw, uint32_t T, const uint32_t *part, uint32_t n, float *out,
                    hipStream_t s);
void launch_mt_commit(int D, MtState *st, const uint32_t *raw, const uint32_t *part, uint32_t T, uint32_t m,
                      hipStream_t s);

// shade.hip
void launch_lighting(const FrameUniforms &u, const float *gbuf, const float *ao, int ao_is_one, const NetAo &net,
                     const float *shadow_map, const ShadowSamples &ss, float *fb, hipStream_t s);
// Render::DoSSAO with the stream's hemisphere samples (32 (x, y, z) per
// pixel in row order); pow3[c] = powf(1 - c / 32, 3) from the host libm.
void launch_ssao_reference(const FrameUniforms &u, const float *samples, const float *pow3,
                           const float *depth, const float *gbuf, float *ao, hipStream_t s);
// Offsets of the PCSS filter's samples for P shading points in the order
// they shade (a point takes n_blocker pairs, plus n_pcf when its blocker
// search finds a blocker); g holds each point's normal, world position and
// light NDC as planes of stride P (the first nine G-buffer planes, or the
// forward pass's fragment planes).  Classify every point as decided by the
// shadow map alone or not (cnt: its pairs when decided, n_blocker otherwise;
// unc: 1 when undecided), compact the undecided points after a scan of unc
// (upos), walk them in order from the scanned cnt (base; launch_pcss_points,
// launch_pcss_segment), then
// offs[i] = base[i] + n_pcf * (blockers found by undecided points before i)
// for i = 0..P.
void launch_pcss_classify(const FrameUniforms &u, const float *g, size_t P, const float *shadow_map, uint32_t *cnt,
                          uint32_t *unc, hipStream_t s);
void launch_pcss_compact(const uint32_t *upos, size_t P, uint32_t *list, hipStream_t s);
// The n undecided points' stream-independent inputs (shadow coordinate and
// bias, sample base), then segments [s0, s0 + L) of the ordered walk from
// M0 misses before s0 (rate: the miss rate so far; L <= kPcssSegmentMax):
// scratch pred (2 L + 2 words), scan_tmp (scan_tmp_words(L)), masks (L
// 64-bit words), fbits (kPcssChainFbits 64-bit words) and res
// (kPcssChainRes words) for the walks of the segment's blocks; out = (first
// unsettled point, misses before it).
constexpr uint32_t kPcssSegmentMax = 32768, kPcssChainBlock = 1024;
constexpr int kPcssBandWidth = 63;
constexpr size_t kPcssChainFbits = (size_t)(kPcssSegmentMax / kPcssChainBlock) * kPcssBandWidth * (kPcssChainBlock / 64);
constexpr size_t kPcssChainRes = 2 * (size_t)(kPcssSegmentMax / kPcssChainBlock) * kPcssBandWidth;
void launch_pcss_points(const FrameUniforms &u, const float *g, size_t P, const uint32_t *base, const uint32_t *list,
                        uint32_t n, float4 *pt, uint32_t *pbase, hipStream_t s);
void launch_pcss_segment(const FrameUniforms &u, const float *shadow_map, const float *pairs, const float4 *pt,
                         const uint32_t *pbase, uint32_t s0, uint32_t L, uint32_t M0, float rate, uint32_t *pred,
                         uint32_t *scan_tmp, unsigned long long *masks, unsigned long long *fbits, uint32_t *res,
                         uint32_t *found, uint32_t *out, hipStream_t s);
void launch_pcss_offsets(const uint32_t *base, const uint32_t *upos, const uint32_t *fpos, uint32_t n_pcf, size_t P,
                         uint32_t *offs, hipStream_t s);
void launch_ssao(const FrameUniforms &u, const SsaoArgs &sa, const float *depth,
                 const float *gbuf, float *ao, hipStream_t s);
void launch_blur_recursive(float *ao, int W, int H, hipStream_t s);
// Separable Gaussian of radius R (1..8), horizontal then vertical, clamped
// taps, in one launch from src into dst (distinct buffers).
void launch_blur_separable(const float *src, float *dst, int W, int H, const float *w, int R, hipStream_t s);
// counts[0..2] (zeroed by the caller) += floats compared, mismatches of
// recip() against 1.f / w, floats on recip's fast path.
void launch_check_recip(unsigned long long *counts, hipStream_t s);
void launch_resolve(const float *fb, int W, int H, int k, uint8_t *img, hipStream_t s);
void launch_copy(const void *src, void *dst, size_t bytes, hipStream_t s);

}  // namespace fgl
