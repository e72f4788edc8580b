// nao.hip -- neural ambient occlusion for gfx950: a small convolutional
// network that stands in for SSAO hemisphere sampling plus blur.
//
// The network runs on a grid `scale` times coarser than the frame (scale 1,
// 2 or 4; the shipped network uses 4): each network pixel stands for a
// scale x scale block of frame pixels.
//   down  per network pixel: the mean over the block's covered pixels of 7
//         view-space features (depth, normal xyz, position xyz) and the
//         covered fraction of the block -- 8 inputs, stored as 16 BF16
//         channels (8 zero); a per-frame-pixel coverage mask
//   L1    conv3x3 16->32 + ReLU      L2..L4 conv3x3 32->32 + ReLU
//   L5    conv1x1 32->1 + sigmoid (fused into L4)
//         the 3x3 layers are dilated by 1, 2, 6, 12: their taps reach every
//         offset in [-21, 21] (receptive field 43 x 43 network pixels, 172 x
//         172 frame pixels at scale 4)
//   up    bilinear interpolation of the network's AO between block centres
//         at every covered frame pixel; AO 1 elsewhere
// Convolutions are implicit GEMMs on MFMA (v_mfma_f32_16x16x32_bf16):
// M = pixels, N = 32 output channels, K = 9 taps x Cin, BF16 operands, FP32
// accumulation, BF16 NHWC activations.  Backward: the data gradient is the
// same kernel with tap-flipped, transposed weights and a ReLU-mask epilogue;
// the weight gradient is an MFMA GEMM over pixels (co x pixels times
// pixels x (tap, ci)) with per-block partial sums reduced in a fixed order
// (bitwise reproducible training, no atomics).  Adam keeps FP32 master
// weights.  Training runs on network-grid crops whose targets are the block
// means of the plain pipeline's AO.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "ctx_internal.hpp"
#include "fgl_math.hpp"

namespace fgl {
namespace nao {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int C0 = 16;    // stored input channels (8 real)
constexpr int NIN = 8;    // real input channels: 7 features + covered fraction
constexpr int CH = 32;    // hidden channels
constexpr int K1 = 9 * C0, K2 = 9 * CH;
// flat FP32 parameter layout
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + CH * K1;
constexpr int OFF_W2 = OFF_B1 + CH, OFF_B2 = OFF_W2 + CH * K2;
constexpr int OFF_W3 = OFF_B2 + CH, OFF_B3 = OFF_W3 + CH * K2;
constexpr int OFF_W4 = OFF_B3 + CH, OFF_B4 = OFF_W4 + CH * K2;
constexpr int OFF_W5 = OFF_B4 + CH, OFF_B5 = OFF_W5 + CH;
constexpr int NPARAM = OFF_B5 + 1;
constexpr int WG_ROWS = 16;   // rows of pixels per weight-gradient workgroup
// Dilation of the four 3x3 convolutions (taps at -d, 0, +d): sums of one tap
// offset per layer cover every offset in [-21, 21] without gaps, a 43 x 43
// receptive field at the cost of four plain 3x3 layers.  Every dilation
// above 1 is even (the tile order pairs x-phases 2m, 2m + 1).
constexpr int kMaxDil = 12;
__host__ __device__ constexpr int layer_dil(int l) { return l == 0 ? 1 : l == 1 ? 2 : l == 2 ? 6 : 12; }

__host__ __device__ __forceinline__ uint16_t f2bf(float f)
{
#if defined(__HIP_DEVICE_COMPILE__)
    // v_cvt_pk_bf16_f32 (round to nearest even): the same bits as the integer
    // rounding below for every non-NaN float (an exhaustive comparison of all
    // 2^32 patterns on gfx950 found no difference), in one instruction
    return __builtin_bit_cast(uint16_t, (__bf16)f);
#else
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
#endif
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// ---------------- features ----------------

// Network coverage of one frame pixel: something was drawn (depth below
// 3e38) and its G-buffer normal and position are finite (a degenerate
// tangent frame makes the mapped normal NaN).  Uncovered pixels take no part
// in the inputs or the loss and get AO 1 -- what kernel SSAO yields there
// too -- so one such pixel cannot poison a weight gradient.
__device__ __forceinline__ bool px_covered(float d, const float g[6])
{
    bool covered = d < 3.0e38f;
#pragma unroll
    for (int k = 0; k < 6; ++k) covered = covered && isfinite(g[k]);
    return covered;
}

// View-space features of a block from the means m of its covered pixels'
// normal N (m[0..2]) and position P (m[3..5]): depth -z_view / 10, V3 N,
// V (P, 1) / 10.  The view transform is affine, so these are the means of
// the pixels' own features up to rounding, at one transform per block.
__device__ __forceinline__ void block_view_features(const float *V, const float m[6], float f[7])
{
    const Vec4 vp = mv4(V, vec4(m[3], m[4], m[5], 1.f));
    const Vec3 N = vec3(m[0], m[1], m[2]);
    f[0] = -vp.z * 0.1f;
    f[1] = dot3(vec3(V[0], V[1], V[2]), N);
    f[2] = dot3(vec3(V[4], V[5], V[6]), N);
    f[3] = dot3(vec3(V[8], V[9], V[10]), N);
    f[4] = vp.x * 0.1f;
    f[5] = vp.y * 0.1f;
    f[6] = vp.z * 0.1f;
}

struct Mat16 {
    float m[16];
};

// Inputs of one network pixel (X, Y) from the block of frame pixels
// [X*f, X*f+f) x [Y*f, Y*f+f) clipped to the W x H frame: the covered
// pixels' normals and positions summed in row-major order from 0, times 1/n
// (n covered), through block_view_features; feat[7] = covered / block
// pixels in the frame; zeros when none is covered.  fmask (optional)
// receives each frame pixel's coverage; ao (optional) gives *ao_mean, the
// covered pixels' mean AO (1 when none).  Returns n.
__device__ __forceinline__ void block_finish(const float *V, const float acc[6], int n, int m, float feat[NIN])
{
    const float inv = n ? 1.f / (float)n : 0.f;
    float mean[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) mean[k] = acc[k] * inv;
    float fv[7];
    block_view_features(V, mean, fv);
#pragma unroll
    for (int k = 0; k < 7; ++k) feat[k] = n ? fv[k] : 0.f;
    feat[7] = m ? (float)n / (float)m : 0.f;
}

// A whole 4 x 4 block with 16-byte aligned rows: the 28 (32 with AO) row
// loads are all issued before the first is used; the same sums, in the same
// order, as the general path below.
__device__ __forceinline__ int block_features_4x4(const float *V, const float *__restrict__ depth,
                                                  const float *__restrict__ g, size_t P, int W, size_t r0,
                                                  float feat[NIN], uint8_t *__restrict__ fmask,
                                                  const float *__restrict__ ao, float *ao_mean)
{
    float4 q[4][7], qa[4];
#pragma unroll
    for (int dy = 0; dy < 4; ++dy) {
        const size_t r = r0 + (size_t)dy * W;
        q[dy][0] = *(const float4 *)(depth + r);
#pragma unroll
        for (int k = 0; k < 6; ++k) q[dy][k + 1] = *(const float4 *)(g + (size_t)k * P + r);
        if (ao) qa[dy] = *(const float4 *)(ao + r);
    }
    float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float aos = 0.f;
    int n = 0;
#pragma unroll
    for (int dy = 0; dy < 4; ++dy) {
        uint32_t mk = 0;
        const float *c[7];
        float a[4] = {qa[dy].x, qa[dy].y, qa[dy].z, qa[dy].w};
#pragma unroll
        for (int k = 0; k < 7; ++k) c[k] = &q[dy][k].x;
#pragma unroll
        for (int dx = 0; dx < 4; ++dx) {
            const float px[6] = {c[1][dx], c[2][dx], c[3][dx], c[4][dx], c[5][dx], c[6][dx]};
            const bool cov = px_covered(c[0][dx], px);
            if (cov) {
                ++n;
#pragma unroll
                for (int k = 0; k < 6; ++k) acc[k] += px[k];
                if (ao) aos += a[dx];
            }
            mk |= (cov ? 1u : 0u) << (8 * dx);
        }
        if (fmask) *(uint32_t *)(fmask + r0 + (size_t)dy * W) = mk;
    }
    block_finish(V, acc, n, 16, feat);
    if (ao_mean) *ao_mean = n ? aos * (1.f / (float)n) : 1.f;
    return n;
}

__device__ __forceinline__ int block_features(const float *V, const float *__restrict__ depth,
                                              const float *__restrict__ g, int W, int H, int X, int Y, int f,
                                              float feat[NIN], uint8_t *__restrict__ fmask,
                                              const float *__restrict__ ao, float *ao_mean)
{
    const size_t P = (size_t)W * H;
    float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float aos = 0.f;
    int n = 0, m = 0;
    const int x0 = X * f;
    // (W % 4 == 0 keeps every row start, plane and ao offset on the same
    // 16-byte alignment as `depth`; fmask then shares it at 4 bytes)
    const bool vec4 = f == 4 && (W & 3) == 0 && x0 + 3 < W && ((uintptr_t)depth & 15) == 0 &&
                      ((uintptr_t)g & 15) == 0 && (!ao || ((uintptr_t)ao & 15) == 0) &&
                      (!fmask || ((uintptr_t)fmask & 3) == 0);
    if (vec4 && Y * 4 + 3 < H)
        return block_features_4x4(V, depth, g, P, W, (size_t)(Y * 4) * W + x0, feat, fmask, ao, ao_mean);
    for (int dy = 0; dy < f; ++dy) {
        const int y = Y * f + dy;
        if (y >= H) break;
        const size_t r = (size_t)y * W + x0;
        const int cnt = W - x0 < f ? W - x0 : f;
        for (int dx = 0; dx < cnt; ++dx) {
            const size_t i = r + dx;
            const float px[6] = {g[i], g[P + i], g[2 * P + i], g[3 * P + i], g[4 * P + i], g[5 * P + i]};
            const bool cov = px_covered(depth[i], px);
            ++m;
            if (cov) {
                ++n;
                for (int k = 0; k < 6; ++k) acc[k] += px[k];
                if (ao) aos += ao[i];
            }
            if (fmask) fmask[i] = cov ? 1 : 0;
        }
    }
    block_finish(V, acc, n, m, feat);
    if (ao_mean) *ao_mean = n ? aos * (1.f / (float)n) : 1.f;
    return n;
}

// The 16 stored BF16 input channels of one network pixel; with
// zero_pad = false only the 8 real ones (the buffer's channels 8-15 were
// zeroed when it was allocated and are never written).
__device__ __forceinline__ void store_inputs(uint16_t *__restrict__ dst, const float feat[NIN], bool zero_pad = true)
{
    uint16_t b[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) b[c] = f2bf(feat[c]);
    uint4 *o = (uint4 *)dst;
    o[0] = make_uint4(b[0] | ((uint32_t)b[1] << 16), b[2] | ((uint32_t)b[3] << 16), b[4] | ((uint32_t)b[5] << 16),
                      b[6] | ((uint32_t)b[7] << 16));
    if (zero_pad) o[1] = make_uint4(0u, 0u, 0u, 0u);
}

// Per-view pixel layout: the view's frame pixels start at px_off, its
// network pixels at lo_off; frame blocks of 64 x 4 pixels are blocks
// blk_base.. of the batch's frame block table, network blocks of 64 x 4
// network pixels blocks lblk_base.. of its network block table.
__device__ __forceinline__ const BatchPixView &view_of_block(const BatchPixView *__restrict__ views,
                                                             const uint8_t *__restrict__ blk_view, int b)
{
    return views[blk_view[b]];
}

// Down pass: network inputs, network-pixel coverage and frame-pixel
// coverage.  One thread per network pixel, 64 x 4 network pixels per
// workgroup.  Single frame: `one` describes it; batch: `views` and the
// network block table.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_nao_down(BatchPixView one, const BatchPixView *__restrict__ views,
                                                  const uint8_t *__restrict__ lblk_view, int f,
                                                  const float *__restrict__ depth, const float *__restrict__ gbuf,
                                                  uint16_t *__restrict__ feat, uint8_t *__restrict__ lmask,
                                                  uint8_t *__restrict__ fmask)
{
    const BatchPixView &v = BATCH ? view_of_block(views, lblk_view, blockIdx.x) : one;
    const int k = BATCH ? (int)(blockIdx.x - v.lblk_base) : (int)blockIdx.x;
    const int bx = (v.lW + 63) / 64;
    const int X = (k % bx) * 64 + (int)(threadIdx.x & 63), Y = (k / bx) * 4 + (int)(threadIdx.x >> 6);
    if (X >= v.lW || Y >= v.lH) return;
    float in[NIN];
    const int n = block_features(v.view, depth + v.px_off, gbuf + (size_t)kGbufPlanes * v.px_off, v.W, v.H, X, Y, f,
                                 in, fmask + v.px_off, nullptr, nullptr);
    const size_t o = v.lo_off + (size_t)Y * v.lW + X;
    store_inputs(feat + o * C0, in, false);
    lmask[o] = n > 0 ? 1 : 0;
}

// Up pass: AO of every frame pixel -- 1 where the pixel is uncovered, else
// the network AO interpolated bilinearly between the centres of the blocks
// around the pixel's centre (block (X, Y) has its centre at frame
// coordinate f X + f/2; edges clamp).  One wavefront per 64 x 4 frame-pixel
// block, four consecutive pixels of a row per lane (one 4-byte mask load and
// one 16-byte store where the row is aligned).
__device__ __forceinline__ float up_value(const BatchPixView &v, const float *__restrict__ L, int f, int x, int y)
{
    return net_ao_interp(L, v.lW, v.lH, f, x, y);
}

// Four pixels x .. x+3 of row y, x a multiple of 4, f = 4: the sample
// points share the row weights and fall between network columns X - 1 and X
// (first two) or X and X + 1 (last two), X = x / 4 -- six network values
// instead of sixteen, and the same products and sums as net_ao_interp.
__device__ __forceinline__ void up_value4(const BatchPixView &v, const float *__restrict__ L, int x, int y,
                                          float out[4])
{
    const float inv = 1.f / 4.f;
    const float w = ((float)y + 0.5f) * inv - 0.5f;
    const float fw = floorf(w);
    const float ty = w - fw;
    const int Y0 = (int)fw;
    const int ya = imax(0, imin(v.lH - 1, Y0)), yb = imax(0, imin(v.lH - 1, Y0 + 1));
    const int X = x >> 2;
    const int c0 = imax(0, imin(v.lW - 1, X - 1)), c1 = imax(0, imin(v.lW - 1, X)), c2 = imax(0, imin(v.lW - 1, X + 1));
    const float *ra = L + (size_t)ya * v.lW, *rb = L + (size_t)yb * v.lW;
    const float a0 = ra[c0], a1 = ra[c1], a2 = ra[c2], b0 = rb[c0], b1 = rb[c1], b2 = rb[c2];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const float u = ((float)(x + e) + 0.5f) * inv - 0.5f;
        const float tx = u - floorf(u);
        const float al = e < 2 ? a0 : a1, ar = e < 2 ? a1 : a2, bl = e < 2 ? b0 : b1, br = e < 2 ? b1 : b2;
        const float top = al * (1.f - tx) + ar * tx;
        const float bot = bl * (1.f - tx) + br * tx;
        out[e] = top * (1.f - ty) + bot * ty;
    }
}

// Four waves per workgroup, one 64 x 4 frame-pixel block each.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_nao_up(BatchPixView one, const BatchPixView *__restrict__ views,
                                                const uint8_t *__restrict__ blk_view, int n_blocks, int f,
                                                const float *__restrict__ lao, const uint8_t *__restrict__ fmask,
                                                float *__restrict__ ao)
{
    const int b = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (b >= n_blocks) return;
    const BatchPixView &v = BATCH ? view_of_block(views, blk_view, b) : one;
    const int k = BATCH ? (int)(b - v.blk_base) : b;
    const int lane = threadIdx.x & 63;
    const int bx = (v.W + 63) / 64;
    const int x = (k % bx) * 64 + 4 * (lane & 15), y = (k / bx) * 4 + (lane >> 4);
    if (x >= v.W || y >= v.H) return;
    const size_t i = v.px_off + (size_t)y * v.W + x;
    const float *L = lao + v.lo_off;
    if (x + 3 < v.W && (i & 3) == 0) {
        const uint32_t mk = *(const uint32_t *)(fmask + i);
        float u[4];
        if (f == 4) {
            up_value4(v, L, x, y, u);
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) u[e] = up_value(v, L, f, x + e, y);
        }
        float4 o;
        o.x = (mk & 0xffu) ? u[0] : 1.f;
        o.y = (mk & 0xff00u) ? u[1] : 1.f;
        o.z = (mk & 0xff0000u) ? u[2] : 1.f;
        o.w = (mk & 0xff000000u) ? u[3] : 1.f;
        *(float4 *)(ao + i) = o;
        return;
    }
    for (int e = 0; e < 4 && x + e < v.W; ++e) ao[i + e] = fmask[i + e] ? up_value(v, L, f, x + e, y) : 1.f;
}

// Receptive-field radius of the network in network pixels (sum of the
// dilations): a crop pixel closer than this to a crop edge that lies inside
// the frame sees zero padding where the frame continues.
constexpr int kReach = 1 + 2 + 6 + 12;

// One crop of the current frame's network grid into batch slot `slot`:
// inputs, target (the block mean of the frame's AO over covered pixels) and
// mask: the block has a covered pixel and lies at least kReach network
// pixels inside every crop edge that is not a frame edge (the loss then
// only sees pixels whose inputs the crop holds as the whole frame does).
// Crop origin (x0, y0) and size S x S in network pixels, inside the lW x lH
// network grid.
__global__ __launch_bounds__(256) void k_crop(Mat16 V, const float *__restrict__ depth,
                                              const float *__restrict__ g,
                                              const float *__restrict__ ao, int W, int H, int f, int lW, int lH,
                                              int x0, int y0, int S, int slot,
                                              uint16_t *__restrict__ feat,
                                              float *__restrict__ target,
                                              uint8_t *__restrict__ mask)
{
    const int cx = blockIdx.x * 64 + (threadIdx.x & 63), cy = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (cx >= S || cy >= S) return;
    const size_t o = (size_t)slot * S * S + (size_t)cy * S + cx;
    float in[NIN], t;
    const int n = block_features(V.m, depth, g, W, H, x0 + cx, y0 + cy, f, in, nullptr, ao, &t);
    store_inputs(feat + o * C0, in);
    target[o] = t;
    const bool inside = (x0 == 0 || cx >= kReach) && (x0 + S == lW || cx < S - kReach) &&
                        (y0 == 0 || cy >= kReach) && (y0 + S == lH || cy < S - kReach);
    mask[o] = n > 0 && inside ? 1 : 0;
}

// ---------------- 3x3 convolution on MFMA ----------------

// Head pre-activation: conv1x1 32 -> 1 over one pixel's BF16 activations, in
// channel order.
// z = b5 + ((p0 + p1) + (p2 + p3)), p_g the sum, in channel order, over the
// channels g*4 .. g*4+3 and 16+g*4 .. 16+g*4+3 -- the channels one lane group
// holds after an MFMA, so the fused inference kernels form the same sum with
// two lane exchanges.
__device__ __forceinline__ float head_z(const float *w5, float b5, const uint16_t *a)
{
    float p[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        float sg = 0.f;
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int c = cb * 16 + g * 4 + i;
                sg += w5[c] * bf2f(a[c]);
            }
        p[g] = sg;
    }
    return b5 + ((p[0] + p[1]) + (p[2] + p[3]));
}

// Head of the network fused into the last convolution (inference): w5 = the
// 32 head weights followed by the bias; pixels whose mask is 0 keep AO 1.
struct HeadArgs {
    const float *w5;
    const uint8_t *mask;
    float *ao;
};

// MODE 0: out = bf16(relu(conv(in) + bias)).
// MODE 1 (data gradient): out = bf16(conv(in) * (act > 0)) with tap-flipped,
// transposed weights; `in` is the layer's output gradient.
// MODE 2 (last layer of inference): as MODE 0, but each row of outputs goes
// to the wavefront's LDS row buffer and the head -- head_z, then
// 1 / (1 + expf(-z)), AO 1 where the mask is 0 (the background keeps
// AmbientOcclusionGBuffer's initial value) -- writes AO; the layer's
// activations never reach HBM.
//
// Implicit GEMM on v_mfma_f32_16x16x32_bf16 with D = W x patches: M = 16
// output channels, N = 16 pixels, K = 32 (one tap of 32 channels, or two taps
// of 16).  A (weights) stays in registers for the wavefront's whole life; B
// (patches) fragments are 16-byte loads -- lane l: pixel l % 16, channels
// 8 * (l / 16) .. + 8 of its tap -- so a wavefront reads 1 KB contiguous per
// fragment; each lane ends with 4 consecutive output channels of one pixel,
// stored as one 8-byte write.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <int CIN>
struct ConvWeights {
    static constexpr int K = 9 * CIN, KC = (K + 31) / 32;
    v8bf w[KC][2];  // [k chunk][16-channel block]
};

template <int CIN>
__device__ __forceinline__ void load_weights(ConvWeights<CIN> &cw, const uint16_t *__restrict__ Wt)
{
    constexpr int K = ConvWeights<CIN>::K, KC = ConvWeights<CIN>::KC;
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int c = 0; c < KC; ++c)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            const int k0 = c * 32 + lq * 8;
            if (k0 < K) cw.w[c][cb] = *(const v8bf *)(Wt + (size_t)(cb * 16 + lr) * K + k0);
            else cw.w[c][cb] = v8bf{};
        }
}

// ---- phase-tiled convolution with an LDS input tile ----
//
// A 3x3 convolution dilated by d only connects pixels of one phase (x mod d,
// y mod d): on that phase's grid it is a plain 3x3 convolution.  A workgroup
// takes a 64 x 8 tile of one phase grid, stages the tile plus a one-pixel
// halo (66 x 10 pixels, all channels) in LDS once -- every input pixel is
// then read from HBM/L2 about 1.3 times instead of 9 -- and its four
// wavefronts compute two rows each from LDS.  Pixels are padded to 80 bytes
// in LDS so the 16-byte fragment reads of 16 consecutive pixels spread over
// the banks.
constexpr int kTileW = 64, kTileH = 8;
constexpr int kStageW = kTileW + 2, kStageH = kTileH + 2;
constexpr int kTilePB = kTileW / 16;   // 16-pixel blocks per tile row
constexpr int kStagePix = 40;  // bf16 per staged pixel (32 channels + 8 pad)

// Tile geometry: image base pixel, phase, tile column / row.
struct ConvTile {
    size_t img;
    int W, H, phx, phy, bx, by;
};

template <int CIN>
struct ConvStage {
    static constexpr int CPP = CIN / 8;                                  // 16-byte chunks per pixel
    static constexpr int N = (kStageH * kStageW * CPP + 255) / 256;      // chunks per thread
    uint4 v[N];
};

// Global -> registers: this thread's chunks of the tile plus halo (zeros
// outside the phase grid).
template <int CIN>
__device__ __forceinline__ void stage_load(ConvStage<CIN> &st, const uint16_t *__restrict__ in,
                                           const ConvTile &t, int dil)
{
    constexpr int CPP = ConvStage<CIN>::CPP, TOT = kStageH * kStageW * CPP;
    const int Wp = (t.W - t.phx + dil - 1) / dil, Hp = (t.H - t.phy + dil - 1) / dil;
    const int i0 = t.bx * kTileW - 1, j0 = t.by * kTileH - 1;
#pragma unroll
    for (int q = 0; q < ConvStage<CIN>::N; ++q) {
        const int e = (int)threadIdx.x + 256 * q;
        st.v[q] = make_uint4(0u, 0u, 0u, 0u);
        if (e < TOT) {
            const int pix = e / CPP, c = e % CPP;
            const int i = i0 + pix % kStageW, j = j0 + pix / kStageW;
            if (i >= 0 && i < Wp && j >= 0 && j < Hp)
                st.v[q] = *(const uint4 *)(in + (t.img + (size_t)(t.phy + j * dil) * t.W + t.phx + i * dil) * CIN +
                                           c * 8);
        }
    }
}

template <int CIN>
__device__ __forceinline__ void stage_store(const ConvStage<CIN> &st, uint16_t *s_in)
{
    constexpr int CPP = ConvStage<CIN>::CPP, TOT = kStageH * kStageW * CPP;
#pragma unroll
    for (int q = 0; q < ConvStage<CIN>::N; ++q) {
        const int e = (int)threadIdx.x + 256 * q;
        if (e < TOT) *(uint4 *)(s_in + (e / CPP) * kStagePix + (e % CPP) * 8) = st.v[q];
    }
}

// MODE 2 epilogue of one output row (64 pixels of the phase grid row y): the
// BF16 activations bf16(relu(acc + bias)) of 32 pixels at a time go to this
// wavefront's LDS rows (32 x 80 B), then lanes 0-31 take one pixel each
// through the head.  LDS operations of one wavefront execute in order.
__device__ __forceinline__ void head_row(const v4f (&acc)[kTilePB][2], const float *__restrict__ bias,
                                         const HeadArgs &hd, uint16_t *sh, const ConvTile &t, int y, int Wp,
                                         int dil)
{
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int h = 0; h < kTilePB / 2; ++h) {
#pragma unroll
        for (int q2 = 0; q2 < 2; ++q2) {
            const int pb = 2 * h + q2;
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                const int co = cb * 16 + lq * 4;
                const float4 bv = *(const float4 *)(bias + co);
                const float bb[4] = {bv.x, bv.y, bv.z, bv.w};
                uint16_t rr[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float v = acc[pb][cb][q] + bb[q];
                    rr[q] = f2bf(v > 0.f ? v : 0.f);
                }
                *(uint64_t *)(sh + (q2 * 16 + lr) * kStagePix + co) =
                    (uint64_t)rr[0] | ((uint64_t)rr[1] << 16) | ((uint64_t)rr[2] << 32) | ((uint64_t)rr[3] << 48);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int i = t.bx * kTileW + 32 * h + lane;
        if (lane < 32 && i < Wp) {
            const size_t p = t.img + (size_t)y * t.W + t.phx + (size_t)i * dil;
            const float z = head_z(hd.w5, hd.w5[CH], sh + lane * kStagePix);
            float v = 1.f / (1.f + expf(-z));
            if (!hd.mask[p]) v = 1.f;
            hd.ao[p] = v;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// The staged tile through the 3x3 taps: wave w computes rows 2w, 2w+1.
template <int CIN, int MODE>
__device__ __forceinline__ void conv_tile_compute(const ConvWeights<CIN> &cw, const uint16_t *s_in,
                                                  const float *__restrict__ bias, const uint16_t *__restrict__ act,
                                                  uint16_t *__restrict__ out, const ConvTile &t, int dil,
                                                  const HeadArgs &hd, uint16_t *s_head)
{
    constexpr int K = ConvWeights<CIN>::K, KC = ConvWeights<CIN>::KC;
    const int Wp = (t.W - t.phx + dil - 1) / dil, Hp = (t.H - t.phy + dil - 1) / dil;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lq = lane >> 4;
#pragma unroll 1
    for (int r = 0; r < kTileH / 4; ++r) {
        const int jr = wave * (kTileH / 4) + r, j = t.by * kTileH + jr;
        if (j >= Hp) break;
        v4f acc[kTilePB][2];
#pragma unroll
        for (int pb = 0; pb < kTilePB; ++pb)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) acc[pb][cb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            const int k0 = c * 32 + lq * 8;
            const int tap = imin(k0 / CIN, 8), ci0 = k0 % CIN;
            const int srow = jr + tap / 3, scol = lr + tap % 3;
            v8bf b[kTilePB];
#pragma unroll
            for (int pb = 0; pb < kTilePB; ++pb) {
                b[pb] = v8bf{};
                if (k0 < K) b[pb] = *(const v8bf *)(s_in + (srow * kStageW + pb * 16 + scol) * kStagePix + ci0);
            }
#pragma unroll
            for (int pb = 0; pb < kTilePB; ++pb) {
                acc[pb][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cw.w[c][0], b[pb], acc[pb][0], 0, 0, 0);
                acc[pb][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cw.w[c][1], b[pb], acc[pb][1], 0, 0, 0);
            }
        }
        const int y = t.phy + j * dil;
        if (MODE == 2) {
            head_row(acc, bias, hd, s_head + wave * (32 * kStagePix), t, y, Wp, dil);
            continue;
        }
#pragma unroll
        for (int pb = 0; pb < kTilePB; ++pb) {
            const int i = t.bx * kTileW + pb * 16 + lr;
            if (i >= Wp) continue;
            const int x = t.phx + i * dil;
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                const int co = cb * 16 + lq * 4;
                const size_t o = (t.img + (size_t)y * t.W + x) * CH + co;
                uint16_t rr[4];
                if (MODE == 0) {
                    const float4 bv = *(const float4 *)(bias + co);
                    const float bb[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float v = acc[pb][cb][q] + bb[q];
                        rr[q] = f2bf(v > 0.f ? v : 0.f);
                    }
                } else {
                    const uint64_t am = *(const uint64_t *)(act + o);
#pragma unroll
                    for (int q = 0; q < 4; ++q) rr[q] = f2bf((uint16_t)(am >> (16 * q)) != 0 ? acc[pb][cb][q] : 0.f);
                }
                *(uint64_t *)(out + o) = (uint64_t)rr[0] | ((uint64_t)rr[1] << 16) | ((uint64_t)rr[2] << 32) |
                                         ((uint64_t)rr[3] << 48);
            }
        }
    }
}

// Tile t of a single-image-set launch: tiles_x x tiles_y tiles per phase
// grid, dil^2 phases per image.
// Order of the tiles of one image: a 128-byte line holds two neighbouring
// pixels of a row (64 B each), which belong to x-phases 2m and 2m+1 when the
// convolution is dilated.  Tile u and tile u + 8 take the same spatial tile of
// those two phases: the persistent workgroups stride over tiles by a multiple
// of 8, so both run on the same XCD at about the same time and the line is
// fetched from HBM once (in phase order the second half of every line came
// back a whole phase later, out of L2: 2x the read traffic).  Returns the
// phase (phy * dil + phx) and the tile within the phase grid.
__host__ __device__ __forceinline__ void tile_order(int u, int per_phase, int dil, int &ph, int &k)
{
    if (dil == 1) {
        ph = 0;
        k = u;
        return;
    }
    const int M = per_phase * dil * dil / 2, full = M / 8 * 8;   // tiles of one phase pair
    int L, lo;
    if (u < 2 * full) {
        const int w = u % 16;
        lo = w / 8;
        L = 8 * (u / 16) + w % 8;
    } else {
        const int v = u - 2 * full;
        lo = v % 2;
        L = full + v / 2;
    }
    const int half = dil / 2, hp = L / per_phase;
    k = L % per_phase;
    ph = (hp / half) * dil + 2 * (hp % half) + lo;
}

__device__ __forceinline__ ConvTile single_tile(int t, int W, int H, int dil, int tiles_x, int tiles_y)
{
    const int per_phase = tiles_x * tiles_y, per_img = per_phase * dil * dil;
    const int n = t / per_img;
    int ph, k;
    tile_order(t % per_img, per_phase, dil, ph, k);
    ConvTile ct;
    ct.img = (size_t)n * H * W;
    ct.W = W;
    ct.H = H;
    ct.phx = ph % dil;
    ct.phy = ph / dil;
    ct.bx = k % tiles_x;
    ct.by = k / tiles_x;
    return ct;
}

__device__ __forceinline__ ConvTile batch_tile(const BatchPixView *__restrict__ views, const uint2 *__restrict__ items,
                                               int t, int dil)
{
    const uint2 it = items[t];
    const BatchPixView &bv = views[it.x & 0xff];
    const int ph = (int)(it.x >> 8);
    ConvTile ct;   // on the view's network grid
    ct.img = bv.lo_off;
    ct.W = bv.lW;
    ct.H = bv.lH;
    ct.phx = ph % dil;
    ct.phy = ph / dil;
    ct.bx = (int)(it.y & 0xffff);
    ct.by = (int)(it.y >> 16);
    return ct;
}

// Persistent workgroups stride over the tiles; the next tile's global reads
// are in flight while the current tile is computed from LDS.
template <int CIN, int MODE, bool BATCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_conv3x3_tiled(const uint16_t *__restrict__ in,
                                                       const uint16_t *__restrict__ Wt,
                                                       const float *__restrict__ bias,
                                                       const uint16_t *__restrict__ act,
                                                       uint16_t *__restrict__ out, int W, int H, int dil,
                                                       int n_tiles, int tiles_x, int tiles_y,
                                                       const BatchPixView *__restrict__ views,
                                                       const uint2 *__restrict__ items, HeadArgs hd)
{
    __shared__ __attribute__((aligned(16))) uint16_t s_in[kStageH * kStageW * kStagePix];
    __shared__ __attribute__((aligned(16))) uint16_t s_head[MODE == 2 ? 4 * 32 * kStagePix : 8];
    ConvWeights<CIN> cw;
    load_weights<CIN>(cw, Wt);
    int t = blockIdx.x;
    if (t >= n_tiles) return;
    ConvTile cur = BATCH ? batch_tile(views, items, t, dil) : single_tile(t, W, H, dil, tiles_x, tiles_y);
    ConvStage<CIN> st;
    stage_load<CIN>(st, in, cur, dil);
    for (;;) {
        __syncthreads();   // the previous tile is no longer read from LDS
        stage_store<CIN>(st, s_in);
        __syncthreads();
        const int nt = t + (int)gridDim.x;
        ConvTile nxt = cur;
        if (nt < n_tiles) {
            nxt = BATCH ? batch_tile(views, items, nt, dil) : single_tile(nt, W, H, dil, tiles_x, tiles_y);
            stage_load<CIN>(st, in, nxt, dil);
        }
        conv_tile_compute<CIN, MODE>(cw, s_in, bias, act, out, cur, dil, hd, s_head);
        if (nt >= n_tiles) break;
        t = nt;
        cur = nxt;
    }
}

static inline void conv_tiles(int W, int H, int N, int dil, int &tx, int &ty, int &n)
{
    const int Wp = (W + dil - 1) / dil, Hp = (H + dil - 1) / dil;
    tx = (Wp + kTileW - 1) / kTileW;
    ty = (Hp + kTileH - 1) / kTileH;
    n = tx * ty * dil * dil * N;
}

// Persistent grid: `per_cu` workgroups per compute unit -- three for the
// 16-channel first layer (52 KB of LDS, 152 registers: three fit), two for
// the 32-channel layers (220-233 registers allow two); more resident
// workgroups keep more tile loads in flight.
static int g_cus = 0;

static inline unsigned conv_workgroups(int n_tiles, int per_cu)
{
    if (g_cus == 0) {
        int dev = 0, cu = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu < 1) cu = 256;
        g_cus = cu;
    }
    const int wgs = per_cu * g_cus;
    return (unsigned)(n_tiles < wgs ? n_tiles : wgs);
}

template <int CIN, int MODE>
static void launch_conv(hipStream_t s, const uint16_t *in, const uint16_t *Wt, const float *bias,
                        const uint16_t *act, uint16_t *out, int W, int H, int N, int dil)
{
    int tx, ty, n;
    conv_tiles(W, H, N, dil, tx, ty, n);
    if (n == 0) return;
    hipLaunchKernelGGL((k_conv3x3_tiled<CIN, MODE, false>), dim3(conv_workgroups(n, CIN == C0 ? 3 : 2)), dim3(256), 0, s, in, Wt, bias,
                       act, out, W, H, dil, n, tx, ty, (const BatchPixView *)nullptr, (const uint2 *)nullptr,
                       HeadArgs{nullptr, nullptr, nullptr});
}

// Last layer with the head fused (single image set of N images).
static void launch_conv_head(hipStream_t s, const uint16_t *in, const uint16_t *Wt, const float *bias,
                             const HeadArgs &hd, int W, int H, int N, int dil)
{
    int tx, ty, n;
    conv_tiles(W, H, N, dil, tx, ty, n);
    if (n == 0) return;
    hipLaunchKernelGGL((k_conv3x3_tiled<CH, 2, false>), dim3(conv_workgroups(n, 2)), dim3(256), 0, s, in, Wt, bias,
                       (const uint16_t *)nullptr, (uint16_t *)nullptr, W, H, dil, n, tx, ty,
                       (const BatchPixView *)nullptr, (const uint2 *)nullptr, hd);
}

template <int CIN, int MODE = 0>
static void launch_conv_batch(hipStream_t s, const BatchPixView *views, const uint2 *items, int n, const uint16_t *in,
                              const uint16_t *Wt, const float *bias, uint16_t *out, int dil,
                              const HeadArgs &hd = HeadArgs{nullptr, nullptr, nullptr})
{
    if (n == 0) return;
    hipLaunchKernelGGL((k_conv3x3_tiled<CIN, MODE, true>), dim3(conv_workgroups(n, CIN == C0 ? 3 : 2)), dim3(256), 0, s, in, Wt, bias,
                       (const uint16_t *)nullptr, out, 0, 0, dil, n, 0, 0, views, items, hd);
}

// ---------------- fused inference: layers 1+2, then 3+4 with the head ----------------
//
// Inference needs no activation in HBM but the one between the two fused
// kernels.  Layers 1 (dilation 1) and 2 (dilation 2) run on one tile of the
// network grid: a workgroup stages its 16-channel inputs with a 3-pixel halo,
// computes layer 1 on the tile plus a 2-pixel halo into LDS (zero where that
// pixel lies outside the frame: the padding layer 2 sees), then layer 2 on the
// tile.  Layers 3 (dilation 6) and 4 (dilation 12) only connect pixels of one
// phase of the 6 x 6 grid, where they are 3x3 convolutions dilated by 1 and 2:
// a workgroup takes a tile of one phase grid the same way and ends with the
// head, summed across lanes.  Every output is the same sum in the same order
// as the layer-by-layer kernels (same K chunks, same BF16 rounding between
// layers, head_z's order): the results are identical bit for bit.  Each
// layer is computed over a flat list of its region's pixels, 16 per MFMA
// column block.  One tile per workgroup, two workgroups per CU (LDS): the
// tile sizes put a 1920 x 1080 frame's grids (480 x 270) in one round of the
// 256 CUs (432 and 504 tiles).
constexpr int kF12W = 64, kF12H = 5;                    // layer-2 output tile
constexpr int kF12OW = kF12W + 4, kF12OH = kF12H + 4;   // layer-1 outputs it reads
constexpr int kF12IW = kF12W + 6, kF12IH = kF12H + 6;   // layer-1 inputs
constexpr int kF34W = 40, kF34H = 7;                    // layer-4 output tile on a phase grid
constexpr int kF34OW = kF34W + 4, kF34OH = kF34H + 4;   // layer-3 outputs it reads
constexpr int kF34IW = kF34W + 6, kF34IH = kF34H + 6;   // layer-3 inputs
constexpr int kF34Dil = 6;                              // phase grid of layers 3 and 4

// LDS pixel layouts of the fused kernels: 16 bf16 per 16-byte chunk, no
// padding; the chunks of pixel q are permuted by an XOR of bits of q, so the
// 16-byte reads of one chunk of 16 consecutive pixels (an MFMA B fragment)
// touch distinct banks.  Offsets in bf16.
template <int CPP>
__device__ __forceinline__ int lds_px(int q, int c)
{
    return q * (CPP * 8) + c * 8;
}

// Staging of a (RW x RH) region of a (phase) grid, global -> registers ->
// LDS: pixel (i, j) of the region is grid pixel (i0 + i, j0 + j), i.e.
// network pixel (phx + step * (i0 + i), phy + step * (j0 + j)); zeros outside
// the Wg x Hg grid.  CPP 16-byte chunks per pixel.  All loads are issued
// before the first store.
// LOAD = the leading chunks of each pixel read from memory (the rest are the
// input's zero channels and staged as zeros).
template <int RW, int RH, int CPP, int LOAD = CPP>
__device__ __forceinline__ void stage_region(uint16_t *s, const uint16_t *__restrict__ in, size_t base, int lW,
                                             int phx, int phy, int step, int i0, int j0, int Wg, int Hg)
{
    constexpr int TOT = RW * RH * LOAD, N = (TOT + 255) / 256;
    uint4 v[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const int e = (int)threadIdx.x + 256 * q;
        v[q] = make_uint4(0u, 0u, 0u, 0u);
        if (e < TOT) {
            const int pix = e / LOAD, c = e % LOAD;
            const int i = i0 + pix % RW, j = j0 + pix / RW;
            if (i >= 0 && i < Wg && j >= 0 && j < Hg)
                v[q] = *(const uint4 *)(in + (base + (size_t)(phy + j * step) * lW + phx + (size_t)i * step) *
                                                 (CPP * 8) +
                                        c * 8);
        }
    }
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const int e = (int)threadIdx.x + 256 * q;
        if (e < TOT) {
            *(uint4 *)(s + lds_px<CPP>(e / LOAD, e % LOAD)) = v[q];
#pragma unroll
            for (int z = LOAD; z < CPP; ++z) *(uint4 *)(s + lds_px<CPP>(e / LOAD, z)) = make_uint4(0u, 0u, 0u, 0u);
        }
    }
}

// bf16(relu(acc + bias)) of one lane's 2 x 4 outputs (channels cb*16 + lq*4
// + 0..3), packed.
__device__ __forceinline__ uint64_t relu_pack(const v4f &acc, const float (&bb)[4])
{
    uint16_t rr[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float v = acc[q] + bb[q];
        rr[q] = f2bf(v > 0.f ? v : 0.f);
    }
    return (uint64_t)rr[0] | ((uint64_t)rr[1] << 16) | ((uint64_t)rr[2] << 32) | ((uint64_t)rr[3] << 48);
}

// One layer's 16-pixel column block: output pixel p of a region of row pitch
// RW reads input pixel (p / RW + dy * D, p % RW + dx * D) of an IW-wide LDS
// region for tap (dy, dx).  Returns the two 16-channel accumulators.
template <int CIN, int RW, int NPIX, int IW, int D>
__device__ __forceinline__ void conv_block(const ConvWeights<CIN> &cw, const uint16_t *s_in, int p, v4f (&acc)[2])
{
    constexpr int K = ConvWeights<CIN>::K, KC = ConvWeights<CIN>::KC, CPP = CIN / 8;
    const int lq = (threadIdx.x & 63) >> 4;
    const int pc = p < NPIX ? p : NPIX - 1;
    const int ry = pc / RW, rx = pc % RW;
    acc[0] = v4f{0.f, 0.f, 0.f, 0.f};
    acc[1] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < KC; ++c) {
        const int k0 = c * 32 + lq * 8;
        const int tap = imin(k0 / CIN, 8), ci0 = k0 % CIN;
        v8bf b = v8bf{};
        if (k0 < K) b = *(const v8bf *)(s_in + lds_px<CPP>((ry + (tap / 3) * D) * IW + rx + (tap % 3) * D, ci0 / 8));
        acc[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cw.w[c][0], b, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cw.w[c][1], b, acc[1], 0, 0, 0);
    }
}

// A layer over the NPIX pixels of a region into LDS (32 channels, lds_px
// layout), zero where `inside` says the pixel is off the grid.  Wavefront w
// takes column blocks w, w + 4, ...
template <int CIN, int RW, int NPIX, int IW, int D, typename Inside>
__device__ __forceinline__ void layer_to_lds(const ConvWeights<CIN> &cw, const uint16_t *s_in,
                                             const float *__restrict__ bias, uint16_t *s_out, Inside inside)
{
    constexpr int NBLK = (NPIX + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lq = lane >> 4;
    float bb[2][4];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        const float4 bv = *(const float4 *)(bias + cb * 16 + lq * 4);
        bb[cb][0] = bv.x; bb[cb][1] = bv.y; bb[cb][2] = bv.z; bb[cb][3] = bv.w;
    }
#pragma unroll 1
    for (int blk = wave; blk < NBLK; blk += 4) {
        const int p = blk * 16 + lr;
        v4f acc[2];
        conv_block<CIN, RW, NPIX, IW, D>(cw, s_in, p, acc);
        if (p >= NPIX) continue;
        const bool in = inside(p / RW, p % RW);
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
            *(uint64_t *)(s_out + lds_px<4>(p, cb * 2 + (lq >> 1)) + (lq & 1) * 4) = in ? relu_pack(acc[cb], bb[cb]) : 0ull;
    }
}

// Tile of a fused launch: image base, grid size, phase, tile column / row.
struct FusedTile {
    size_t base;
    int lW, lH, phx, phy, bx, by;
};

template <bool BATCH>
__device__ __forceinline__ FusedTile fused_tile(const BatchPixView &one, const BatchPixView *__restrict__ views,
                                                const uint2 *__restrict__ items, int t, int tiles_x, int tiles_y,
                                                int dil)
{
    FusedTile ft;
    if (BATCH) {
        const uint2 it = items[t];
        const BatchPixView &bv = views[it.x & 0xff];
        const int ph = (int)(it.x >> 8);
        ft.base = bv.lo_off;
        ft.lW = bv.lW;
        ft.lH = bv.lH;
        ft.phx = ph % dil;
        ft.phy = ph / dil;
        ft.bx = (int)(it.y & 0xffff);
        ft.by = (int)(it.y >> 16);
    } else {
        int ph, k;
        tile_order(t, tiles_x * tiles_y, dil, ph, k);
        ft.base = 0;
        ft.lW = one.lW;
        ft.lH = one.lH;
        ft.phx = ph % dil;
        ft.phy = ph / dil;
        ft.bx = k % tiles_x;
        ft.by = k / tiles_x;
    }
    return ft;
}

// Layers 1 and 2, one tile per workgroup: `in` 16-channel inputs -> `out`
// layer-2 activations.
template <bool BATCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_nao_l12(
    BatchPixView one, const BatchPixView *__restrict__ views, const uint2 *__restrict__ items, int n_tiles,
    int tiles_x, int tiles_y, const uint16_t *__restrict__ in, const uint16_t *__restrict__ W1,
    const float *__restrict__ b1, const uint16_t *__restrict__ W2, const float *__restrict__ b2,
    uint16_t *__restrict__ out)
{
    __shared__ __attribute__((aligned(16))) uint16_t s_x[kF12IW * kF12IH * C0];
    __shared__ __attribute__((aligned(16))) uint16_t s_a[kF12OW * kF12OH * CH];
    constexpr int NPIX = kF12W * kF12H, NBLK = (NPIX + 15) / 16;
    if ((int)blockIdx.x >= n_tiles) return;
    const FusedTile t = fused_tile<BATCH>(one, views, items, (int)blockIdx.x, tiles_x, tiles_y, 1);
    const int x0 = t.bx * kF12W, y0 = t.by * kF12H, lW = t.lW, lH = t.lH;
    ConvWeights<C0> w1;
    load_weights<C0>(w1, W1);
    stage_region<kF12IW, kF12IH, C0 / 8, NIN / 8>(s_x, in, t.base, lW, 0, 0, 1, x0 - 3, y0 - 3, lW, lH);
    ConvWeights<CH> w2;
    load_weights<CH>(w2, W2);
    __syncthreads();
    layer_to_lds<C0, kF12OW, kF12OW * kF12OH, kF12IW, 1>(w1, s_x, b1, s_a, [=](int ry, int rx) {
        const int gx = x0 - 2 + rx, gy = y0 - 2 + ry;
        return gx >= 0 && gx < lW && gy >= 0 && gy < lH;
    });
    __syncthreads();
    // layer 2 (dilation 2) over the tile's pixels: output (oy, ox) reads
    // layer-1 region pixel (oy + 2 dy, ox + 2 dx)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lq = lane >> 4;
    float bb[2][4];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        const float4 bv = *(const float4 *)(b2 + cb * 16 + lq * 4);
        bb[cb][0] = bv.x; bb[cb][1] = bv.y; bb[cb][2] = bv.z; bb[cb][3] = bv.w;
    }
#pragma unroll 1
    for (int blk = wave; blk < NBLK; blk += 4) {
        const int p = blk * 16 + lr;
        v4f acc[2];
        conv_block<CH, kF12W, NPIX, kF12OW, 2>(w2, s_a, p, acc);
        const int gx = x0 + p % kF12W, gy = y0 + p / kF12W;
        if (p >= NPIX || gx >= lW || gy >= lH) continue;
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
            *(uint64_t *)(out + (t.base + (size_t)gy * lW + gx) * CH + cb * 16 + lq * 4) = relu_pack(acc[cb], bb[cb]);
    }
}

// Layers 3 and 4 and the head, one tile of a phase grid per workgroup: `in`
// layer-2 activations -> network-grid AO (1 where the network pixel has no
// covered frame pixel).  Layer 4's weights are loaded while layer 3 runs.
template <bool BATCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_nao_l34(
    BatchPixView one, const BatchPixView *__restrict__ views, const uint2 *__restrict__ items, int n_tiles,
    int tiles_x, int tiles_y, const uint16_t *__restrict__ in, const uint16_t *__restrict__ W3,
    const float *__restrict__ b3, const uint16_t *__restrict__ W4, const float *__restrict__ b4, HeadArgs hd)
{
    __shared__ __attribute__((aligned(16))) uint16_t s_x[kF34IW * kF34IH * CH];
    __shared__ __attribute__((aligned(16))) uint16_t s_a[kF34OW * kF34OH * CH];
    constexpr int NPIX = kF34W * kF34H, NBLK = (NPIX + 15) / 16;
    if ((int)blockIdx.x >= n_tiles) return;
    const FusedTile t = fused_tile<BATCH>(one, views, items, (int)blockIdx.x, tiles_x, tiles_y, kF34Dil);
    const int Wp = (t.lW - t.phx + kF34Dil - 1) / kF34Dil, Hp = (t.lH - t.phy + kF34Dil - 1) / kF34Dil;
    const int i0 = t.bx * kF34W, j0 = t.by * kF34H;
    ConvWeights<CH> w3;
    load_weights<CH>(w3, W3);
    stage_region<kF34IW, kF34IH, CH / 8>(s_x, in, t.base, t.lW, t.phx, t.phy, kF34Dil, i0 - 3, j0 - 3, Wp, Hp);
    ConvWeights<CH> w4;
    load_weights<CH>(w4, W4);
    __syncthreads();
    layer_to_lds<CH, kF34OW, kF34OW * kF34OH, kF34IW, 1>(w3, s_x, b3, s_a, [=](int ry, int rx) {
        const int i = i0 - 2 + rx, j = j0 - 2 + ry;
        return i >= 0 && i < Wp && j >= 0 && j < Hp;
    });
    __syncthreads();
    // layer 4 (dilation 2 on the phase grid) over the tile's pixels, then the
    // head: each lane sums its 8 channels, two exchanges across the lane
    // groups give head_z's sum
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lq = lane >> 4;
    float bb[2][4], w5[2][4];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        const float4 bv = *(const float4 *)(b4 + cb * 16 + lq * 4);
        bb[cb][0] = bv.x; bb[cb][1] = bv.y; bb[cb][2] = bv.z; bb[cb][3] = bv.w;
        const float4 wv = *(const float4 *)(hd.w5 + cb * 16 + lq * 4);
        w5[cb][0] = wv.x; w5[cb][1] = wv.y; w5[cb][2] = wv.z; w5[cb][3] = wv.w;
    }
    const float b5 = hd.w5[CH];
#pragma unroll 1
    for (int blk = wave; blk < NBLK; blk += 4) {
        const int p = blk * 16 + lr;
        v4f acc[2];
        conv_block<CH, kF34W, NPIX, kF34OW, 2>(w4, s_a, p, acc);
        float sg = 0.f;
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float v = acc[cb][q] + bb[cb][q];
                sg += w5[cb][q] * bf2f(f2bf(v > 0.f ? v : 0.f));
            }
        const float t1 = sg + __shfl_xor(sg, 16);
        const float t2 = t1 + __shfl_xor(t1, 32);
        const int i = i0 + p % kF34W, j = j0 + p / kF34W;
        if (lq == 0 && p < NPIX && i < Wp && j < Hp) {
            const size_t q = t.base + (size_t)(t.phy + j * kF34Dil) * t.lW + t.phx + (size_t)i * kF34Dil;
            const float z = b5 + t2;
            float v = 1.f / (1.f + expf(-z));
            if (!hd.mask[q]) v = 1.f;
            hd.ao[q] = v;
        }
    }
}

static inline void fused12_tiles(int lW, int lH, int &tx, int &ty)
{
    tx = (lW + kF12W - 1) / kF12W;
    ty = (lH + kF12H - 1) / kF12H;
}

// Tiles per phase grid of layers 3-4: the widest phase grid (phase 0) sets them.
static inline void fused34_tiles(int lW, int lH, int &tx, int &ty)
{
    tx = ((lW + kF34Dil - 1) / kF34Dil + kF34W - 1) / kF34W;
    ty = ((lH + kF34Dil - 1) / kF34Dil + kF34H - 1) / kF34H;
}


// Weight gradient: dW[co][tap*CIN+ci] = sum_p dout[p][co] * in[p+off(tap)][ci]
// and db[co] = sum_p dout[p][co], per workgroup over a 64 x WG_ROWS pixel
// tile, written as partials [block][...] for a fixed-order reduction.  Tap
// offsets are (dy, dx) * dil.  Two rows of pixels per stage: their output
// gradients, and the six input rows the three tap rows reach, each with a
// dil-pixel halo left and right.
template <int CIN>
__global__ __launch_bounds__(256) void k_conv_wgrad(const uint16_t *__restrict__ in,
                                                    const uint16_t *__restrict__ dout,
                                                    float *__restrict__ part_dw,
                                                    float *__restrict__ part_db, int W, int H, int dil)
{
    constexpr int K = 9 * CIN, NB = K / 16;
    constexpr int NBW = (NB + 3) / 4;  // n-blocks per wavefront
    constexpr int SR = 2;              // pixel rows per stage
    __shared__ uint16_t s_out[SR * 64 * CH];
    __shared__ uint16_t s_in[3 * SR * (64 + 2 * kMaxDil) * CIN];
    const int SW = 64 + 2 * dil;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lr = lane & 15, lq = lane >> 4;
    const int xb = blockIdx.x * 64, yb = blockIdx.y * WG_ROWS, n = blockIdx.z;
    const size_t img = (size_t)n * H * W;
    v4f acc[2][NBW];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int j = 0; j < NBW; ++j) acc[mb][j] = v4f{0.f, 0.f, 0.f, 0.f};
    float dbacc = 0.f;
    for (int r0 = 0; r0 < WG_ROWS; r0 += SR) {
        const int y0 = yb + r0;
        // output gradients: SR rows x 64 px x 32 ch
        for (int e = tid; e < SR * 64 * CH / 4; e += 256) {
            const int pix = e / (CH / 4), c4 = e % (CH / 4);
            const int yy = y0 + pix / 64, xx = xb + pix % 64;
            uint64_t v = 0;
            if (yy < H && xx < W) v = *(const uint64_t *)(dout + (img + (size_t)yy * W + xx) * CH + c4 * 4);
            *(uint64_t *)(s_out + pix * CH + c4 * 4) = v;
        }
        // inputs: stage row (dyi * SR + r) is image row y0 + r + (dyi - 1) * dil,
        // column c is image column xb - dil + c
        for (int e = tid; e < 3 * SR * SW * CIN / 4; e += 256) {
            const int pix = e / (CIN / 4), c4 = e % (CIN / 4);
            const int srow = pix / SW, col = pix % SW;
            const int yy = y0 + srow % SR + (srow / SR - 1) * dil, xx = xb - dil + col;
            uint64_t v = 0;
            if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                v = *(const uint64_t *)(in + (img + (size_t)yy * W + xx) * CIN + c4 * 4);
            *(uint64_t *)(s_in + pix * CIN + c4 * 4) = v;
        }
        __syncthreads();
        if (tid < CH)
            for (int p = 0; p < SR * 64; ++p) dbacc += bf2f(s_out[p * CH + tid]);
        // K dimension of this GEMM = SR * 64 pixels in steps of 16
        for (int ps = 0; ps < SR * 4; ++ps) {
            v4s a[2];
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    a[mb][j] = (short)s_out[(ps * 16 + lq * 4 + j) * CH + mb * 16 + lr];
#pragma unroll
            for (int jb = 0; jb < NBW; ++jb) {
                const int nb = wave + 4 * jb;
                if (nb >= NB) break;
                const int k = nb * 16 + lr, tap = k / CIN, ci = k % CIN;
                const int dyi = tap / 3, dxi = tap % 3;
                v4s b;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int pix = ps * 16 + lq * 4 + j;
                    const int pr = pix / 64, pc = pix % 64;
                    b[j] = (short)s_in[((dyi * SR + pr) * SW + pc + dxi * dil) * CIN + ci];
                }
#pragma unroll
                for (int mb = 0; mb < 2; ++mb)
                    acc[mb][jb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[mb], b, acc[mb][jb], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const size_t blk = ((size_t)n * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    float *pd = part_dw + blk * CH * K;
#pragma unroll
    for (int jb = 0; jb < NBW; ++jb) {
        const int nb = wave + 4 * jb;
        if (nb >= NB) break;
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int i = 0; i < 4; ++i) pd[(size_t)(mb * 16 + lq * 4 + i) * K + nb * 16 + lr] = acc[mb][jb][i];
    }
    if (tid < CH) part_db[blk * CH + tid] = dbacc;
}

// out[i] = sum over b of part[b*len + i], b ascending (fixed order).
__global__ __launch_bounds__(256) void k_reduce_cols(const float *__restrict__ part, int nb, int len,
                                                     float *__restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= len) return;
    float s = 0.f;
    for (int b = 0; b < nb; ++b) s += part[(size_t)b * len + i];
    out[i] = s;
}

// ---------------- head backward (training) ----------------

// Loss = sum over covered pixels of (y - t)^2 / n_total.  Writes the data
// gradient of a4 and per-block partials of [dW5 (32), db5, loss, |err|, count].
constexpr int HEAD_PART = CH + 4;
__global__ __launch_bounds__(256) void k_head_bwd(const float *__restrict__ params,
                                                  const uint16_t *__restrict__ a4,
                                                  const float *__restrict__ target,
                                                  const uint8_t *__restrict__ mask, size_t n,
                                                  uint16_t *__restrict__ da4,
                                                  float *__restrict__ part)
{
    __shared__ float red[256][HEAD_PART + 1];
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    float loc[HEAD_PART];
    for (int k = 0; k < HEAD_PART; ++k) loc[k] = 0.f;
    if (p < n) {
        const uint16_t *a = a4 + p * CH;
        const float *w5 = params + OFF_W5;
        const float z = head_z(w5, params[OFF_B5], a);
        const float y = 1.f / (1.f + expf(-z));
        const bool m = mask[p] != 0;
        const float diff = y - target[p];
        const float g = m ? 2.f * diff / (float)n : 0.f;
        const float dz = g * y * (1.f - y);
        for (int c = 0; c < CH; ++c) {
            const float av = bf2f(a[c]);
            da4[p * CH + c] = f2bf(av > 0.f ? dz * w5[c] : 0.f);
            loc[c] = dz * av;
        }
        loc[CH] = dz;
        loc[CH + 1] = m ? diff * diff : 0.f;
        loc[CH + 2] = m ? fabsf(diff) : 0.f;
        loc[CH + 3] = m ? 1.f : 0.f;
    }
    for (int k = 0; k < HEAD_PART; ++k) red[threadIdx.x][k] = loc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < HEAD_PART; ++k) red[threadIdx.x][k] += red[threadIdx.x + s][k];
        __syncthreads();
    }
    if (threadIdx.x < HEAD_PART) part[(size_t)blockIdx.x * HEAD_PART + threadIdx.x] = red[0][threadIdx.x];
}

// ---------------- weights, Adam ----------------

// FP32 master -> BF16 forward layouts [co][tap*Cin+ci] and the data-gradient
// layouts [ci][tap'*32+co] with tap' = 8 - tap (layers 2..4).
__global__ __launch_bounds__(256) void k_prep_weights(const float *__restrict__ params,
                                                      uint16_t *__restrict__ w1,
                                                      uint16_t *__restrict__ w234,
                                                      uint16_t *__restrict__ wt234)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    // layer 1: only the 8 real input channels; the padding channels' weights
    // are held at zero (their inputs are zero)
    if (i < CH * K1) w1[i] = (i % K1) % C0 < NIN ? f2bf(params[OFF_W1 + i]) : (uint16_t)0;
    if (i < 3 * CH * K2) {
        const int l = i / (CH * K2), r = i % (CH * K2);
        const int offs[3] = {OFF_W2, OFF_W3, OFF_W4};
        const float v = params[offs[l] + r];
        w234[i] = f2bf(v);
        const int co = r / K2, k = r % K2, tap = k / CH, ci = k % CH;
        wt234[(size_t)l * CH * K2 + (size_t)ci * K2 + (8 - tap) * CH + co] = f2bf(v);
    }
}

struct AdamArgs {
    float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
};

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, int n,
                                              AdamArgs a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = a.b1 * m[i] + (1.f - a.b1) * gi;
    const float vi = a.b2 * v[i] + (1.f - a.b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / a.c1, vh = vi / a.c2;
    w[i] -= a.lr * mh / (sqrtf(vh) + a.eps);
}

}  // namespace nao
}  // namespace fgl

// ======================= host object and C ABI =======================

using namespace fgl;
using namespace fgl::nao;

struct fgl_nao {
    fgl_ctx *ctx = nullptr;
    int batch = 0, crop = 0;
    int scale = 4;   // frame pixels per network pixel, each way
    int step = 0;
    // parameters (FP32 master), gradients, Adam moments
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *w1 = nullptr, *w234 = nullptr, *wt234 = nullptr;
    // training batch
    uint16_t *feat = nullptr, *a[4] = {nullptr, nullptr, nullptr, nullptr};
    uint16_t *d[4] = {nullptr, nullptr, nullptr, nullptr};
    float *target = nullptr, *y = nullptr;
    uint8_t *mask = nullptr;
    float *part = nullptr, *part_db = nullptr, *head_part = nullptr, *head_red = nullptr;
    size_t part_cap = 0;
    // inference buffers (grow-only): network-grid inputs, activations,
    // coverage and AO; frame-pixel coverage
    uint16_t *ifeat = nullptr, *ia = nullptr, *ib = nullptr;
    uint8_t *imask = nullptr, *fmask = nullptr;
    float *lao = nullptr;
    size_t icap = 0, fcap = 0;
    uint64_t igen = 0;   // bumped whenever the inference buffers move
    int fused = 1;       // single-frame inference: fused kernels (1) or one kernel per layer (0)
    float *h_stats = nullptr;
};

static int nao_alloc(void **p, size_t bytes)
{
    return hipMalloc(p, bytes < 256 ? 256 : bytes) == hipSuccess ? FGL_OK : FGL_ERR_NOMEM;
}

static inline dim3 g2(int W, int H) { return dim3((W + 63) / 64, (H + 3) / 4); }

namespace fgl {

int nao_reserve(fgl_nao *n, size_t frame_px, size_t net_px)
{
    if (net_px <= n->icap && frame_px <= n->fcap) return FGL_OK;
    int rc = ctx_settle_network_ao(n->ctx, n);   // its buffers are about to move
    if (rc) return rc;
    if (hipStreamSynchronize(ctx_stream(n->ctx)) != hipSuccess) return FGL_ERR_DEVICE;
    ++n->igen;
    if (net_px > n->icap) {
        void *ps[] = {n->ifeat, n->ia, n->ib, n->imask, n->lao};
        for (void *q : ps)
            if (q) (void)hipFree(q);
        n->ifeat = n->ia = n->ib = nullptr;
        n->imask = nullptr;
        n->lao = nullptr;
        n->icap = 0;
        // network inputs, activations of layers 1-3, coverage, AO
        if (nao_alloc((void **)&n->ifeat, net_px * C0 * 2) || nao_alloc((void **)&n->ia, net_px * CH * 2) ||
            nao_alloc((void **)&n->ib, net_px * CH * 2) || nao_alloc((void **)&n->imask, net_px) ||
            nao_alloc((void **)&n->lao, net_px * 4))
            return FGL_ERR_NOMEM;
        // input channels 8-15 stay zero: the down pass writes channels 0-7 only
        if (hipMemset(n->ifeat, 0, net_px * C0 * 2) != hipSuccess) return FGL_ERR_DEVICE;
        n->icap = net_px;
    }
    if (frame_px > n->fcap) {
        if (n->fmask) (void)hipFree(n->fmask);
        n->fmask = nullptr;
        n->fcap = 0;
        if (nao_alloc((void **)&n->fmask, frame_px)) return FGL_ERR_NOMEM;
        n->fcap = frame_px;
    }
    return FGL_OK;
}

int nao_scale(const fgl_nao *n) { return n->scale; }

uint64_t nao_buffers_generation(const fgl_nao *n) { return n->igen; }

void nao_conv_items(int view, int W, int H, std::vector<uint2> items[2])
{
    // items[0]: tiles of the fused layers 1-2 (network grid), items[1]: of
    // layers 3-4 (phase grids), in the single-frame launches' tile order
    for (int l = 0; l < 2; ++l) {
        const int d = l == 0 ? 1 : kF34Dil;
        int tx, ty;
        if (l == 0) fused12_tiles(W, H, tx, ty);
        else fused34_tiles(W, H, tx, ty);
        const int n = tx * ty * d * d;
        for (int u = 0; u < n; ++u) {
            int ph, k;
            tile_order(u, tx * ty, d, ph, k);
            const int bx = k % tx, by = k / tx;
            items[l].push_back(make_uint2((uint32_t)view | ((uint32_t)ph << 8), (uint32_t)bx | ((uint32_t)by << 16)));
        }
    }
}

void nao_launch_up(const NetAo &net, int W, int H, float *ao, hipStream_t s)
{
    if (W <= 0 || H <= 0) return;
    BatchPixView one{};
    one.W = W;
    one.H = H;
    one.lW = net.lW;
    one.lH = net.lH;
    const int nb = ((W + 63) / 64) * ((H + 3) / 4);
    hipLaunchKernelGGL(k_nao_up<false>, dim3((nb + 3) / 4), dim3(256), 0, s, one, (const BatchPixView *)nullptr,
                       (const uint8_t *)nullptr, nb, net.f, net.lao, net.fmask, ao);
}

int nao_infer_batch(fgl_nao *n, hipStream_t s, const NaoBatchTables &t, const float *depth, const float *gbuf,
                    float *ao)
{
    if (t.total_px > n->fcap || t.total_lpx > n->icap) return FGL_ERR_INVALID;
    // the batch overwrites the network's buffers, which the context's AO may still be
    if (int rc = ctx_settle_network_ao(n->ctx, n)) return rc;
    if (t.n_blocks == 0) return FGL_OK;
    const int f = n->scale;
    const BatchPixView none{};
    hipLaunchKernelGGL(k_nao_down<true>, dim3(t.n_lblocks), dim3(256), 0, s, none, t.views, t.lblk_view, f, depth,
                       gbuf, n->ifeat, n->imask, n->fmask);
    if (t.n_items[0])
        hipLaunchKernelGGL(k_nao_l12<true>, dim3(t.n_items[0]), dim3(256), 0, s, none,
                           t.views, t.items[0], (int)t.n_items[0], 0, 0, (const uint16_t *)n->ifeat, n->w1,
                           n->params + OFF_B1, n->w234, n->params + OFF_B2, n->ia);
    // layers 3-4 with the head, into the network-grid AO
    if (t.n_items[1])
        hipLaunchKernelGGL(k_nao_l34<true>, dim3(t.n_items[1]), dim3(256), 0, s, none,
                           t.views, t.items[1], (int)t.n_items[1], 0, 0, (const uint16_t *)n->ia,
                           n->w234 + (size_t)CH * K2, n->params + OFF_B3, n->w234 + (size_t)2 * CH * K2,
                           n->params + OFF_B4, HeadArgs{n->params + OFF_W5, n->imask, n->lao});
    hipLaunchKernelGGL(k_nao_up<true>, dim3((t.n_blocks + 3) / 4), dim3(256), 0, s, none, t.views, t.blk_view,
                       (int)t.n_blocks, f, n->lao, n->fmask, ao);
    return hipGetLastError() == hipSuccess ? FGL_OK : FGL_ERR_DEVICE;
}

void nao_detach(void *p) { static_cast<fgl_nao *>(p)->ctx = nullptr; }

}  // namespace fgl

extern "C" {

int fgl_nao_param_count(void) { return NPARAM; }

int fgl_nao_create(fgl_ctx *ctx, int batch, int crop, int scale, fgl_nao **out)
{
    if (!ctx || !out || batch < 1 || crop < 16 || crop % 64 != 0 || crop % WG_ROWS != 0 ||
        (scale != 1 && scale != 2 && scale != 4))
        return FGL_ERR_INVALID;
    fgl_nao *n = new fgl_nao();
    n->ctx = ctx;
    ctx_register_nao(ctx, n);
    n->batch = batch;
    n->crop = crop;
    n->scale = scale;
    const size_t px = (size_t)batch * crop * crop;
    int rc = FGL_OK;
    rc |= nao_alloc((void **)&n->params, NPARAM * 4);
    rc |= nao_alloc((void **)&n->grads, NPARAM * 4);
    rc |= nao_alloc((void **)&n->m, NPARAM * 4);
    rc |= nao_alloc((void **)&n->v, NPARAM * 4);
    rc |= nao_alloc((void **)&n->w1, CH * K1 * 2);
    rc |= nao_alloc((void **)&n->w234, 3 * CH * K2 * 2);
    rc |= nao_alloc((void **)&n->wt234, 3 * CH * K2 * 2);
    rc |= nao_alloc((void **)&n->feat, px * C0 * 2);
    for (int l = 0; l < 4; ++l) {
        rc |= nao_alloc((void **)&n->a[l], px * CH * 2);
        rc |= nao_alloc((void **)&n->d[l], px * CH * 2);
    }
    rc |= nao_alloc((void **)&n->target, px * 4);
    rc |= nao_alloc((void **)&n->y, px * 4);
    rc |= nao_alloc((void **)&n->mask, px);
    const size_t nblk = (size_t)(crop / 64) * (crop / WG_ROWS) * batch;
    n->part_cap = nblk;
    rc |= nao_alloc((void **)&n->part, nblk * CH * K2 * 4);
    rc |= nao_alloc((void **)&n->part_db, nblk * CH * 4);
    const size_t hb = (px + 255) / 256;
    rc |= nao_alloc((void **)&n->head_part, hb * HEAD_PART * 4);
    rc |= nao_alloc((void **)&n->head_red, HEAD_PART * 4);
    if (hipHostMalloc((void **)&n->h_stats, 64 * 4, hipHostMallocDefault) != hipSuccess) rc = FGL_ERR_NOMEM;
    if (rc != FGL_OK) {
        fgl_nao_destroy(n);   // frees what was allocated, leaves the context
        return FGL_ERR_NOMEM;
    }
    if (hipMemset(n->m, 0, NPARAM * 4) != hipSuccess || hipMemset(n->v, 0, NPARAM * 4) != hipSuccess ||
        hipMemset(n->params, 0, NPARAM * 4) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        fgl_nao_destroy(n);
        return FGL_ERR_DEVICE;
    }
    *out = n;
    return FGL_OK;
}

void fgl_nao_destroy(fgl_nao *n)
{
    if (!n) return;
    if (n->ctx) {
        // the context's AO may still be this network's grid (teardown: a
        // failure here has no caller to go to)
        (void)ctx_settle_network_ao(n->ctx, n);
        ctx_unregister_nao(n->ctx, n);
    }
    (void)hipDeviceSynchronize();
    void *ps[] = {n->params, n->grads, n->m, n->v, n->w1, n->w234, n->wt234, n->feat, n->target, n->y,
                  n->mask, n->part, n->part_db, n->head_part, n->head_red, n->ifeat, n->ia, n->ib,
                  n->imask, n->fmask, n->lao};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    for (int l = 0; l < 4; ++l) {
        if (n->a[l]) (void)hipFree(n->a[l]);
        if (n->d[l]) (void)hipFree(n->d[l]);
    }
    if (n->h_stats) (void)hipHostFree(n->h_stats);
    delete n;
}

static int prep(fgl_nao *n, hipStream_t s)
{
    const int total = 3 * CH * K2;
    hipLaunchKernelGGL(k_prep_weights, dim3((total + 255) / 256), dim3(256), 0, s, n->params, n->w1,
                       n->w234, n->wt234);
    return hipGetLastError() == hipSuccess ? FGL_OK : FGL_ERR_DEVICE;
}

// Parameters: FP32 master weights in the flat layout of fgl_nao_param_count();
// state = [params | adam m | adam v] (3 * count floats) + step.
int fgl_nao_set_state(fgl_nao *n, const float *state, int with_adam, int step)
{
    if (!n || !state) return FGL_ERR_INVALID;
    if (hipMemcpy(n->params, state, NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess) return FGL_ERR_DEVICE;
    hipError_t e;
    if (with_adam) {
        e = hipMemcpy(n->m, state + NPARAM, NPARAM * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(n->v, state + 2 * NPARAM, NPARAM * 4, hipMemcpyHostToDevice);
        n->step = step;
    } else {
        e = hipMemset(n->m, 0, NPARAM * 4);
        if (e == hipSuccess) e = hipMemset(n->v, 0, NPARAM * 4);
        n->step = 0;
    }
    if (e != hipSuccess || hipDeviceSynchronize() != hipSuccess) return FGL_ERR_DEVICE;
    return prep(n, 0) == FGL_OK && hipDeviceSynchronize() == hipSuccess ? FGL_OK : FGL_ERR_DEVICE;
}

int fgl_nao_get_state(fgl_nao *n, float *state, int *step)
{
    if (!n || !n->ctx || !state) return FGL_ERR_INVALID;
    if (hipStreamSynchronize(ctx_stream(n->ctx)) != hipSuccess ||
        hipMemcpy(state, n->params, NPARAM * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(state + NPARAM, n->m, NPARAM * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(state + 2 * NPARAM, n->v, NPARAM * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return FGL_ERR_DEVICE;
    if (step) *step = n->step;
    return FGL_OK;
}

// Copies crops of the context's current frame (depth, G-buffer, AO as the
// target) on the network grid into batch slots slot0 .. slot0+count-1; xy
// holds crop origins in network pixels.
int fgl_nao_stage_crops(fgl_nao *n, int slot0, int count, const int *xy)
{
    if (!n || !n->ctx || !xy || slot0 < 0 || count < 0 || slot0 + count > n->batch) return FGL_ERR_INVALID;
    CtxView cv;
    int rc = ctx_view(n->ctx, &cv);
    if (rc) return rc;
    Mat16 V;
    std::memcpy(V.m, cv.view, sizeof(V.m));
    const int f = n->scale, lW = (cv.W + f - 1) / f, lH = (cv.H + f - 1) / f;
    for (int i = 0; i < count; ++i) {
        const int x0 = xy[2 * i], y0 = xy[2 * i + 1];
        if (x0 < 0 || y0 < 0 || x0 + n->crop > lW || y0 + n->crop > lH)
            return ctx_fail(n->ctx, FGL_ERR_INVALID, "crop outside the frame's network grid");
        hipLaunchKernelGGL(k_crop, g2(n->crop, n->crop), dim3(256), 0, cv.stream, V, cv.depth, cv.gbuf, cv.ao,
                           cv.W, cv.H, f, lW, lH, x0, y0, n->crop, slot0 + i, n->feat, n->target, n->mask);
    }
    return hipGetLastError() == hipSuccess ? FGL_OK : FGL_ERR_DEVICE;
}

// Forward of conv layers 1..4 on images [N][H][W] in place of the activations.
static void forward_convs(fgl_nao *n, hipStream_t s, const uint16_t *feat, uint16_t **acts, int W, int H,
                          int N)
{
    launch_conv<C0, 0>(s, feat, n->w1, n->params + OFF_B1, nullptr, acts[0], W, H, N, layer_dil(0));
    const int boffs[3] = {OFF_B2, OFF_B3, OFF_B4};
    for (int l = 0; l < 3; ++l)
        launch_conv<CH, 0>(s, acts[l], n->w234 + (size_t)l * CH * K2, n->params + boffs[l], nullptr, acts[l + 1], W, H,
                           N, layer_dil(l + 1));
}

static int run_batch(fgl_nao *n, bool update, float lr, float *stats)
{
    CtxView cv;
    int rc = ctx_view(n->ctx, &cv);
    if (rc) return rc;
    hipStream_t s = cv.stream;
    const int S = n->crop, N = n->batch;
    const size_t px = (size_t)N * S * S;
    forward_convs(n, s, n->feat, n->a, S, S, N);
    // head forward + loss + data gradient of a4
    const int hb = (int)((px + 255) / 256);
    hipLaunchKernelGGL(k_head_bwd, dim3(hb), dim3(256), 0, s, n->params, n->a[3], n->target, n->mask, px,
                       n->d[3], n->head_part);
    hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, s, n->head_part, hb, HEAD_PART, n->head_red);
    if (hipMemcpyAsync(n->grads + OFF_W5, n->head_red, CH * 4, hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(n->grads + OFF_B5, n->head_red + CH, 4, hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(n->h_stats, n->head_red + CH + 1, 3 * 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        return FGL_ERR_DEVICE;
    // layers 4..1: weight gradient from (input, output gradient), then the
    // data gradient of the layer input (not needed for layer 1)
    const dim3 wg(S / 64, S / WG_ROWS, N);
    const int nblk = (int)(wg.x * wg.y * wg.z);
    const int woffs[4] = {OFF_W1, OFF_W2, OFF_W3, OFF_W4}, boffs[4] = {OFF_B1, OFF_B2, OFF_B3, OFF_B4};
    for (int l = 3; l >= 0; --l) {
        const int K = l == 0 ? K1 : K2;
        if (l == 0)
            hipLaunchKernelGGL((k_conv_wgrad<C0>), wg, dim3(256), 0, s, n->feat, n->d[0], n->part, n->part_db, S, S,
                               layer_dil(0));
        else
            hipLaunchKernelGGL((k_conv_wgrad<CH>), wg, dim3(256), 0, s, n->a[l - 1], n->d[l], n->part, n->part_db, S, S,
                               layer_dil(l));
        hipLaunchKernelGGL(k_reduce_cols, dim3((CH * K + 255) / 256), dim3(256), 0, s, n->part, nblk, CH * K,
                           n->grads + woffs[l]);
        hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, s, n->part_db, nblk, CH, n->grads + boffs[l]);
        if (l > 0)
            launch_conv<CH, 1>(s, n->d[l], n->wt234 + (size_t)(l - 1) * CH * K2, nullptr, n->a[l - 1], n->d[l - 1], S,
                               S, N, layer_dil(l));
    }
    if (update) {
        n->step += 1;
        AdamArgs a;
        a.lr = lr;
        a.b1 = 0.9f;
        a.b2 = 0.999f;
        a.eps = 1e-8f;
        a.c1 = (float)(1.0 - std::pow(0.9, n->step));
        a.c2 = (float)(1.0 - std::pow(0.999, n->step));
        hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, s, n->params, n->grads, n->m, n->v,
                           NPARAM, a);
        if ((rc = prep(n, s))) return rc;
    }
    if (hipStreamSynchronize(s) != hipSuccess) return FGL_ERR_DEVICE;
    if (stats) {
        stats[0] = n->h_stats[0];  // sum of squared errors over covered pixels
        stats[1] = n->h_stats[1];  // sum of |errors|
        stats[2] = n->h_stats[2];  // covered pixel count
    }
    return FGL_OK;
}

// One Adam step on the staged batch.  stats: [sum sq err, sum abs err, covered].
int fgl_nao_train_step(fgl_nao *n, float lr, float *stats)
{
    if (!n || !n->ctx) return FGL_ERR_INVALID;
    return run_batch(n, true, lr, stats);
}

// Gradients of the staged batch without updating (for gradient checks).
int fgl_nao_gradients(fgl_nao *n, float *grads, float *stats)
{
    if (!n || !n->ctx || !grads) return FGL_ERR_INVALID;
    int rc = run_batch(n, false, 0.f, stats);
    if (rc) return rc;
    return hipMemcpy(grads, n->grads, NPARAM * 4, hipMemcpyDeviceToHost) == hipSuccess ? FGL_OK : FGL_ERR_DEVICE;
}

// Copies the staged batch to the host: features (N*S*S*16 BF16), target, mask.
int fgl_nao_read_batch(fgl_nao *n, uint16_t *feat, float *target, uint8_t *mask)
{
    if (!n) return FGL_ERR_INVALID;
    if (hipDeviceSynchronize() != hipSuccess) return FGL_ERR_DEVICE;
    const size_t px = (size_t)n->batch * n->crop * n->crop;
    if ((feat && hipMemcpy(feat, n->feat, px * C0 * 2, hipMemcpyDeviceToHost) != hipSuccess) ||
        (target && hipMemcpy(target, n->target, px * 4, hipMemcpyDeviceToHost) != hipSuccess) ||
        (mask && hipMemcpy(mask, n->mask, px, hipMemcpyDeviceToHost) != hipSuccess))
        return FGL_ERR_DEVICE;
    return FGL_OK;
}

// Inference kernels: fused (two kernels for the four convolutions and the
// head) or one kernel per layer.
int fgl_nao_set_fused(fgl_nao *n, int fused)
{
    if (!n || (fused != 0 && fused != 1)) return FGL_ERR_INVALID;
    n->fused = fused;
    return FGL_OK;
}

// Neural AO of the context's current frame (depth + G-buffer): down pass,
// the convolutions and the head; on a coarser grid the AO stays there as the
// context's AO target (the lighting pass interpolates it, a read of the
// target writes it out).
int fgl_nao_infer(fgl_nao *n)
{
    if (!n || !n->ctx) return FGL_ERR_INVALID;
    ctx_drop_network_ao(n->ctx, n);   // replaced below: no need to write it out
    CtxView cv;
    // a coarser grid leaves the AO target alone (it becomes the network's);
    // the grid of the frame itself writes it
    int rc = ctx_view(n->ctx, &cv, n->scale == 1);
    if (rc) return rc;
    const int f = n->scale, lW = (cv.W + f - 1) / f, lH = (cv.H + f - 1) / f;
    if ((rc = nao_reserve(n, (size_t)cv.W * cv.H, (size_t)lW * lH))) return rc;
    BatchPixView one{};
    one.W = cv.W;
    one.H = cv.H;
    one.lW = lW;
    one.lH = lH;
    std::memcpy(one.view, cv.view, sizeof(one.view));
    {
        CtxTimer t(n->ctx, "nao_conv");
        hipLaunchKernelGGL(k_nao_down<false>, dim3(((lW + 63) / 64) * ((lH + 3) / 4)), dim3(256), 0, cv.stream, one,
                           (const BatchPixView *)nullptr, (const uint8_t *)nullptr, f, cv.depth, cv.gbuf, n->ifeat,
                           n->imask, n->fmask);
        // the network-grid AO (the frame's AO itself when the grid is the frame)
        float *dst = f == 1 ? cv.ao : n->lao;
        const HeadArgs hd{n->params + OFF_W5, n->imask, dst};
        if (n->fused) {
            int tx, ty;
            fused12_tiles(lW, lH, tx, ty);
            int nt = tx * ty;
            hipLaunchKernelGGL(k_nao_l12<false>, dim3(nt), dim3(256), 0, cv.stream, one,
                               (const BatchPixView *)nullptr, (const uint2 *)nullptr, nt, tx, ty,
                               (const uint16_t *)n->ifeat, n->w1, n->params + OFF_B1, n->w234, n->params + OFF_B2,
                               n->ia);
            fused34_tiles(lW, lH, tx, ty);
            nt = tx * ty * kF34Dil * kF34Dil;
            hipLaunchKernelGGL(k_nao_l34<false>, dim3(nt), dim3(256), 0, cv.stream, one,
                               (const BatchPixView *)nullptr, (const uint2 *)nullptr, nt, tx, ty,
                               (const uint16_t *)n->ia, n->w234 + (size_t)CH * K2, n->params + OFF_B3,
                               n->w234 + (size_t)2 * CH * K2, n->params + OFF_B4, hd);
        } else {
            launch_conv<C0, 0>(cv.stream, n->ifeat, n->w1, n->params + OFF_B1, nullptr, n->ia, lW, lH, 1,
                               layer_dil(0));
            launch_conv<CH, 0>(cv.stream, n->ia, n->w234, n->params + OFF_B2, nullptr, n->ib, lW, lH, 1,
                               layer_dil(1));
            launch_conv<CH, 0>(cv.stream, n->ib, n->w234 + (size_t)CH * K2, n->params + OFF_B3, nullptr, n->ia,
                               lW, lH, 1, layer_dil(2));
            // layer 4 with the head fused
            launch_conv_head(cv.stream, n->ia, n->w234 + (size_t)2 * CH * K2, n->params + OFF_B4, hd, lW, lH, 1,
                             layer_dil(3));
        }
    }
    // a coarser grid stays there: the lighting pass interpolates it, and the
    // up pass writes the AO target only when the target itself is read
    if (f > 1) ctx_ao_from_network(n->ctx, NetAo{n->lao, n->fmask, lW, lH, f}, n);
    else ctx_ao_written(n->ctx);
    return hipGetLastError() == hipSuccess ? FGL_OK : FGL_ERR_DEVICE;
}

}  // extern "C"
