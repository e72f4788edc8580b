// sort.hip -- device-wide exclusive scan and stable LSD radix sort for gfx950
// (64-lane wavefronts).
//
// Binning sorts (tile, triangle) pairs by tile.  The rasteriser's depth
// resolve does not depend on the order inside a tile's list, but the sort is
// stable anyway (pairs come out in submission order per tile), so the lists
// are reproducible.  Each radix pass ranks the items of a 2048-item block
// with wave ballots (eight ballots give the mask of lanes holding the same
// 8-bit digit; popcount of the lower lanes is the rank inside the wave) and
// LDS per-wave digit counts for the order across waves and rounds -- no
// atomics decide any position (device-scope atomics on shared counters are
// several times slower here than the three sort passes).
#include "fgl_kernels.hpp"

namespace fgl {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;  // 2048

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x, int lane)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *lds4,
                                                         uint32_t &total)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t inc = wave_inclusive_scan(v, lane);
    if (lane == 63) lds4[wave] = inc;
    __syncthreads();
    uint32_t prefix = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / 64; ++w) {
        uint32_t c = lds4[w];
        if (w < wave) prefix += c;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return prefix + inc - v;
}

// A thread's kScanItems consecutive values: two 16-byte loads when all are
// in range and the array is 16-byte aligned, else one load per value.
__device__ __forceinline__ void scan_load8(const uint32_t *in, size_t n, size_t base, bool vec, uint32_t *v)
{
    if (vec && base + kScanItems <= n) {
        const uint4 a = *(const uint4 *)(in + base), b = *(const uint4 *)(in + base + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) v[i] = (base + i < n) ? in[base + i] : 0u;
    }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t *__restrict__ in,
                                                            size_t n, uint32_t *__restrict__ part)
{
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    scan_load8(in, n, base, ((uintptr_t)in & 15u) == 0, v);
    // the block's sum in 64 bits; a sum that does not fit is stored as
    // 0xffffffff, which makes the total saturate in k_scan_partials
    unsigned long long s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) s += v[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ unsigned long long lds8[4];
    if ((threadIdx.x & 63) == 0) lds8[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = lds8[0] + lds8[1] + lds8[2] + lds8[3];
        part[blockIdx.x] = t >= 0xffffffffull ? 0xffffffffu : (uint32_t)t;
    }
}

// Single workgroup: exclusive scan of the block partials, total to out[n].
// A total of 2^32 - 1 or more is written as 0xffffffff (the offsets have
// wrapped; callers treat it as an overflow).
__global__ __launch_bounds__(kScanBlock) void k_scan_partials(uint32_t *__restrict__ part,
                                                              int nparts,
                                                              uint32_t *__restrict__ out_total)
{
    __shared__ uint32_t lds4[4];
    __shared__ unsigned long long lds8[4];
    uint32_t carry = 0;
    unsigned long long carry64 = 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int base = 0; base < nparts; base += kScanBlock) {
        int i = base + threadIdx.x;
        uint32_t v = i < nparts ? part[i] : 0u;
        uint32_t total;
        uint32_t ex = block_exclusive_scan(v, lds4, total);
        if (i < nparts) part[i] = carry + ex;
        carry += total;
        // the chunk's sum again in 64 bits (the 32-bit one may wrap)
        unsigned long long w = v;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o, 64);
        if (lane == 0) lds8[wave] = w;
        __syncthreads();
        carry64 += lds8[0] + lds8[1] + lds8[2] + lds8[3];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out_total = carry64 >= 0xffffffffull ? 0xffffffffu : carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_down(const uint32_t *in, uint32_t *out,
                                                          size_t n,
                                                          const uint32_t *__restrict__ part)
{
    __shared__ uint32_t lds4[4];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    const bool vec = (((uintptr_t)in | (uintptr_t)out) & 15u) == 0;
    uint32_t v[kScanItems];
    scan_load8(in, n, base, vec, v);
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) s += v[i];
    uint32_t total;
    uint32_t ex = block_exclusive_scan(s, lds4, total) + part[blockIdx.x];
    uint32_t o[kScanItems];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        o[i] = ex;
        ex += v[i];
    }
    if (vec && base + kScanItems <= n) {
        *(uint4 *)(out + base) = make_uint4(o[0], o[1], o[2], o[3]);
        *(uint4 *)(out + base + 4) = make_uint4(o[4], o[5], o[6], o[7]);
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i)
            if (base + i < n) out[base + i] = o[i];
    }
}

size_t scan_tmp_words(size_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void launch_exclusive_scan(const uint32_t *in, uint32_t *out, size_t n, uint32_t *tmp,
                           hipStream_t s)
{
    int nb = (int)((n + kScanTile - 1) / kScanTile);
    if (n == 0) {
        hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanBlock), 0, s, tmp, 0, out + n);
        return;
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, s, in, n, tmp);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanBlock), 0, s, tmp, nb, out + n);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kScanBlock), 0, s, in, out, n, tmp);
}

// ---------------- radix sort ----------------

constexpr int kRadixBlock = 256;
constexpr int kRadixRounds = 8;
constexpr int kRadixTile = kRadixBlock * kRadixRounds;  // 2048 items, striped by round

// n is the item count, or the capacity when n_dev holds the device-side
// count (then min(n, *n_dev) items are sorted).
__global__ __launch_bounds__(kRadixBlock) void k_radix_hist(const uint32_t *__restrict__ keys,
                                                            uint32_t n, const uint32_t *__restrict__ n_dev,
                                                            int shift, int nb,
                                                            uint32_t *__restrict__ hist)
{
    if (n_dev) n = min(n, *n_dev);
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kRadixTile;
#pragma unroll
    for (int r = 0; r < kRadixRounds; ++r) {
        uint32_t idx = base + r * kRadixBlock + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// Scatter of one 2048-item block: ranks by wave ballots as above, but the
// block first reorders its items by digit in LDS and then writes each
// digit's run with consecutive lanes on consecutive addresses.  Writing
// straight from the ranks leaves runs of a few items per digit and round
// (28-byte pieces of 128-byte lines when a block's items spread over a few
// dozen digits), which costs 3-4x the time of whole-line writes.  The
// block's count of a digit is the difference of two scanned histogram
// entries (hist[256 * nb] holds the total).
__global__ __launch_bounds__(kRadixBlock) void k_radix_scatter(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, uint32_t n,
    const uint32_t *__restrict__ n_dev, int shift, int nb, const uint32_t *__restrict__ hist)
{
    if (n_dev) n = min(n, *n_dev);
    __shared__ uint32_t goff[256];
    __shared__ uint32_t lstart[256];
    __shared__ uint32_t dbase[256];
    __shared__ uint32_t wcnt[kRadixBlock / 64][256];
    __shared__ uint32_t lds4[4];
    __shared__ uint32_t sk[kRadixTile], sv[kRadixTile];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t hi = (size_t)tid * nb + blockIdx.x;
    const uint32_t g0 = hist[hi];
    const uint32_t g1 = (blockIdx.x + 1 < (unsigned)nb) ? hist[hi + 1] : hist[(size_t)(tid + 1) * nb];
    goff[tid] = g0;
    uint32_t total;
    lstart[tid] = block_exclusive_scan(g1 - g0, lds4, total);
    dbase[tid] = 0;
#pragma unroll
    for (int w = 0; w < kRadixBlock / 64; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t base = blockIdx.x * kRadixTile;
    for (int r = 0; r < kRadixRounds; ++r) {
        const uint32_t idx = base + r * kRadixBlock + tid;
        const bool valid = idx < n;
        const uint32_t k = valid ? kin[idx] : 0u;
        const uint32_t v = valid ? vin[idx] : 0u;
        const uint32_t d = (k >> shift) & 255u;
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool on = (d >> bit) & 1u;
            const unsigned long long bm = __ballot(on);
            m &= on ? bm : ~bm;
        }
        const uint32_t rank = (uint32_t)__popcll(m & lt);
        if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pos = lstart[d] + dbase[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
            sk[pos] = k;
            sv[pos] = v;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kRadixBlock / 64; ++w) {
            add += wcnt[w][tid];
            wcnt[w][tid] = 0;
        }
        dbase[tid] += add;
        __syncthreads();
    }
    const uint32_t nvalid = n > base ? min((uint32_t)kRadixTile, n - base) : 0u;
    for (uint32_t i = tid; i < nvalid; i += kRadixBlock) {
        const uint32_t k = sk[i];
        const uint32_t d = (k >> shift) & 255u;
        const uint32_t pos = goff[d] + (i - lstart[d]);
        kout[pos] = k;
        vout[pos] = sv[i];
    }
}

size_t radix_tmp_words(size_t n)
{
    size_t nb = (n + kRadixTile - 1) / kRadixTile;
    size_t h = 256 * nb;
    return h + 1 + scan_tmp_words(h);
}

// Returns 0 when the sorted result is in (keys, vals), 1 when in the alt
// buffers.  With n_dev, n is the capacity and *n_dev the item count.
int launch_radix_sort_impl(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                           size_t n, const uint32_t *n_dev, int bits, uint32_t *tmp, hipStream_t s)
{
    if (n <= 1 || bits <= 0) return 0;
    int nb = (int)((n + kRadixTile - 1) / kRadixTile);
    size_t h = 256 * (size_t)nb;
    uint32_t *hist = tmp;
    uint32_t *stmp = tmp + h + 1;
    uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
    int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kRadixBlock), 0, s, ki, (uint32_t)n, n_dev,
                           shift, nb, hist);
        launch_exclusive_scan(hist, hist, h, stmp, s);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kRadixBlock), 0, s, ki, vi, ko, vo,
                           (uint32_t)n, n_dev, shift, nb, hist);
        uint32_t *t;
        t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
    return (passes & 1) ? 1 : 0;
}

void launch_radix_sort(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                       size_t n, int bits, uint32_t *tmp, hipStream_t s)
{
    if (launch_radix_sort_impl(keys, vals, keys_alt, vals_alt, n, nullptr, bits, tmp, s)) {
        // a failure here is the stream's last error, which the caller checks
        (void)hipMemcpyAsync(keys, keys_alt, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(vals, vals_alt, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    }
}

}  // namespace fgl
