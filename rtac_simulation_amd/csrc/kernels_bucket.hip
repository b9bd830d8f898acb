// Binning of the grid (MFMA) form: one counting pass that groups the samples
// by blocks of R grid rows (BucketArgs, rtsim_internal.h).
//
//   bucket counts : per tile, the items of every bucket (a sample with a
//                   Binner key goes to the buckets of its two grid rows) and
//                   the largest finite |re|, |im| of the tile's binned
//                   samples; made by the streaming kernel itself when it
//                   writes every sample of a ping (kernels_bin.hip), else by
//                   k_bucket_hist from the records
//   bucket scan   : k_sort_scan over the rows [bucket][tile] (exclusive, the
//                   bucket totals), its first workgroup also reduces the tile
//                   maxima to the deposit's fixed-point scale
//   bucket scatter: k_bucket_scatter writes every binned sample as an item
//                   {re, im, travel, grid column} into its buckets' ranges
//                   (LDS cursors per bucket) and the tile's Binner-key
//                   occupancy bits
// Items of one bucket land in an order that depends on the LDS atomics; the
// deposit sums them in exact fixed point, so the grid does not depend on it.
// The Binner keys of the reference (Binner.cu:24-49) decide which samples are
// binned (bin_key, the same function the sort uses) and the occupancy bits
// give the windows' emptiness (Binner.cu:82-106) without a sort.
#include "device_common.h"
#include "rtsim_internal.h"

namespace rtsim {

constexpr int kBucketThreads = 256;
constexpr int kBucketU = 4;   // samples per thread with their loads in flight together

// One tile's bucket counts into LDS `hist` and its running maximum `m`.
__device__ __forceinline__ void bucket_count(const BucketArgs& a, const float4& r, uint32_t* hist,
                                             float& m)
{
    if (bin_key(a.key, r.z) == kOutOfRange) return;
    int b0, b1;
    bucket_pair(a, g_row_of(a.rows, r.z), b0, b1);
    if (b0 >= 0) atomicAdd(&hist[b0], 1u);
    if (b1 >= 0) atomicAdd(&hist[b1], 1u);
    const float c = fmaxf(fabsf(r.x), fabsf(r.y));
    if (c <= 3.402823466e38f) m = fmaxf(m, c);
}

// k_bucket_hist: tile `tile` of the records (blockDim = kBucketThreads).
__device__ __forceinline__ void bucket_hist_body(const BucketArgs& a, uint32_t tile)
{
    extern __shared__ uint32_t hist[];   // [NB]
    __shared__ float wmax[kBucketThreads / 64];
    const uint32_t t = threadIdx.x;
    for (int b = (int)t; b < a.NB; b += kBucketThreads) hist[b] = 0u;
    __syncthreads();
    float m = 0.f;
    const uint32_t i0 = tile * a.tile, i1 = min(i0 + a.tile, a.n);
    for (uint32_t ib = i0; ib < i1; ib += kBucketU * kBucketThreads) {
        float4 r[kBucketU];
#pragma unroll
        for (int u = 0; u < kBucketU; u++) r[u] = a.rec[min(ib + u * kBucketThreads + t, i1 - 1)];
#pragma unroll
        for (int u = 0; u < kBucketU; u++)
            if (ib + u * kBucketThreads + t < i1) bucket_count(a, r[u], hist, m);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((t & 63) == 0) wmax[t >> 6] = m;
    __syncthreads();
    for (int b = (int)t; b < a.NB; b += kBucketThreads) a.counts[(size_t)b * a.tiles + tile] = hist[b];
    if (t == 0) {
        float mm = wmax[0];
        for (int k = 1; k < kBucketThreads / 64; k++) mm = fmaxf(mm, wmax[k]);
        a.tmax[tile] = __float_as_uint(mm);
    }
}

// k_bucket_scatter: tile `tile`'s items into their buckets, its occupancy bits.
__device__ __forceinline__ void bucket_scatter_body(const BucketArgs& a, uint32_t tile)
{
    extern __shared__ uint32_t lds[];    // cursor[NB], then occupancy[occ_words]
    uint32_t* cur = lds;
    uint32_t* locc = lds + a.NB;
    __shared__ uint32_t ws[4];
    const uint32_t t = threadIdx.x;
    // bucket b starts at the sum of the totals of the buckets before it; this
    // tile's part of it after the earlier tiles' (the scanned counts)
    const int per = (a.NB + kBucketThreads - 1) / kBucketThreads;
    const int b0 = (int)t * per, b1 = min(b0 + per, a.NB);
    uint32_t s = 0;
    for (int b = b0; b < b1; b++) s += a.totals[b];
    uint32_t run = sort_block_scan(s, ws, nullptr);
    for (int b = b0; b < b1; b++) {
        cur[b] = run + a.counts[(size_t)b * a.tiles + tile];
        run += a.totals[b];
    }
    for (int w = (int)t; w < a.occ_words; w += kBucketThreads) locc[w] = 0u;
    __syncthreads();
    const uint32_t i0 = tile * a.tile, i1 = min(i0 + a.tile, a.n);
    for (uint32_t ib = i0; ib < i1; ib += kBucketU * kBucketThreads) {
        float4 r[kBucketU];
#pragma unroll
        for (int u = 0; u < kBucketU; u++) r[u] = a.rec[min(ib + u * kBucketThreads + t, i1 - 1)];
#pragma unroll
        for (int u = 0; u < kBucketU; u++) {
            if (ib + u * kBucketThreads + t >= i1) continue;
            const uint32_t k = bin_key(a.key, r[u].z);
            if (k == kOutOfRange) continue;
            int c0, c1;
            bucket_pair(a, g_row_of(a.rows, r[u].z), c0, c1);
            const float4 item = make_float4(r[u].x, r[u].y, r[u].z, grid_x_of(a, r[u].w));
            if (c0 >= 0) a.items[atomicAdd(&cur[c0], 1u)] = item;
            if (c1 >= 0) a.items[atomicAdd(&cur[c1], 1u)] = item;
            atomicOr(&locc[k >> 5], 1u << (k & 31));
        }
    }
    __syncthreads();
    for (int w = (int)t; w < a.occ_words; w += kBucketThreads)
        a.occ[(size_t)tile * a.occ_words + w] = locc[w];
}

__global__ __launch_bounds__(kBucketThreads) void k_bucket_hist(BucketArgs a) { bucket_hist_body(a, blockIdx.x); }

__global__ __launch_bounds__(kBucketThreads) void k_bucket_scatter(BucketArgs a)
{
    bucket_scatter_body(a, xcd_contiguous_tile(blockIdx.x, gridDim.x));
}

__global__ __launch_bounds__(kBucketThreads) void k_bucket_hist_batch(const BucketArgs* __restrict__ args,
                                                                      BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const BucketArgs a = args[p];   // a private copy: the body's stores cannot alias it
    bucket_hist_body(a, blk);
}

__global__ __launch_bounds__(kBucketThreads) void k_bucket_scatter_batch(
    const BucketArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const BucketArgs a = args[p];
    bucket_scatter_body(a, xcd_contiguous_tile(blk, nblk));
}

size_t bucket_lds_bytes(const BucketArgs& b) { return sizeof(uint32_t) * ((size_t)b.NB + b.occ_words); }

hipError_t launch_bucket_hist(const BucketArgs& b, hipStream_t s)
{
    if (b.tiles)
        hipLaunchKernelGGL(k_bucket_hist, dim3(b.tiles), dim3(kBucketThreads), sizeof(uint32_t) * b.NB, s, b);
    return hipGetLastError();
}

hipError_t launch_bucket_scan(const BucketArgs& b, hipStream_t s)
{
    return launch_scan_rows(ScanArgs{b.counts, b.tiles, b.totals, b.tmax, b.vmax}, (uint32_t)b.NB, s);
}

hipError_t launch_bucket_scatter(const BucketArgs& b, hipStream_t s)
{
    if (b.tiles)
        hipLaunchKernelGGL(k_bucket_scatter, dim3(b.tiles), dim3(kBucketThreads), bucket_lds_bytes(b), s, b);
    return hipGetLastError();
}

hipError_t launch_bucket_hist_batch(const BucketArgs* args, BatchMap m, uint32_t blocks, size_t shm,
                                    hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_bucket_hist_batch, dim3(blocks), dim3(kBucketThreads), shm, s, args, m);
    return hipGetLastError();
}

hipError_t launch_bucket_scatter_batch(const BucketArgs* args, BatchMap m, uint32_t blocks, size_t shm,
                                       hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_bucket_scatter_batch, dim3(blocks), dim3(kBucketThreads), shm, s, args, m);
    return hipGetLastError();
}

}  // namespace rtsim
