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
#include "bucket_body.h"
#include "device_common.h"
#include "rtsim_internal.h"

namespace rtsim {

constexpr int kBucketThreads = 256;
constexpr int kBucketU = 4;   // samples per thread with their loads in flight together

__global__ __launch_bounds__(kBucketThreads) void k_bucket_hist(BucketArgs a)
{
    extern __shared__ uint32_t hist[];   // [NB]
    bucket_hist_body<kBucketThreads, kBucketU>(a, blockIdx.x, hist);
}

template <bool kStaged>
__global__ __launch_bounds__(kBucketThreads) void k_bucket_scatter(BucketArgs a)
{
    extern __shared__ uint32_t lds[];
    bucket_scatter_body<kStaged, kBucketThreads, kBucketU>(a, xcd_contiguous_tile(blockIdx.x, gridDim.x), lds);
}

__global__ __launch_bounds__(kBucketThreads) void k_bucket_hist_batch(const BucketArgs* __restrict__ args,
                                                                      BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const BucketArgs a = args[p];   // a private copy: the body's stores cannot alias it
    extern __shared__ uint32_t hist[];
    bucket_hist_body<kBucketThreads, kBucketU>(a, blk, hist);
}

template <bool kStaged>
__global__ __launch_bounds__(kBucketThreads) void k_bucket_scatter_batch(
    const BucketArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const BucketArgs a = args[p];
    extern __shared__ uint32_t lds[];
    bucket_scatter_body<kStaged, kBucketThreads, kBucketU>(a, xcd_contiguous_tile(blk, nblk), lds);
}

namespace {
int g_scatter_lds_limit = 64 * 1024;   // dynamic LDS a scatter workgroup may take (set at init)
}

// Staging pays when a pass spreads its items thinly over many buckets (runs
// of a few items); with a few hundred buckets the direct stores already
// fill their lines in L2 (measured on one ping: 1M samples, 264 buckets:
// 12.6 us direct, ~16 us staged; 16M samples, 1386 buckets: 309 us direct,
// ~255 us staged; the 64-ping batch with up to 724 buckets was faster direct).
constexpr int kStagedMinBuckets = 1024;

bool bucket_scatter_staged(const BucketArgs& b)
{
    return b.NB >= kStagedMinBuckets && bucket_staged_bytes(b) <= (size_t)g_scatter_lds_limit;
}

size_t bucket_lds_bytes(const BucketArgs& b)
{
    return bucket_scatter_staged(b) ? bucket_staged_bytes(b) : sizeof(uint32_t) * ((size_t)b.NB + b.occ_words);
}

void init_bucket_attributes(int max_lds)
{
    g_scatter_lds_limit = max_lds - 64;   // the kernels' static LDS
    set_dynamic_lds_limit((const void*)k_bucket_scatter<true>, max_lds);
    set_dynamic_lds_limit((const void*)k_bucket_scatter_batch<true>, max_lds);
}

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
    if (!b.tiles) return hipGetLastError();
    if (bucket_scatter_staged(b))
        hipLaunchKernelGGL(k_bucket_scatter<true>, dim3(b.tiles), dim3(kBucketThreads), bucket_lds_bytes(b), s, b);
    else
        hipLaunchKernelGGL(k_bucket_scatter<false>, dim3(b.tiles), dim3(kBucketThreads), bucket_lds_bytes(b), s, b);
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
                                       bool staged, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    if (staged)
        hipLaunchKernelGGL(k_bucket_scatter_batch<true>, dim3(blocks), dim3(kBucketThreads), shm, s, args, m);
    else
        hipLaunchKernelGGL(k_bucket_scatter_batch<false>, dim3(blocks), dim3(kBucketThreads), shm, s, args, m);
    return hipGetLastError();
}

}  // namespace rtsim
