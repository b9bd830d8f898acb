// Streaming sample weighting, sort by travel, keys/segments and windows.
//
// Reference semantics (rtac_simulation source tree):
//   k_stream   : RayCaster.cu:62-74 (phase, spreading) + Receiver.h:51-60
//                (set_sample) + Directivity.h:30-34 (lookup) + Sample.h:144-160
//   sort       : Receiver.cu:9-42 (thrust::sort by travel), whose last pass
//                writes the keys (Binner.cu:24-49) and bounds the segments
//                (Binner.cu:51-80), kernels_sort.hip
//   k_windows  : Binner.cu:82-136 (views with the empty-window rule)
// The streaming body itself is in stream_body.h.
#include "device_common.h"
#include "stream_body.h"

namespace rtsim {

// ---------------------------------------------------------------------------
// streaming kernel: 24 B in (range, dir, amplitude) -> 16 B record out; one
// workgroup of 1024 threads per tile of 1024 hits
// ---------------------------------------------------------------------------
constexpr uint32_t kStreamThreads = kStreamTile;

__device__ __forceinline__ void stream_body(const StreamArgs& p, const BucketArgs& bk, int with_hist,
                                            uint32_t blk)
{
    extern __shared__ uint32_t hist[];   // [NB] when counting
    stream_tile<(int)kStreamThreads>(p, bk, with_hist, blk, hist);
}

__global__ __launch_bounds__(kStreamThreads) void k_stream(StreamArgs p, BucketArgs b, int with_hist)
{
    stream_body(p, b, with_hist, xcd_contiguous_tile(blockIdx.x, gridDim.x));
}

// bargs[q]: ping q's bucket counts to make on the way (tiles == 0: none)
__global__ __launch_bounds__(kStreamThreads) void k_stream_batch(const StreamArgs* __restrict__ args,
                                                                 BatchMap m,
                                                                 const BucketArgs* __restrict__ bargs)
{
    uint32_t blk, nblk;
    const uint32_t q = batch_locate(m, blk, nblk);
    const StreamArgs a = args[q];   // private copies: the body's stores cannot alias them
    const BucketArgs b = bargs[q];
    stream_body(a, b, b.tiles != 0, xcd_contiguous_tile(blk, nblk));
}

uint32_t stream_blocks(const StreamArgs& a) { return (a.n + kStreamThreads - 1) / kStreamThreads; }

hipError_t launch_stream(const StreamArgs& a, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_stream, dim3(stream_blocks(a)), dim3(kStreamThreads), 0, s, a, BucketArgs{}, 0);
    return hipGetLastError();
}

hipError_t launch_stream_batch(const StreamArgs* args, BatchMap m, uint32_t blocks,
                               const BucketArgs* bargs, size_t shm, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_stream_batch, dim3(blocks), dim3(kStreamThreads), shm, s, args, m, bargs);
    return hipGetLastError();
}

bool stream_bucket_fusable(const StreamArgs& a, const BucketArgs& b)
{
    return a.n != 0 && b.tiles != 0 && b.tile == kStreamThreads && stream_blocks(a) == b.tiles &&
           b.rec == a.out && b.n == a.n;
}

hipError_t launch_stream_bucket(const StreamArgs& a, const BucketArgs& b, hipStream_t s)
{
    if (!stream_bucket_fusable(a, b)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_stream, dim3(b.tiles), dim3(kStreamThreads), sizeof(uint32_t) * b.NB, s, a, b, 1);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// fill
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    size_t blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, s, p, value, count);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// windows: do_make_bins (Binner.cu:82-106); the segments (Binner.cu:51-80)
// were bounded by the sort's last pass (kernels_sort.hip)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_windows(BinArgs a) { windows_body(a, blockIdx.x); }

__global__ __launch_bounds__(256) void k_windows_batch(const BinArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const BinArgs a = args[p];
    windows_body(a, blk);
}

uint32_t windows_blocks(const BinArgs& a) { return (a.N + 255) / 256; }

hipError_t launch_windows(const BinArgs& a, hipStream_t s)
{
    if (windows_blocks(a)) hipLaunchKernelGGL(k_windows, dim3(windows_blocks(a)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_windows_batch(const BinArgs* args, BatchMap m, uint32_t blocks, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_windows_batch, dim3(blocks), dim3(256), 0, s, args, m);
    return hipGetLastError();
}

}  // namespace rtsim
