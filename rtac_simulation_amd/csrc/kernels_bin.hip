// Streaming sample weighting, sort by travel, keys/segments and windows.
//
// Reference semantics (rtac_simulation source tree):
//   k_stream   : RayCaster.cu:62-74 (phase, spreading) + Receiver.h:51-60
//                (set_sample) + Directivity.h:30-34 (lookup) + Sample.h:144-160
//   sort       : Receiver.cu:9-42 (thrust::sort by travel), whose last pass
//                writes the keys (Binner.cu:24-49) and bounds the segments
//                (Binner.cu:51-80), kernels_sort.hip
//   k_windows  : Binner.cu:82-136 (views with the empty-window rule)
#include "device_common.h"

namespace rtsim {

// ---------------------------------------------------------------------------
// streaming kernel: 24 B in (range, dir, amplitude) -> 16 B record out
// ---------------------------------------------------------------------------
__device__ __forceinline__ float mirror_coord(float x)
{
    float f = floorf(x);
    float fr = x - f;
    int fi = (int)f;
    return (fi & 1) ? 1.0f - fr : fr;
}

__device__ __forceinline__ int mirror_index(int i, int n)
{
    if (i < 0) i = -1 - i;
    if (i >= n) i = 2 * n - 1 - i;
    return min(max(i, 0), n - 1);
}

// tex2D<float> with linear filtering, mirror addressing, normalised
// coordinates (Directivity.cpp:109-113), weights in full FP32; `az` is
// atan2(ly, lx).
__device__ __forceinline__ float directivity_lookup(const StreamArgs& p, float az, float lx,
                                                   float lz)
{
    float u = mirror_coord(p.dir_xs * az);
    float v = mirror_coord(p.dir_ys * atan2f(lz, lx));
    float xb = u * (float)p.dir_w - 0.5f, yb = v * (float)p.dir_h - 0.5f;
    float fi = floorf(xb), fj = floorf(yb);
    float a = xb - fi, b = yb - fj;
    int i0 = (int)fi, j0 = (int)fj;
    int c0 = mirror_index(i0, p.dir_w), c1 = mirror_index(i0 + 1, p.dir_w);
    int r0 = mirror_index(j0, p.dir_h), r1 = mirror_index(j0 + 1, p.dir_h);
    const float* t = p.dir;
    float t00 = t[r0 * p.dir_w + c0], t10 = t[r0 * p.dir_w + c1];
    float t01 = t[r1 * p.dir_w + c0], t11 = t[r1 * p.dir_w + c1];
    float top = t00 + a * (t10 - t00);
    float bot = t01 + a * (t11 - t01);
    return top + b * (bot - top);
}

// set_sample (Receiver.h:51-60) of one ray hit
__device__ __forceinline__ float4 stream_sample(const StreamArgs& p, float r, float d0, float d1,
                                                float d2, float2 A)
{
    if (r < 1.0e-4f) {
        // set_null_sample (Receiver.h:62-65): value 0, travel -1
        return make_float4(0.f, 0.f, -1.f, 0.f);
    }
    float lx = p.R[0] * d0 + p.R[3] * d1 + p.R[6] * d2;
    float ly = p.R[1] * d0 + p.R[4] * d1 + p.R[7] * d2;
    float lz = p.R[2] * d0 + p.R[5] * d1 + p.R[8] * d2;
    const float az = atan2f(ly, lx);
    float D = p.dir ? directivity_lookup(p, az, lx, lz) : 1.0f;
    // round-trip phase 2 k r reduced modulo one wavelength in FP64
    double q = (double)r * p.k2;
    q -= floor(q);
    float sn, cs;
    sincospif(2.0f * (float)q, &sn, &cs);
    float vr = A.x * cs - A.y * sn;
    float vi = A.x * sn + A.y * cs;
    if (p.spreading) {
        float g = 1.0f / (r * r);
        vr *= g;
        vi *= g;
    }
    float4 o;
    o.x = vr * D;
    o.y = vi * D;
    o.z = r;
    o.w = p.is2d ? -az : 0.0f;   // atan2(-ly, lx): atan2 is odd in y
    return o;
}

// One workgroup per tile of 1024 hits, one hit per thread (more per thread
// measured 5-10 % slower: fewer resident waves hide the sample's dependent
// table lookups less well).  With b (the grid form's bucketing) the
// workgroup also makes that tile's bucket counts and its largest finite
// |re|, |im| of the binned samples, as k_bucket_hist makes them from the same
// records (counts and a maximum: the thread mapping is irrelevant).  Every
// streaming launch (one ping, fused or not, and batched) runs this one body
// with the counts switched at run time (a compile-time switch lets the
// compiler fuse multiply-adds differently, and the records would differ in
// their last bits between launches).
constexpr uint32_t kStreamThreads = 1024;
constexpr int kStreamWaves = (int)kStreamThreads / 64;

__device__ __forceinline__ void stream_body(const StreamArgs& p, const BucketArgs& bk, int with_hist,
                                            uint32_t blk)
{
    const BucketArgs* b = with_hist ? &bk : nullptr;
    extern __shared__ uint32_t hist[];   // [NB] when counting
    __shared__ float wmax[kStreamWaves];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (b)
        for (int k = (int)t; k < b->NB; k += kStreamThreads) hist[k] = 0u;
    const uint32_t i = blk * kStreamThreads + t;
    const uint32_t ic = min(i, p.n - 1);
    const float r = p.range[ic], d0 = p.dx[ic], d1 = p.dy[ic], d2 = p.dz[ic];
    const float2 A = p.amp[ic];
    if (b) __syncthreads();   // counts cleared
    float m = 0.f;
    if (i < p.n) {
        const float4 o = stream_sample(p, r, d0, d1, d2, A);
        p.out[i] = o;
        if (b && bin_key(b->key, o.z) != kOutOfRange) {
            int b0, b1;
            bucket_pair(*b, g_row_of(b->rows, o.z), b0, b1);
            if (b0 >= 0) atomicAdd(&hist[b0], 1u);
            if (b1 >= 0) atomicAdd(&hist[b1], 1u);
            const float c = fmaxf(fabsf(o.x), fabsf(o.y));
            if (c <= 3.402823466e38f) m = fmaxf(m, c);
        }
    }
    if (!b) return;
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if (lane == 0) wmax[w] = m;
    __syncthreads();
    for (int k = (int)t; k < b->NB; k += kStreamThreads) b->counts[(size_t)k * b->tiles + blk] = hist[k];
    if (t == 0) {
        float mm = wmax[0];
        for (int k = 1; k < kStreamWaves; k++) mm = fmaxf(mm, wmax[k]);
        b->tmax[blk] = __float_as_uint(mm);
    }
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
