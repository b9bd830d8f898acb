// Plain FP32 sparse PSF convolution: the reference's per-sample form.
//
//   2D: do_sparse_convolve_2d_f (SensorInstance2D.cu:44-83)
//       out(r, b) = sum_{s in view(r)} v_s K(x_s - x_b, t_s - r_r)
//   1D: do_sparse_convolve_1d_f (SensorInstance1D.cu:35-72)
//       out(r)    = sum_{s in view(r)} v_s K(t_s - r_r)
//
// K is the bilinear (texture linear filter, border 0) interpolant of the PSF
// taps (ReductionKernel.h:44-73) with full-FP32 weights.  The reference
// gathers every window sample into every bearing column; here each sample
// of view(r) scatters only into the bearing columns where K can be non-zero
// (|x_s - x_b| < (Tb + 1)/2 tap pitches), which skips exactly the zero terms.
// Output coordinates are double-float (hi, lo) pairs so that t_s - r_r and
// x_s - x_b keep full FP32 relative accuracy.
#include "device_common.h"

namespace rtsim {

// PSF value at tap coordinates (xB along bearing taps, yB along range taps).
// Caller guarantees xB in (-1, Tb) and yB in (-1, Tr).
__device__ __forceinline__ float2 psf_bilinear(const float2* __restrict__ psf, int Tr, int Tb,
                                               float xB, float yB)
{
    const float fi = floorf(xB), fj = floorf(yB);
    const float ax = xB - fi, ay = yB - fj;
    const int i0 = (int)fi, j0 = (int)fj;
    float2 acc = make_float2(0.f, 0.f);
    if (j0 >= 0) {
        const float2* row = psf + j0 * Tb;
        float2 lo = i0 >= 0 ? row[i0] : make_float2(0.f, 0.f);
        float2 hi = i0 + 1 < Tb ? row[i0 + 1] : make_float2(0.f, 0.f);
        const float wy = 1.0f - ay;
        acc.x += wy * (lo.x + ax * (hi.x - lo.x));
        acc.y += wy * (lo.y + ax * (hi.y - lo.y));
    }
    if (j0 + 1 < Tr) {
        const float2* row = psf + (j0 + 1) * Tb;
        float2 lo = i0 >= 0 ? row[i0] : make_float2(0.f, 0.f);
        float2 hi = i0 + 1 < Tb ? row[i0 + 1] : make_float2(0.f, 0.f);
        acc.x += ay * (lo.x + ax * (hi.x - lo.x));
        acc.y += ay * (lo.y + ax * (hi.y - lo.y));
    }
    return acc;
}

size_t plain2d_lds_bytes(const ConvArgs& a)
{
    return sizeof(float) * 2 * a.W + sizeof(float2) * a.W +
           (a.psf_in_lds ? sizeof(float2) * a.Tr * a.Tb : 0);
}

// One workgroup per output range row.  LDS holds the row accumulator, the
// bearing coordinates and (when it fits) the PSF table.
__device__ __forceinline__ void conv2d_plain_body(const ConvArgs& a, int r)
{
    extern __shared__ float smem[];
    const int W = a.W;
    float* acc = smem;                                  // 2 W
    float2* bear = reinterpret_cast<float2*>(smem + 2 * W); // W
    // PSF table in LDS when it fits, else read through the caches
    const float2* psf = a.psf_in_lds ? bear + W : a.psf;
    const int t = threadIdx.x;
    for (int b = t; b < W; b += blockDim.x) {
        acc[2 * b] = 0.f;
        acc[2 * b + 1] = 0.f;
        bear[b] = a.bear_c[b];
    }
    if (a.psf_in_lds)
        for (int k = t; k < a.Tr * a.Tb; k += blockDim.x) (bear + W)[k] = a.psf[k];
    __syncthreads();

    const uint2 win = a.win[r];
    const float2 rc = a.range_c[r];
    for (uint32_t k = t; k < win.y; k += blockDim.x) {
        const uint32_t s = win.x + k;
        const float4 rec = a.sorted[s];
        const float dr = (rec.z - rc.x) - rc.y;
        const float yB = dr * a.inv_pr + a.half_r;
        if (!(yB > -1.0f && yB < (float)a.Tr)) continue;
        const float2 xc = a.coord[s];
        // first bearing whose coordinate exceeds x_s - supp (binary search)
        const float lo_x = xc.x - a.supp_b * 1.0001f - 1e-30f;
        int l = 0, h = W;
        while (l < h) {
            int m = (l + h) >> 1;
            if (bear[m].x < lo_x) l = m + 1; else h = m;
        }
        const float hi_x = xc.x + a.supp_b * 1.0001f + 1e-30f;
        for (int b = l; b < W && bear[b].x <= hi_x; b++) {
            const float dx = (xc.x - bear[b].x) + (xc.y - bear[b].y);
            const float xB = dx * a.inv_pb + a.half_b;
            if (!(xB > -1.0f && xB < (float)a.Tb)) continue;
            const float2 K = psf_bilinear(psf, a.Tr, a.Tb, xB, yB);
            atomicAdd(&acc[2 * b], rec.x * K.x - rec.y * K.y);
            atomicAdd(&acc[2 * b + 1], rec.x * K.y + rec.y * K.x);
        }
    }
    __syncthreads();
    float2* out = a.out + (size_t)r * W;
    for (int b = t; b < W; b += blockDim.x) out[b] = make_float2(acc[2 * b], acc[2 * b + 1]);
}

__global__ __launch_bounds__(256) void k_conv2d_plain(ConvArgs a) { conv2d_plain_body(a, blockIdx.x); }

__global__ __launch_bounds__(256) void k_conv2d_plain_batch(const ConvArgs* __restrict__ args,
                                                            BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const ConvArgs a = args[p];   // a private copy: the body's stores cannot alias it
    conv2d_plain_body(a, (int)blk);
}

uint32_t conv2d_plain_blocks(const ConvArgs& a) { return a.W ? (uint32_t)a.H : 0u; }

hipError_t launch_conv2d_plain(const ConvArgs& a, hipStream_t s)
{
    if (a.H == 0 || a.W == 0) return hipSuccess;
    size_t shm = plain2d_lds_bytes(a);
    hipLaunchKernelGGL(k_conv2d_plain, dim3(a.H), dim3(256), shm, s, a);
    return hipGetLastError();
}

hipError_t launch_conv2d_plain_batch(const ConvArgs* args, BatchMap m, uint32_t blocks, size_t shm,
                                     hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_conv2d_plain_batch, dim3(blocks), dim3(256), shm, s, args, m);
    return hipGetLastError();
}

void init_plain_attributes(int max_lds)
{
    set_dynamic_lds_limit((const void*)k_conv2d_plain, max_lds);
    set_dynamic_lds_limit((const void*)k_conv2d_plain_batch, max_lds);
}

// 1D: one wave per output range bin, FP32 partial sums reduced across the
// wave (the reference reduces a 512-thread block in shared memory).
__device__ __forceinline__ void conv1d_plain_body(const ConvArgs& a, uint32_t blk)
{
    const int lane = threadIdx.x & 63;
    const int r = (int)blk * 4 + (threadIdx.x >> 6);
    if (r >= a.H) return;
    const uint2 win = a.win[r];
    const float2 rc = a.range_c[r];
    float ar = 0.f, ai = 0.f;
    for (uint32_t k = lane; k < win.y; k += 64) {
        const float4 rec = a.sorted[win.x + k];
        const float dr = (rec.z - rc.x) - rc.y;
        const float yB = dr * a.inv_pr + a.half_r;
        if (!(yB > -1.0f && yB < (float)a.Tr)) continue;
        const float2 K = psf_bilinear(a.psf, a.Tr, 1, 0.0f, yB);
        ar += rec.x * K.x - rec.y * K.y;
        ai += rec.x * K.y + rec.y * K.x;
    }
    for (int o = 32; o > 0; o >>= 1) {
        ar += __shfl_xor(ar, o, 64);
        ai += __shfl_xor(ai, o, 64);
    }
    if (lane == 0) a.out[r] = make_float2(ar, ai);
}

__global__ __launch_bounds__(256) void k_conv1d_plain(ConvArgs a) { conv1d_plain_body(a, blockIdx.x); }

__global__ __launch_bounds__(256) void k_conv1d_plain_batch(const ConvArgs* __restrict__ args,
                                                            BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const ConvArgs a = args[p];
    conv1d_plain_body(a, blk);
}

uint32_t conv1d_plain_blocks(const ConvArgs& a) { return (uint32_t)(a.H + 3) / 4; }

hipError_t launch_conv1d_plain(const ConvArgs& a, hipStream_t s)
{
    if (a.H == 0) return hipSuccess;
    hipLaunchKernelGGL(k_conv1d_plain, dim3(conv1d_plain_blocks(a)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_conv1d_plain_batch(const ConvArgs* args, BatchMap m, uint32_t blocks,
                                     hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_conv1d_plain_batch, dim3(blocks), dim3(256), 0, s, args, m);
    return hipGetLastError();
}

}  // namespace rtsim
