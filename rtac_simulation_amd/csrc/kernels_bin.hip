// Streaming sample weighting, sort by travel, keys/segments and windows.
//
// Reference semantics (rtac_simulation source tree):
//   k_stream   : RayCaster.cu:62-74 (phase, spreading) + Receiver.h:51-60
//                (set_sample) + Directivity.h:30-34 (lookup) + Sample.h:144-160
//   sort       : Receiver.cu:9-42 (thrust::sort by travel)
//   k_gather   : Binner.cu:24-49 (keys) + Binner.cu:51-80 (segments)
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
// coordinates (Directivity.cpp:109-113), weights in full FP32.
__device__ __forceinline__ float directivity_lookup(const StreamArgs& p, float lx, float ly,
                                                   float lz)
{
    float u = mirror_coord(p.dir_xs * atan2f(ly, lx));
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

constexpr uint32_t kStreamMaxBlocks = 256 * 64;

__device__ __forceinline__ void stream_body(const StreamArgs& p, uint32_t blk, uint32_t nblk)
{
    const uint32_t stride = nblk * 256;
    for (uint32_t i = blk * 256 + threadIdx.x; i < p.n; i += stride) {
        float r = p.range[i];
        float4 o;
        if (r < 1.0e-4f) {
            // set_null_sample (Receiver.h:62-65): value 0, travel -1
            o = make_float4(0.f, 0.f, -1.f, 0.f);
        } else {
            float d0 = p.dx[i], d1 = p.dy[i], d2 = p.dz[i];
            float lx = p.R[0] * d0 + p.R[3] * d1 + p.R[6] * d2;
            float ly = p.R[1] * d0 + p.R[4] * d1 + p.R[7] * d2;
            float lz = p.R[2] * d0 + p.R[5] * d1 + p.R[8] * d2;
            float D = p.dir ? directivity_lookup(p, lx, ly, lz) : 1.0f;
            // round-trip phase 2 k r reduced modulo one wavelength in FP64
            double q = (double)r * p.k2;
            q -= floor(q);
            float sn, cs;
            sincospif(2.0f * (float)q, &sn, &cs);
            float2 A = p.amp[i];
            float vr = A.x * cs - A.y * sn;
            float vi = A.x * sn + A.y * cs;
            if (p.spreading) {
                float g = 1.0f / (r * r);
                vr *= g;
                vi *= g;
            }
            o.x = vr * D;
            o.y = vi * D;
            o.z = r;
            o.w = p.is2d ? atan2f(-ly, lx) : 0.0f;
        }
        p.out[i] = o;
    }
}

__global__ __launch_bounds__(256) void k_stream(StreamArgs p) { stream_body(p, blockIdx.x, gridDim.x); }

__global__ __launch_bounds__(256) void k_stream_batch(const StreamArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t q = batch_locate(m, blk, nblk);
    stream_body(args[q], blk, nblk);
}

uint32_t stream_blocks(const StreamArgs& a)
{
    const uint32_t blocks = (a.n + 255) / 256;
    return blocks > kStreamMaxBlocks ? kStreamMaxBlocks : blocks;
}

hipError_t launch_stream(const StreamArgs& a, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_stream, dim3(stream_blocks(a)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_stream_batch(const StreamArgs* args, BatchMap m, uint32_t blocks, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_stream_batch, dim3(blocks), dim3(256), 0, s, args, m);
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
// segments (Binner.cu:51-80) and the grid row table, over the sorted keys
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bounds_body(const BinArgs& a, uint32_t blk)
{
    const uint32_t i = blk * 256 + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t k = a.skeys[i];
    if (k != kOutOfRange) {
        if (i == 0)
            a.seg[0].x = 0;  // Binner.cu:56-58 (equal to the zero fill)
        else if (a.skeys[i - 1] != k)
            a.seg[k].x = i;
        if (i == a.n - 1)
            a.seg[k].y = a.n;
        else if (a.skeys[i + 1] != k)
            a.seg[k].y = i + 1;
    }
    if (a.rows.row_start) {
        // first sample of each G row present (the table holds n elsewhere)
        auto row = [&](uint32_t j) {
            return a.img_is_row ? (int)a.img[j] - 2 : g_row_of(a.rows, image_travel(a.img[j]));
        };
        const int r = row(i);
        const int rp = i == 0 ? -3 : row(i - 1);
        if (r != rp && r >= -1 && r <= a.rows.GH + 1) a.rows.row_start[r + 1] = i;
    }
}

// ---------------------------------------------------------------------------
// windows: do_make_bins (Binner.cu:82-106); with a row table, the last
// workgroup gives every row the minimum over itself and the rows above it
// (rows without samples then start where the next row that has some starts)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void windows_body(const BinArgs& a, uint32_t blk)
{
    const uint32_t win_blocks = (a.N + 255) / 256;
    if (blk < win_blocks) {
        const int tid = blk * 256 + threadIdx.x;
        if (tid >= (int)a.N) return;
        const uint2* seg = a.seg;
        int istart = max(0, tid - a.ov);
        int iend = min((int)a.N - 1, tid + a.ov);
        while (istart < iend && seg[istart].y == seg[istart].x) istart++;
        while (iend > istart && seg[iend].y == seg[iend].x) iend--;
        if (istart >= iend)
            a.win[tid] = make_uint2(0u, 0u);
        else
            a.win[tid] = make_uint2(seg[istart].x, seg[iend].y - seg[istart].x);
        return;
    }
    // reverse inclusive min-scan of the GH + 3 row-table entries
    __shared__ uint32_t part[256];
    uint32_t* rs = a.rows.row_start;
    const int count = a.rows.GH + 3;
    const int t = threadIdx.x;
    const int per = (count + 255) / 256;
    const int beg = t * per, end = min(beg + per, count);
    uint32_t m = 0xffffffffu;
    for (int k = end - 1; k >= beg; k--) m = min(m, rs[k]);
    part[t] = m;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {   // suffix minimum of the chunk minima
        const uint32_t v = t + o < 256 ? part[t + o] : 0xffffffffu;
        __syncthreads();
        part[t] = min(part[t], v);
        __syncthreads();
    }
    uint32_t run = t + 1 < 256 ? part[t + 1] : 0xffffffffu;
    for (int k = end - 1; k >= beg; k--) {
        run = min(run, rs[k]);
        rs[k] = run;
    }
}

__global__ __launch_bounds__(256) void k_bounds(BinArgs a) { bounds_body(a, blockIdx.x); }
__global__ __launch_bounds__(256) void k_windows(BinArgs a) { windows_body(a, blockIdx.x); }

__global__ __launch_bounds__(256) void k_bounds_batch(const BinArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    bounds_body(args[p], blk);
}

__global__ __launch_bounds__(256) void k_windows_batch(const BinArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    windows_body(args[p], blk);
}

uint32_t bounds_blocks(const BinArgs& a) { return (a.n + 255) / 256; }
uint32_t windows_blocks(const BinArgs& a)
{
    return (a.N + 255) / 256 + (a.rows.row_start ? 1u : 0u);
}

hipError_t launch_bins(const BinArgs& a, hipStream_t s)
{
    if (bounds_blocks(a)) hipLaunchKernelGGL(k_bounds, dim3(bounds_blocks(a)), dim3(256), 0, s, a);
    if (windows_blocks(a)) hipLaunchKernelGGL(k_windows, dim3(windows_blocks(a)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bounds_batch(const BinArgs* args, BatchMap m, uint32_t blocks, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_bounds_batch, dim3(blocks), dim3(256), 0, s, args, m);
    return hipGetLastError();
}

hipError_t launch_windows_batch(const BinArgs* args, BatchMap m, uint32_t blocks, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_windows_batch, dim3(blocks), dim3(256), 0, s, args, m);
    return hipGetLastError();
}

}  // namespace rtsim
