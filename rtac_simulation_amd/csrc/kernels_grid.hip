// Grid (MFMA) form of the PSF convolution.
//
// Restates do_sparse_convolve_2d_f (SensorInstance2D.cu:44-83) as
//   1. a bilinear (cloud-in-cell) deposit of every in-range sample onto the
//      output grid extended by the PSF half-widths (k_deposit), and
//   2. a "valid" 2D correlation of that grid with the PSF taps,
//        out(r, b) = sum_{i < Tr, j < Tb} T[i][j] G[r + i][b + j],
//      computed as complex Toeplitz implicit GEMMs on MFMA (k_correlate),
//   3. with the empty-window rule of do_make_bins (Binner.cu:96-97) applied
//      as a row mask in the epilogue.
// In exact arithmetic this equals the gather whenever the output axes are
// uniform at the tap pitch and the Binner overlap covers the PSF support.
// This file holds the deposit; the correlation is in kernels_corr.hip.
#include "device_common.h"
#include "grid_internal.h"

namespace rtsim {

// ---------------------------------------------------------------------------
// deposit
// ---------------------------------------------------------------------------
// grid range coordinate of a travel; one explicit fma so that the row table
// and the deposit round identically
__device__ __forceinline__ float grid_y(const DepositArgs& a, float t)
{
    return __builtin_fmaf(t - a.range0, a.inv_dr, a.off_r);
}

__device__ __forceinline__ float grid_row_of(const DepositArgs& a, float t)
{
    return floorf(grid_y(a, t));
}

// G row (grid row + fr) of a sorted sample, clamped to [-2, GH + 2]; the
// travel is sorted so it is non-decreasing in the sample index (NaN travel
// sorts last and maps to the top).
__device__ __forceinline__ int g_row(const DepositArgs& a, float t)
{
    const float r = grid_row_of(a, t) + (float)a.fr;
    if (!(r >= -2.0f)) return r != r ? a.GH + 2 : -2;
    return r > (float)(a.GH + 2) ? a.GH + 2 : (int)r;
}

// row_start[rho + 1] = first sorted sample whose G row is >= rho, for rho in
// [-1, GH + 1].  Built in two steps: k_row_marks writes, for every row that
// has samples, the index of its first sample (the table is pre-filled with
// n); k_row_suffix_min then gives every row the minimum over itself and the
// rows above it.
__global__ __launch_bounds__(256) void k_row_marks(DepositArgs a)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int r = g_row(a, image_travel(a.img[i]));
    const int rp = i == 0 ? -3 : g_row(a, image_travel(a.img[i - 1]));
    if (r != rp && r >= -1 && r <= a.GH + 1) a.row_start[r + 1] = i;
}

// one workgroup of 1024 threads, reverse inclusive min-scan of GH + 3 entries
__global__ __launch_bounds__(1024) void k_row_suffix_min(uint32_t* __restrict__ rs, int count)
{
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    const int per = (count + 1023) / 1024;
    const int beg = t * per, end = min(beg + per, count);
    uint32_t m = 0xffffffffu;
    for (int k = end - 1; k >= beg; k--) m = min(m, rs[k]);
    part[t] = m;
    __syncthreads();
    // suffix minimum of the chunk minima (Hillis-Steele, toward lower t)
    for (int o = 1; o < 1024; o <<= 1) {
        uint32_t v = t + o < 1024 ? part[t + o] : 0xffffffffu;
        __syncthreads();
        part[t] = min(part[t], v);
        __syncthreads();
    }
    uint32_t run = t + 1 < 1024 ? part[t + 1] : 0xffffffffu;
    for (int k = end - 1; k >= beg; k--) {
        run = min(run, rs[k]);
        rs[k] = run;
    }
}

__global__ __launch_bounds__(256) void k_fill_u32_grid(uint32_t* p, uint32_t v, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_deposit(DepositArgs a)
{
    extern __shared__ float sacc[];                   // [kDepositRows][GW][2]
    const int t = threadIdx.x;
    const int rho0 = blockIdx.x * kDepositRows;       // first G row of this block
    const int GW = a.GW;
    for (int k = t; k < kDepositRows * GW * 2; k += blockDim.x) sacc[k] = 0.f;
    __syncthreads();
    // samples whose G row is in [rho0 - 1, rho0 + R): they touch this block
    const uint32_t s0 = a.row_start[rho0 - 1 + 1];
    const uint32_t s1 = a.row_start[min(rho0 + kDepositRows, a.GH + 1) + 1];
    constexpr int U = 4;   // samples per thread whose loads are in flight together
    for (uint32_t sb = s0; sb < s1; sb += U * blockDim.x) {
        uint32_t key[U];
        float4 rec[U];
        float2 xc[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t s = sb + u * blockDim.x + t;
            const bool ok = s < s1;
            key[u] = ok ? a.skeys[s] : kOutOfRange;
            rec[u] = ok ? a.sorted[s] : make_float4(0.f, 0.f, 0.f, 0.f);
            xc[u] = ok ? a.coord[s] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (key[u] == kOutOfRange) continue;
            const float y = grid_y(a, rec[u].z);
            const float x = ((xc[u].x - a.bear0_hi) + (xc[u].y - a.bear0_lo)) * a.inv_db + a.off_b;
            const float fy = floorf(y), fx = floorf(x);
            const float wy1 = y - fy, wx1 = x - fx;
            const int gy = (int)fy + a.fr, gx = (int)fx + a.fb;
#pragma unroll
            for (int dy = 0; dy < 2; dy++) {
                const int lr = gy + dy - rho0;
                if (lr < 0 || lr >= kDepositRows) continue;
                const float wy = dy ? wy1 : 1.0f - wy1;
#pragma unroll
                for (int dx = 0; dx < 2; dx++) {
                    const int c = gx + dx;
                    if (c < 0 || c >= GW) continue;
                    const float w = wy * (dx ? wx1 : 1.0f - wx1);
                    atomicAdd(&sacc[2 * (lr * GW + c)], w * rec[u].x);
                    atomicAdd(&sacc[2 * (lr * GW + c) + 1], w * rec[u].y);
                }
            }
        }
    }
    __syncthreads();
    for (int k = t; k < kDepositRows * GW; k += blockDim.x) {
        const int lr = k / GW, c = k - lr * GW;
        const int rho = rho0 + lr;
        if (rho < a.GH)
            a.grid[(size_t)rho * GW + c] = make_float2(sacc[2 * k], sacc[2 * k + 1]);
    }
    // max |re|, |im| per 16 columns: the correlation's tile scales
    const int CG = GW / 16;
    for (int k = t; k < kDepositRows * CG; k += blockDim.x) {
        const int lr = k / CG, g = k - lr * CG;
        const int rho = rho0 + lr;
        if (rho >= a.GH) continue;
        const float* v = sacc + 2 * (lr * GW + 16 * g);
        float mx = 0.f;
#pragma unroll
        for (int c = 0; c < 32; c++) mx = fmaxf(mx, fabsf(v[c]));
        a.colmax[(size_t)rho * CG + g] = mx;
    }
}

hipError_t launch_deposit(const DepositArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u32_grid, dim3((a.GH + 3 + 255) / 256), dim3(256), 0, s,
                       a.row_start, a.n, (uint32_t)(a.GH + 3));
    if (a.n) hipLaunchKernelGGL(k_row_marks, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_row_suffix_min, dim3(1), dim3(1024), 0, s, a.row_start, a.GH + 3);
    const int blocks = (a.GH + kDepositRows - 1) / kDepositRows;
    const size_t shm = sizeof(float) * 2 * kDepositRows * a.GW;
    hipLaunchKernelGGL(k_deposit, dim3(blocks), dim3(256), shm, s, a);
    return hipGetLastError();
}

void init_grid_attributes(int max_lds)
{
    set_dynamic_lds_limit((const void*)k_deposit, max_lds);
    init_corr_attributes(max_lds);
}

}  // namespace rtsim
