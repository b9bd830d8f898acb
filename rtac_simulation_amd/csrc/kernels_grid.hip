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
// grid range coordinate of a travel (the row table uses the same rounding)
__device__ __forceinline__ float grid_y(const DepositArgs& a, float t)
{
    return grid_y_of(t, a.range0, a.inv_dr, a.off_r);
}

// First entry below n at index >= j0 of the G-row table (rows without samples
// hold n), and the same from j0 + 2 (n when there is none): the first samples
// of G rows >= j0 - 1 and >= j0 + 1.  Each wave scans 64 entries per step.
__device__ __forceinline__ void run_bounds(const uint32_t* __restrict__ marks, int count,
                                           uint32_t n, int j0, uint32_t& s0, uint32_t& s1)
{
    const uint32_t lane = lane_id();
    s0 = n;
    s1 = n;
    bool need0 = true;
    for (int base = j0; base < count; base += 64) {
        const int j = base + (int)lane;
        const uint32_t m = j < count ? marks[j] : n;
        const uint64_t hit = __ballot(m < n);
        if (need0 && hit) {
            s0 = __shfl(m, __ffsll((long long)hit) - 1, 64);
            need0 = false;
        }
        const uint64_t hit1 = base == j0 ? hit & ~3ull : hit;
        if (hit1) {
            s1 = __shfl(m, __ffsll((long long)hit1) - 1, 64);
            break;
        }
    }
}

// One workgroup per G row rho.  Its samples are the sorted run whose G row
// is rho - 1 or rho (each sample's range weights go to rows gy and gy + 1);
// each adds w_y w_x v to two columns of the row.  The row is accumulated in
// LDS as 64-bit fixed point: integer adds are exact and independent of their
// order, so the grid is bit-reproducible (and integer LDS atomics run at the
// rate of plain LDS traffic, unlike FP32 ones).  With M the largest finite
// |re|, |im| over the ping's in-range samples (found by the sort), products
// are scaled by 2^k with (run length) M 2^k < 2^62, so no partial sum can
// overflow; each product is rounded once, in FP32, then summed exactly.  A
// run holding a non-finite in-range value yields a NaN row.
constexpr int kDepositU = 4;   // samples per thread whose loads are in flight together
constexpr int kDepositThreads = 256;   // 512 and 1024 measured 0.6 and 2.5 us slower

__device__ __forceinline__ void deposit_body(const DepositArgs& a, int rho)
{
    extern __shared__ long long sacc[];               // [2][GW]: re plane, im plane
    const int t = threadIdx.x;
    const int GW = a.GW;
    unsigned long long* sre = reinterpret_cast<unsigned long long*>(sacc);
    unsigned long long* sim = sre + GW;
    for (int k = t; k < 2 * GW; k += kDepositThreads) sacc[k] = 0;
    uint32_t s0, s1;   // first samples with G row >= rho - 1 and >= rho + 1
    run_bounds(a.row_start, a.GH + 3, a.n, rho, s0, s1);
    const float M = __uint_as_float(a.vmax[0]);
    int em = 0;
    frexpf(M, &em);                                   // M < 2^em (em = 0 for M = 0)
    const int ec = 32 - __clz(max(s1 - s0, 1u));     // run length < 2^ec
    const int k = 62 - em - ec;
    constexpr int U = kDepositU;
    __syncthreads();                                  // LDS cleared
    int bad = 0;
    float2* row = a.grid + (size_t)rho * GW;

    for (uint32_t sb = s0; sb < s1; sb += U * kDepositThreads) {
        float4 rec[U];
        float xg[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t s = min(sb + u * kDepositThreads + t, s1 - 1);
            rec[u] = a.sorted[s];
            xg[u] = a.xg[s];
        }
        // weights for every slot first (no branch between the loads and their
        // use), then the atomics; a slot past the run or out of range gets
        // column -2 and adds nothing
        int col[U];
        float wx[U], vr[U], vi[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const float y = grid_y(a, rec[u].z);
            const float x = xg[u];
            const float fy = floorf(y), fx = floorf(x);
            const float wy1 = y - fy;
            const int gy = (int)fy + a.fr;
            // this row receives weight 1 - wy1 from gy == rho, wy1 from gy == rho - 1
            const float wy = gy == rho ? 1.0f - wy1 : (gy == rho - 1 ? wy1 : 0.0f);
            const bool ok = sb + u * kDepositThreads + t < s1 && rec[u].z >= a.t_lo && rec[u].z <= a.t_hi;
            bad |= ok && !(fmaxf(fabsf(rec[u].x), fabsf(rec[u].y)) <= 3.402823466e38f &&
                           rec[u].x == rec[u].x && rec[u].y == rec[u].y);
            wx[u] = x - fx;
            col[u] = ok ? (int)fx + a.fb : -2;
            vr[u] = wy * rec[u].x;
            vi[u] = wy * rec[u].y;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int gx = col[u];
            if (gx >= 0 && gx < GW) {
                const float w = 1.0f - wx[u];
                atomicAdd(&sre[gx], (unsigned long long)(long long)ldexpf(w * vr[u], k));
                atomicAdd(&sim[gx], (unsigned long long)(long long)ldexpf(w * vi[u], k));
            }
            if (gx + 1 >= 0 && gx + 1 < GW) {
                atomicAdd(&sre[gx + 1], (unsigned long long)(long long)ldexpf(wx[u] * vr[u], k));
                atomicAdd(&sim[gx + 1], (unsigned long long)(long long)ldexpf(wx[u] * vi[u], k));
            }
        }
    }
    if (__syncthreads_or(bad)) {
        for (int c = t; c < GW; c += kDepositThreads) row[c] = make_float2(NAN, NAN);
        return;
    }
    for (int c = t; c < GW; c += kDepositThreads)
        row[c] = make_float2(ldexpf((float)sacc[c], -k), ldexpf((float)sacc[GW + c], -k));
}

// workgroups [0, GH) deposit G rows, the rest compute the Binner windows
// (do_make_bins, Binner.cu:82-106) for the correlation's empty-row mask
__device__ __forceinline__ void deposit_or_windows(const DepositArgs& a, uint32_t blk)
{
    if (blk < (uint32_t)a.GH) deposit_body(a, (int)blk);
    else windows_body(a.bins, blk - (uint32_t)a.GH);
}

__global__ __launch_bounds__(kDepositThreads) void k_deposit(DepositArgs a) { deposit_or_windows(a, blockIdx.x); }

__global__ __launch_bounds__(kDepositThreads) void k_deposit_batch(const DepositArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const DepositArgs a = args[p];   // a private copy: the body's stores cannot alias it
    deposit_or_windows(a, blk);
}

uint32_t deposit_blocks(const DepositArgs& a)
{
    return (uint32_t)a.GH + (a.bins.N + kDepositThreads - 1) / kDepositThreads;   // + windows
}
size_t deposit_lds_bytes(const DepositArgs& a) { return sizeof(long long) * 2 * a.GW; }

hipError_t launch_deposit(const DepositArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_deposit, dim3(deposit_blocks(a)), dim3(kDepositThreads), deposit_lds_bytes(a), s, a);
    return hipGetLastError();
}

hipError_t launch_deposit_batch(const DepositArgs* args, BatchMap m, uint32_t blocks, size_t shm,
                                hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_deposit_batch, dim3(blocks), dim3(kDepositThreads), shm, s, args, m);
    return hipGetLastError();
}

void init_grid_attributes(int max_lds)
{
    set_dynamic_lds_limit((const void*)k_deposit, max_lds);
    set_dynamic_lds_limit((const void*)k_deposit_batch, max_lds);
    init_corr_attributes(max_lds);
}

}  // namespace rtsim
