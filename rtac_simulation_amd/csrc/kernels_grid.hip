// Grid (MFMA) form of the PSF convolution.
//
// Restates do_sparse_convolve_2d_f (SensorInstance2D.cu:44-83) as
//   1. a bilinear (cloud-in-cell) deposit of every sample with a Binner key
//      onto the output grid extended by the PSF half-widths (k_deposit, from
//      the items the bucketing pass grouped by blocks of grid rows,
//      kernels_bucket.hip), and
//   2. a "valid" 2D correlation of that grid with the PSF taps,
//        out(r, b) = sum_{i < Tr, j < Tb} T[i][j] G[r + i][b + j],
//      computed as complex Toeplitz implicit GEMMs on MFMA (k_correlate),
//   3. with the empty-window rule of do_make_bins (Binner.cu:96-97) applied
//      as a row mask in the epilogue; the deposit launch's last workgroups
//      find which windows are empty from the bucketing's key occupancy.
// In exact arithmetic this equals the gather whenever the output axes are
// uniform at the tap pitch and the Binner overlap covers the PSF support.
// This file holds the deposit; the correlation is in kernels_corr.hip.
#include "deposit_body.h"
#include "device_common.h"
#include "grid_internal.h"

namespace rtsim {

// 16 waves per bucket: with one workgroup per bucket (a few hundred on a
// 1024-row grid) fewer waves leave the CUs idle while the LDS atomics wait
constexpr int kDepositThreads = 1024;

__global__ __launch_bounds__(kDepositThreads) void k_deposit(DepositArgs a)
{
    extern __shared__ long long sacc[];
    deposit_or_windows<kDepositThreads>(a, blockIdx.x, sacc);
}

__global__ __launch_bounds__(kDepositThreads) void k_deposit_batch(const DepositArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const DepositArgs a = args[p];   // a private copy: the body's stores cannot alias it
    extern __shared__ long long sacc[];
    deposit_or_windows<kDepositThreads>(a, blk, sacc);
}

uint32_t deposit_blocks(const DepositArgs& a)
{
    return (uint32_t)a.NB + (a.N + kWindowBins - 1) / kWindowBins;   // + windows
}

size_t deposit_lds_bytes(const DepositArgs& a)
{
    const size_t dep = sizeof(long long) * 2 * (size_t)a.GW * a.R;
    const size_t win = sizeof(uint32_t) * ((size_t)a.occ_words + 2);
    return dep > win ? dep : win;
}

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
    init_bucket_attributes(max_lds);
    init_corr_attributes(max_lds);
}

}  // namespace rtsim
