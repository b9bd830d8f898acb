// Grid-form correlation on MFMA: out(r, b) = sum_{i < Tr, j < Tb} T[i][j] G[r + i][b + j]
// (the extended grid G comes from k_deposit, kernels_grid.hip).
//
// For each bearing tap j the range taps form a Toeplitz GEMM:
//   out(r0 + m, b0 + n) += sum_k A_j[m][k] B_j[k][n],
//   A_j[m][k] = T[k - m][j] (zero outside [0, Tr)),  B_j[k][n] = G[r0 + k][b0 + n + j],
// on v_mfma_f32_32x32x16_{bf16,f16} (wave64): M = 32 output rows, N = 32
// bearings, K = 16 grid rows per instruction, KR = ceil((31 + Tr) / 16) K blocks
// per tap column.  Lane l holds A[l % 32][8 (l / 32) + p], B[8 (l / 32) + p][l % 32]
// and D[8 (e / 4) + 4 (l / 32) + e % 4][l % 32] (element e of 16).  On gfx950 this
// instruction sustains 1.7x the flop rate of v_mfma_f32_16x16x32 (measured with
// independent chains: 2.21 vs 1.32 PFLOP/s in FP16), which pays for the wider
// Toeplitz band (31 taps over 64 K rows).
// Complex products use the real embedding (re = ArBr - AiBi, im = ArBi + AiBr);
// every real product is split hi+lo, X Y ~ XhYh + XlYh + XhYl, with FP32
// accumulation in six independent chains (so that no MFMA waits for the
// previous one's result).
// Scaling: each workgroup scales its G tile by a power of two (max |G| in
// [2^13, 2^14)), found from the staged values themselves, and the taps carry a
// host power of two, so FP16 operands stay in range; the epilogue divides both
// out exactly.  Rows whose reference window is empty (Binner.cu:96-97) are
// written as zero.
//
// Tile shapes (template TR x TC output tiles of 32 x 32 per workgroup, JS =
// NW / (TR TC) waves per tile for NW = 4 or 8 waves per workgroup, each taking
// an equal share of the tile's Tb KR (tap, K block) steps; partial sums are
// added through LDS): chosen per sensor by timing (Sensor::autotune).
// LDS (16-bit elements):
//   G tile, transposed: [4 planes re_hi, re_lo, im_hi, im_lo][column][row],
//     row stride RS = (rows rounded to 16) + 8, so that the 16-byte B fragment
//     reads of 8 consecutive lanes (consecutive columns) fall in distinct banks;
//   tap table, copied from the prebuilt [2 shifts][4 planes][Tb][L] (shift c
//     holds Tpad[x + c], Tpad[x] = T[x - 31][j]; the shifts are
//     corr_copy_stride apart): lane (m, h) of K block kb reads
//     Tpad[16 kb + 8 h - m + 31 + p], p < 8, from the shift that makes the
//     read 4-byte aligned.
// Staging: every thread issues all its grid loads (2 x 2 cells per item, two
// float4 loads; each column's two rows pack into one 32-bit LDS word per
// plane) and the table copy at once,
// so a workgroup pays one memory round trip before its MFMA loop, where every
// step's fragments are read while the previous step's MFMAs run.
#include <algorithm>

#include "corr_body.h"
#include "device_common.h"
#include "grid_internal.h"

namespace rtsim {

template <bool FP16, int TR, int TC, bool W8>
__global__ __launch_bounds__(W8 ? 512 : 256) void k_correlate(CorrArgs a)
{
    correlate_body<FP16, TR, TC, W8>(a, (int)blockIdx.x, (a.W + 32 * TC - 1) / (32 * TC));
}

template <bool FP16, int TR, int TC, bool W8>
__global__ __launch_bounds__(W8 ? 512 : 256) void k_correlate_batch(const CorrArgs* __restrict__ args,
                                                                  BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const CorrArgs a = args[p];   // a private copy: the body's stores cannot alias it
    const int gx = (a.W + 32 * TC - 1) / (32 * TC);
    correlate_body<FP16, TR, TC, W8>(a, (int)blk, gx);
}

// tile shapes {TR, TC}: rows = 32 TR, cols = 32 TC; 4 / (TR TC) waves per tile
static const int kShapes[kCorrShapes][2] = {{1, 1}, {1, 2}, {2, 1}, {2, 2}};

void corr_tile(int shape, int* rows, int* cols)
{
    *rows = 32 * kShapes[shape][0];
    *cols = 32 * kShapes[shape][1];
}

int corr_copy_stride_host(int Tb, int L) { return corr_copy_stride(Tb, L); }

int corr_waves8(const CorrArgs& a, int max_lds)
{
    CorrArgs b = a;
    b.waves8 = 0;
    return 2 * correlate_lds_bytes(b) > (size_t)max_lds ? 1 : 0;
}

int corr_row_stride(int shape, int KR)
{
    const int rows = 32 * (kShapes[shape][0] - 1) + 16 * KR;   // a multiple of 16
    return rows + 8;
}

size_t correlate_lds_bytes(const CorrArgs& a)
{
    int rows, cols;
    corr_tile(a.shape, &rows, &cols);
    const size_t gt = (size_t)4 * corr_cols(cols / 32, a.Tb) * a.RS;
    const size_t tt = (size_t)2 * corr_copy_stride(a.Tb, a.L);
    const int nw = a.waves8 ? 8 : 4;
    const size_t part = (size_t)nw * 32 * 64 * 2;   // partial sums (floats, as 16-bit units)
    return sizeof(uint16_t) * std::max(gt + tt, part);
}

template <bool FP16, bool W8>
static void launch_shape(const CorrArgs& a, dim3 grid, size_t shm, hipStream_t s)
{
    switch (a.shape) {
    case 0: hipLaunchKernelGGL((k_correlate<FP16, 1, 1, W8>), grid, dim3(W8 ? 512 : 256), shm, s, a); break;
    case 1: hipLaunchKernelGGL((k_correlate<FP16, 1, 2, W8>), grid, dim3(W8 ? 512 : 256), shm, s, a); break;
    case 2: hipLaunchKernelGGL((k_correlate<FP16, 2, 1, W8>), grid, dim3(W8 ? 512 : 256), shm, s, a); break;
    default: hipLaunchKernelGGL((k_correlate<FP16, 2, 2, W8>), grid, dim3(W8 ? 512 : 256), shm, s, a); break;
    }
}

template <bool FP16, bool W8>
static void launch_shape_batch(int shape, const CorrArgs* args, BatchMap m, dim3 grid, size_t shm,
                               hipStream_t s)
{
    const dim3 block(W8 ? 512 : 256);
    switch (shape) {
    case 0: hipLaunchKernelGGL((k_correlate_batch<FP16, 1, 1, W8>), grid, block, shm, s, args, m); break;
    case 1: hipLaunchKernelGGL((k_correlate_batch<FP16, 1, 2, W8>), grid, block, shm, s, args, m); break;
    case 2: hipLaunchKernelGGL((k_correlate_batch<FP16, 2, 1, W8>), grid, block, shm, s, args, m); break;
    default: hipLaunchKernelGGL((k_correlate_batch<FP16, 2, 2, W8>), grid, block, shm, s, args, m); break;
    }
}

uint32_t correlate_blocks(const CorrArgs& a)
{
    int rows, cols;
    corr_tile(a.shape, &rows, &cols);
    return (uint32_t)(((a.W + cols - 1) / cols) * ((a.H + rows - 1) / rows));
}

hipError_t launch_correlate_batch(bool fp16, int shape, bool waves8, const CorrArgs* args, BatchMap m,
                                  uint32_t blocks, size_t shm, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    const dim3 g(blocks);
    if (fp16) {
        if (waves8) launch_shape_batch<true, true>(shape, args, m, g, shm, s);
        else launch_shape_batch<true, false>(shape, args, m, g, shm, s);
    } else {
        if (waves8) launch_shape_batch<false, true>(shape, args, m, g, shm, s);
        else launch_shape_batch<false, false>(shape, args, m, g, shm, s);
    }
    return hipGetLastError();
}

hipError_t launch_correlate(const CorrArgs& a, bool fp16, hipStream_t s)
{
    const dim3 grid(correlate_blocks(a));
    const size_t shm = correlate_lds_bytes(a);
    if (fp16) {
        if (a.waves8) launch_shape<true, true>(a, grid, shm, s);
        else launch_shape<true, false>(a, grid, shm, s);
    } else {
        if (a.waves8) launch_shape<false, true>(a, grid, shm, s);
        else launch_shape<false, false>(a, grid, shm, s);
    }
    return hipGetLastError();
}

void init_corr_attributes(int max_lds)
{
    set_dynamic_lds_limit((const void*)k_correlate<true, 1, 1, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<true, 1, 2, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<true, 2, 1, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<true, 2, 2, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<false, 1, 1, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<false, 1, 2, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<false, 2, 1, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<false, 2, 2, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<true, 1, 1, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<true, 1, 2, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<true, 2, 1, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<true, 2, 2, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<false, 1, 1, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<false, 1, 2, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<false, 2, 1, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<false, 2, 2, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<true, 1, 1, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<true, 1, 2, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<true, 2, 1, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<true, 2, 2, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<true, 1, 1, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<true, 1, 2, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<true, 2, 1, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<true, 2, 2, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<false, 1, 1, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<false, 1, 2, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<false, 2, 1, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<false, 2, 2, false>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<false, 1, 1, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<false, 1, 2, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<false, 2, 1, true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<false, 2, 2, true>, max_lds);
}

}  // namespace rtsim
