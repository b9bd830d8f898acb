// Grid (MFMA) form: kernel arguments and tile constants.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "rtsim_internal.h"

namespace rtsim {

constexpr int kCorrRows = 64;     // largest tile height of the correlation shapes
constexpr int kCorrCols = 64;     // largest tile width of the correlation shapes
constexpr int kCorrShapes = 4;    // {32x32, 32x64, 64x32, 64x64} (rows x bearings)

// rows x bearings of correlation tile shape `shape`
void corr_tile(int shape, int* rows, int* cols);
// LDS row stride (16-bit elements per staged grid column) of `shape`
int corr_row_stride(int shape, int KR);
// stride (16-bit elements) between the two shifted copies of the tap table
int corr_copy_stride_host(int Tb, int L);

// Extended grid G: rho = grid row + fr, kappa = grid column + fb, so that
// out(r, b) = sum T[i][j] G[r + i][b + j].
constexpr int kMaxBucketRows = 16;   // grid rows per bucket of the deposit
// bucket-count words of one ping (pings whose 1024-sample tiles would need
// more take larger tiles) and the deposit's LDS per workgroup (its rows)
constexpr size_t kBucketCountBudget = (size_t)2 << 20;
constexpr size_t kDepositLdsBudget = 64 * 1024;
struct DepositArgs {
    const float4* items;       // {re, im, travel, grid column}, grouped by bucket
    uint32_t items_cap;        // items the buffer holds (reads stop there)
    const uint32_t* totals;    // [NB] items per bucket
    int NB, R;                 // buckets of R grid rows
    const uint32_t* vmax;      // 1 word: the largest finite binned |re|, |im| (FP32 bits)
    float2* grid;              // [GH][GW]
    int GH, GW;
    int fb;                    // floor((Tb - 1) / 2)
    GridRows rows;             // G row of a travel (the bucketing's map)
    // windows (the launch's last workgroups): key occupancy per bucketing tile
    const uint32_t* occ;       // [tiles][occ_words]
    int occ_words;
    uint32_t tiles;
    uint32_t N;                // range bins
    int ov;                    // binOverlap
    uint2* win;                // [N]: (0, 1) window with samples, (0, 0) empty
};

struct CorrArgs {
    const float2* grid;        // [GH + zero rows][GW]
    const uint16_t* toep;      // tap columns [2 shifts][4 planes][Tb][L]
    const uint2* win;          // empty window -> masked row
    float2* out;               // [H][W]
    int H, W, Tr, Tb;
    int GH, GW;
    int KR;                    // 16-row K blocks per 32-row output tile: ceil((31 + Tr) / 16)
    int L;                     // length of one padded tap column (elements): 16 KR + 32
    int shape;                 // tile shape index (corr_tile)
    int RS;                    // LDS row stride of the staged grid columns (corr_row_stride)
    float tap_scale;           // power of two applied to the taps
    int waves8;                // 1: 8 waves per workgroup (corr_waves8)
};

hipError_t launch_deposit(const DepositArgs& a, hipStream_t s);
hipError_t launch_correlate(const CorrArgs& a, bool fp16, hipStream_t s);
// batched launches (args are device arrays, one entry per ping; every ping of
// a correlation launch uses the same form and tile shape)
uint32_t deposit_blocks(const DepositArgs& a);
size_t deposit_lds_bytes(const DepositArgs& a);
hipError_t launch_deposit_batch(const DepositArgs* args, BatchMap m, uint32_t blocks, size_t shm,
                                hipStream_t s);
uint32_t correlate_blocks(const CorrArgs& a);
hipError_t launch_correlate_batch(bool fp16, int shape, bool waves8, const CorrArgs* args, BatchMap m,
                                  uint32_t blocks, size_t shm, hipStream_t s);
size_t correlate_lds_bytes(const CorrArgs& a);
// 1 when the shape's LDS leaves room for one workgroup per CU only: the
// workgroup then has 8 waves (two per SIMD), else 4
int corr_waves8(const CorrArgs& a, int max_lds);
void init_corr_attributes(int max_lds);

}  // namespace rtsim
