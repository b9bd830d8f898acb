// Grid (MFMA) form: kernel arguments and tile constants.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "rtsim_internal.h"

namespace rtsim {

constexpr int kCorrRows = 64;     // largest tile height of the correlation shapes
constexpr int kCorrCols = 64;     // largest tile width of the correlation shapes
constexpr int kCorrShapes = 4;    // {64x32, 32x32, 32x64, 64x16} (rows x bearings)
constexpr size_t kCorrLdsTarget = 80 * 1024;   // two workgroups per CU

// rows x bearings of correlation tile shape `shape`
void corr_tile(int shape, int* rows, int* cols);

// Extended grid G: rho = grid row + fr, kappa = grid column + fb, so that
// out(r, b) = sum T[i][j] G[r + i][b + j].
struct DepositArgs {
    const float4* sorted;
    const float2* coord;       // bearing-axis coordinate (hi, lo) per sorted sample
    const uint32_t* skeys;     // OutOfRange samples are skipped
    uint32_t n;
    const uint32_t* row_start; // GH + 3 entries: first sample of each G row present, else n
    const uint32_t* vmax;      // kVmaxSlots words whose maximum bounds |re|, |im| (FP32 bits)
    BinArgs bins;              // the launch's last workgroups compute the windows
    float2* grid;              // [GH][GW]
    int GH, GW;
    int fr, fb;                // floor((Tr - 1) / 2), floor((Tb - 1) / 2)
    float range0, inv_dr, off_r;        // y = (t - range0) / dr + off_r
    float bear0_hi, bear0_lo, inv_db, off_b;  // x = (x_s - bear0) / db + off_b
};

struct CorrArgs {
    const float2* grid;        // [GH][GW]
    const uint16_t* toep;      // Toeplitz tap rows [2 shifts][4 planes][Tr][BL]
    const uint2* win;          // empty window -> masked row
    float2* out;               // [H][W]
    int H, W, Tr, Tb;
    int GH, GW;
    int KB;                    // 32-wide K blocks per 16-wide N tile
    int shape;                 // tile shape index (corr_tile)
    int AS;                    // LDS row stride of the A planes (elements)
    int BL;                    // length of one padded tap row (elements)
    int TCH;                   // range taps whose Toeplitz rows are resident in LDS at once
    float tap_scale;           // power of two applied to the taps
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
hipError_t launch_correlate_batch(bool fp16, int shape, const CorrArgs* args, BatchMap m,
                                  uint32_t blocks, size_t shm, hipStream_t s);
size_t correlate_lds_bytes(const CorrArgs& a);
// Tap chunk for a.shape: all Tr taps when that fits kCorrLdsTarget, else
// the largest chunk that does, else the largest that fits max_lds; 0 when
// even one tap row does not fit.
int correlate_chunk(const CorrArgs& a, size_t max_lds);
void init_corr_attributes(int max_lds);

}  // namespace rtsim
