// Grid (MFMA) form: kernel arguments and tile constants.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rtsim {

constexpr int kDepositRows = 4;   // G rows per deposit workgroup
constexpr int kCorrRows = 64;     // output rows per correlation workgroup (4 waves x 16)
constexpr int kCorrCols = 32;     // output bearings per workgroup (2 MFMA N tiles)
constexpr int kTapChunk = 16;     // range taps staged in LDS at a time

// Extended grid G: rho = grid row + fr, kappa = grid column + fb, so that
// out(r, b) = sum T[i][j] G[r + i][b + j].
struct DepositArgs {
    const float4* sorted;
    const float2* coord;       // bearing-axis coordinate (hi, lo) per sorted sample
    const uint32_t* skeys;     // OutOfRange samples are skipped
    uint32_t n;
    uint32_t* row_start;       // GH + 3 entries
    float2* grid;              // [GH][GW]
    int GH, GW;
    int fr, fb;                // floor((Tr - 1) / 2), floor((Tb - 1) / 2)
    float range0, inv_dr, off_r;        // y = (t - range0) / dr + off_r
    float bear0_hi, bear0_lo, inv_db, off_b;  // x = (x_s - bear0) / db + off_b
};

struct CorrArgs {
    const float2* grid;        // [GH][GW]
    const uint16_t* taps;      // [4 planes][Tr][Tb]: re_hi, re_lo, im_hi, im_lo
    const uint2* win;          // empty window -> masked row
    float2* out;               // [H][W]
    int H, W, Tr, Tb;
    int GH, GW;
    int KB;                    // 32-wide K blocks per 16-wide N tile
    int AS;                    // LDS row stride of the A planes (elements)
    int BL;                    // LDS length of one padded tap row (elements)
    float tap_scale;           // power of two applied to the taps
};

hipError_t launch_deposit(const DepositArgs& a, hipStream_t s);
hipError_t launch_correlate(const CorrArgs& a, bool fp16, hipStream_t s);
size_t correlate_lds_bytes(const CorrArgs& a);

}  // namespace rtsim
