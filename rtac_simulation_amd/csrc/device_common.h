// Device helpers shared by the HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "rtsim_internal.h"

namespace rtsim {

// Order-preserving 32-bit image of a travel value: equal travels map to equal
// images (-0 == +0), every NaN maps above +inf.
__device__ __forceinline__ uint32_t travel_image(float t)
{
    uint32_t u = __float_as_uint(t);
    if (t != t) u = 0x7fc00000u;
    if (u == 0x80000000u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Binner::Key::operator() (Binner.h:47-53): k = (int)fmaf(a, t, b) truncated
// toward zero (saturating), clamps compare the index with the margin bounds.
// NaN travel is out of range.
__device__ __forceinline__ uint32_t bin_key(const KeyParams& k, float t)
{
    if (t != t) return kOutOfRange;
    float v = __builtin_fmaf(k.a, t, k.b);
    if (v != v) return kOutOfRange;
    int res;
    if (v >= 2147483648.0f)       res = INT_MAX;
    else if (v <= -2147483648.0f) res = INT_MIN;
    else                          res = (int)v;
    if (res < 0) return ((float)res >= k.lower) ? 0u : kOutOfRange;
    if ((uint32_t)res >= k.n) return ((float)res <= k.upper) ? k.n - 1u : kOutOfRange;
    return (uint32_t)res;
}

// sin(x) in FP64 for bearing-axis coordinates.  |x| <= pi/2 (every sensor
// bearing): the Taylor polynomial of degree 19, whose truncation error is
// below (pi/2)^21 / 21! < 3e-16, in Horner form on x^2 (11 FP64 operations
// instead of a full range-reduced sine); larger |x| and NaN use sin().
__device__ __forceinline__ double sin_bearing(double x)
{
    if (!(fabs(x) <= 1.5707963267948966)) return sin(x);
    const double z = x * x;
    double p = -1.0 / 121645100408832000.0;            // -1/19!
    p = __builtin_fma(p, z, 1.0 / 355687428096000.0);  //  1/17!
    p = __builtin_fma(p, z, -1.0 / 1307674368000.0);   // -1/15!
    p = __builtin_fma(p, z, 1.0 / 6227020800.0);       //  1/13!
    p = __builtin_fma(p, z, -1.0 / 39916800.0);        // -1/11!
    p = __builtin_fma(p, z, 1.0 / 362880.0);           //  1/9!
    p = __builtin_fma(p, z, -1.0 / 5040.0);            // -1/7!
    p = __builtin_fma(p, z, 1.0 / 120.0);              //  1/5!
    p = __builtin_fma(p, z, -1.0 / 6.0);               // -1/3!
    return __builtin_fma(p * z, x, x);
}

// Grid (MFMA) form: range coordinate of a travel in grid rows, one explicit
// fma so that the bucketing and the deposit round identically.
__device__ __forceinline__ float grid_y_of(float t, float range0, float inv_dr, float off_r)
{
    return __builtin_fmaf(t - range0, inv_dr, off_r);
}

// G row (grid row of the first range weight) of a travel, clamped to
// [-2, GH + 1] (beyond, both of its rows are outside the grid).
__device__ __forceinline__ int g_row_of(const GridRows& m, float t)
{
    float r = floorf(grid_y_of(t, m.range0, m.inv_dr, m.off_r)) + (float)m.fr;
    r = fminf(fmaxf(r, -2.0f), (float)m.GH + 1.0f);
    return (int)r;
}

// The buckets of a sample with G row gy: b0 = the bucket of grid row gy, b1
// = the bucket of row gy + 1 when it is another one; -1 where a row is
// outside the grid [0, GH) or the bucket repeats.
__device__ __forceinline__ void bucket_pair(const BucketArgs& a, int gy, int& b0, int& b1)
{
    b0 = gy >= 0 && gy < a.rows.GH ? gy / a.R : -1;
    const int c1 = gy + 1 >= 0 && gy + 1 < a.rows.GH ? (gy + 1) / a.R : -1;
    b1 = c1 != b0 ? c1 : -1;
}

// Grid column coordinate of a bearing: sin(bearing) in FP64 (axis 1) or the
// bearing, mapped to (x - bear0) inv_db + off_b.
__device__ __forceinline__ float grid_x_of(const BucketArgs& a, float bearing)
{
    const double x = a.axis == 1 ? sin_bearing((double)bearing) : (double)bearing;
    return (float)((x - a.bear0) * (double)a.inv_db) + a.off_b;
}

// Inclusive wave scan (wave64) of a 32-bit value.
__device__ __forceinline__ uint32_t sort_wave_scan(uint32_t x, uint32_t lane)
{
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    return x;
}

// Exclusive block scan (NT threads, a multiple of 64) of one value per
// thread; `ws` holds NT / 64 words.  Ends with the last wave's partial in ws:
// a second call needs its own ws (or a barrier first).
template <int NT = 256>
__device__ __forceinline__ uint32_t sort_block_scan(uint32_t x, uint32_t* ws, uint32_t* total)
{
    constexpr int NW = NT / 64;
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t incl = sort_wave_scan(x, lane);
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    uint32_t off = 0, all = 0;
#pragma unroll
    for (int k = 0; k < NW; k++) {
        off += (uint32_t)k < w ? ws[k] : 0u;
        all += ws[k];
    }
    if (total) *total = all;
    return off + incl - x;
}

// do_make_bins (Binner.cu:82-106) for output bin blk * 256 + threadIdx.x.
// A segment is empty when x >= y: the sort's last pass leaves (~0, 0) on keys
// it never writes.  Each thread writes its own empty segment back as (0, 0),
// the reference's memset value (Binner.cu:56-58); other threads may read that
// entry in either state, and both read as empty.
__device__ __forceinline__ void windows_body(const BinArgs& a, uint32_t blk)
{
    const int tid = blk * blockDim.x + threadIdx.x;
    if (tid >= (int)a.N) return;
    const uint2* seg = a.seg;
    auto empty = [&](int k) {
        const uint2 s = seg[k];
        return s.x >= s.y;
    };
    int istart = max(0, tid - a.ov);
    int iend = min((int)a.N - 1, tid + a.ov);
    while (istart < iend && empty(istart)) istart++;
    while (iend > istart && empty(iend)) iend--;
    if (istart >= iend)
        a.win[tid] = make_uint2(0u, 0u);
    else
        a.win[tid] = make_uint2(seg[istart].x, seg[iend].y - seg[istart].x);
    if (empty(tid)) a.seg[tid] = make_uint2(0u, 0u);
}

// Number of set bits of `mask` in lanes strictly below the calling lane.
__device__ __forceinline__ uint32_t popc_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t lane_id()
{
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Tile of workgroup `blk` of `nblk` such that each XCD runs a contiguous
// range of tiles.  Workgroups are dealt to the 8 XCDs round robin (XCD
// blk mod 8); XCD x gets the tiles [x q + min(x, r), ...) with q = nblk / 8,
// r = nblk mod 8.  Scatters whose neighbouring tiles write neighbouring
// addresses then fill their cache lines in one L2 instead of several.
__device__ __forceinline__ uint32_t xcd_contiguous_tile(uint32_t blk, uint32_t nblk)
{
    const uint32_t q = nblk >> 3, r = nblk & 7u, x = blk & 7u, k = blk >> 3;
    return x * q + min(x, r) + k;
}

// Batched launch: the ping of this workgroup (largest p with first[p] <=
// blockIdx.x, so pings without workgroups are skipped), its local workgroup
// index and its workgroup count.
__device__ __forceinline__ uint32_t batch_locate(const BatchMap& m, uint32_t& block,
                                                 uint32_t& nblocks)
{
    const uint32_t bx = blockIdx.x;
    uint32_t lo = 0, hi = m.nping;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (m.first[mid] <= bx) lo = mid;
        else hi = mid;
    }
    block = bx - m.first[lo];
    nblocks = m.first[lo + 1] - m.first[lo];
    return lo;
}

}  // namespace rtsim
