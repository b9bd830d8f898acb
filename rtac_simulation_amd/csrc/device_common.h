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

__device__ __forceinline__ float image_travel(uint32_t img)
{
    uint32_t u = (img & 0x80000000u) ? (img & 0x7fffffffu) : ~img;
    return __uint_as_float(u);
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

}  // namespace rtsim
