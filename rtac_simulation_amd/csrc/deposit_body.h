// Grid (MFMA) form deposit bodies: the bilinear fixed-point deposit of one
// bucket's items on its grid rows, and the windows' emptiness from the key
// occupancy.  Shared by the deposit launch (kernels_grid.hip) and the
// persistent launch (kernels_persist.hip).
#pragma once

#include "device_common.h"
#include "grid_internal.h"

namespace rtsim {

// One workgroup per bucket b: grid rows [b R, b R + R).  Its items are every
// binned sample whose grid row gy or gy + 1 lies in those rows (the bucketing
// put a sample on a bucket's last row into both buckets); each adds
// w_y w_x v to two columns of the rows of the bucket it reaches.  The rows are
// accumulated in LDS as 64-bit fixed point: integer adds are exact and
// independent of their order, so the grid is bit-reproducible whatever order
// the bucketing left the items in (and integer LDS atomics run at the rate of
// plain LDS traffic, unlike FP32 ones).  With M the largest finite |re|, |im|
// over the ping's binned samples (from the bucket scan), products are scaled
// by 2^k with (items in the bucket) M 2^k < 2^62, so no partial sum can
// overflow, and M 2^k < 2^51 (the exact conversion below); each product is rounded once, in FP32, then converted to fixed
// point (rounded to the nearest integer) and summed exactly.  A row reached
// by a non-finite value is NaN.
constexpr int kDepositU = 4;          // items per thread whose loads are in flight together (1024 threads)

// v 2^k rounded to the nearest integer, as a 64-bit two's complement word,
// for |v 2^k| < 2^51: the FMA lands v 2^k + 1.5 2^52 in the binade where the
// double's unit in the last place is 1, so its mantissa bits minus the
// constant's are the integer (four instructions instead of the generic
// float -> int64 conversion's dozen)
__device__ __forceinline__ unsigned long long fixed_of(float v, double scale)
{
    constexpr double kMagic = 6755399441055744.0;   // 1.5 * 2^52
    const double d = __builtin_fma((double)v, scale, kMagic);
    return (unsigned long long)(__double_as_longlong(d) - __double_as_longlong(kMagic));
}
// NT threads per workgroup (the deposit launch: 1024; the persistent launch:
// 512); sacc is the dynamic LDS ([R][2][GW]: re plane, im plane per row).
template <int NT, int U = kDepositU>
__device__ __forceinline__ void deposit_body(const DepositArgs& a, int b, long long* sacc)
{
    __shared__ uint32_t ws[NT / 64];
    __shared__ uint32_t bad[kMaxBucketRows];
    const int t = threadIdx.x;
    const int GW = a.GW;
    const int r0 = b * a.R;
    const int rows = min(a.R, a.GH - r0);
    for (int k = t; k < 2 * GW * rows; k += NT) sacc[k] = 0;
    if (t < kMaxBucketRows) bad[t] = 0u;
    // the bucket's items start after those of the buckets before it
    uint32_t s = 0;
    for (int i = t; i < b; i += NT) s += a.totals[i];
    uint32_t part;
    (void)sort_block_scan<NT>(s, ws, &part);
    // bounded by the items buffer (the bucketing never fills it past its end)
    const uint32_t cnt = a.totals[b];
    const uint32_t s0 = min(part, a.items_cap), s1 = min(s0 + cnt, a.items_cap);
    const float M = __uint_as_float(a.vmax[0]);
    int em = 0;
    frexpf(M, &em);                                   // M < 2^em (em = 0 for M = 0)
    const int ec = 32 - __clz(max(cnt, 1u));          // items < 2^ec
    const int k = min(62 - em - ec, 51 - em);         // sums < 2^62, products < 2^51
    const double scale = ldexp(1.0, k);
    __syncthreads();                                  // LDS cleared
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(sacc);

    for (uint32_t sb = s0; sb < s1; sb += U * NT) {
        float4 it[U];
#pragma unroll
        for (int u = 0; u < U; u++) it[u] = a.items[min(sb + u * NT + t, s1 - 1)];
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (sb + u * NT + t >= s1) continue;
            const float y = grid_y_of(it[u].z, a.rows.range0, a.rows.inv_dr, a.rows.off_r);
            const float wy1 = y - floorf(y);
            const int l0 = g_row_of(a.rows, it[u].z) - r0;   // rows l0 (1 - wy1) and l0 + 1 (wy1)
            const float x = it[u].w;
            const float fx = floorf(x);
            const float wx = x - fx;
            const int gx = (int)fx + a.fb;
            const bool nonfinite = !(fmaxf(fabsf(it[u].x), fabsf(it[u].y)) <= 3.402823466e38f &&
                                     it[u].x == it[u].x && it[u].y == it[u].y);
            const bool c0 = (unsigned)gx < (unsigned)GW, c1 = (unsigned)(gx + 1) < (unsigned)GW;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int l = l0 + h;
                if ((unsigned)l >= (unsigned)rows) continue;
                if (nonfinite) bad[l] = 1u;
                const float wy = h ? wy1 : 1.0f - wy1;
                const float vr = wy * it[u].x, vi = wy * it[u].y;
                // 32-bit LDS offsets: row l's re plane, its im plane GW further
                unsigned long long* re = acc + (__mul24(l, 2 * GW) + gx);
                if (c0) {
                    const float w = 1.0f - wx;
                    atomicAdd(re, fixed_of(w * vr, scale));
                    atomicAdd(re + GW, fixed_of(w * vi, scale));
                }
                if (c1) {
                    atomicAdd(re + 1, fixed_of(wx * vr, scale));
                    atomicAdd(re + GW + 1, fixed_of(wx * vi, scale));
                }
            }
        }
    }
    __syncthreads();
    for (int l = 0; l < rows; l++) {
        float2* row = a.grid + (size_t)(r0 + l) * GW;
        const long long* re = sacc + (size_t)l * 2 * GW;
        if (bad[l]) {
            for (int c = t; c < GW; c += NT) row[c] = make_float2(NAN, NAN);
        } else {
            for (int c = t; c < GW; c += NT)
                row[c] = make_float2(ldexpf((float)re[c], -k), ldexpf((float)re[GW + c], -k));
        }
    }
}

// Windows of output bins [wb B, wb B + B), B = kWindowBins (do_make_bins,
// Binner.cu:82-106): a window holds samples when at least two of the bins
// [r - ov, r + ov] (inside [0, N)) hold samples; the occupancy is the OR of
// the bucketing's per-tile key bits.  Only the emptiness is kept:
// win[r] = (0, 1) for a window with samples, (0, 0) for an empty one (the
// correlation masks rows whose window size is 0).  A workgroup covers few bins
// so that its OR over every tile is a single round of loads per thread.
constexpr int kWindowBins = 64;

template <int NT>
__device__ __forceinline__ void occupancy_windows(const DepositArgs& a, int wb, uint32_t* lw)
{
    static_assert(kWindowBins <= NT, "one thread per bin of the group");
    const int t = threadIdx.x;
    const int N = (int)a.N, ov = a.ov;
    const int first = wb * kWindowBins;
    const int lo_bin = max(0, first - ov), hi_bin = min(N - 1, first + kWindowBins - 1 + ov);
    const int wlo = lo_bin >> 5, nw = (hi_bin >> 5) - wlo + 1;
    for (int w = t; w < nw; w += NT) lw[w] = 0u;
    __syncthreads();
    // thread t ORs word t mod nw_e of every phases-th tile in a register (its
    // loads are independent, so many are in flight), then once into LDS
    const int nw_e = min(nw, NT), phases = NT / nw_e;
    const int slot = t % nw_e, phase = t / nw_e;
    if (phase < phases) {
        for (int w = slot; w < nw; w += nw_e) {
            const uint32_t* col = a.occ + wlo + w;
            uint32_t v = 0u;
#pragma unroll 8
            for (int tile = phase; tile < (int)a.tiles; tile += phases) v |= col[(size_t)tile * a.occ_words];
            if (v) atomicOr(&lw[w], v);
        }
    }
    __syncthreads();
    const int r = first + t;
    if (t >= kWindowBins || r >= N) return;
    const int lo = max(0, r - ov), hi = min(N - 1, r + ov);
    int c = 0;
    for (int w = lo >> 5; w <= (hi >> 5) && c < 2; w++) {
        uint32_t v = lw[w - wlo];
        const int wb0 = w << 5;
        if (wb0 < lo) v &= ~0u << (lo - wb0);
        if (wb0 + 31 > hi) v &= ~0u >> (wb0 + 31 - hi);
        c += __popc(v);
    }
    a.win[r] = make_uint2(0u, c >= 2 ? 1u : 0u);
}

// workgroups [0, NB) deposit buckets, the rest compute the windows
template <int NT, int U = kDepositU>
__device__ __forceinline__ void deposit_or_windows(const DepositArgs& a, uint32_t blk, void* lds)
{
    if (blk < (uint32_t)a.NB) deposit_body<NT, U>(a, (int)blk, static_cast<long long*>(lds));
    else occupancy_windows<NT>(a, (int)(blk - (uint32_t)a.NB), static_cast<uint32_t*>(lds));
}

}  // namespace rtsim
