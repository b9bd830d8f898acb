// Grid (MFMA) form bucketing bodies (BucketArgs, rtsim_internal.h): the
// bucket counts of a tile, and the scatter of a tile's items into their
// buckets with its key-occupancy bits.  Shared by the bucketing launches
// (kernels_bucket.hip) and the persistent launch (kernels_persist.hip).
#pragma once

#include "device_common.h"
#include "rtsim_internal.h"

namespace rtsim {

// NT threads per workgroup, U samples per thread with their loads in flight
// together: a pass covers U NT samples (the launches of kernels_bucket.hip:
// 256 x 4; the persistent launch: 512 x 2)

// One tile's bucket counts into LDS `hist` and its running maximum `m`.
__device__ __forceinline__ void bucket_count(const BucketArgs& a, const float4& r, uint32_t* hist,
                                             float& m)
{
    if (bin_key(a.key, r.z) == kOutOfRange) return;
    int b0, b1;
    bucket_pair(a, g_row_of(a.rows, r.z), b0, b1);
    if (b0 >= 0) atomicAdd(&hist[b0], 1u);
    if (b1 >= 0) atomicAdd(&hist[b1], 1u);
    const float c = fmaxf(fabsf(r.x), fabsf(r.y));
    if (c <= 3.402823466e38f) m = fmaxf(m, c);
}

// k_bucket_hist: tile `tile` of the records, counts in LDS `hist` ([NB]).
template <int NT, int U>
__device__ __forceinline__ void bucket_hist_body(const BucketArgs& a, uint32_t tile, uint32_t* hist)
{
    __shared__ float wmax[NT / 64];
    const uint32_t t = threadIdx.x;
    for (int b = (int)t; b < a.NB; b += NT) hist[b] = 0u;
    __syncthreads();
    float m = 0.f;
    const uint32_t i0 = tile * a.tile, i1 = min(i0 + a.tile, a.n);
    for (uint32_t ib = i0; ib < i1; ib += U * NT) {
        float4 r[U];
#pragma unroll
        for (int u = 0; u < U; u++) r[u] = a.rec[min(ib + u * NT + t, i1 - 1)];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (ib + u * NT + t < i1) bucket_count(a, r[u], hist, m);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((t & 63) == 0) wmax[t >> 6] = m;
    __syncthreads();
    for (int b = (int)t; b < a.NB; b += NT) a.counts[(size_t)b * a.tiles + tile] = hist[b];
    if (t == 0) {
        float mm = wmax[0];
        for (int k = 1; k < NT / 64; k++) mm = fmaxf(mm, wmax[k]);
        a.tmax[tile] = __float_as_uint(mm);
    }
}

// k_bucket_scatter: tile `tile`'s items into their buckets, its occupancy bits.
// The cursor of bucket b starts at the sum of the totals of the buckets
// before it plus this tile's place among the tiles (the scanned counts).
//
// Staged form (kStaged): each pass of kScatterPass samples first ranks its
// items per bucket in LDS, stages them in bucket order with their global
// positions, then writes them out in that order: the items of one bucket
// run go to consecutive addresses from consecutive lanes, instead of every
// lane of a store hitting another bucket (partial cache lines, written back
// half empty).  Unstaged form: every item is stored at its bucket's cursor
// directly (pings with too many buckets for the staging LDS).
constexpr int kScatterPass = 1024;                // samples per pass (U NT)
constexpr int kStageItems = 2 * kScatterPass;     // a sample makes at most two items

template <int NT>
__device__ __forceinline__ void bucket_cursors(const BucketArgs& a, uint32_t tile, uint32_t* cur, uint32_t* ws)
{
    const uint32_t t = threadIdx.x;
    const int per = (a.NB + NT - 1) / NT;
    const int b0 = (int)t * per, b1 = min(b0 + per, a.NB);
    uint32_t s = 0;
    for (int b = b0; b < b1; b++) s += a.totals[b];
    uint32_t run = sort_block_scan<NT>(s, ws, nullptr);
    for (int b = b0; b < b1; b++) {
        cur[b] = run + a.counts[(size_t)b * a.tiles + tile];
        run += a.totals[b];
    }
}

// the item of record r and its buckets (c0 = c1 = -1: not binned)
__device__ __forceinline__ void bucket_item(const BucketArgs& a, const float4& r, bool valid, int& c0, int& c1,
                                            float4& item, uint32_t& key)
{
    c0 = c1 = -1;
    key = valid ? bin_key(a.key, r.z) : kOutOfRange;
    if (key == kOutOfRange) return;
    bucket_pair(a, g_row_of(a.rows, r.z), c0, c1);
    item = make_float4(r.x, r.y, r.z, grid_x_of(a, r.w));
}

// lds: the dynamic LDS (bucket_lds_bytes)
template <bool kStaged, int NT, int U>
__device__ __forceinline__ void bucket_scatter_body(const BucketArgs& a, uint32_t tile, uint32_t* lds)
{
    // the staged form's LDS holds the items of kScatterPass samples
    static_assert(!kStaged || U * NT == kScatterPass, "a staged pass is kScatterPass samples");
    uint32_t* cur = lds;                                  // [NB] global cursor
    uint32_t* locc = lds + a.NB;                          // [occ_words]
    uint32_t* lcnt = locc + a.occ_words;                  // staged: [NB] items of the pass
    uint32_t* lofs = lcnt + a.NB;                         // staged: [NB] their offset in the stage
    float4* sitem = reinterpret_cast<float4*>(lds + ((2 * a.NB + a.occ_words + a.NB + 3) & ~3));
    uint32_t* spos = reinterpret_cast<uint32_t*>(sitem + kStageItems);
    __shared__ uint32_t ws[NT / 64], ws2[NT / 64];
    const uint32_t t = threadIdx.x;
    bucket_cursors<NT>(a, tile, cur, ws);
    for (int w = (int)t; w < a.occ_words; w += NT) locc[w] = 0u;
    if (kStaged)
        for (int b = (int)t; b < a.NB; b += NT) lcnt[b] = 0u;
    __syncthreads();
    const int per = (a.NB + NT - 1) / NT;
    const int rb0 = (int)t * per, rb1 = min(rb0 + per, a.NB);   // this thread's buckets
    const uint32_t i0 = tile * a.tile, i1 = min(i0 + a.tile, a.n);
    for (uint32_t ib = i0; ib < i1; ib += U * NT) {
        float4 r[U];
#pragma unroll
        for (int u = 0; u < U; u++) r[u] = a.rec[min(ib + u * NT + t, i1 - 1)];
        if constexpr (!kStaged) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                int c0, c1;
                float4 item;
                uint32_t k;
                bucket_item(a, r[u], ib + u * NT + t < i1, c0, c1, item, k);
                if (k == kOutOfRange) continue;
                if (c0 >= 0) {
                    const uint32_t q = atomicAdd(&cur[c0], 1u);
                    if (q < a.items_cap) a.items[q] = item;
                }
                if (c1 >= 0) {
                    const uint32_t q = atomicAdd(&cur[c1], 1u);
                    if (q < a.items_cap) a.items[q] = item;
                }
                atomicOr(&locc[k >> 5], 1u << (k & 31));
            }
        } else {
            int c0[U], c1[U];
            uint32_t k0[U], k1[U];
            float4 item[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                uint32_t k;
                bucket_item(a, r[u], ib + u * NT + t < i1, c0[u], c1[u], item[u], k);
                k0[u] = c0[u] >= 0 ? atomicAdd(&lcnt[c0[u]], 1u) : 0u;
                k1[u] = c1[u] >= 0 ? atomicAdd(&lcnt[c1[u]], 1u) : 0u;
                if (k != kOutOfRange) atomicOr(&locc[k >> 5], 1u << (k & 31));
            }
            __syncthreads();                              // the pass's counts
            uint32_t s = 0, total;
            for (int b = rb0; b < rb1; b++) s += lcnt[b];
            uint32_t run = sort_block_scan<NT>(s, ws2, &total);
            for (int b = rb0; b < rb1; b++) {
                lofs[b] = run;
                run += lcnt[b];
            }
            __syncthreads();                              // stage offsets
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (c0[u] >= 0) {
                    const uint32_t q = lofs[c0[u]] + k0[u];
                    sitem[q] = item[u];
                    spos[q] = cur[c0[u]] + k0[u];
                }
                if (c1[u] >= 0) {
                    const uint32_t q = lofs[c1[u]] + k1[u];
                    sitem[q] = item[u];
                    spos[q] = cur[c1[u]] + k1[u];
                }
            }
            __syncthreads();                              // staged
            for (int b = rb0; b < rb1; b++) {
                cur[b] += lcnt[b];
                lcnt[b] = 0u;
            }
            for (uint32_t q = t; q < total; q += NT)
                if (spos[q] < a.items_cap) a.items[spos[q]] = sitem[q];
            __syncthreads();                              // counts cleared, stage free
        }
    }
    __syncthreads();
    for (int w = (int)t; w < a.occ_words; w += NT)
        a.occ[(size_t)tile * a.occ_words + w] = locc[w];
}


// Dynamic LDS of the staged scatter: cursors, occupancy, pass counts and
// offsets, then the staged items and their positions.
__host__ __device__ inline size_t bucket_staged_bytes(const BucketArgs& b)
{
    const size_t words = ((3 * (size_t)b.NB + b.occ_words + 3) & ~(size_t)3);
    return sizeof(uint32_t) * words + (sizeof(float4) + sizeof(uint32_t)) * kStageItems;
}

}  // namespace rtsim
