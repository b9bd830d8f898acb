// Binning of the grid (MFMA) form: one counting pass that groups the samples
// by blocks of R grid rows (BucketArgs, rtsim_internal.h).
//
//   bucket counts : per tile, the items of every bucket (a sample with a
//                   Binner key goes to the buckets of its two grid rows) and
//                   the largest finite |re|, |im| of the tile's binned
//                   samples; made by the streaming kernel itself when it
//                   writes every sample of a ping (kernels_bin.hip), else by
//                   k_bucket_hist from the records
//   bucket scan   : k_sort_scan over the rows [bucket][tile] (exclusive, the
//                   bucket totals), its first workgroup also reduces the tile
//                   maxima to the deposit's fixed-point scale
//   bucket scatter: k_bucket_scatter writes every binned sample as an item
//                   {re, im, travel, grid column} into its buckets' ranges
//                   (LDS cursors per bucket) and the tile's Binner-key
//                   occupancy bits
// Items of one bucket land in an order that depends on the LDS atomics; the
// deposit sums them in exact fixed point, so the grid does not depend on it.
// The Binner keys of the reference (Binner.cu:24-49) decide which samples are
// binned (bin_key, the same function the sort uses) and the occupancy bits
// give the windows' emptiness (Binner.cu:82-106) without a sort.
#include "device_common.h"
#include "rtsim_internal.h"

namespace rtsim {

constexpr int kBucketThreads = 256;
constexpr int kBucketU = 4;   // samples per thread with their loads in flight together

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

// k_bucket_hist: tile `tile` of the records (blockDim = kBucketThreads).
__device__ __forceinline__ void bucket_hist_body(const BucketArgs& a, uint32_t tile)
{
    extern __shared__ uint32_t hist[];   // [NB]
    __shared__ float wmax[kBucketThreads / 64];
    const uint32_t t = threadIdx.x;
    for (int b = (int)t; b < a.NB; b += kBucketThreads) hist[b] = 0u;
    __syncthreads();
    float m = 0.f;
    const uint32_t i0 = tile * a.tile, i1 = min(i0 + a.tile, a.n);
    for (uint32_t ib = i0; ib < i1; ib += kBucketU * kBucketThreads) {
        float4 r[kBucketU];
#pragma unroll
        for (int u = 0; u < kBucketU; u++) r[u] = a.rec[min(ib + u * kBucketThreads + t, i1 - 1)];
#pragma unroll
        for (int u = 0; u < kBucketU; u++)
            if (ib + u * kBucketThreads + t < i1) bucket_count(a, r[u], hist, m);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((t & 63) == 0) wmax[t >> 6] = m;
    __syncthreads();
    for (int b = (int)t; b < a.NB; b += kBucketThreads) a.counts[(size_t)b * a.tiles + tile] = hist[b];
    if (t == 0) {
        float mm = wmax[0];
        for (int k = 1; k < kBucketThreads / 64; k++) mm = fmaxf(mm, wmax[k]);
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
constexpr int kScatterPass = kBucketU * kBucketThreads;   // samples per pass
constexpr int kStageItems = 2 * kScatterPass;              // a sample makes at most two items

__device__ __forceinline__ void bucket_cursors(const BucketArgs& a, uint32_t tile, uint32_t* cur, uint32_t* ws)
{
    const uint32_t t = threadIdx.x;
    const int per = (a.NB + kBucketThreads - 1) / kBucketThreads;
    const int b0 = (int)t * per, b1 = min(b0 + per, a.NB);
    uint32_t s = 0;
    for (int b = b0; b < b1; b++) s += a.totals[b];
    uint32_t run = sort_block_scan(s, ws, nullptr);
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

template <bool kStaged>
__device__ __forceinline__ void bucket_scatter_body(const BucketArgs& a, uint32_t tile)
{
    extern __shared__ uint32_t lds[];
    uint32_t* cur = lds;                                  // [NB] global cursor
    uint32_t* locc = lds + a.NB;                          // [occ_words]
    uint32_t* lcnt = locc + a.occ_words;                  // staged: [NB] items of the pass
    uint32_t* lofs = lcnt + a.NB;                         // staged: [NB] their offset in the stage
    float4* sitem = reinterpret_cast<float4*>(lds + ((2 * a.NB + a.occ_words + a.NB + 3) & ~3));
    uint32_t* spos = reinterpret_cast<uint32_t*>(sitem + kStageItems);
    __shared__ uint32_t ws[4], ws2[4];
    const uint32_t t = threadIdx.x;
    bucket_cursors(a, tile, cur, ws);
    for (int w = (int)t; w < a.occ_words; w += kBucketThreads) locc[w] = 0u;
    if (kStaged)
        for (int b = (int)t; b < a.NB; b += kBucketThreads) lcnt[b] = 0u;
    __syncthreads();
    const int per = (a.NB + kBucketThreads - 1) / kBucketThreads;
    const int rb0 = (int)t * per, rb1 = min(rb0 + per, a.NB);   // this thread's buckets
    const uint32_t i0 = tile * a.tile, i1 = min(i0 + a.tile, a.n);
    for (uint32_t ib = i0; ib < i1; ib += kScatterPass) {
        float4 r[kBucketU];
#pragma unroll
        for (int u = 0; u < kBucketU; u++) r[u] = a.rec[min(ib + u * kBucketThreads + t, i1 - 1)];
        if constexpr (!kStaged) {
#pragma unroll
            for (int u = 0; u < kBucketU; u++) {
                int c0, c1;
                float4 item;
                uint32_t k;
                bucket_item(a, r[u], ib + u * kBucketThreads + t < i1, c0, c1, item, k);
                if (k == kOutOfRange) continue;
                if (c0 >= 0) a.items[atomicAdd(&cur[c0], 1u)] = item;
                if (c1 >= 0) a.items[atomicAdd(&cur[c1], 1u)] = item;
                atomicOr(&locc[k >> 5], 1u << (k & 31));
            }
        } else {
            int c0[kBucketU], c1[kBucketU];
            uint32_t k0[kBucketU], k1[kBucketU];
            float4 item[kBucketU];
#pragma unroll
            for (int u = 0; u < kBucketU; u++) {
                uint32_t k;
                bucket_item(a, r[u], ib + u * kBucketThreads + t < i1, c0[u], c1[u], item[u], k);
                k0[u] = c0[u] >= 0 ? atomicAdd(&lcnt[c0[u]], 1u) : 0u;
                k1[u] = c1[u] >= 0 ? atomicAdd(&lcnt[c1[u]], 1u) : 0u;
                if (k != kOutOfRange) atomicOr(&locc[k >> 5], 1u << (k & 31));
            }
            __syncthreads();                              // the pass's counts
            uint32_t s = 0, total;
            for (int b = rb0; b < rb1; b++) s += lcnt[b];
            uint32_t run = sort_block_scan(s, ws2, &total);
            for (int b = rb0; b < rb1; b++) {
                lofs[b] = run;
                run += lcnt[b];
            }
            __syncthreads();                              // stage offsets
#pragma unroll
            for (int u = 0; u < kBucketU; u++) {
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
            for (uint32_t q = t; q < total; q += kBucketThreads) a.items[spos[q]] = sitem[q];
            __syncthreads();                              // counts cleared, stage free
        }
    }
    __syncthreads();
    for (int w = (int)t; w < a.occ_words; w += kBucketThreads)
        a.occ[(size_t)tile * a.occ_words + w] = locc[w];
}

__global__ __launch_bounds__(kBucketThreads) void k_bucket_hist(BucketArgs a) { bucket_hist_body(a, blockIdx.x); }

template <bool kStaged>
__global__ __launch_bounds__(kBucketThreads) void k_bucket_scatter(BucketArgs a)
{
    bucket_scatter_body<kStaged>(a, xcd_contiguous_tile(blockIdx.x, gridDim.x));
}

__global__ __launch_bounds__(kBucketThreads) void k_bucket_hist_batch(const BucketArgs* __restrict__ args,
                                                                      BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const BucketArgs a = args[p];   // a private copy: the body's stores cannot alias it
    bucket_hist_body(a, blk);
}

template <bool kStaged>
__global__ __launch_bounds__(kBucketThreads) void k_bucket_scatter_batch(
    const BucketArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const BucketArgs a = args[p];
    bucket_scatter_body<kStaged>(a, xcd_contiguous_tile(blk, nblk));
}

namespace {
int g_scatter_lds_limit = 64 * 1024;   // dynamic LDS a scatter workgroup may take (set at init)
}

static size_t staged_bytes(const BucketArgs& b)
{
    const size_t words = ((3 * (size_t)b.NB + b.occ_words + 3) & ~(size_t)3);
    return sizeof(uint32_t) * words + (sizeof(float4) + sizeof(uint32_t)) * kStageItems;
}

// Staging pays when a pass spreads its items thinly over many buckets (runs
// of a few items); with a few hundred buckets the direct stores already
// fill their lines in L2 (measured on one ping: 1M samples, 264 buckets:
// 12.6 us direct, ~16 us staged; 16M samples, 1386 buckets: 309 us direct,
// ~255 us staged; the 64-ping batch with up to 724 buckets was faster direct).
constexpr int kStagedMinBuckets = 1024;

bool bucket_scatter_staged(const BucketArgs& b)
{
    return b.NB >= kStagedMinBuckets && staged_bytes(b) <= (size_t)g_scatter_lds_limit;
}

size_t bucket_lds_bytes(const BucketArgs& b)
{
    return bucket_scatter_staged(b) ? staged_bytes(b) : sizeof(uint32_t) * ((size_t)b.NB + b.occ_words);
}

void init_bucket_attributes(int max_lds)
{
    g_scatter_lds_limit = max_lds - 64;   // the kernels' static LDS
    set_dynamic_lds_limit((const void*)k_bucket_scatter<true>, max_lds);
    set_dynamic_lds_limit((const void*)k_bucket_scatter_batch<true>, max_lds);
}

hipError_t launch_bucket_hist(const BucketArgs& b, hipStream_t s)
{
    if (b.tiles)
        hipLaunchKernelGGL(k_bucket_hist, dim3(b.tiles), dim3(kBucketThreads), sizeof(uint32_t) * b.NB, s, b);
    return hipGetLastError();
}

hipError_t launch_bucket_scan(const BucketArgs& b, hipStream_t s)
{
    return launch_scan_rows(ScanArgs{b.counts, b.tiles, b.totals, b.tmax, b.vmax}, (uint32_t)b.NB, s);
}

hipError_t launch_bucket_scatter(const BucketArgs& b, hipStream_t s)
{
    if (!b.tiles) return hipGetLastError();
    if (bucket_scatter_staged(b))
        hipLaunchKernelGGL(k_bucket_scatter<true>, dim3(b.tiles), dim3(kBucketThreads), bucket_lds_bytes(b), s, b);
    else
        hipLaunchKernelGGL(k_bucket_scatter<false>, dim3(b.tiles), dim3(kBucketThreads), bucket_lds_bytes(b), s, b);
    return hipGetLastError();
}

hipError_t launch_bucket_hist_batch(const BucketArgs* args, BatchMap m, uint32_t blocks, size_t shm,
                                    hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_bucket_hist_batch, dim3(blocks), dim3(kBucketThreads), shm, s, args, m);
    return hipGetLastError();
}

hipError_t launch_bucket_scatter_batch(const BucketArgs* args, BatchMap m, uint32_t blocks, size_t shm,
                                       bool staged, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    if (staged)
        hipLaunchKernelGGL(k_bucket_scatter_batch<true>, dim3(blocks), dim3(kBucketThreads), shm, s, args, m);
    else
        hipLaunchKernelGGL(k_bucket_scatter_batch<false>, dim3(blocks), dim3(kBucketThreads), shm, s, args, m);
    return hipGetLastError();
}

}  // namespace rtsim
