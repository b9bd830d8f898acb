// Stable sort of the receiver samples by travel (Receiver.cu:9-42): an LSD
// radix sort of (travel image, sample index) in four 8-bit passes over tiles
// of 1024 items (one workgroup of 256 lanes x 4 items per tile).  A compute
// that needs only the Binner segments and windows sorts by the Binner key
// instead (out-of-range samples after every bin): the same segments and
// windows, since the key is a non-decreasing function of travel, in one or
// two passes (samples of one bin stay in input order).  Per pass:
//
//   k_sort_hist    : per-tile digit counts -> counts[digit][tile]
//                    (pass 0 reads the records themselves)
//   k_sort_scan    : one workgroup per digit scans its row over the tiles
//                    (exclusive, in place) and writes the digit total
//   k_sort_scatter : each tile ranks its items stably per wave (ballots over
//                    the digit bits give each item's peers in its round; a
//                    per-wave LDS row carries the running digit counts, and
//                    since the LDS operations of one wave complete in order
//                    no block barrier is needed between rounds), stages the
//                    tile in LDS in digit order and writes every digit run to
//                    consecutive global positions.  The last pass also
//                    gathers the records to their sorted positions, writes
//                    their bearing-axis coordinates and bounds the Binner
//                    segments (the Binner keys and the permutation only when
//                    asked for).
// Kernel boundaries order the passes; no workgroup ever waits for another.
// Items of a tile are ordered (wave, round, lane) = ascending index, ranks
// count earlier rounds, lower lanes and earlier waves, tiles are scanned in
// order: every pass is stable, so the result is the stable order of the
// travel images.  The scan kernel also serves the grid form's bucket counts
// (one row per bucket instead of per digit).
#include "device_common.h"
#include "rtsim_internal.h"

namespace rtsim {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;                  // 4
constexpr int kSortItems = (int)kSortTileSmall / kSortThreads;  // 4 items per thread

// the sort key of a record: its travel image, or its Binner key with the
// out-of-range ones after every bin (k.n)
__device__ __forceinline__ uint32_t sort_key(float t, bool by_key, const KeyParams& k)
{
    if (!by_key) return travel_image(t);
    const uint32_t b = bin_key(k, t);
    return b == kOutOfRange ? k.n : b;
}

template <bool FIRST>
__device__ __forceinline__ void load_tile(const float4* __restrict__ rec,
                                          const uint2* __restrict__ kvin, uint32_t n,
                                          uint32_t wbase, uint32_t lane, uint32_t* key,
                                          uint32_t* val, bool by_key, const KeyParams& kp)
{
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const uint32_t i = min(wbase + j * 64 + lane, n - 1);
        if (FIRST) {
            key[j] = sort_key(rec[i].z, by_key, kp);
            if (val) val[j] = i;
        } else {
            const uint2 kv = kvin[i];
            key[j] = kv.x;
            if (val) val[j] = kv.y;
        }
    }
}

// ---- per-tile digit counts ------------------------------------------------------
template <bool FIRST>
__device__ __forceinline__ void sort_hist_body(const HistArgs& a, uint32_t blk, uint32_t nblk)
{
    __shared__ uint32_t h[256];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    h[t] = 0u;
    if (FIRST)
        for (size_t i = (size_t)blk * 256 + t; i < a.seg_count; i += (size_t)nblk * 256)
            a.seg_init[i] = make_uint2(0xffffffffu, 0u);
    if (a.n == 0 || blk >= a.tiles) return;
    const uint32_t wbase = blk * kSortTileSmall + w * (64 * kSortItems);
    uint32_t key[kSortItems];
    load_tile<FIRST>(a.rec, a.kvin, a.n, wbase, lane, key, nullptr, a.by_key != 0, a.key);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; j++)
        if (wbase + j * 64 + lane < a.n) atomicAdd(&h[(key[j] >> a.shift) & a.mask], 1u);
    __syncthreads();
    a.counts[(size_t)t * a.tiles + blk] = h[t];
}

template <bool FIRST>
__global__ __launch_bounds__(256) void k_sort_hist(HistArgs a)
{
    sort_hist_body<FIRST>(a, blockIdx.x, gridDim.x);
}

template <bool FIRST>
__global__ __launch_bounds__(256) void k_sort_hist_batch(const HistArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const HistArgs a = args[p];   // a private copy: the body's stores cannot alias it
    sort_hist_body<FIRST>(a, blk, nblk);
}

// ---- one workgroup per row: exclusive scan over tiles, row total -------------
constexpr int kScanPer = 64;   // tiles per thread: up to 16384 tiles per row

__device__ __forceinline__ void sort_scan_body(const ScanArgs& a, uint32_t rowi)
{
    __shared__ uint32_t ws[4];
    __shared__ uint32_t srow[256 * kScanPer];     // the row of tile counts
    const uint32_t t = threadIdx.x;
    const uint32_t tiles = a.tiles;
    uint32_t* row = a.counts + (size_t)rowi * tiles;
    // thread t scans tiles [t per, (t + 1) per): its sum, the block scan, its
    // prefixes.  Long rows go through LDS (coalesced in and out; strided
    // global reads of up to 64 words per thread measured 60 % slower), short
    // ones are read in place (the extra LDS round trip cost more)
    const uint32_t per = (tiles + 255) / 256;
    const bool staged = per > 8;
    uint32_t* src = staged ? srow : row;
    if (staged) {
        for (uint32_t i = t; i < tiles; i += 256) srow[i] = row[i];
        __syncthreads();
    }
    const uint32_t i0 = t * per, i1 = min(i0 + per, tiles);
    uint32_t s = 0;
    for (uint32_t i = i0; i < i1; i++) s += src[i];
    uint32_t total;
    uint32_t run = sort_block_scan(s, ws, &total);
    for (uint32_t i = i0; i < i1; i++) {
        const uint32_t v = src[i];
        src[i] = run;
        run += v;
    }
    if (staged) {
        __syncthreads();
        for (uint32_t i = t; i < tiles; i += 256) row[i] = srow[i];
    }
    if (t == 0) a.totals[rowi] = total;
    if (a.tmax && rowi == 0) {
        // the largest tile maximum (non-negative floats)
        __shared__ float wm[4];
        float m = 0.f;
        for (uint32_t i = t; i < tiles; i += 256) m = fmaxf(m, __uint_as_float(a.tmax[i]));
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        if ((t & 63) == 0) wm[t >> 6] = m;
        __syncthreads();
        if (t == 0) a.vmax[0] = __float_as_uint(fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3])));
    }
}

__global__ __launch_bounds__(256) void k_sort_scan(ScanArgs a) { sort_scan_body(a, blockIdx.x); }

__global__ __launch_bounds__(256) void k_sort_scan_batch(const ScanArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const ScanArgs a = args[p];
    sort_scan_body(a, blk);
}

// the scan kernels serve the grid form's bucket rows too (kernels_bucket.hip)
hipError_t launch_scan_rows(const ScanArgs& a, uint32_t rows, hipStream_t s)
{
    if (rows) hipLaunchKernelGGL(k_sort_scan, dim3(rows), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---- rank + stage + scatter ----------------------------------------------------
template <bool FIRST, bool LAST>
__device__ __forceinline__ void sort_scatter_body(const ScatterArgs& a, uint32_t tile)
{
    constexpr uint32_t TILE = kSortTileSmall;
    constexpr int LI = kSortItems;
    __shared__ uint32_t wcnt[kSortWaves][256];   // per-wave digit counts, then staging bases
    __shared__ uint32_t toff[256];               // digit run offsets in the staging tile
    __shared__ uint32_t gpos[256];               // global position of each digit run
    __shared__ uint32_t ws[4], ws2[4];
    __shared__ uint32_t skey[TILE];
    __shared__ uint32_t sval[TILE];
    // last pass: the Binner key of each staged item (read by its neighbours)
    __shared__ uint32_t scmp[LAST ? TILE : 1];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t tb = tile * TILE;
    const int shift = a.shift;
    const uint32_t mask = a.mask;

    // global run start of digit t: digit base (scan of the totals) + earlier tiles
    const uint32_t tot = a.totals[t];
    const uint32_t pre = a.counts[(size_t)t * a.tiles + tile];
#pragma unroll
    for (int k = 0; k < kSortWaves; k++) wcnt[k][t] = 0u;

    const uint32_t wbase = tb + w * (64 * LI);
    uint32_t key[LI], val[LI], lrank[LI];
    load_tile<FIRST>(a.rec, a.kvin, a.n, wbase, lane, key, val, a.by_key != 0, a.out.key);
    __syncthreads();

    // stable ranks inside each wave
#pragma unroll
    for (int j = 0; j < LI; j++) {
        const bool ok = wbase + j * 64 + lane < a.n;
        const uint32_t d = (key[j] >> shift) & mask;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint32_t bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t r = popc_below(peers);
        const uint32_t old = wcnt[w][d];
        lrank[j] = old + r;
        if (ok && r == 0) wcnt[w][d] = old + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // thread t = digit t: per-wave offsets, tile-local run offset, global run
    uint32_t c[kSortWaves], cnt = 0;
#pragma unroll
    for (int k = 0; k < kSortWaves; k++) {
        c[k] = wcnt[k][t];
        cnt += c[k];
    }
    const uint32_t o = sort_block_scan(cnt, ws, nullptr);
    const uint32_t base = sort_block_scan(tot, ws2, nullptr);
    toff[t] = o;
    gpos[t] = base + pre;
    {
        uint32_t run = o;
#pragma unroll
        for (int k = 0; k < kSortWaves; k++) {
            wcnt[k][t] = run;
            run += c[k];
        }
    }
    __syncthreads();

    // stage the tile in digit order
#pragma unroll
    for (int j = 0; j < LI; j++) {
        if (wbase + j * 64 + lane < a.n) {
            const uint32_t d = (key[j] >> shift) & mask;
            const uint32_t p = wcnt[w][d] + lrank[j];
            skey[p] = key[j];
            sval[p] = val[j];
        }
    }
    __syncthreads();

    // each digit run of the staged tile goes to consecutive global positions;
    // the last pass also gathers the records and writes their keys / coordinates
    const uint32_t tn = min(TILE, a.n - tb);
    uint32_t kq[LI], vq[LI], pq[LI];
#pragma unroll
    for (int u = 0; u < LI; u++) {
        const uint32_t q = t + u * kSortThreads;
        kq[u] = 0u;
        vq[u] = 0u;
        pq[u] = 0xffffffffu;
        if (q < tn) {
            kq[u] = skey[q];
            vq[u] = sval[q];
            const uint32_t d = (kq[u] >> shift) & mask;
            pq[u] = gpos[d] + (q - toff[d]);
        }
    }
    float4 rq[LAST ? LI : 1];
    uint32_t cq[LAST ? LI : 1];
    if (LAST) {
#pragma unroll
        for (int u = 0; u < LI; u++) rq[u] = a.rec[pq[u] < a.n ? vq[u] : 0u];
        // Binner key of every held item
#pragma unroll
        for (int u = 0; u < LI; u++) {
            cq[u] = bin_key(a.out.key, rq[u].z);
            if (pq[u] < a.n) scmp[t + u * kSortThreads] = cq[u];
        }
        // run bounds, before the stores so that the atomics' round trips
        // overlap them: the staged tile holds one run per digit (its items
        // land at consecutive positions); inside a run, an item whose key
        // differs from its predecessor's starts a key run, one whose key
        // differs from its successor's ends one.  The keys are non-decreasing
        // over the sorted order, so the smallest start and the largest end
        // over all workgroups bound each segment.  Thread t examines the
        // staged items q = LI t .. LI t + LI - 1: consecutive, so their
        // neighbours are its own but at both ends; a digit run starts where
        // the digit changes (the runs are in digit order).
        __syncthreads();
        uint32_t* seg = reinterpret_cast<uint32_t*>(a.out.seg);
        const uint32_t q0 = t * (uint32_t)LI;
        if (q0 < tn) {
            uint32_t kk[LI + 2], cc[LI + 2];   // items q0 - 1 .. q0 + LI
            {
                const uint4 k4 = *reinterpret_cast<const uint4*>(&skey[q0]);
                const uint4 c4 = *reinterpret_cast<const uint4*>(&scmp[q0]);
                kk[1] = k4.x; kk[2] = k4.y; kk[3] = k4.z; kk[4] = k4.w;
                cc[1] = c4.x; cc[2] = c4.y; cc[3] = c4.z; cc[4] = c4.w;
            }
            kk[0] = q0 > 0 ? skey[q0 - 1] : 0u;
            cc[0] = q0 > 0 ? scmp[q0 - 1] : 0u;
            kk[LI + 1] = q0 + LI < tn ? skey[q0 + LI] : 0u;
            cc[LI + 1] = q0 + LI < tn ? scmp[q0 + LI] : 0u;
#pragma unroll
            for (int j = 0; j < LI; j++) {
                const uint32_t q = q0 + j;
                if (q >= tn) break;
                const uint32_t d = (kk[j + 1] >> shift) & mask;
                const bool first = q == 0 || ((kk[j] >> shift) & mask) != d;
                const bool last = q + 1 == tn || ((kk[j + 2] >> shift) & mask) != d;
                const uint32_t k = cc[j + 1];
                const bool kstart = k != kOutOfRange && (first || cc[j] != k);
                const bool kend = k != kOutOfRange && (last || cc[j + 2] != k);
                if (kstart || kend) {
                    const uint32_t pos = gpos[d] + (q - toff[d]);
                    if (kstart) atomicMin(&seg[2 * k], pos);
                    if (kend) atomicMax(&seg[2 * k + 1], pos + 1);
                }
            }
        }
    }
#pragma unroll
    for (int u = 0; u < LI; u++) {
        const uint32_t pos = pq[u];
        if (pos < a.n) {   // holds for every staged item
            if (!LAST) {
                a.kvout[pos] = make_uint2(kq[u], vq[u]);
            } else {
                const float4 r = rq[u];
                a.out.sorted[pos] = r;
                // the permutation and the Binner keys only when asked for
                // (nothing downstream reads them)
                if (a.order) a.order[pos] = vq[u];
                if (a.out.skeys) a.out.skeys[pos] = cq[u];
                const double x = a.out.axis == 1 ? sin_bearing((double)r.w) : (double)r.w;
                const float xh = (float)x;
                a.out.coord[pos] = make_float2(xh, (float)(x - (double)xh));
            }
        }
    }
}

template <bool FIRST, bool LAST>
__global__ __launch_bounds__(256) void k_sort_scatter(ScatterArgs a)
{
    sort_scatter_body<FIRST, LAST>(a, xcd_contiguous_tile(blockIdx.x, gridDim.x));
}

template <bool FIRST, bool LAST>
__global__ __launch_bounds__(256) void k_sort_scatter_batch(const ScatterArgs* __restrict__ args,
                                                            BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    sort_scatter_body<FIRST, LAST>(args[p], xcd_contiguous_tile(blk, nblk));
}

hipError_t make_sort_plan(const float4* records, uint32_t n, const SortBuffers& b, SortPlan* p)
{
    // by Binner key: keys 0 .. N (N: out of range) in 8-bit digits
    const bool by_key = b.by_key && b.out.key.n < 65535u;
    const int passes = !by_key ? 4 : b.out.key.n < 255u ? 1 : 2;
    p->passes = passes;
    const uint32_t tiles = (n + kSortTileSmall - 1) / kSortTileSmall;
    if (tiles > b.tiles_cap || tiles > 256u * kScanPer) return hipErrorInvalidValue;
    int cur = 0;   // buffer pair holding the current order (pass 0 reads the records)
    for (int pass = 0; pass < passes; pass++) {
        p->tile[pass] = kSortTileSmall;
        p->hist_blocks[pass] = n ? tiles : (pass == 0 && b.seg_count ? 1u : 0u);
        p->pass_blocks[pass] = n ? tiles : 0u;
        HistArgs& h = p->hist[pass];
        h.rec = records;
        h.kvin = b.kv[cur];
        h.n = n;
        h.shift = 8 * pass;
        h.mask = 0xffu;
        h.counts = b.counts;
        h.tiles = tiles;
        h.seg_init = pass == 0 ? b.seg_init : nullptr;
        h.seg_count = pass == 0 ? b.seg_count : 0;
        h.by_key = by_key;
        h.key = b.out.key;
        p->scan[pass] = ScanArgs{b.counts, tiles, b.totals, nullptr, nullptr};
        ScatterArgs& c = p->scatter[pass];
        c.rec = records;
        c.kvin = b.kv[cur];
        c.kvout = b.kv[cur ^ 1];
        c.order = pass == passes - 1 ? b.order : nullptr;
        c.n = n;
        c.shift = 8 * pass;
        c.mask = 0xffu;
        c.counts = b.counts;
        c.totals = b.totals;
        c.tiles = tiles;
        c.out = b.out;
        c.by_key = by_key;
        cur ^= 1;
    }
    p->order = b.order;
    return hipSuccess;
}

template <bool FIRST>
static void launch_scatter(bool last, dim3 g, const ScatterArgs& a, hipStream_t s)
{
    if (last) hipLaunchKernelGGL((k_sort_scatter<FIRST, true>), g, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_sort_scatter<FIRST, false>), g, dim3(256), 0, s, a);
}

hipError_t launch_sort_plan(const SortPlan& p, hipStream_t s)
{
    for (int pass = 0; pass < p.passes; pass++) {
        const uint32_t hb = p.hist_blocks[pass];
        if (hb) {
            if (pass == 0) hipLaunchKernelGGL(k_sort_hist<true>, dim3(hb), dim3(256), 0, s, p.hist[pass]);
            else hipLaunchKernelGGL(k_sort_hist<false>, dim3(hb), dim3(256), 0, s, p.hist[pass]);
        }
        if (!p.pass_blocks[pass]) continue;
        hipLaunchKernelGGL(k_sort_scan, dim3(p.hist[pass].mask + 1), dim3(256), 0, s, p.scan[pass]);
        const dim3 g(p.pass_blocks[pass]);
        const bool last = pass == p.passes - 1;
        if (pass == 0) launch_scatter<true>(last, g, p.scatter[pass], s);
        else launch_scatter<false>(last, g, p.scatter[pass], s);
    }
    return hipGetLastError();
}

hipError_t launch_sort_hist_batch(bool first_pass, const HistArgs* args, BatchMap m, uint32_t blocks,
                                  hipStream_t s)
{
    if (!blocks) return hipSuccess;
    if (first_pass) hipLaunchKernelGGL(k_sort_hist_batch<true>, dim3(blocks), dim3(256), 0, s, args, m);
    else hipLaunchKernelGGL(k_sort_hist_batch<false>, dim3(blocks), dim3(256), 0, s, args, m);
    return hipGetLastError();
}

hipError_t launch_sort_scan_batch(const ScanArgs* args, BatchMap m, uint32_t blocks, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_sort_scan_batch, dim3(blocks), dim3(256), 0, s, args, m);
    return hipGetLastError();
}

hipError_t launch_sort_scatter_batch(bool first_pass, bool last_pass, const ScatterArgs* args,
                                     BatchMap m, uint32_t blocks, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    const dim3 g(blocks);
    if (first_pass) {
        if (last_pass) hipLaunchKernelGGL((k_sort_scatter_batch<true, true>), g, dim3(256), 0, s, args, m);
        else hipLaunchKernelGGL((k_sort_scatter_batch<true, false>), g, dim3(256), 0, s, args, m);
    } else {
        if (last_pass) hipLaunchKernelGGL((k_sort_scatter_batch<false, true>), g, dim3(256), 0, s, args, m);
        else hipLaunchKernelGGL((k_sort_scatter_batch<false, false>), g, dim3(256), 0, s, args, m);
    }
    return hipGetLastError();
}

}  // namespace rtsim
