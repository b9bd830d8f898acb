// Stable sort of the receiver samples by travel (Receiver.cu:9-42): an LSD
// radix sort of (travel image, sample index) in four 8-bit passes over tiles
// of 2048 items (one workgroup of 256 lanes x 8 items per tile).  Per pass:
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
//                    consecutive global positions.  The last pass also moves
//                    the records to their sorted positions and writes their
//                    bearing-axis coordinates (the Binner keys and the
//                    permutation only when asked for).
// Kernel boundaries order the passes; no workgroup ever waits for another.
// Items of a tile are ordered (wave, round, lane) = ascending index, ranks
// count earlier rounds, lower lanes and earlier waves, tiles are scanned in
// order: every pass is stable, so the result is the stable order of the
// travel images.  The grid (MFMA) form sorts by the grouping key K of
// GroupKey instead (G row + bin slot, below 2^16 for grids up to ~32k range
// rows), which takes two passes: same keys, segments, windows and G-row
// table, order permuted only inside runs of equal (bin, G row).
#include "device_common.h"
#include "rtsim_internal.h"

namespace rtsim {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;                  // 4
// grouping key: split its bits evenly over the passes (else 8-bit digits)
constexpr bool kGroupEvenDigits = true;

// Inclusive wave scan (wave64) of a 32-bit value.
__device__ __forceinline__ uint32_t sort_wave_scan(uint32_t x, uint32_t lane)
{
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    return x;
}

// Exclusive block scan (256 threads) of one value per thread; `ws` holds 4 words.
__device__ __forceinline__ uint32_t sort_block_scan(uint32_t x, uint32_t* ws, uint32_t* total)
{
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t incl = sort_wave_scan(x, lane);
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t k = 0; k < w; k++) off += ws[k];
    if (total) *total = ws[0] + ws[1] + ws[2] + ws[3];
    return off + incl - x;
}

template <bool FIRST, int LI>
__device__ __forceinline__ void load_tile(const float4* __restrict__ rec,
                                          const uint2* __restrict__ kvin, const GroupKey& gk,
                                          uint32_t n, uint32_t wbase, uint32_t lane, uint32_t* key,
                                          uint32_t* val)
{
#pragma unroll
    for (int j = 0; j < LI; j++) {
        const uint32_t i = min(wbase + j * 64 + lane, n - 1);
        if (FIRST) {
            const float t = rec[i].z;
            key[j] = gk.on ? group_key(gk, t) : travel_image(t);
            if (val) val[j] = i;
        } else {
            const uint2 kv = kvin[i];
            key[j] = kv.x;
            if (val) val[j] = kv.y;
        }
    }
}

// ---- per-tile digit counts (tiles of TILE items) ------------------------------
template <bool FIRST, uint32_t TILE>
__device__ __forceinline__ void sort_hist_body(const HistArgs& a, uint32_t blk, uint32_t nblk)
{
    constexpr int LI = (int)TILE / kSortThreads;   // items per thread
    __shared__ uint32_t h[256];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    h[t] = 0u;
    if (FIRST) {
        for (size_t i = (size_t)blk * 256 + t; i < a.seg_count; i += (size_t)nblk * 256)
            a.seg_init[i] = make_uint2(0xffffffffu, 0u);
        for (size_t i = (size_t)blk * 256 + t; i < a.fill_count; i += (size_t)nblk * 256)
            a.fill[i] = a.fill_value;
        if (a.gk.on && blk == 0 && t == 0) a.gk.vmax[0] = 0u;   // the first scan sets it (n > 0)
    }
    if (a.n == 0 || blk >= a.tiles) return;
    const uint32_t wbase = blk * TILE + w * (64 * LI);
    uint32_t key[LI];
    __shared__ float wmax[kSortWaves];
    if (FIRST && a.gk.on) {
        // grouping keys, and the largest finite |re|, |im| of the in-range
        // samples (the record's components share the travel's cache line)
        float m = 0.f;
#pragma unroll
        for (int j = 0; j < LI; j++) {
            const float4 r = a.rec[min(wbase + j * 64 + lane, a.n - 1)];
            const uint32_t slot = bin_slot(a.gk.key, r.z);
            key[j] = (uint32_t)(g_row_of(a.gk.rows, r.z) + 2) + slot;
            const float c = fmaxf(fabsf(r.x), fabsf(r.y));
            if (wbase + j * 64 + lane < a.n && slot - 1u < a.gk.key.n && c <= 3.402823466e38f)
                m = fmaxf(m, c);
        }
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        if (lane == 0) wmax[w] = m;
    } else {
        load_tile<FIRST, LI>(a.rec, a.kvin, a.gk, a.n, wbase, lane, key, nullptr);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < LI; j++)
        if (wbase + j * 64 + lane < a.n) atomicAdd(&h[(key[j] >> a.shift) & a.mask], 1u);
    __syncthreads();
    if (t <= a.mask) a.counts[(size_t)t * a.tiles + blk] = h[t];   // digits beyond the mask: none
    if (FIRST && a.gk.on && t == 0)
        a.gk.tmax[blk] = __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3])));
}

template <bool FIRST, uint32_t TILE>
__global__ __launch_bounds__(256) void k_sort_hist(HistArgs a)
{
    sort_hist_body<FIRST, TILE>(a, blockIdx.x, gridDim.x);
}

template <bool FIRST, uint32_t TILE>
__global__ __launch_bounds__(256) void k_sort_hist_batch(const HistArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const HistArgs a = args[p];   // a private copy: the body's stores cannot alias it
    sort_hist_body<FIRST, TILE>(a, blk, nblk);
}

// ---- one workgroup per digit: exclusive scan over tiles, digit total ---------
constexpr int kScanPer = 64;   // tiles per thread: up to 16384 tiles (16.7M items of small tiles)

__device__ __forceinline__ void sort_scan_body(const ScanArgs& a, uint32_t digit)
{
    __shared__ uint32_t ws[4];
    __shared__ uint32_t srow[256 * kScanPer];     // the digit's row of tile counts
    const uint32_t t = threadIdx.x;
    const uint32_t tiles = a.tiles;
    uint32_t* row = a.counts + (size_t)digit * tiles;
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
    if (t == 0) a.totals[digit] = total;
    if (a.tmax && digit == 0) {
        // deposit scale: the largest tile maximum (non-negative floats)
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

// Binner key from the value the last pass compares: a bin slot (grouping key)
// or the key itself
__device__ __forceinline__ uint32_t binner_key(bool grp, uint32_t c, uint32_t n)
{
    return grp ? (c - 1u < n ? c - 1u : kOutOfRange) : c;
}

// ---- rank + stage + scatter ----------------------------------------------------
// MOVE: the records travel with their keys through every pass (staged in LDS,
// written in digit runs), so the last pass reads them in order; otherwise the
// last pass gathers them by sample index.
template <bool FIRST, bool LAST, bool MOVE, bool BATCHED>
__device__ __forceinline__ void sort_scatter_body(const ScatterArgs& a, uint32_t tile)
{
    constexpr uint32_t TILE = sort_pass_tile(LAST, MOVE, BATCHED);
    constexpr int LI = (int)TILE / kSortThreads;   // items per thread
    __shared__ uint32_t wcnt[kSortWaves][256];   // per-wave digit counts, then staging bases
    __shared__ uint32_t toff[256];               // digit run offsets in the staging tile
    __shared__ uint32_t gpos[256];               // global position of each digit run
    __shared__ uint32_t ws[4], ws2[4];
    __shared__ uint32_t skey[TILE];
    __shared__ uint32_t sval[TILE];
    // Records that move: the first pass stages them in digit order from
    // registers, a middle pass reads them at the staging step, and the last
    // pass copies them into LDS in input order at its start (their loads
    // overlap the key loads; held in registers across the ranking they would
    // spill) and takes each staged item's record by its input slot.
    constexpr bool RIN = MOVE && LAST && !FIRST;
    __shared__ float4 srec[MOVE ? TILE : 1];
    // last pass: the input slot of each staged item (RIN), then the value its
    // neighbours compare (bin slot or Binner key); each entry is read and
    // rewritten by the thread that owns it
    __shared__ uint32_t scmp[LAST ? TILE : 1];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t tb = tile * TILE;
    const int shift = a.shift;
    const uint32_t mask = a.mask;

    // global run start of digit t: digit base (scan of the totals) + earlier tiles
    const uint32_t tot = t <= mask ? a.totals[t] : 0u;
    const uint32_t pre = t <= mask ? a.counts[(size_t)t * a.tiles + tile] : 0u;
#pragma unroll
    for (int k = 0; k < kSortWaves; k++) wcnt[k][t] = 0u;

    const uint32_t wbase = tb + w * (64 * LI);
    uint32_t key[LI], val[LI], lrank[LI];
    float4 rv[MOVE ? LI : 1];
    if (MOVE && FIRST) {
#pragma unroll
        for (int j = 0; j < LI; j++) {
            const uint32_t i = min(wbase + j * 64 + lane, a.n - 1);
            rv[j] = a.rec[i];
            key[j] = a.gk.on ? group_key(a.gk, rv[j].z) : travel_image(rv[j].z);
            val[j] = i;
        }
    } else {
        load_tile<FIRST, LI>(a.rec, a.kvin, a.gk, a.n, wbase, lane, key, val);
        if (RIN) {
#pragma unroll
            for (int j = 0; j < LI; j++)
                srec[w * (64 * LI) + j * 64 + lane] = a.rin[min(wbase + j * 64 + lane, a.n - 1)];
        }
    }
    __syncthreads();

    // stable ranks inside each wave
#pragma unroll
    for (int j = 0; j < LI; j++) {
        const bool ok = wbase + j * 64 + lane < a.n;
        const uint32_t d = (key[j] >> shift) & mask;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {   // bits above dbits are zero: no-op ballots
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
            if (RIN) scmp[p] = w * (64 * LI) + j * 64 + lane;
            else if (MOVE) srec[p] = FIRST ? rv[j] : a.rin[wbase + j * 64 + lane];
        }
    }
    __syncthreads();

    // each digit run of the staged tile goes to consecutive global positions;
    // the last pass also moves the records and writes their keys / coordinates
    const uint32_t tn = min(TILE, a.n - tb);
    constexpr int Q = LI;
    uint32_t kq[Q], vq[Q], pq[Q];
#pragma unroll
    for (int u = 0; u < Q; u++) {
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
    // the last pass holds its records (keys, bounds and stores use them); an
    // earlier pass reads each one from LDS at its store
    float4 rq[LAST ? Q : 1];
    uint32_t cq[LAST ? Q : 1];
    if (LAST && MOVE) {
#pragma unroll
        for (int u = 0; u < Q; u++) {
            const uint32_t q = t + u * kSortThreads;
            rq[u] = srec[RIN ? (pq[u] < a.n ? scmp[q] : 0u) : min(q, TILE - 1)];
        }
    } else if (LAST) {
#pragma unroll
        for (int u = 0; u < Q; u++) rq[u] = a.rec[pq[u] < a.n ? vq[u] : 0u];
    }
    if (LAST) {
        // Binner key of every held item.  A grouping key K = (G row + 2) +
        // bin slot decodes with the G row of the travel (the same function
        // of the same float as in the first pass); the travel sort evaluates
        // the key.  cq is what neighbours compare: the bin slot (grouping)
        // or the Binner key.
        const bool grp = a.gk.on;
#pragma unroll
        for (int u = 0; u < Q; u++) {
            cq[u] = grp ? kq[u] - (uint32_t)(g_row_of(a.gk.rows, rq[u].z) + 2)
                        : bin_key(a.out.key, rq[u].z);
            if (pq[u] < a.n) scmp[t + u * kSortThreads] = cq[u];
        }
        // run bounds, before the stores so that the atomics' round trips
        // overlap them: the staged tile holds one run per digit (its items
        // land at consecutive positions); inside a run, an item whose bin (G
        // row) differs from its predecessor's starts a key (row) run, one
        // whose bin differs from its successor's ends one.  The Binner keys
        // are non-decreasing over the sorted order, so the smallest start and
        // the largest end over all workgroups bound each segment.  Thread t
        // examines the staged items q = LI t .. LI t + LI - 1: consecutive,
        // so their neighbours are its own but at both ends; a digit run
        // starts where the digit changes (the runs are in digit order).  The
        // G row of a staged item is its K minus (its slot + 2).
        uint32_t* rows = grp ? a.gk.rows.row_start : nullptr;
        __syncthreads();
        uint32_t* seg = reinterpret_cast<uint32_t*>(a.out.seg);
        const bool marks = a.out.seg_marks != 0;
        const uint32_t q0 = t * (uint32_t)LI;
        if (q0 < tn) {
            uint32_t kk[LI + 2], cc[LI + 2];   // items q0 - 1 .. q0 + LI
#pragma unroll
            for (int j = 0; j < LI; j += 4) {
                const uint4 k4 = *reinterpret_cast<const uint4*>(&skey[q0 + j]);
                const uint4 c4 = *reinterpret_cast<const uint4*>(&scmp[q0 + j]);
                kk[j + 1] = k4.x; kk[j + 2] = k4.y; kk[j + 3] = k4.z; kk[j + 4] = k4.w;
                cc[j + 1] = c4.x; cc[j + 2] = c4.y; cc[j + 3] = c4.z; cc[j + 4] = c4.w;
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
                const uint32_t c = cc[j + 1];
                const uint32_t k = binner_key(grp, c, a.out.key.n);
                const bool kstart = k != kOutOfRange && (first || cc[j] != c);
                const bool kend = k != kOutOfRange && (last || cc[j + 2] != c);
                const int g = (int)(kk[j + 1] - c) - 2;
                const bool gstart = rows && g >= -1 && g <= a.gk.rows.GH + 1 &&
                                    (first || (int)(kk[j] - cc[j]) - 2 != g);
                if (marks ? kstart || gstart : kstart || kend || gstart) {
                    const uint32_t pos = gpos[d] + (q - toff[d]);
                    if (marks) {
                        // every tile holding key k stores the same mark
                        if (kstart) reinterpret_cast<uint2*>(seg)[k] = make_uint2(0u, 1u);
                    } else {
                        if (kstart) atomicMin(&seg[2 * k], pos);
                        if (kend) atomicMax(&seg[2 * k + 1], pos + 1);
                    }
                    if (gstart) atomicMin(&rows[g + 1], pos);
                }
            }
        }
    }
#pragma unroll
    for (int u = 0; u < Q; u++) {
        const uint32_t pos = pq[u];
        if (pos < a.n) {   // holds for every staged item unless the records changed
            if (!LAST) a.kvout[pos] = make_uint2(kq[u], vq[u]);
            if (MOVE && !LAST) a.rout[pos] = srec[t + u * kSortThreads];
            if (LAST) {
                const float4 r = rq[u];
                a.out.sorted[pos] = r;
                // the permutation and the Binner keys only when asked for
                // (nothing downstream reads them)
                if (a.order) a.order[pos] = vq[u];
                if (a.out.skeys) a.out.skeys[pos] = binner_key(a.gk.on, cq[u], a.out.key.n);
                const double x = a.out.axis == 1 ? sin_bearing((double)r.w) : (double)r.w;
                if (a.out.xg) {
                    a.out.xg[pos] = (float)((x - a.out.xg_b0) * (double)a.out.xg_inv) + a.out.xg_off;
                } else {
                    const float xh = (float)x;
                    a.out.coord[pos] = make_float2(xh, (float)(x - (double)xh));
                }
            }
        }
    }
}

template <bool FIRST, bool LAST, bool MOVE>
__global__ __launch_bounds__(256) void k_sort_scatter(ScatterArgs a)
{
    sort_scatter_body<FIRST, LAST, MOVE, false>(a, blockIdx.x);
}

template <bool FIRST, bool LAST, bool MOVE>
__global__ __launch_bounds__(256) void k_sort_scatter_batch(const ScatterArgs* __restrict__ args,
                                                            BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    sort_scatter_body<FIRST, LAST, MOVE, true>(args[p], blk);
}

hipError_t make_sort_plan(const float4* records, uint32_t n, const SortBuffers& b,
                          const GroupKey* gk, SortPlan* p, bool batched)
{
    GroupKey g{};
    if (gk && gk->on) g = *gk;
    // 8-bit digits: 4 passes for travel images; ceil(bits(K) / 8) passes for a
    // grouping key, its bits split evenly over them (a tile's digit runs are
    // then long in every pass: 2048 items over at most 2^dbits digits)
    int passes = 4, dbits = 8;
    if (g.on) {
        int bits = 1;
        while (bits < 32 && (group_key_max(g) >> bits)) bits++;
        passes = (bits + 7) / 8;
        dbits = kGroupEvenDigits ? (bits + passes - 1) / passes : 8;
    }
    p->passes = passes;
    const uint32_t mask = (1u << dbits) - 1u;
    const bool clears = b.seg_count || b.fill_count;
    int cur = 0;   // buffer pair holding the current order (pass 0 reads the records)
    for (int pass = 0; pass < 4; pass++) {
        const uint32_t tile = sort_pass_tile(pass == passes - 1, g.on, batched);
        const uint32_t tiles = (n + tile - 1) / tile;
        if (pass < passes && (tiles > b.tiles_cap || tiles > 256u * kScanPer)) return hipErrorInvalidValue;
        p->tile[pass] = tile;
        p->hist_blocks[pass] = n ? tiles : (pass == 0 && clears ? 1u : 0u);
        p->pass_blocks[pass] = n ? tiles : 0u;
        HistArgs& h = p->hist[pass];
        h.rec = records;
        h.kvin = b.kv[cur];
        h.n = n;
        h.shift = dbits * pass;
        h.mask = mask;
        h.gk = g;
        h.counts = b.counts;
        h.tiles = tiles;
        h.seg_init = pass == 0 ? b.seg_init : nullptr;
        h.seg_count = pass == 0 ? b.seg_count : 0;
        h.fill = pass == 0 ? b.fill : nullptr;
        h.fill_value = b.fill_value;
        h.fill_count = pass == 0 ? b.fill_count : 0;
        p->scan[pass] = ScanArgs{b.counts, tiles, b.totals, pass == 0 && g.on ? g.tmax : nullptr,
                                 pass == 0 && g.on ? g.vmax : nullptr};
        ScatterArgs& c = p->scatter[pass];
        c.rec = records;
        c.kvin = b.kv[cur];
        c.kvout = b.kv[cur ^ 1];
        c.order = pass == passes - 1 ? b.order : nullptr;
        c.n = n;
        c.shift = dbits * pass;
        c.mask = mask;
        c.dbits = dbits;
        c.counts = b.counts;
        c.totals = b.totals;
        c.tiles = tiles;
        c.gk = g;
        c.out = b.out;
        // grouping: the records move through every pass, ending in out.sorted
        // (the pass before writes tmp, the one before that out.sorted, ...)
        c.move = g.on;
        c.rin = pass == 0 ? records : ((passes - pass) % 2 ? b.tmp : b.out.sorted);
        c.rout = (passes - 1 - pass) % 2 ? b.tmp : b.out.sorted;
        if (pass < passes) cur ^= 1;
    }
    p->order = b.order;
    return hipSuccess;
}

template <bool FIRST, bool MOVE>
static void launch_scatter(bool last, dim3 g, const ScatterArgs& a, hipStream_t s)
{
    if (last) hipLaunchKernelGGL((k_sort_scatter<FIRST, true, MOVE>), g, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_sort_scatter<FIRST, false, MOVE>), g, dim3(256), 0, s, a);
}

template <uint32_t TILE>
static void launch_hist(bool first, dim3 g, const HistArgs& a, hipStream_t s)
{
    if (first) hipLaunchKernelGGL((k_sort_hist<true, TILE>), g, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_sort_hist<false, TILE>), g, dim3(256), 0, s, a);
}

hipError_t launch_sort_plan(const SortPlan& p, hipStream_t s, bool skip_hist0)
{
    for (int pass = 0; pass < p.passes; pass++) {
        const uint32_t hb = pass == 0 && skip_hist0 ? 0u : p.hist_blocks[pass];
        if (hb) {
            if (p.tile[pass] == kSortTile) launch_hist<kSortTile>(pass == 0, dim3(hb), p.hist[pass], s);
            else launch_hist<kSortTileSmall>(pass == 0, dim3(hb), p.hist[pass], s);
        }
        if (!p.pass_blocks[pass]) continue;
        hipLaunchKernelGGL(k_sort_scan, dim3(p.hist[pass].mask + 1), dim3(256), 0, s, p.scan[pass]);
        const dim3 g(p.pass_blocks[pass]);
        const bool last = pass == p.passes - 1;
        const ScatterArgs& c = p.scatter[pass];
        if (pass == 0) {
            if (c.move) launch_scatter<true, true>(last, g, c, s);
            else launch_scatter<true, false>(last, g, c, s);
        } else {
            if (c.move) launch_scatter<false, true>(last, g, c, s);
            else launch_scatter<false, false>(last, g, c, s);
        }
    }
    return hipGetLastError();
}

hipError_t launch_sort_hist_batch(bool first_pass, bool big_tiles, const HistArgs* args, BatchMap m,
                                  uint32_t blocks, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    const dim3 g(blocks);
    if (big_tiles) {
        if (first_pass) hipLaunchKernelGGL((k_sort_hist_batch<true, kSortTile>), g, dim3(256), 0, s, args, m);
        else hipLaunchKernelGGL((k_sort_hist_batch<false, kSortTile>), g, dim3(256), 0, s, args, m);
    } else {
        if (first_pass)
            hipLaunchKernelGGL((k_sort_hist_batch<true, kSortTileSmall>), g, dim3(256), 0, s, args, m);
        else
            hipLaunchKernelGGL((k_sort_hist_batch<false, kSortTileSmall>), g, dim3(256), 0, s, args, m);
    }
    return hipGetLastError();
}

hipError_t launch_sort_scan_batch(const ScanArgs* args, BatchMap m, uint32_t blocks, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_sort_scan_batch, dim3(blocks), dim3(256), 0, s, args, m);
    return hipGetLastError();
}

template <bool FIRST, bool MOVE>
static void launch_scatter_batch(bool last, dim3 g, const ScatterArgs* args, BatchMap m, hipStream_t s)
{
    if (last) hipLaunchKernelGGL((k_sort_scatter_batch<FIRST, true, MOVE>), g, dim3(256), 0, s, args, m);
    else hipLaunchKernelGGL((k_sort_scatter_batch<FIRST, false, MOVE>), g, dim3(256), 0, s, args, m);
}

hipError_t launch_sort_scatter_batch(bool first_pass, bool last_pass, bool move, const ScatterArgs* args,
                                     BatchMap m, uint32_t blocks, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    const dim3 g(blocks);
    if (first_pass) {
        if (move) launch_scatter_batch<true, true>(last_pass, g, args, m, s);
        else launch_scatter_batch<true, false>(last_pass, g, args, m, s);
    } else {
        if (move) launch_scatter_batch<false, true>(last_pass, g, args, m, s);
        else launch_scatter_batch<false, false>(last_pass, g, args, m, s);
    }
    return hipGetLastError();
}

}  // namespace rtsim
