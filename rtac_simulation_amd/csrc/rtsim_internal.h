// Internal declarations shared by the host runtime (capi.cpp) and the HIP
// kernels (kernels_*.hip).  Not part of the public ABI (include/rtsim.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace rtsim {

constexpr uint32_t kOutOfRange = 0xFFFFFFFFu;

// Binner::Key constants (Binner.h:25-45), computed on the host in FP32.
struct KeyParams {
    uint32_t n;
    float a, b, lower, upper;
};

// ---- batched launches ---------------------------------------------------------
// One launch covers the work of many pings: first[p] is the first workgroup of
// ping p (exclusive prefix of workgroups per ping, nping + 1 entries) and each
// workgroup finds its ping by binary search, then runs the single-ping body
// with that ping's arguments and its local workgroup index.
struct BatchMap {
    const uint32_t* first;
    uint32_t nping;
};

// ---- streaming kernel: ReceiverView::set_sample + raygen phase -------------
struct StreamArgs {
    const float* range;
    const float* dx;
    const float* dy;
    const float* dz;
    const float2* amp;
    float4* out;          // records {re, im, travel, bearing}
    uint32_t n;
    float R[9];           // row-major pose rotation; local = R^T dir
    double k2;            // 2 f / c (cycles per metre of one-way travel)
    int spreading;
    int is2d;
    const float* dir;     // [dir_h][dir_w] or nullptr
    int dir_w, dir_h;
    float dir_xs, dir_ys;
};
hipError_t launch_stream(const StreamArgs& a, hipStream_t s);
struct HistArgs;
// the streaming kernel fused with the grid form's first sort histogram h
// (tiles of kSortTileSmall samples; a.n == h.n, a.out == h.rec)
hipError_t launch_stream_hist(const StreamArgs& a, const HistArgs& h, hipStream_t s);
uint32_t stream_blocks(const StreamArgs& a);
hipError_t launch_stream_batch(const StreamArgs* args, BatchMap m, uint32_t blocks, hipStream_t s);

// G-row map of the grid (MFMA) form: rho = floor(fma(t - range0, inv_dr, off_r)) + fr.
// row_start[rho + 1] = first sorted sample whose G row is rho, rho in [-1, GH + 1]
// (n when the row has none).
struct RowMap {
    float range0, inv_dr, off_r;
    int fr, GH;
    uint32_t* row_start;  // GH + 3 entries, pre-filled with n; nullptr: no row table
    // grid column coordinate of a bearing-axis value x: (x - bear0) inv_db + off_b
    double bear0;
    float inv_db, off_b;
};

// Sort key of the grid (MFMA) form.  K = (G row + 2) + bin slot, where the
// bin slot is 0 below the Binner's range, key + 1 inside it and n + 1 above
// it (NaN included).  Both terms are non-decreasing in travel, so samples of
// equal K share their bin and their G row, and the stable order of K is the
// travel order up to a permutation inside each run of equal (bin, G row):
// keys, segments, windows and the G-row table come out exactly as after the
// travel sort, in ceil(bits(K) / 8) passes instead of 4.
// With a grouping key the first histogram also writes tmax[tile], the largest
// finite |re|, |im| of the tile's in-range samples (FP32 bits), and the first
// scan reduces them into vmax[0]: the deposit's fixed-point scale (no
// atomics: a few dozen on one address serialise at the end of a kernel).
struct GroupKey {
    int on;               // 0: sort by travel image
    RowMap rows;
    KeyParams key;
    uint32_t* vmax;       // 1 word
    uint32_t* tmax;       // one word per tile
};
// largest K of a grouping (the sort runs as many 8-bit passes as it needs)
inline uint32_t group_key_max(const GroupKey& g) { return (uint32_t)g.rows.GH + 4 + g.key.n + 1; }

// ---- sort by travel: stable LSD radix sort of (travel image, index) ---------
// Items per sort workgroup (256 lanes x 8 or x 4).  The last pass of one
// ping's sort whose records move uses the large tiles (at 1M samples its
// tiles fit the chip's workgroup slots either way, and twice the small
// tiles' run-bound atomics cost more than the large tiles' lower
// parallelism), every other pass the small ones (more workgroups hide the
// latency of their loads).  Batched launches use the small tiles in every
// pass: a batch queues many workgroups per slot, and the small tiles' five
// resident workgroups per CU (against two) finish them sooner.
constexpr uint32_t kSortTile = 2048;
constexpr uint32_t kSortTileSmall = 1024;
constexpr uint32_t sort_pass_tile(bool last, bool move, bool batched = false)
{
    return last && move && !batched ? kSortTile : kSortTileSmall;
}
// per-tile digit counts of one pass for `capacity` items
inline size_t sort_count_words(size_t capacity)
{
    return 256 * ((capacity + kSortTileSmall - 1) / kSortTileSmall);
}
// Written by the last pass for every sorted position i: the record and its
// bearing-axis coordinate, as (hi, lo) (coord) or, for the grid form, as a
// grid column coordinate (xg); when skeys is set, also its Binner key
// (Binner.cu:24-49).
// The last pass also bounds the runs of equal keys (Binner.cu:51-80): each
// workgroup lowers seg[k].x to the first and raises seg[k].y past the last
// position of every key run it writes (atomic min / max; the runs of one key
// are contiguous, so the result is the segment), and with a grouping key it
// lowers row_start[rho + 1] to the first position of every G row rho it
// writes (rows without samples keep the fill value n).
struct SortOutputs {
    float4* sorted;
    uint32_t* skeys;      // may be null (the keys are not needed downstream)
    float2* coord;        // used when xg is null
    float* xg;            // (x - xg_b0) xg_inv + xg_off of x = the bearing-axis value
    double xg_b0;
    float xg_inv, xg_off;
    KeyParams key;
    int axis;             // 1: coordinate = sin(bearing) in FP64, else the bearing
    uint2* seg;           // [key.n], initialised to (~0, 0) by the first pass
    // 1: seg only marks the keys that hold samples, as (0, 1) (a window's
    // emptiness is all the grid form reads); 0: the Binner segments
    int seg_marks;
};
struct SortBuffers {
    uint2* kv[2];         // (key, sample index) ping-pong between passes
    uint32_t* order;      // the permutation, written by the last pass (may be null)
    uint32_t* counts;     // [256 digits][tiles_cap]: tile counts, then their exclusive scan
    uint32_t* totals;     // [256] digit totals of the current pass
    uint32_t tiles_cap;
    float4* tmp;          // records between passes when they move with their keys
    // prepared by the first sort kernel (before any later kernel reads them)
    uint2* seg_init;      // seg_init[0..seg_count) = (~0, 0) (Binner segments, min / max)
    size_t seg_count;
    uint32_t* fill;       // fill[0..fill_count) = fill_value (grid row table)
    uint32_t fill_value;
    size_t fill_count;
    SortOutputs out;
};
// One pass of the sort: per-tile digit counts, their scan, the scatter.
struct HistArgs {
    const float4* rec;        // pass 0 reads the records themselves
    const uint2* kvin;        // later passes: (key, index) of the previous pass
    uint32_t n;
    int shift;
    uint32_t mask;            // digit = (key >> shift) & mask
    GroupKey gk;              // pass 0 key: K when gk.on, else the travel image
    uint32_t* counts;         // [256 digits][tiles]
    uint32_t tiles;
    uint2* seg_init;          // pass 0 only: initialised / filled arrays
    size_t seg_count;
    uint32_t* fill;
    uint32_t fill_value;
    size_t fill_count;
};
struct ScanArgs {
    uint32_t* counts;
    uint32_t tiles;
    uint32_t* totals;
    const uint32_t* tmax;     // first pass of a grouping sort: vmax[0] = max of tmax[0..tiles)
    uint32_t* vmax;
};
struct ScatterArgs {
    const float4* rec;
    const uint2* kvin;        // (key, index) in the previous pass's order (not pass 0)
    uint2* kvout;             // not the last pass: (key, index) in this pass's order
    uint32_t* order;          // last pass: the permutation (null: not written)
    uint32_t n;
    int shift;
    uint32_t mask;            // digit = (key >> shift) & mask, mask = 2^dbits - 1
    int dbits;
    const uint32_t* counts;   // scanned [digit][tile]
    const uint32_t* totals;
    uint32_t tiles;
    GroupKey gk;              // as HistArgs
    int move;                 // records move with their keys (grouping key)
    const float4* rin;        // move: records in the order of kvin (pass 0: the input)
    float4* rout;             // move, not the last pass: records in the order of kvout
    SortOutputs out;          // last pass only
};
// Arguments and workgroup counts of the 3 launches per pass of one sort.
struct SortPlan {
    int passes;               // 4 for travel images, fewer for a grouping key
    HistArgs hist[4];
    ScanArgs scan[4];
    ScatterArgs scatter[4];
    uint32_t tile[4];         // items per tile of each pass (sort_pass_tile)
    uint32_t hist_blocks[4];  // per pass: tiles (pass 0: >= 1 when it clears arrays)
    uint32_t pass_blocks[4];  // scan: 256 per pass, scatter: tiles (0 when n = 0)
    const uint32_t* order;    // permutation after the sort
};
// gk == nullptr or !gk->on: sort by travel.  batched: the plan is for the
// batched launches (their tiles, sort_pass_tile)
hipError_t make_sort_plan(const float4* records, uint32_t n, const SortBuffers& b,
                          const GroupKey* gk, SortPlan* p, bool batched = false);
// skip_hist0: the first pass's histogram (and its clears) are already made
// (launch_stream_hist)
hipError_t launch_sort_plan(const SortPlan& p, hipStream_t s, bool skip_hist0 = false);
// batched sort passes (args are device arrays, one entry per ping)
hipError_t launch_sort_hist_batch(bool first_pass, bool big_tiles, const HistArgs* args, BatchMap m,
                                  uint32_t blocks, hipStream_t s);
hipError_t launch_sort_scan_batch(const ScanArgs* args, BatchMap m, uint32_t blocks,
                                  hipStream_t s);
hipError_t launch_sort_scatter_batch(bool first_pass, bool last_pass, bool move, const ScatterArgs* args,
                                     BatchMap m, uint32_t blocks, hipStream_t s);

// ---- windows after the sort ------------------------------------------------------
// do_make_bins (Binner.cu:82-106) over the segments the sort's last pass
// bounded; every thread also writes its own empty segment back as (0, 0).
struct BinArgs {
    uint32_t N;
    int ov;
    uint2* seg;
    uint2* win;
};
hipError_t launch_windows(const BinArgs& a, hipStream_t s);
uint32_t windows_blocks(const BinArgs& a);
hipError_t launch_windows_batch(const BinArgs* args, BatchMap m, uint32_t blocks, hipStream_t s);

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s);


// ---- plain FP32 PSF convolution (SensorInstance{1D,2D}.cu) ------------------
struct ConvArgs {
    const float4* sorted;
    const float2* coord;
    const uint2* win;
    const float2* range_c;   // per output row: (hi, lo) of the range in metres
    const float2* bear_c;    // per output bearing: (hi, lo) of the axis coordinate
    const float2* psf;       // [Tr][Tb] complex
    float2* out;             // [H][W]
    int H, W, Tr, Tb;
    float inv_pr, half_r;    // range tap coordinate yB = dr * inv_pr + half_r
    float inv_pb, half_b;    // bearing tap coordinate xB = dx * inv_pb + half_b
    float supp_b;            // |dx| >= supp_b  =>  PSF weight is exactly 0
    int psf_in_lds;          // stage the PSF table in LDS (else read it from global)
};
size_t plain2d_lds_bytes(const ConvArgs& a);
void init_plain_attributes(int max_lds);
void init_grid_attributes(int max_lds);

// Raises a kernel's dynamic LDS limit to what the device allows next to its
// static LDS; clears the error state so a refusal cannot leak into a later
// hipGetLastError().
inline void set_dynamic_lds_limit(const void* fn, int max_lds)
{
    hipFuncAttributes attr{};
    int limit = max_lds;
    if (hipFuncGetAttributes(&attr, fn) == hipSuccess) limit -= (int)attr.sharedSizeBytes;
    if (limit > 0) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, limit);
    (void)hipGetLastError();
}
hipError_t launch_conv2d_plain(const ConvArgs& a, hipStream_t s);
hipError_t launch_conv1d_plain(const ConvArgs& a, hipStream_t s);
uint32_t conv2d_plain_blocks(const ConvArgs& a);
uint32_t conv1d_plain_blocks(const ConvArgs& a);
hipError_t launch_conv2d_plain_batch(const ConvArgs* args, BatchMap m, uint32_t blocks,
                                     size_t shm, hipStream_t s);
hipError_t launch_conv1d_plain_batch(const ConvArgs* args, BatchMap m, uint32_t blocks,
                                     hipStream_t s);

}  // namespace rtsim
