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

struct BucketArgs;

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
uint32_t stream_blocks(const StreamArgs& a);
// batched: bargs[q] (parallel to args) holds ping q's bucket counts to make
// on the way, or tiles == 0; shm = 4 x the largest NB among them
hipError_t launch_stream_batch(const StreamArgs* args, BatchMap m, uint32_t blocks,
                               const BucketArgs* bargs, size_t shm, hipStream_t s);

// ---- grid (MFMA) form binning: samples bucketed by blocks of grid rows ------
// The deposit of the grid form sums every sample exactly (fixed point), so it
// needs the samples grouped by grid row, in any order inside a group, and
// from the Binner only which keys hold samples (the windows' emptiness).
// One counting pass does that: a sample with a Binner key whose range
// coordinate y = fma(t - range0, inv_dr, off_r) gives G row gy = floor(y) + fr
// adds weight 1 - frac(y) to grid row gy and frac(y) to row gy + 1; it goes
// to the bucket of each of those rows inside the grid (buckets of R rows, a
// sample on a bucket's last row goes to two buckets), as an item
// {re, im, travel, grid column}.
struct GridRows {
    float range0, inv_dr, off_r;
    int fr, GH;
};
struct BucketArgs {
    const float4* rec;    // records (unsorted)
    uint32_t n;
    uint32_t tile;        // samples per tile (a multiple of 1024), one workgroup each
    uint32_t tiles;
    KeyParams key;        // Binner key: a sample without one is not binned
    GridRows rows;
    int R, NB;            // grid rows per bucket, buckets = ceil(GH / R)
    uint32_t* counts;     // [NB][tiles]: items per (bucket, tile), then their exclusive scan over tiles
    uint32_t* totals;     // [NB] items per bucket
    uint32_t* tmax;       // [tiles] largest finite |re|, |im| of the tile's binned samples (FP32 bits)
    uint32_t* vmax;       // [1] the largest tile maximum: the deposit's fixed-point scale
    float4* items;        // {re, im, travel, grid column} of every binned sample, by bucket
    uint32_t items_cap;   // items the buffer holds (stores beyond it are dropped)
    uint32_t* occ;        // [tiles][occ_words]: bit k set when Binner key k holds a sample of the tile
    int occ_words;        // ceil(N / 32)
    // grid column of a sample: (x - bear0) inv_db + off_b, x = sin(bearing)
    // (axis 1, in FP64) or the bearing
    double bear0;
    float inv_db, off_b;
    int axis;
};
// the streaming kernel fused with the bucket counts of the grid form (tiles
// of 1024 samples; b.rec == a.out, b.n == a.n)
hipError_t launch_stream_bucket(const StreamArgs& a, const BucketArgs& b, hipStream_t s);
// whether launch_stream_bucket accepts (a, b): one tile per streaming workgroup
bool stream_bucket_fusable(const StreamArgs& a, const BucketArgs& b);
// bucket counts and tile maxima (from the records), their scan, the scatter
// of the items; blocks per launch: tiles, NB, tiles
hipError_t launch_bucket_hist(const BucketArgs& b, hipStream_t s);
hipError_t launch_bucket_scan(const BucketArgs& b, hipStream_t s);
hipError_t launch_bucket_scatter(const BucketArgs& b, hipStream_t s);
size_t bucket_lds_bytes(const BucketArgs& b);   // the scatter's dynamic LDS
// whether the scatter stages its items in LDS (false: too many buckets for it)
bool bucket_scatter_staged(const BucketArgs& b);
void init_bucket_attributes(int max_lds);
// (batched: the bucket scan runs k_sort_scan's batched launch over ScanArgs)
hipError_t launch_bucket_hist_batch(const BucketArgs* args, BatchMap m, uint32_t blocks, size_t shm,
                                    hipStream_t s);
// staged: every ping of the launch has bucket_scatter_staged
hipError_t launch_bucket_scatter_batch(const BucketArgs* args, BatchMap m, uint32_t blocks, size_t shm,
                                       bool staged, hipStream_t s);
constexpr int kMaxBuckets = 8192;        // buckets of one ping (LDS counters)
constexpr int kMaxOccWords = 2048;       // Binner keys of a grid-form ping: 65536

// ---- sort by travel: stable LSD radix sort of (travel image, index) ---------
// Four 8-bit passes over tiles of 1024 items (256 lanes x 4).  Used by the
// plain form (its gather reads the samples of each window in travel order)
// and by Binner introspection.
constexpr uint32_t kSortTileSmall = 1024;
constexpr uint32_t kSortTile = kSortTileSmall;
constexpr uint32_t sort_pass_tile(bool, bool, bool = false) { return kSortTileSmall; }
// per-tile digit counts of one pass for `capacity` items
inline size_t sort_count_words(size_t capacity)
{
    return 256 * ((capacity + kSortTileSmall - 1) / kSortTileSmall);
}
// Written by the last pass for every sorted position i: the record and its
// bearing-axis coordinate as (hi, lo) (coord); when skeys is set, also its
// Binner key (Binner.cu:24-49).  The last pass also bounds the runs of equal
// keys (Binner.cu:51-80): each workgroup lowers seg[k].x to the first and
// raises seg[k].y past the last position of every key run it writes (atomic
// min / max; the runs of one key are contiguous, so the result is the segment).
struct SortOutputs {
    float4* sorted;
    uint32_t* skeys;      // may be null (the keys are not needed downstream)
    float2* coord;
    KeyParams key;
    int axis;             // 1: coordinate = sin(bearing) in FP64, else the bearing
    uint2* seg;           // [key.n], initialised to (~0, 0) by the first pass
};
struct SortBuffers {
    uint2* kv[2];         // (key, sample index) ping-pong between passes
    uint32_t* order;      // the permutation, written by the last pass (may be null)
    uint32_t* counts;     // [256 digits][tiles_cap]: tile counts, then their exclusive scan
    uint32_t* totals;     // [256] digit totals of the current pass
    uint32_t tiles_cap;
    uint2* seg_init;      // seg_init[0..seg_count) = (~0, 0) (Binner segments, min / max),
    size_t seg_count;     // prepared by the first histogram
    SortOutputs out;
    // sort by Binner key (out-of-range last) instead of by travel: the same
    // segments and windows in 1 or 2 passes instead of 4 (input order inside
    // a bin); used when nothing asks for the travel order
    bool by_key;
};
// One pass of the sort: per-tile digit counts, their scan, the scatter.
struct HistArgs {
    const float4* rec;        // pass 0 reads the records themselves
    const uint2* kvin;        // later passes: (key, index) of the previous pass
    uint32_t n;
    int shift;
    uint32_t mask;            // digit = (key >> shift) & mask
    uint32_t* counts;         // [256 digits][tiles]
    uint32_t tiles;
    uint2* seg_init;          // pass 0 only
    size_t seg_count;
    int by_key;               // pass 0: the sort key is the Binner key (key.n: out of range)
    KeyParams key;
};
// Exclusive scan of rows of tile counts: one workgroup per row (a digit, or a
// bucket of the grid form), the row total in totals[row]; with tmax, row 0's
// workgroup also writes vmax[0] = max of tmax[0..tiles).
struct ScanArgs {
    uint32_t* counts;
    uint32_t tiles;
    uint32_t* totals;
    const uint32_t* tmax;
    uint32_t* vmax;
};
struct ScatterArgs {
    const float4* rec;
    const uint2* kvin;        // (key, index) in the previous pass's order (not pass 0)
    uint2* kvout;             // not the last pass: (key, index) in this pass's order
    uint32_t* order;          // last pass: the permutation (null: not written)
    uint32_t n;
    int shift;
    uint32_t mask;            // digit = (key >> shift) & mask
    const uint32_t* counts;   // scanned [digit][tile]
    const uint32_t* totals;
    uint32_t tiles;
    SortOutputs out;          // last pass only (out.key: the Binner key of every pass 0 sort key)
    int by_key;               // pass 0: as HistArgs::by_key
};
// Arguments and workgroup counts of the 3 launches per pass of one sort.
struct SortPlan {
    int passes;               // 4 by travel, 1 or 2 by Binner key
    HistArgs hist[4];
    ScanArgs scan[4];
    ScatterArgs scatter[4];
    uint32_t tile[4];         // items per tile of each pass
    uint32_t hist_blocks[4];  // per pass: tiles (pass 0: >= 1 when it clears arrays)
    uint32_t pass_blocks[4];  // scan: 256 per pass, scatter: tiles (0 when n = 0)
    const uint32_t* order;    // permutation after the sort
};
hipError_t make_sort_plan(const float4* records, uint32_t n, const SortBuffers& b, SortPlan* p);
hipError_t launch_sort_plan(const SortPlan& p, hipStream_t s);
// k_sort_scan over `rows` rows of a count matrix (also the grid form's buckets)
hipError_t launch_scan_rows(const ScanArgs& a, uint32_t rows, hipStream_t s);
// batched sort passes (args are device arrays, one entry per ping)
hipError_t launch_sort_hist_batch(bool first_pass, const HistArgs* args, BatchMap m, uint32_t blocks,
                                  hipStream_t s);
hipError_t launch_sort_scan_batch(const ScanArgs* args, BatchMap m, uint32_t blocks,
                                  hipStream_t s);
hipError_t launch_sort_scatter_batch(bool first_pass, bool last_pass, const ScatterArgs* args,
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
    const float4* psf_cells; // 2D: bilinear cells [Tr + 2][Tb + 2], j, i from -1, in two planes:
                             // (A, B) of every cell, then (C, D) of every cell, with
                             // K(j + ay, i + ax) = A + ax B + ay (C + ax D),
                             // taps T zero outside [0, Tr) x [0, Tb)
    float2* out;             // [H][W]
    int H, W, Tr, Tb;
    float inv_pr, half_r;    // range tap coordinate yB = dr * inv_pr + half_r
    float inv_pb, half_b;    // bearing tap coordinate xB = dx * inv_pb + half_b
    float supp_b;            // |dx| >= supp_b  =>  PSF weight is exactly 0
    int psf_in_lds;          // 2D: stage the (A, B) cells in LDS (else read them from global)
    int chunk3;              // 2D: stage 3072 samples per chunk (else 2048)
    uint32_t* work;          // 2D: band lengths and band order (plain2d_work_words)
    // 2D strip lists (null: every tile stages its band's whole window):
    // strip j's list = the sorted samples its tiles keep, in window order, as
    // (sample index, bearing tap offset xs as bits)
    uint32_t n;              // sorted samples
    uint2* strip_ent;
    uint32_t* strip_cnt;     // [strips] totals, then [strips][strip_blocks(n)] counts / offsets
    size_t strip_cap;        // entries strip_ent holds
};
size_t plain2d_lds_bytes(const ConvArgs& a);
size_t plain2d_stage_bytes(int chunk_per);   // staged samples of one chunk (chunk_per x 1024)
// dynamic LDS of a 2D gather workgroup when two of them share a CU
constexpr size_t kPlainLdsBudget = 81280;
// cells of the largest tap plane the 2D gather stages in LDS (two 1024-thread
// workgroups per CU: 42 KiB of staged samples + the plane each)
constexpr size_t kPlainLdsCells = 2380;
// strip lists: a strip is the 16 bearing columns of one tile column; the
// lists are built per block of kStripBlock sorted samples
constexpr uint32_t kStripCols = 16, kStripBlock = 1024, kMaxStrips = 512;
constexpr uint32_t plain_strips(uint32_t W) { return (W + kStripCols - 1) / kStripCols; }
constexpr uint32_t strip_blocks(uint32_t n) { return (n + kStripBlock - 1) / kStripBlock; }
constexpr size_t strip_cnt_words(uint32_t W, uint32_t n) { return (size_t)plain_strips(W) * (1 + strip_blocks(n)); }
// the strips a sample can reach at most, for bearing-axis coordinates x
// (ascending) and a PSF bearing support of `supp` (0: no lists: the strips
// are too many, or a sample reaches too many of them to save staging)
uint32_t strip_multiplicity(const double* x, uint32_t W, double supp, double pitch);
// the plain form's bins stage: the windows, and for a 2D ping with active
// strip lists the lists and the band ranking (launch_conv2d_plain then only
// gathers)
bool plain_lists_active(const ConvArgs& a);
hipError_t launch_plain_bins(const BinArgs& w, const ConvArgs& a, hipStream_t s);
hipError_t launch_strip_count_batch(const ConvArgs* args, BatchMap m, uint32_t blocks, hipStream_t s);
hipError_t launch_strip_scan_batch(const ConvArgs* args, BatchMap m, uint32_t blocks, hipStream_t s);
hipError_t launch_strip_scatter_batch(const ConvArgs* args, BatchMap m, uint32_t blocks, hipStream_t s);
// the 2D gather's bands of output rows, and its work words for `rows` rows
constexpr uint32_t kPlainBandRows = 16;
constexpr size_t plain2d_work_words(uint32_t rows) { return 2 * (((size_t)rows + kPlainBandRows - 1) / kPlainBandRows); }
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
// variant: bit 0 taps in LDS, bit 1 3072-sample chunks (for every ping of the launch)
hipError_t launch_conv2d_plain_batch(const ConvArgs* args, BatchMap m, uint32_t blocks,
                                     size_t shm, int variant, hipStream_t s);
// the band ordering that precedes the 2D gather: one workgroup per ping
hipError_t launch_plain_rank_batch(const ConvArgs* args, BatchMap m, uint32_t blocks, hipStream_t s);
hipError_t launch_conv1d_plain_batch(const ConvArgs* args, BatchMap m, uint32_t blocks,
                                     hipStream_t s);

}  // namespace rtsim
