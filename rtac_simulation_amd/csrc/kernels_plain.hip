// Plain FP32 sparse PSF convolution: the reference's per-sample form.
//
//   2D: do_sparse_convolve_2d_f (SensorInstance2D.cu:44-83)
//       out(r, b) = sum_{s in view(r)} v_s K(x_s - x_b, t_s - r_r)
//   1D: do_sparse_convolve_1d_f (SensorInstance1D.cu:35-72)
//       out(r)    = sum_{s in view(r)} v_s K(t_s - r_r)
//
// K is the bilinear (texture linear filter, border 0) interpolant of the PSF
// taps (ReductionKernel.h:44-73) with full-FP32 weights.  The reference
// gathers every window sample into every bearing column; here a cell sums
// only the window samples inside the PSF support (|x_s - x_b| < (Tb + 1)/2
// tap pitches and the range tap coordinate inside (-1, Tr)), which skips
// exactly the zero terms.  Output coordinates are double-float (hi, lo)
// pairs so that t_s - r_r and x_s - x_b keep full FP32 relative accuracy.
#include "device_common.h"

#include <algorithm>

namespace rtsim {

// PSF value at tap coordinates (xB along bearing taps, yB along range taps).
// Caller guarantees xB in (-1, Tb) and yB in (-1, Tr).
__device__ __forceinline__ float2 psf_bilinear(const float2* __restrict__ psf, int Tr, int Tb,
                                               float xB, float yB)
{
    const float fi = floorf(xB), fj = floorf(yB);
    const float ax = xB - fi, ay = yB - fj;
    const int i0 = (int)fi, j0 = (int)fj;
    float2 acc = make_float2(0.f, 0.f);
    if (j0 >= 0) {
        const float2* row = psf + j0 * Tb;
        float2 lo = i0 >= 0 ? row[i0] : make_float2(0.f, 0.f);
        float2 hi = i0 + 1 < Tb ? row[i0 + 1] : make_float2(0.f, 0.f);
        const float wy = 1.0f - ay;
        acc.x += wy * (lo.x + ax * (hi.x - lo.x));
        acc.y += wy * (lo.y + ax * (hi.y - lo.y));
    }
    if (j0 + 1 < Tr) {
        const float2* row = psf + (j0 + 1) * Tb;
        float2 lo = i0 >= 0 ? row[i0] : make_float2(0.f, 0.f);
        float2 hi = i0 + 1 < Tb ? row[i0 + 1] : make_float2(0.f, 0.f);
        acc.x += ay * (lo.x + ax * (hi.x - lo.x));
        acc.y += ay * (lo.y + ax * (hi.y - lo.y));
    }
    return acc;
}

// ---- 2D: register-accumulating gather over output tiles ---------------------
// A workgroup owns a tile of 16 x 16 output cells made of four 8 x 8 blocks;
// four waves share a block, and each lane of a wave owns one cell of it and
// sums its terms in two registers.  The samples that can reach one of the
// tile's blocks (in the union of the block's rows' windows, bearing coordinate
// within the PSF's bearing support of its columns) are staged in LDS one chunk
// at a time, in window order, with a bit per block they reach: from the
// strip's sample list when the ping has strip lists (below), else from every
// sample of the window.  Each block's four waves share the staged samples
// that reach the block (the k-th of them goes to wave
// k mod 4), and every lane adds v_s K(x_s - x_b, t_s - r_r) when sample s is
// in its row's window and inside the PSF support (the reference's border
// mode makes every other term 0).  At the end the four partial sums of a cell
// are added in a fixed order, so the ping is bit-reproducible, and the same
// whichever source staged the samples.
//
// The taps are read as bilinear cells with a zero border (A, B, C, D:
// K = A + ax B + ay (C + ax D) inside the cell), in two planes of cells,
// (A, B) and (C, D): a lookup is two 16-byte reads, the first from the (A, B)
// plane staged in LDS, the second from the (C, D) plane through the vector L1
// cache, so the two go to different units (the 8 lanes of a block row read 8
// consecutive cells, 128 contiguous bytes: no LDS bank conflict, one cache
// line), and three FMAs per component.  A term outside the support or the
// window reads only zeros, so the lanes need no branch.  (Terms of a
// non-finite sample value therefore reach every lane of the blocks it is
// staged for: 0 x inf is NaN, as in the reference, whose gather multiplies
// every window sample.)
//
// Work per tile follows the samples, which crowd into some ranges: a first
// kernel orders the bands of 16 rows by the length of their union window,
// heaviest first, and deals them to the 8 XCDs in turn; each XCD runs its
// bands in that order (every tile of a band on the same XCD, which reads the
// band's samples through its own L2), so the longest tiles start first.
constexpr int kCellRows = 8, kCellCols = 8;        // a block
constexpr int kTileRows = (int)kPlainBandRows, kTileCols = 16;   // 2 x 2 blocks
static_assert(kTileCols == (int)kStripCols, "a strip is one tile column");
static_assert(kTileRows == 2 * kCellRows, "a band is two blocks tall");
constexpr int kTileBlocks = 4;
constexpr int kWavesPerBlock = 4;
constexpr int kTileWaves = kTileBlocks * kWavesPerBlock;
constexpr int kTileThreads = 64 * kTileWaves;      // 1024
// samples per thread per staging pass: 3 (chunks of 3072) when the staged
// (A, B) plane leaves room for them at two workgroups per CU
// (ConvArgs::chunk3), else 2; fewer chunks mean fewer barriers (the ping is
// the same either way)
constexpr int kRankThreads = 1024;
constexpr int kRankLds = 4096;                     // band lengths staged per pass
static_assert(sizeof(float2) * kTileThreads <= (sizeof(float4) + sizeof(float)) * 2 * kTileThreads,
              "partial sums reuse the staging space");
static_assert(3 * kTileWaves <= 64, "one wave's count per lane");

size_t plain2d_stage_bytes(int chunk_per)
{
    return (sizeof(float4) + sizeof(float) + sizeof(uint8_t)) * kTileThreads * (size_t)chunk_per;
}

// staged samples + the (A, B) plane when ConvArgs::psf_in_lds
size_t plain2d_lds_bytes(const ConvArgs& a)
{
    const size_t taps = a.psf_in_lds ? sizeof(float4) * (size_t)(a.Tr + 2) * (a.Tb + 2) : 0;
    return plain2d_stage_bytes(a.chunk3 ? 3 : 2) + taps;
}

__host__ __device__ inline uint32_t plain2d_bands(const ConvArgs& a) { return (uint32_t)(a.H + kTileRows - 1) / kTileRows; }

uint32_t conv2d_plain_blocks(const ConvArgs& a)
{
    if (a.H <= 0 || a.W <= 0) return 0u;
    const uint32_t per_xcd = (plain2d_bands(a) + 7) / 8;
    return 8u * per_xcd * (uint32_t)((a.W + kTileCols - 1) / kTileCols);
}

// work[0, B): band lengths; work[B, 2B): band of each rank (longest first,
// ties by band index).  One workgroup per ping.
__device__ __forceinline__ void plain_rank_body(const ConvArgs& a, uint32_t* work)
{
    __shared__ uint32_t lens[kRankLds];
    const int B = (int)plain2d_bands(a);
    const int t = threadIdx.x;
    static_assert(64 % kTileRows == 0, "a band's rows are lanes of one wave");
    // one row per thread; a band's union over its 16 lanes
    for (int r0 = 0; r0 < B * kTileRows; r0 += kRankThreads) {
        const int r = r0 + t;
        uint32_t u0 = 0xFFFFFFFFu, u1 = 0u;
        if (r < a.H) {
            const uint2 w = a.win[r];
            if (w.y) {
                u0 = w.x;
                u1 = w.x + w.y;
            }
        }
        for (int m = 1; m < kTileRows; m <<= 1) {
            u0 = min(u0, (uint32_t)__shfl_xor((int)u0, m, 64));
            u1 = max(u1, (uint32_t)__shfl_xor((int)u1, m, 64));
        }
        if (r % kTileRows == 0 && r / kTileRows < B) work[r / kTileRows] = u1 > u0 ? u1 - u0 : 0u;
    }
    __threadfence_block();
    __syncthreads();
    for (int b0 = 0; b0 < B; b0 += kRankThreads) {
        const int b = b0 + t;
        const uint32_t mine = b < B ? work[b] : 0u;
        uint32_t rank = 0;
        for (int j0 = 0; j0 < B; j0 += kRankLds) {
            const int nj = min(kRankLds, B - j0);
            __syncthreads();
            for (int j = t; j < nj; j += kRankThreads) lens[j] = work[j0 + j];
            __syncthreads();
            for (int j = 0; j < nj; j++) {
                const uint32_t o = lens[j];
                rank += (o > mine || (o == mine && j0 + j < b)) ? 1u : 0u;
            }
        }
        if (b < B) work[B + rank] = (uint32_t)b;
    }
}

__global__ __launch_bounds__(kRankThreads) void k_plain_rank(ConvArgs a, uint32_t* work)
{
    plain_rank_body(a, work);
}

__global__ __launch_bounds__(kRankThreads) void k_plain_rank_batch(const ConvArgs* __restrict__ args,
                                                                   BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const ConvArgs a = args[p];
    plain_rank_body(a, a.work);
}

__device__ __forceinline__ float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ uint32_t wave_umin(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

__device__ __forceinline__ uint32_t wave_umax(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

// VGPR copy of a wave-uniform value: an FMA may read one scalar operand only,
// so constants that meet in one FMA are kept in vector registers
__device__ __forceinline__ float in_vgpr(float x)
{
    float v;
    asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(x));
    return v;
}

// floor(x) as an integer in one instruction (|x| < 2^31)
__device__ __forceinline__ int floor_to_int(float x)
{
    int i;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(i) : "v"(x));
    return i;
}

// ---- bearing staging test ----------------------------------------------------
// The bearing tap coordinate of a term is split into the sample's part and
// the column's: with the origin o_b = the coordinate of the middle column of
// the tile's strip (double-float),
//   xB = (x_s - x_b) / pb + (Tb - 1) / 2 = xs + xc,
//   xs = (x_s - o_b) / pb,   xc = (Tb - 1) / 2 - (x_b - o_b) / pb.
// A sample goes into a strip's list when xs lies in the strip's box: xB in
// (-1, Tb) for one of its columns, widened by 1/64 tap (a term just outside is
// clamped to the zero border and adds exactly 0), and reaches a block of a
// tile when xs lies in the block's box of the same form.  The gather and the
// strip lists evaluate these with the same expressions, so they agree bit for
// bit (a block's box lies inside its strip's).
constexpr float kBoxMargin = 1.0f / 64.0f;

__device__ __forceinline__ float2 strip_origin(const ConvArgs& a, int j)
{
    return a.bear_c[min(j * (int)kStripCols + (int)kStripCols / 2, a.W - 1)];
}
__device__ __forceinline__ float column_xc(const ConvArgs& a, float2 o_b, float2 bx)
{
    return __builtin_fmaf(-((bx.x - o_b.x) + (bx.y - o_b.y)), a.inv_pb, a.half_b);
}
__device__ __forceinline__ float box_lo(float xc) { return -1.0f - xc - kBoxMargin; }
__device__ __forceinline__ float box_hi(const ConvArgs& a, float xc) { return (float)a.Tb - xc + kBoxMargin; }
__device__ __forceinline__ float sample_xs(const ConvArgs& a, float2 o_b, float2 x)
{
    return ((x.x - o_b.x) + (x.y - o_b.y)) * a.inv_pb;
}

template <bool kTapsLds, int kChunkPer>
__device__ __forceinline__ void conv2d_plain_body(const ConvArgs& a, uint32_t blk)
{
    constexpr int kChunk = kTileThreads * kChunkPer;
    extern __shared__ float4 lds4[];
    float4* taps_lds = lds4;                                    // the (A, B) plane when staged
    const int ncells = (a.Tr + 2) * (a.Tb + 2);                 // cells of a plane
    float4* cs_rec = lds4 + (kTapsLds ? ncells : 0);            // [kChunk] (re, im, t, index)
    float* cs_x = reinterpret_cast<float*>(cs_rec + kChunk);    // [kChunk] xs
    uint8_t* cs_m = reinterpret_cast<uint8_t*>(cs_x + kChunk);  // [kChunk] blocks it reaches (bit per block)
    __shared__ uint32_t counts[kChunkPer * kTileWaves];
    __shared__ float4 wave_box[kTileWaves];                     // (index lo, index hi, xs lo, xs hi)
    __shared__ uint32_t strip_ws[kTileWaves];
    __shared__ uint2 strip_range;

    // workgroup -> (band, tile column): XCD x = blk mod 8 runs the bands of
    // rank x, x + 8, ... in that order
    const int B = (int)plain2d_bands(a);
    const int tiles_c = (a.W + kTileCols - 1) / kTileCols;
    const uint32_t k = blk >> 3;
    const int rank = (int)(blk & 7u) + 8 * (int)(k / (uint32_t)tiles_c);
    if (rank >= B) return;                                      // padding of the last round
    const int band = (int)a.work[B + rank];
    const int tc = (int)(k % (uint32_t)tiles_c);

    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int blkw = w / kWavesPerBlock, sub = w % kWavesPerBlock;
    const int r = band * kTileRows + (blkw >> 1) * kCellRows + (lane >> 3);
    const int c = tc * kTileCols + (blkw & 1) * kCellCols + (lane & 7);
    const int stride = a.Tb + 2;                                // cells per row
    const float4* taps = a.psf_cells;
    if constexpr (kTapsLds) {
        for (int j = t; j < ncells; j += kTileThreads) taps_lds[j] = a.psf_cells[j];
        taps = taps_lds;
    }
    taps += stride + 1;                                         // cell of tap (0, 0)
    // its (C, D) cell, through L1 (measured: the gather 11% faster than with
    // both planes in LDS at 512 x 1024 with the 31 x 17 PSF)
    const float4* taps_cd = a.psf_cells + ncells + (stride + 1);

    // xs is computed once per staged sample, each lane's xc once, so a term's
    // bearing coordinate costs one add.  |xs| and |xc| stay within the PSF's
    // bearing span plus 8 grid columns, so their FP32 rounding is of the order
    // of xB's own.  (The range coordinate keeps the exact difference
    // t_s - r_r per term: a grid row can span many range taps, so a tile
    // origin would put a range offset far from 0.)
    const float2 o_b = strip_origin(a, tc);
    const float inv_pr = in_vgpr(a.inv_pr), half_r = in_vgpr(a.half_r);
    const float Tbf = (float)a.Tb, Trf = (float)a.Tr;

    // this lane's cell: its row's window and range, its column's xc
    uint2 win = make_uint2(0u, 0u);
    float2 rc = make_float2(0.f, 0.f), bx = make_float2(0.f, 0.f);
    if (r < a.H) {
        win = a.win[r];
        rc = a.range_c[r];
    }
    const bool col_ok = c < a.W;
    if (col_ok) bx = a.bear_c[c];
    const float xc = column_xc(a, o_b, bx);
    // the range tap coordinate of a term: yB = (t_s - r_r) / pr + (Tr - 1) / 2
    // = (t_s - r_hi) / pr + hr, hr = (Tr - 1) / 2 - r_lo / pr
    const float hr = __builtin_fmaf(-rc.y, inv_pr, half_r);
    // the block: index range of its rows' windows, and its columns' box
    const uint32_t wlo = wave_umin(win.y ? win.x : 0xFFFFFFFFu);
    const uint32_t whi = wave_umax(win.y ? win.x + win.y : 0u);
    const float xlo = wave_min(col_ok ? box_lo(xc) : INFINITY);
    const float xhi = wave_max(col_ok ? box_hi(a, xc) : -INFINITY);
    if (lane == 0) wave_box[w] = make_float4(__uint_as_float(wlo), __uint_as_float(whi), xlo, xhi);
    __syncthreads();
    // every block's box in scalar registers (uniform across the workgroup),
    // and the union of the blocks' window index ranges
    constexpr int kBlocks = kTileWaves / kWavesPerBlock;
    uint32_t blo[kBlocks], bhi[kBlocks];
    float bxlo[kBlocks], bxhi[kBlocks];
    uint32_t u0 = 0xFFFFFFFFu, u1 = 0u;
#pragma unroll
    for (int j = 0; j < kBlocks; j++) {
        const float4 b = wave_box[j * kWavesPerBlock];
        blo[j] = (uint32_t)__builtin_amdgcn_readfirstlane((int)__float_as_uint(b.x));
        bhi[j] = (uint32_t)__builtin_amdgcn_readfirstlane((int)__float_as_uint(b.y));
        bxlo[j] = __uint_as_float((uint32_t)__builtin_amdgcn_readfirstlane((int)__float_as_uint(b.z)));
        bxhi[j] = __uint_as_float((uint32_t)__builtin_amdgcn_readfirstlane((int)__float_as_uint(b.w)));
        u0 = min(u0, blo[j]);
        u1 = max(u1, bhi[j]);
    }

    float acc_re = 0.f, acc_im = 0.f;

    // one term: v K(x_s - x_b, t_s - r_r) when s is in the row's window.
    // Tap coordinates are clamped to [-1, Tb] x [-1, Tr]: a clamped or
    // out-of-window term then lands on a cell of the zero border at a corner
    // (ax or ay = 0) where K is 0, so it adds exactly 0 (for finite v)
    auto term = [&](const float4& e, float xs) {
        const uint32_t s = __float_as_uint(e.w);
        const float xB = __builtin_amdgcn_fmed3f(xs + xc, -1.0f, Tbf);
        const float y0 = __builtin_fmaf(e.z - rc.x, inv_pr, hr);
        const float yB = __builtin_amdgcn_fmed3f(s - win.x < win.y ? y0 : -1.0f, -1.0f, Trf);
        // the cell (floor(yB), floor(xB)) and the weights xB - floor(xB),
        // yB - floor(yB) (fract; its result stays below 1)
        const float ax = __builtin_amdgcn_fractf(xB), ay = __builtin_amdgcn_fractf(yB);
        const int cell = __mul24(floor_to_int(yB), stride) + floor_to_int(xB);
        const float4 ab = taps[cell], cd = taps_cd[cell];       // (A, B), (C, D)
        const float kr = __builtin_fmaf(ay, __builtin_fmaf(ax, cd.z, cd.x), __builtin_fmaf(ax, ab.z, ab.x));
        const float ki = __builtin_fmaf(ay, __builtin_fmaf(ax, cd.w, cd.y), __builtin_fmaf(ax, ab.w, ab.y));
        acc_re = __builtin_fmaf(e.x, kr, acc_re);
        acc_re = __builtin_fmaf(-e.y, ki, acc_re);
        acc_im = __builtin_fmaf(e.x, ki, acc_im);
        acc_im = __builtin_fmaf(e.y, kr, acc_im);
    };

    // The staging source: the tile's strip list entries of its window when
    // the lists are complete, else every sample of the window.  Both give the
    // same staged samples in the same order.
    uint32_t p0 = u0, p1 = u1;
    bool from_list = false;
    if (a.strip_ent != nullptr && u0 < u1) {
        const uint32_t ns = plain_strips((uint32_t)a.W), nb = strip_blocks(a.n);
        const uint32_t tv = (uint32_t)t < ns ? a.strip_cnt[t] : 0u;   // strip totals
        uint32_t total;
        const uint32_t first = sort_block_scan<kTileThreads>(tv, strip_ws, &total);
        if (t == tc) {
            const uint32_t* offs = a.strip_cnt + ns + (size_t)tc * nb;
            const uint32_t blo = u0 / kStripBlock, bhi = (u1 - 1u) / kStripBlock;
            strip_range = make_uint2(first + offs[blo], first + (bhi + 1 < nb ? offs[bhi + 1] : tv));
        }
        __syncthreads();
        if (total <= a.strip_cap) {
            from_list = true;
            p0 = strip_range.x;
            p1 = strip_range.y;
        }
    }

    // relevant samples of the block so far: the k-th goes to wave k mod 4, so
    // every lane adds the same terms in the same order however they are staged
    uint32_t ranked = 0;
    for (uint32_t base = p0; base < p1; base += kChunk) {
        // stage the chunk's samples that can reach the tile, in window order
        float4 rec[kChunkPer];
        float xsu[kChunkPer];
        uint32_t reach = 0u;                                    // 4 bits per u
        uint64_t keep[kChunkPer];
#pragma unroll
        for (int u = 0; u < kChunkPer; u++) {
            const uint32_t q = base + (uint32_t)(u * kTileThreads + t);
            const uint32_t ql = min(q, p1 - 1u);                // in bounds; lanes past p1 keep nothing
            uint32_t si;
            if (from_list) {
                const uint2 en = a.strip_ent[ql];
                si = en.x;
                xsu[u] = __uint_as_float(en.y);
            } else {
                si = ql;
                xsu[u] = sample_xs(a, o_b, a.coord[ql]);
            }
            // the blocks it reaches (window index range and bearing box of
            // each block); only a sample that reaches one is staged, and few
            // samples of a window reach the tile, so only theirs are loaded
            uint32_t bits = 0u;
            if (q < p1) {
#pragma unroll
                for (int b = 0; b < kBlocks; b++) {
                    const bool in = si >= blo[b] && si < bhi[b] && xsu[u] >= bxlo[b] && xsu[u] <= bxhi[b];
                    bits |= (uint32_t)in << b;
                }
            }
            const bool kp = bits != 0u;
            reach |= bits << (4 * u);
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (kp) v = a.sorted[si];
            v.w = __uint_as_float(si);
            keep[u] = __ballot(kp);
            rec[u] = v;
            if (lane == 0) counts[u * kTileWaves + w] = (uint32_t)__popcll(keep[u]);
        }
        __syncthreads();
        // staged position of each (u, wave) run: a scan of the counts in
        // (u, wave) order across the lanes of every wave (every lane takes
        // part in the shuffles: a lane reads nothing from an inactive one)
        const uint32_t cv = lane < kChunkPer * kTileWaves ? counts[lane] : 0u;
        const uint32_t incl = sort_wave_scan(cv, (uint32_t)lane);
        const uint32_t n = (uint32_t)__shfl((int)incl, kChunkPer * kTileWaves - 1, 64);
#pragma unroll
        for (int u = 0; u < kChunkPer; u++) {
            const uint32_t at = (uint32_t)__shfl((int)(incl - cv), u * kTileWaves + w, 64);
            if ((keep[u] >> lane) & 1u) {
                const uint32_t pos = at + popc_below(keep[u]);
                cs_rec[pos] = rec[u];
                cs_x[pos] = xsu[u];
                cs_m[pos] = (uint8_t)((reach >> (4 * u)) & 15u);
            }
        }
        __syncthreads();

        // this wave's share of its block: a quarter of the relevant samples of
        // every staged group of 64 (a chunk of the window often stages a
        // single group: every wave takes part in it), two samples at a time
        for (uint32_t g = 0; g < n; g += 64) {
            const bool rel = g + lane < n && ((cs_m[g + lane] >> blkw) & 1u);
            // this wave's share: the relevant samples of rank sub, sub + 4, ...
            const uint64_t all = __ballot(rel);
            uint64_t m = __ballot(rel && ((ranked + popc_below(all)) & (kWavesPerBlock - 1)) == (uint32_t)sub);
            ranked += (uint32_t)__popcll(all);
            while (m & (m - 1)) {                               // two or more left
                const uint32_t j1 = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                const uint32_t j2 = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                const float4 e1 = cs_rec[g + j1], e2 = cs_rec[g + j2];   // the same address in every lane
                const float x1 = cs_x[g + j1], x2 = cs_x[g + j2];
                term(e1, x1);
                term(e2, x2);
            }
            if (m) {
                const uint32_t j1 = (uint32_t)__builtin_ctzll(m);
                term(cs_rec[g + j1], cs_x[g + j1]);
            }
        }
        // the next chunk's staging overwrites counts only after every wave
        // has read them (before the barrier above), and the staged samples
        // only after its first barrier, which every wave reaches after this loop
    }
    // the block's four partial sums, added in wave order
    __syncthreads();
    float2* part = reinterpret_cast<float2*>(cs_rec);
    part[t] = make_float2(acc_re, acc_im);
    __syncthreads();
    if (sub == 0 && r < a.H && col_ok) {
        float2 v = part[t];
#pragma unroll
        for (int j = 1; j < kWavesPerBlock; j++) {
            const float2 o = part[t + 64 * j];
            v.x += o.x;
            v.y += o.y;
        }
        a.out[(size_t)r * a.W + c] = v;
    }
}

// 8 waves per SIMD (two workgroups per CU): at most 64 VGPRs
template <bool kTapsLds, int kChunkPer>
__global__ __launch_bounds__(kTileThreads, 8) void k_conv2d_plain(ConvArgs a)
{
    conv2d_plain_body<kTapsLds, kChunkPer>(a, blockIdx.x);
}

template <bool kTapsLds, int kChunkPer>
__global__ __launch_bounds__(kTileThreads, 8) void k_conv2d_plain_batch(const ConvArgs* __restrict__ args,
                                                                     BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const ConvArgs a = args[p];   // a private copy: the body's stores cannot alias it
    conv2d_plain_body<kTapsLds, kChunkPer>(a, blk);
}

template <bool kTapsLds>
static void launch_gather(const ConvArgs& a, uint32_t blocks, bool chunk3, hipStream_t s)
{
    const size_t shm = plain2d_lds_bytes(a);
    if (chunk3) hipLaunchKernelGGL((k_conv2d_plain<kTapsLds, 3>), dim3(blocks), dim3(kTileThreads), shm, s, a);
    else hipLaunchKernelGGL((k_conv2d_plain<kTapsLds, 2>), dim3(blocks), dim3(kTileThreads), shm, s, a);
}

hipError_t launch_conv2d_plain(const ConvArgs& a, hipStream_t s)
{
    const uint32_t blocks = conv2d_plain_blocks(a);
    if (!blocks) return hipSuccess;
    // the band ranking ran with the lists when they are active
    if (!plain_lists_active(a)) hipLaunchKernelGGL(k_plain_rank, dim3(1), dim3(kRankThreads), 0, s, a, a.work);
    if (a.psf_in_lds) launch_gather<true>(a, blocks, a.chunk3 != 0, s);
    else launch_gather<false>(a, blocks, a.chunk3 != 0, s);
    return hipGetLastError();
}

hipError_t launch_plain_rank_batch(const ConvArgs* args, BatchMap m, uint32_t blocks, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_plain_rank_batch, dim3(blocks), dim3(kRankThreads), 0, s, args, m);
    return hipGetLastError();
}

hipError_t launch_conv2d_plain_batch(const ConvArgs* args, BatchMap m, uint32_t blocks, size_t shm,
                                     int variant, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    const dim3 g(blocks), b(kTileThreads);
    switch (variant & 3) {
    case 3: hipLaunchKernelGGL((k_conv2d_plain_batch<true, 3>), g, b, shm, s, args, m); break;
    case 2: hipLaunchKernelGGL((k_conv2d_plain_batch<false, 3>), g, b, shm, s, args, m); break;
    case 1: hipLaunchKernelGGL((k_conv2d_plain_batch<true, 2>), g, b, shm, s, args, m); break;
    default: hipLaunchKernelGGL((k_conv2d_plain_batch<false, 2>), g, b, shm, s, args, m); break;
    }
    return hipGetLastError();
}

void init_plain_attributes(int max_lds)
{
    set_dynamic_lds_limit((const void*)k_conv2d_plain<true, 2>, max_lds);
    set_dynamic_lds_limit((const void*)k_conv2d_plain<false, 2>, max_lds);
    set_dynamic_lds_limit((const void*)k_conv2d_plain<true, 3>, max_lds);
    set_dynamic_lds_limit((const void*)k_conv2d_plain<false, 3>, max_lds);
    set_dynamic_lds_limit((const void*)k_conv2d_plain_batch<true, 2>, max_lds);
    set_dynamic_lds_limit((const void*)k_conv2d_plain_batch<false, 2>, max_lds);
    set_dynamic_lds_limit((const void*)k_conv2d_plain_batch<true, 3>, max_lds);
    set_dynamic_lds_limit((const void*)k_conv2d_plain_batch<false, 3>, max_lds);
}

// ---- strip lists ---------------------------------------------------------------
// The tiles of one strip (one tile column) keep the samples whose xs lies in
// the strip's box, about (16 + PSF bearing support) / W of their band's
// window.  The lists give each strip those samples once per ping, in window
// order as (sample index, xs), so a tile stages only them instead of
// scanning the whole window: counts per block of kStripBlock sorted samples and strip, an
// exclusive scan of each strip's counts (its totals at the front), then the
// scatter of the entries.  Entry order is sample order (the scatter ranks
// samples by ballots, not atomics), so the lists are deterministic.
constexpr int kStripThreads = 1024;
static_assert(kStripBlock == (uint32_t)kStripThreads, "one sample per thread");
static_assert(kMaxStrips <= (uint32_t)kStripThreads && kMaxStrips <= (uint32_t)kTileThreads,
              "one thread per strip");
static_assert(kMaxStrips <= 1024, "a strip index fits 10 bits of a pair");

// every strip's (origin hi, origin lo, xs lo, xs hi), as the gather's tiles
// compute theirs (the minimum and maximum over the same 16 columns)
__device__ __forceinline__ void strip_boxes(const ConvArgs& a, uint32_t ns, float4* box)
{
    const uint32_t t = threadIdx.x;
    for (uint32_t c0 = 0; c0 < ns * kStripCols; c0 += kStripThreads) {
        const uint32_t c = c0 + t, j = c / kStripCols;
        float lo = INFINITY, hi = -INFINITY;
        float2 o = make_float2(0.f, 0.f);
        if (j < ns) {
            o = strip_origin(a, (int)j);
            if (c < (uint32_t)a.W) {
                const float xc = column_xc(a, o, a.bear_c[c]);
                lo = box_lo(xc);
                hi = box_hi(a, xc);
            }
        }
        for (int m = 1; m < (int)kStripCols; m <<= 1) {   // the strip's 16 lanes
            lo = fminf(lo, __shfl_xor(lo, m, 64));
            hi = fmaxf(hi, __shfl_xor(hi, m, 64));
        }
        if (j < ns && c % kStripCols == 0) box[j] = make_float4(o.x, o.y, lo, hi);
    }
}

__device__ __forceinline__ bool in_strip(const ConvArgs& a, const float4& bx, float2 x, float& xs)
{
    xs = sample_xs(a, make_float2(bx.x, bx.y), x);
    return xs >= bx.z && xs <= bx.w;
}

// the strips of the block's samples, four at a time (their boxes read from
// LDS together): fn(j, member mask of the wave, xs) for every strip j
template <typename Fn>
__device__ __forceinline__ void for_each_strip(const ConvArgs& a, const float4* box, uint32_t ns, bool ok,
                                               float2 x, Fn&& fn)
{
    for (uint32_t j0 = 0; j0 < ns; j0 += 4) {
        float4 bx[4];
#pragma unroll
        for (int q = 0; q < 4; q++) bx[q] = box[min(j0 + q, ns - 1)];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (j0 + q >= ns) break;
            float xs;
            const bool in = ok && in_strip(a, bx[q], x, xs);
            fn(j0 + q, in, __ballot(in), xs);
        }
    }
}

// block b: counts[j][b] = its samples in strip j
__device__ __forceinline__ void strip_count_body(const ConvArgs& a, uint32_t b)
{
    constexpr int NW = kStripThreads / 64;
    __shared__ float4 box[kMaxStrips];
    __shared__ uint32_t wcnt[NW][kMaxStrips];
    const uint32_t ns = plain_strips((uint32_t)a.W), nb = strip_blocks(a.n);
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t i = b * kStripBlock + t;
    const bool ok = i < a.n;
    const float2 x = a.coord[min(i, a.n - 1u)];
    strip_boxes(a, ns, box);
    __syncthreads();
    for_each_strip(a, box, ns, ok, x, [&](uint32_t j, bool, uint64_t m, float) {
        if (lane == 0) wcnt[w][j] = (uint32_t)__popcll(m);
    });
    __syncthreads();
    if (t < ns) {
        uint32_t c = 0;
        for (int q = 0; q < NW; q++) c += wcnt[q][t];
        a.strip_cnt[ns + (size_t)t * nb + b] = c;
    }
}

// strip j: its counts become offsets within its list, its total goes first
__device__ __forceinline__ void strip_scan_body(const ConvArgs& a, uint32_t j)
{
    __shared__ uint32_t ws[kStripThreads / 64];
    const uint32_t ns = plain_strips((uint32_t)a.W), nb = strip_blocks(a.n);
    uint32_t* row = a.strip_cnt + ns + (size_t)j * nb;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += kStripThreads) {
        const uint32_t b = b0 + threadIdx.x;
        const uint32_t v = b < nb ? row[b] : 0u;
        uint32_t total;
        const uint32_t ex = sort_block_scan<kStripThreads>(v, ws, &total);
        if (b < nb) row[b] = carry + ex;
        carry += total;
        __syncthreads();   // ws is reused by the next round
    }
    if (threadIdx.x == 0) a.strip_cnt[j] = carry;
}

// block b: its entries of every strip, at the strip's base + the block's
// offset, in sample order (waves in order, lanes in order).  The membership
// test runs once: each lane keeps its (strip, rank in its wave) pairs in LDS
// (a sample reaches at most kMaxStripMult strips) and writes them once the
// waves' positions are known.
constexpr int kMaxStripMult = 8;

__device__ __forceinline__ void strip_scatter_body(const ConvArgs& a, uint32_t b)
{
    constexpr int NW = kStripThreads / 64;
    __shared__ float4 box[kMaxStrips];
    __shared__ uint32_t wpos[NW][kMaxStrips];   // per wave: count, then first entry
    __shared__ uint16_t mine[kMaxStripMult][kStripThreads];   // strip << 6 | rank
    __shared__ uint32_t ws[NW];
    const uint32_t ns = plain_strips((uint32_t)a.W), nb = strip_blocks(a.n);
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t i = b * kStripBlock + t;
    const bool ok = i < a.n;
    const float2 x = a.coord[min(i, a.n - 1u)];
    const uint32_t tv = t < ns ? a.strip_cnt[t] : 0u;
    uint32_t total;
    const uint32_t first = sort_block_scan<kStripThreads>(tv, ws, &total);
    if (total > a.strip_cap) return;   // incomplete lists: the gather scans the windows
    strip_boxes(a, ns, box);
    __syncthreads();
    uint32_t k = 0;
    for_each_strip(a, box, ns, ok, x, [&](uint32_t j, bool in, uint64_t m, float) {
        if (lane == 0) wpos[w][j] = (uint32_t)__popcll(m);
        if (in && k < (uint32_t)kMaxStripMult) mine[k++][t] = (uint16_t)(j << 6 | popc_below(m));
    });
    __syncthreads();
    if (t < ns) {
        uint32_t r = first + a.strip_cnt[ns + (size_t)t * nb + b];
        for (int q = 0; q < NW; q++) {
            const uint32_t c = wpos[q][t];
            wpos[q][t] = r;
            r += c;
        }
    }
    __syncthreads();
    for (uint32_t q = 0; q < k; q++) {
        const uint32_t v = mine[q][t], j = v >> 6;
        float xs;
        (void)in_strip(a, box[j], x, xs);
        a.strip_ent[wpos[w][j] + (v & 63u)] = make_uint2(i, __float_as_uint(xs));
    }
}

__global__ __launch_bounds__(kStripThreads) void k_strip_count(ConvArgs a) { strip_count_body(a, blockIdx.x); }
__global__ __launch_bounds__(kStripThreads) void k_strip_scan(ConvArgs a) { strip_scan_body(a, blockIdx.x); }
__global__ __launch_bounds__(kStripThreads) void k_strip_scatter(ConvArgs a) { strip_scatter_body(a, blockIdx.x); }

__global__ __launch_bounds__(kStripThreads) void k_strip_count_batch(const ConvArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const ConvArgs a = args[p];
    strip_count_body(a, blk);
}
__global__ __launch_bounds__(kStripThreads) void k_strip_scan_batch(const ConvArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const ConvArgs a = args[p];
    strip_scan_body(a, blk);
}
__global__ __launch_bounds__(kStripThreads) void k_strip_scatter_batch(const ConvArgs* __restrict__ args, BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const ConvArgs a = args[p];
    strip_scatter_body(a, blk);
}

// One ping's lists with two launches folded in: the Binner windows run as
// extra workgroups of the counts (both only need the sort), the band ranking
// as an extra workgroup of the scan (it needs the windows).
__global__ __launch_bounds__(kStripThreads) void k_strip_count_windows(ConvArgs a, BinArgs w)
{
    const uint32_t nb = strip_blocks(a.n);
    if (blockIdx.x < nb) strip_count_body(a, blockIdx.x);
    else windows_body(w, blockIdx.x - nb);
}
__global__ __launch_bounds__(kStripThreads) void k_strip_scan_rank(ConvArgs a)
{
    if (blockIdx.x < plain_strips((uint32_t)a.W)) strip_scan_body(a, blockIdx.x);
    else plain_rank_body(a, a.work);
}

bool plain_lists_active(const ConvArgs& a) { return a.strip_ent && a.n > 0 && a.W > 0 && a.H > 0; }

hipError_t launch_plain_bins(const BinArgs& w, const ConvArgs& a, hipStream_t s)
{
    if (!plain_lists_active(a)) return launch_windows(w, s);
    const uint32_t nb = strip_blocks(a.n), ns = plain_strips((uint32_t)a.W);
    const uint32_t wb = (w.N + kStripThreads - 1) / kStripThreads;
    hipLaunchKernelGGL(k_strip_count_windows, dim3(nb + wb), dim3(kStripThreads), 0, s, a, w);
    hipLaunchKernelGGL(k_strip_scan_rank, dim3(ns + 1), dim3(kStripThreads), 0, s, a);
    hipLaunchKernelGGL(k_strip_scatter, dim3(nb), dim3(kStripThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_strip_count_batch(const ConvArgs* args, BatchMap m, uint32_t blocks, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_strip_count_batch, dim3(blocks), dim3(kStripThreads), 0, s, args, m);
    return hipGetLastError();
}
hipError_t launch_strip_scan_batch(const ConvArgs* args, BatchMap m, uint32_t blocks, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_strip_scan_batch, dim3(blocks), dim3(kStripThreads), 0, s, args, m);
    return hipGetLastError();
}
hipError_t launch_strip_scatter_batch(const ConvArgs* args, BatchMap m, uint32_t blocks, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_strip_scatter_batch, dim3(blocks), dim3(kStripThreads), 0, s, args, m);
    return hipGetLastError();
}

// Strip boxes in the bearing coordinate x: [x of the first column - supp,
// x of the last + supp], widened by 2 tap pitches (more than the tiles'
// 1/64 tap and FP32 rounding), so the count bounds what the device lists hold.
uint32_t strip_multiplicity(const double* x, uint32_t W, double supp, double pitch)
{
    const uint32_t ns = plain_strips(W);
    if (W == 0 || ns > kMaxStrips) return 0u;
    for (uint32_t c = 1; c < W; c++)
        if (!(x[c] > x[c - 1])) return 0u;   // the boxes are ordered only for ascending columns
    const double wide = supp + 2.0 * pitch;
    uint32_t mult = 0;
    for (uint32_t j = 0, k = 0; j < ns; j++) {
        // the point x_hi(j): boxes j .. k - 1 contain it (boxes start in order)
        const double hi = x[std::min(W, (j + 1) * kStripCols) - 1] + wide;
        k = std::max(k, j);
        while (k < ns && x[k * kStripCols] - wide <= hi) k++;
        mult = std::max(mult, k - j);
    }
    // lists only where they save most of the staging, at a bounded size
    return mult <= (uint32_t)kMaxStripMult && 2 * mult <= ns ? mult : 0u;
}

// 1D: one wave per output range bin, FP32 partial sums reduced across the
// wave (the reference reduces a 512-thread block in shared memory).
__device__ __forceinline__ void conv1d_plain_body(const ConvArgs& a, uint32_t blk)
{
    const int lane = threadIdx.x & 63;
    const int r = (int)blk * 4 + (threadIdx.x >> 6);
    if (r >= a.H) return;
    const uint2 win = a.win[r];
    const float2 rc = a.range_c[r];
    float ar = 0.f, ai = 0.f;
    for (uint32_t k = lane; k < win.y; k += 64) {
        const float4 rec = a.sorted[win.x + k];
        const float dr = (rec.z - rc.x) - rc.y;
        const float yB = dr * a.inv_pr + a.half_r;
        if (!(yB > -1.0f && yB < (float)a.Tr)) continue;
        const float2 K = psf_bilinear(a.psf, a.Tr, 1, 0.0f, yB);
        ar += rec.x * K.x - rec.y * K.y;
        ai += rec.x * K.y + rec.y * K.x;
    }
    for (int o = 32; o > 0; o >>= 1) {
        ar += __shfl_xor(ar, o, 64);
        ai += __shfl_xor(ai, o, 64);
    }
    if (lane == 0) a.out[r] = make_float2(ar, ai);
}

__global__ __launch_bounds__(256) void k_conv1d_plain(ConvArgs a) { conv1d_plain_body(a, blockIdx.x); }

__global__ __launch_bounds__(256) void k_conv1d_plain_batch(const ConvArgs* __restrict__ args,
                                                            BatchMap m)
{
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const ConvArgs a = args[p];
    conv1d_plain_body(a, blk);
}

uint32_t conv1d_plain_blocks(const ConvArgs& a) { return (uint32_t)(a.H + 3) / 4; }

hipError_t launch_conv1d_plain(const ConvArgs& a, hipStream_t s)
{
    if (a.H == 0) return hipSuccess;
    hipLaunchKernelGGL(k_conv1d_plain, dim3(conv1d_plain_blocks(a)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_conv1d_plain_batch(const ConvArgs* args, BatchMap m, uint32_t blocks,
                                     hipStream_t s)
{
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_conv1d_plain_batch, dim3(blocks), dim3(256), 0, s, args, m);
    return hipGetLastError();
}

}  // namespace rtsim
