// Grid-form correlation on MFMA: out(r, b) = sum_{i < Tr, j < Tb} T[i][j] G[r + i][b + j]
// (the extended grid G comes from k_deposit, kernels_grid.hip).
//
// MFMA mapping (v_mfma_f32_16x16x32_{bf16,f16}, wave64; lane l holds
// A[l % 16][8 (l / 16) + p], B[8 (l / 16) + p][l % 16], D[4 (l / 16) + e][l % 16]):
//   M = 16 output range rows, N = 16 output bearings, K = 32 grid columns.
//   For range tap i, A[m][k] = G[r0 + m + i][c + k] and the Toeplitz
//   B_i[k][n] = T[i][k - n] (zero outside [0, Tb)), so out += A B_i.
//   Complex products use the real embedding (re = ArBr - AiBi,
//   im = ArBi + AiBr); every real product is split hi+lo,
//   X Y ~ XhYh + XlYh + XhYl, with FP32 accumulation.
// Scaling: each workgroup scales its G tile by a power of two (max |G| in
// [2^13, 2^14)), found from the staged values themselves, and the taps
// carry a host power of two, so FP16 operands stay in range; the epilogue
// divides both out exactly.  Rows whose reference window is empty
// (Binner.cu:96-97) are written as zero.
//
// Tile shapes (template WR x WC waves, NT 16-wide N tiles per wave): a
// workgroup computes 16 WR output rows x 16 NT WC bearings with 4 waves;
// the shape is chosen per sensor by timing (Sensor::autotune).
// LDS (16-bit elements):
//   A: 4 planes (re_hi, re_lo, im_hi, im_lo) x AR rows x AS stride
//   B: the Toeplitz rows of TCH range taps, copied from the prebuilt table
//      [2 shifts][4 planes][Tr][BL] (shift c holds Tpad[x + c],
//      Tpad[x] = T[x - 15]); the two shifts keep every lane's 8-element B
//      fragment 4-byte aligned.
// Staging: every thread issues all its grid loads (float4 = two complex
// cells) and the B copy at once, so a workgroup pays one memory round trip
// before its MFMA loop.  Per K step a wave reads its A fragments
// (ds_read_b128) and the shared B fragment with pointer increments only;
// fragments are double-buffered so the next step's reads overlap the
// current step's MFMAs.
#include "device_common.h"
#include "grid_internal.h"

namespace rtsim {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kRegCells = 16;   // float4 grid loads per thread held in registers

__device__ __forceinline__ uint16_t bf16_rne(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

template <bool FP16>
__device__ __forceinline__ void split_hi_lo(float x, uint16_t& hi, uint16_t& lo)
{
    if (FP16) {
        _Float16 h = (_Float16)x;
        _Float16 l = (_Float16)(x - (float)h);
        hi = __builtin_bit_cast(uint16_t, h);
        lo = __builtin_bit_cast(uint16_t, l);
    } else {
        hi = bf16_rne(x);
        lo = bf16_rne(x - bf16_f32(hi));
    }
}

template <bool FP16>
__device__ __forceinline__ f32x4 mfma(u32x4 a, u32x4 b, f32x4 c)
{
    if (FP16)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                      __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

template <int NT>
struct Frags {
    u32x4 a[NT][4];  // [n tile][plane]
    u32x4 b[4];      // [plane]
};

template <bool FP16, int NT>
__device__ __forceinline__ void mma_step(const Frags<NT>& f, f32x4 (&re)[NT], f32x4 (&im)[NT])
{
    const u32x4 sgn = {0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    const u32x4 nbh = f.b[2] ^ sgn, nbl = f.b[3] ^ sgn;
    // re += Ar Br - Ai Bi, im += Ar Bi + Ai Br; chains interleaved across n
#pragma unroll
    for (int n = 0; n < NT; n++) { re[n] = mfma<FP16>(f.a[n][0], f.b[0], re[n]); im[n] = mfma<FP16>(f.a[n][0], f.b[2], im[n]); }
#pragma unroll
    for (int n = 0; n < NT; n++) { re[n] = mfma<FP16>(f.a[n][1], f.b[0], re[n]); im[n] = mfma<FP16>(f.a[n][1], f.b[2], im[n]); }
#pragma unroll
    for (int n = 0; n < NT; n++) { re[n] = mfma<FP16>(f.a[n][0], f.b[1], re[n]); im[n] = mfma<FP16>(f.a[n][0], f.b[3], im[n]); }
#pragma unroll
    for (int n = 0; n < NT; n++) { re[n] = mfma<FP16>(f.a[n][2], nbh, re[n]); im[n] = mfma<FP16>(f.a[n][2], f.b[0], im[n]); }
#pragma unroll
    for (int n = 0; n < NT; n++) { re[n] = mfma<FP16>(f.a[n][3], nbh, re[n]); im[n] = mfma<FP16>(f.a[n][3], f.b[0], im[n]); }
#pragma unroll
    for (int n = 0; n < NT; n++) { re[n] = mfma<FP16>(f.a[n][2], nbl, re[n]); im[n] = mfma<FP16>(f.a[n][2], f.b[1], im[n]); }
}

template <int NT>
__device__ __forceinline__ void load_frags(Frags<NT>& f, const uint16_t* aptr, int plane_stride,
                                           const uint16_t* bptr, int bplane_stride)
{
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(bptr + p * bplane_stride);
        f.b[p] = (u32x4){src[0], src[1], src[2], src[3]};
    }
#pragma unroll
    for (int n = 0; n < NT; n++)
#pragma unroll
        for (int p = 0; p < 4; p++)
            f.a[n][p] = *reinterpret_cast<const u32x4*>(aptr + p * plane_stride + 16 * n);
}

// q = floor(k / d) for 0 <= k < 2^20, d >= 1: (k + 0.5) / d lies at least
// 0.5 / d from an integer, and the float product's error is below
// (k + 0.5) / d * 2^-23 < 0.5 / d.
__device__ __forceinline__ int small_div(int k, float inv_d)
{
    return (int)(((float)k + 0.5f) * inv_d);
}

// Copies the Toeplitz rows of range taps [i0, i0 + tc) into LDS B: for each
// (shift, plane) run a contiguous tc * BL elements, moved as uint4.
__device__ __forceinline__ void copy_toeplitz(const CorrArgs& a, uint16_t* B, int i0, int tc, int t)
{
    const int n16 = tc * a.BL / 8;                 // uint4 per (shift, plane) run
    const float inv_n16 = 1.0f / (float)n16;
    const uint4* src = reinterpret_cast<const uint4*>(a.toep);
    uint4* dst = reinterpret_cast<uint4*>(B);
    const int run_src = a.Tr * a.BL / 8, run_dst = a.TCH * a.BL / 8, off = i0 * a.BL / 8;
    for (int k = t; k < 8 * n16; k += 256) {
        const int cp = small_div(k, inv_n16), o = k - cp * n16;
        dst[cp * run_dst + o] = src[cp * run_src + off + o];
    }
}

template <bool FP16, int WR, int NT>
__device__ __forceinline__ void correlate_body(const CorrArgs& a, int bx, int by)
{
    constexpr int WC = 4 / WR;
    constexpr int ROWS = 16 * WR, COLS = 16 * NT * WC;
    extern __shared__ uint16_t lds[];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wr = wave % WR, wc = wave / WR;
    const int m = lane & 15, kg = lane >> 4;
    const int r0 = by * ROWS;
    const int c0 = bx * COLS;
    const int AR = ROWS + a.Tr - 1;
    const int AC = COLS - 16 + 32 * a.KB;   // columns [c0, c0 + COLS - 16 + 32 KB)
    const int PR = AC / 2;                  // float4 loads per A row
    const int total = AR * PR;
    const float inv_pr = 1.0f / (float)PR;
    const int AS = a.AS;
    const int bps = a.TCH * a.BL;           // B plane stride (elements)
    uint16_t* A = lds;
    uint16_t* B = lds + 4 * AR * AS;
    __shared__ float red[4];

    // window sizes of this lane's output rows, for the epilogue mask
    uint32_t wsz[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int rr = r0 + 16 * wr + 4 * kg + e;
        wsz[e] = a.win[min(rr, a.H - 1)].y;
    }

    // ---- grid tile loads: all in flight together ------------------------------
    const float4* gtile = reinterpret_cast<const float4*>(a.grid + (size_t)r0 * a.GW + c0);
    const int grow4 = a.GW / 2;             // float4 per grid row
    const bool in_regs = total <= kRegCells * 256;
    float4 v[kRegCells];
    float mx = 0.f;
    if (in_regs) {
        const int nreg = (total + 255) / 256;
#pragma unroll
        for (int j = 0; j < kRegCells; j++) {
            if (j < nreg) {
                const int k = min(t + 256 * j, total - 1);
                const int rr = small_div(k, inv_pr);
                v[j] = gtile[rr * grow4 + (k - rr * PR)];
            }
        }
        copy_toeplitz(a, B, 0, min(a.TCH, a.Tr), t);
#pragma unroll
        for (int j = 0; j < kRegCells; j++) {
            if (j < nreg && t + 256 * j < total)
                mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v[j].x), fabsf(v[j].y)),
                                     fmaxf(fabsf(v[j].z), fabsf(v[j].w))));
        }
    } else {
        copy_toeplitz(a, B, 0, min(a.TCH, a.Tr), t);
        for (int k = t; k < total; k += 256) {
            const int rr = small_div(k, inv_pr);
            const float4 g = gtile[rr * grow4 + (k - rr * PR)];
            mx = fmaxf(mx, fmaxf(fmaxf(fabsf(g.x), fabsf(g.y)), fmaxf(fabsf(g.z), fabsf(g.w))));
        }
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) red[wave] = mx;
    __syncthreads();
    mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    if (!(mx > 0.f)) {
        for (int k = t; k < ROWS * COLS; k += 256) {
            const int rr = r0 + k / COLS, cc = c0 + k % COLS;
            if (rr < a.H && cc < a.W) a.out[(size_t)rr * a.W + cc] = make_float2(0.f, 0.f);
        }
        return;
    }
    int e;
    frexpf(mx, &e);                                       // mx in [2^(e-1), 2^e)
    const float scale = ldexpf(1.0f, min(14 - e, 126));   // mx * scale in [2^13, 2^14)

    // ---- A: scaled, hi/lo split, two complex cells per 32-bit LDS word pair ----
    {
        const int ps = AR * AS / 2;   // plane stride in words
        auto put = [&](int k, const float4& g) {
            const int rr = small_div(k, inv_pr), cp = k - rr * PR;
            uint16_t h0, l0, h1, l1, h2, l2, h3, l3;
            split_hi_lo<FP16>(g.x * scale, h0, l0);   // re, column 2cp
            split_hi_lo<FP16>(g.z * scale, h1, l1);   // re, column 2cp + 1
            split_hi_lo<FP16>(g.y * scale, h2, l2);   // im, column 2cp
            split_hi_lo<FP16>(g.w * scale, h3, l3);   // im, column 2cp + 1
            uint32_t* base = reinterpret_cast<uint32_t*>(A + rr * AS) + cp;
            base[0] = (uint32_t)h0 | ((uint32_t)h1 << 16);
            base[ps] = (uint32_t)l0 | ((uint32_t)l1 << 16);
            base[2 * ps] = (uint32_t)h2 | ((uint32_t)h3 << 16);
            base[3 * ps] = (uint32_t)l2 | ((uint32_t)l3 << 16);
        };
        if (in_regs) {
#pragma unroll
            for (int j = 0; j < kRegCells; j++)
                if (t + 256 * j < total) put(t + 256 * j, v[j]);
        } else {
            for (int k = t; k < total; k += 256) {
                const int rr = small_div(k, inv_pr);
                put(k, gtile[rr * grow4 + (k - rr * PR)]);
            }
        }
    }

    f32x4 re[NT], im[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) {
        re[n] = (f32x4){0.f, 0.f, 0.f, 0.f};
        im[n] = re[n];
    }
    // per-lane fragment origins (advanced by pointer increments)
    const int cpy = (15 - m) & 1;                       // parity of 32q + 8kg + 15 - m
    const uint16_t* a_lane = A + (16 * wr + m) * AS + 16 * NT * wc + 8 * kg;
    const uint16_t* b_lane = B + cpy * 4 * bps + (8 * kg + 15 - m - cpy);

    for (int i0 = 0; i0 < a.Tr; i0 += a.TCH) {
        const int tc = min(a.TCH, a.Tr - i0);
        if (i0 > 0) {
            __syncthreads();   // previous chunk consumed
            copy_toeplitz(a, B, i0, tc, t);
        }
        __syncthreads();       // A (first chunk) and B staged

        // K steps (tap ii, K block q) of this chunk, ping-pong fragments
        const int nsteps = tc * a.KB;
        const uint16_t* ap = a_lane + i0 * AS;
        const uint16_t* bp = b_lane;
        int q = 0;
        auto advance = [&]() {
            if (++q == a.KB) {
                q = 0;
                ap += AS - 32 * (a.KB - 1);
                bp += a.BL - 32 * (a.KB - 1);
            } else {
                ap += 32;
                bp += 32;
            }
        };
        Frags<NT> f0, f1;
        load_frags<NT>(f0, ap, AR * AS, bp, bps);
        advance();
        for (int step = 0; step < nsteps; step += 2) {
            if (step + 1 < nsteps) {
                load_frags<NT>(f1, ap, AR * AS, bp, bps);
                advance();
            }
            mma_step<FP16, NT>(f0, re, im);
            if (step + 1 < nsteps) {
                if (step + 2 < nsteps) {
                    load_frags<NT>(f0, ap, AR * AS, bp, bps);
                    advance();
                }
                mma_step<FP16, NT>(f1, re, im);
            }
        }
    }

    // ---- epilogue: undo scales, empty-window mask, store ---------------------
    const float inv = 1.0f / (scale * a.tap_scale);
#pragma unroll
    for (int n = 0; n < NT; n++) {
        const int cc = c0 + 16 * (NT * wc + n) + m;
#pragma unroll
        for (int e2 = 0; e2 < 4; e2++) {
            const int rr = r0 + 16 * wr + 4 * kg + e2;
            if (rr < a.H && cc < a.W) {
                float2 v2 = make_float2(re[n][e2] * inv, im[n][e2] * inv);
                if (wsz[e2] == 0) v2 = make_float2(0.f, 0.f);
                a.out[(size_t)rr * a.W + cc] = v2;
            }
        }
    }
}

template <bool FP16, int WR, int NT>
__global__ __launch_bounds__(256) void k_correlate(CorrArgs a)
{
    correlate_body<FP16, WR, NT>(a, blockIdx.x, blockIdx.y);
}

template <bool FP16, int WR, int NT>
__global__ __launch_bounds__(256) void k_correlate_batch(const CorrArgs* __restrict__ args, BatchMap m)
{
    constexpr int COLS = 16 * NT * (4 / WR);
    uint32_t blk, nblk;
    const uint32_t p = batch_locate(m, blk, nblk);
    const CorrArgs& a = args[p];
    const int gx = (a.W + COLS - 1) / COLS;
    correlate_body<FP16, WR, NT>(a, (int)blk % gx, (int)blk / gx);
}

// tile shapes: {WR, NT}; rows = 16 WR, cols = 16 NT (4 / WR):
// 64x32, 32x32, 32x64, 64x16
static const int kShapes[kCorrShapes][2] = {{4, 2}, {2, 1}, {2, 2}, {4, 1}};

void corr_tile(int shape, int* rows, int* cols)
{
    const int WR = kShapes[shape][0], NT = kShapes[shape][1];
    *rows = 16 * WR;
    *cols = 16 * NT * (4 / WR);
}

size_t correlate_lds_bytes(const CorrArgs& a)
{
    int rows, cols;
    corr_tile(a.shape, &rows, &cols);
    const int AR = rows + a.Tr - 1;
    return sizeof(uint16_t) * ((size_t)4 * AR * a.AS + (size_t)8 * a.TCH * a.BL);
}

int correlate_chunk(const CorrArgs& a, size_t max_lds)
{
    CorrArgs c = a;
    auto largest = [&](size_t budget) {
        int lo = 0;
        for (int tc = a.Tr; tc >= 1; tc--) {
            c.TCH = tc;
            if (correlate_lds_bytes(c) <= budget) { lo = tc; break; }
        }
        return lo;
    };
    int tc = largest(kCorrLdsTarget);
    if (tc == 0) tc = largest(max_lds);
    return tc;
}

template <bool FP16>
static void launch_shape(const CorrArgs& a, dim3 grid, size_t shm, hipStream_t s)
{
    switch (a.shape) {
    case 0: hipLaunchKernelGGL((k_correlate<FP16, 4, 2>), grid, dim3(256), shm, s, a); break;
    case 1: hipLaunchKernelGGL((k_correlate<FP16, 2, 1>), grid, dim3(256), shm, s, a); break;
    case 2: hipLaunchKernelGGL((k_correlate<FP16, 2, 2>), grid, dim3(256), shm, s, a); break;
    default: hipLaunchKernelGGL((k_correlate<FP16, 4, 1>), grid, dim3(256), shm, s, a); break;
    }
}

template <bool FP16>
static void launch_shape_batch(int shape, const CorrArgs* args, BatchMap m, dim3 grid, size_t shm,
                               hipStream_t s)
{
    switch (shape) {
    case 0: hipLaunchKernelGGL((k_correlate_batch<FP16, 4, 2>), grid, dim3(256), shm, s, args, m); break;
    case 1: hipLaunchKernelGGL((k_correlate_batch<FP16, 2, 1>), grid, dim3(256), shm, s, args, m); break;
    case 2: hipLaunchKernelGGL((k_correlate_batch<FP16, 2, 2>), grid, dim3(256), shm, s, args, m); break;
    default: hipLaunchKernelGGL((k_correlate_batch<FP16, 4, 1>), grid, dim3(256), shm, s, args, m); break;
    }
}

uint32_t correlate_blocks(const CorrArgs& a)
{
    int rows, cols;
    corr_tile(a.shape, &rows, &cols);
    return (uint32_t)(((a.W + cols - 1) / cols) * ((a.H + rows - 1) / rows));
}

hipError_t launch_correlate_batch(bool fp16, int shape, const CorrArgs* args, BatchMap m,
                                  uint32_t blocks, size_t shm, hipStream_t s)
{
    if (!blocks) return hipSuccess;
    if (fp16) launch_shape_batch<true>(shape, args, m, dim3(blocks), shm, s);
    else launch_shape_batch<false>(shape, args, m, dim3(blocks), shm, s);
    return hipGetLastError();
}

hipError_t launch_correlate(const CorrArgs& a, bool fp16, hipStream_t s)
{
    int rows, cols;
    corr_tile(a.shape, &rows, &cols);
    dim3 grid((a.W + cols - 1) / cols, (a.H + rows - 1) / rows);
    const size_t shm = correlate_lds_bytes(a);
    if (fp16) launch_shape<true>(a, grid, shm, s);
    else launch_shape<false>(a, grid, shm, s);
    return hipGetLastError();
}

void init_corr_attributes(int max_lds)
{
    set_dynamic_lds_limit((const void*)k_correlate<true, 4, 2>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<true, 2, 1>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<true, 2, 2>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<true, 4, 1>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<false, 4, 2>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<false, 2, 1>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<false, 2, 2>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<false, 4, 1>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<true, 4, 2>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<true, 2, 1>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<true, 2, 2>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<true, 4, 1>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<false, 4, 2>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<false, 2, 1>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<false, 2, 2>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate_batch<false, 4, 1>, max_lds);
}

}  // namespace rtsim
