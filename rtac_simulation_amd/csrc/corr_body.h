// Grid-form correlation body on MFMA (out(r, b) = sum T[i][j] G[r + i][b + j]):
// the Toeplitz implicit GEMMs of one output tile, with staging, hi+lo
// split products, partial-sum reduction and the masked epilogue.  The
// algorithm, LDS layout and tile shapes are described in kernels_corr.hip.
// Shared by the correlation launches (kernels_corr.hip) and the persistent
// launch (kernels_persist.hip).
#pragma once

#include "device_common.h"
#include "grid_internal.h"

namespace rtsim {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kRegQuads = 8;    // staged 2 x 2 cell items per thread held in registers
constexpr int kChains = 6;      // independent accumulators (3 re, 3 im)

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
__device__ __forceinline__ f32x16 mfma(u32x4 a, u32x4 b, f32x16 c)
{
    if (FP16)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a),
                                                      __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// One K step: A = taps (planes a_hi, a_lo, b_hi, b_lo of T = a + ib), B = grid
// (c_hi, c_lo, d_hi, d_lo of G = c + id); acc[0..2] collect re, acc[3..5] im.
template <bool FP16>
__device__ __forceinline__ void mma_step(const u32x4 (&A)[4], const u32x4 (&B)[4],
                                         f32x16 (&acc)[kChains])
{
    const u32x4 sgn = {0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    const u32x4 nbh = A[2] ^ sgn, nbl = A[3] ^ sgn;
    acc[0] = mfma<FP16>(A[0], B[0], acc[0]);   // ah ch
    acc[3] = mfma<FP16>(A[0], B[2], acc[3]);   // ah dh
    acc[1] = mfma<FP16>(A[1], B[0], acc[1]);   // al ch
    acc[4] = mfma<FP16>(A[1], B[2], acc[4]);   // al dh
    acc[2] = mfma<FP16>(A[0], B[1], acc[2]);   // ah cl
    acc[5] = mfma<FP16>(A[0], B[3], acc[5]);   // ah dl
    acc[0] = mfma<FP16>(nbh, B[2], acc[0]);    // -bh dh
    acc[3] = mfma<FP16>(A[2], B[0], acc[3]);   // bh ch
    acc[1] = mfma<FP16>(nbl, B[2], acc[1]);    // -bl dh
    acc[4] = mfma<FP16>(A[3], B[0], acc[4]);   // bl ch
    acc[2] = mfma<FP16>(nbh, B[3], acc[2]);    // -bh dl
    acc[5] = mfma<FP16>(A[2], B[1], acc[5]);   // bh cl
}

// q = floor(k / d) for 0 <= k < 2^20, d >= 1: (k + 0.5) / d lies at least
// 0.5 / d from an integer, and the float product's error is below
// (k + 0.5) / d * 2^-23 < 0.5 / d.
__device__ __forceinline__ int small_div(int k, float inv_d)
{
    return (int)(((float)k + 0.5f) * inv_d);
}

// staged grid columns of a workgroup of TC 32-column tiles: 32 TC + Tb - 1,
// rounded up to even (the grid is read two columns at a time)
__host__ __device__ inline int corr_cols(int TC, int Tb)
{
    const int c = 32 * TC + Tb - 1;
    return c + (c & 1);
}

// Stride (elements) between the two shifted copies of the tap table, padded to
// 16 dwords modulo the 32 LDS banks: lanes 2k and 2k + 1 read the same dword
// offset of the two copies, which then falls in distinct banks.
__host__ __device__ inline int corr_copy_stride(int Tb, int L)
{
    const int dw = (4 * Tb * L / 2 + 31) / 32 * 32 + 16;
    return 2 * dw;
}

// NW waves per workgroup: 4, or 8 (W8) when the shape's LDS leaves room for
// one workgroup per CU only: two waves per SIMD then hide each other's LDS
// latency in the MFMA loop (one wave alone ran it at about half the MFMA rate)
template <int TR, int TC, bool W8>
struct CorrGeom {
    static constexpr int NW = W8 ? 8 : 4;
    static constexpr int NT = 64 * NW;               // threads
    static constexpr int JS = NW / (TR * TC);        // waves per output tile
};

// Output tile tt at (tt % gx, tt / gx) in units of 32 TC bearings x 32 TR rows.
template <bool FP16, int TR, int TC, bool W8>
__device__ __forceinline__ void correlate_body(const CorrArgs& a, int tt, int gx)
{
    constexpr int JS = CorrGeom<TR, TC, W8>::JS;
    constexpr int NW = CorrGeom<TR, TC, W8>::NW;
    constexpr int NT = CorrGeom<TR, TC, W8>::NT;
    extern __shared__ __align__(16) uint16_t lds[];
    __shared__ float red[NW];
    __shared__ uint32_t s_win[32 * TR];          // window sizes of the tile's output rows
    // the wave index in a scalar register: everything derived from it (tile,
    // tap subset, step range, LDS offsets) stays scalar
    const int t = threadIdx.x, lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int tile = wave / JS, js = wave % JS;   // this wave's output tile and tap subset
    const int tr = tile / TC, tc = tile % TC;
    const int KR = a.KR, RS = a.RS, L = a.L;
    const int SR = 32 * (TR - 1) + 16 * KR;      // staged grid rows [r0, r0 + SR)
    const int CW = corr_cols(TC, a.Tb);          // staged grid columns [c0, c0 + CW)
    const int gps = CW * RS;                     // G plane stride (elements)
    uint16_t* Gt = lds;
    uint16_t* Tt = lds + 4 * gps;                // tap table [2][4][Tb][L]
    const int cs = corr_copy_stride(a.Tb, L);   // elements between the two shifted copies
    const int tsize = 2 * cs;
    const int hh = lane >> 5, col = lane & 31;

    // staging: 2 x 2 cells (rows r, r + 1, columns c, c + 1) per item, all in flight
    const int CQ = CW / 2;                       // column pairs
    const int nquads = (SR / 2) * CQ;
    const float inv_cq = 1.0f / (float)CQ;
    const int grow4 = a.GW / 2;                  // float4 per grid row
    const bool in_regs = nquads <= kRegQuads * NT;
    float4 v0[kRegQuads], v1[kRegQuads];
    auto quad_of = [&](int k, int& rp, int& cq) {
        rp = small_div(k, inv_cq);
        cq = k - rp * CQ;
    };
    auto quad_max = [](const float4& g0, const float4& g1) {
        return fmaxf(fmaxf(fmaxf(fabsf(g0.x), fabsf(g0.y)), fmaxf(fabsf(g0.z), fabsf(g0.w))),
                     fmaxf(fmaxf(fabsf(g1.x), fabsf(g1.y)), fmaxf(fabsf(g1.z), fabsf(g1.w))));
    };
    const int r0 = (tt / gx) * 32 * TR, c0 = (tt % gx) * 32 * TC;
    const float4* gtile = reinterpret_cast<const float4*>(a.grid + (size_t)r0 * a.GW + c0);
    auto load_regs = [&]() {
#pragma unroll
        for (int u = 0; u < kRegQuads; u++) {
            const int k = min(t + NT * u, nquads - 1);
            int rp, cq;
            quad_of(k, rp, cq);
            v0[u] = gtile[(size_t)(2 * rp) * grow4 + cq];
            v1[u] = gtile[(size_t)(2 * rp + 1) * grow4 + cq];
        }
    };
    if (in_regs) load_regs();   // before the table copy
    // the output rows' window sizes (the epilogue's mask), read under the staging
    if (t < 32 * TR) s_win[t] = a.win[min(r0 + t, a.H - 1)].y;
    {   // tap table: a contiguous copy, uint4 at a time
        const uint4* src = reinterpret_cast<const uint4*>(a.toep);
        uint4* dst = reinterpret_cast<uint4*>(Tt);
        for (int k = t; k < tsize / 8; k += NT) dst[k] = src[k];
    }

    const int m = lane & 31, h = lane >> 5;
    // A: lane (m, h) reads Tpad[16 kb + 8 h - m + 31 + p] of tap column j
    const int s0 = 8 * h - m + 31;
    const int cpy = s0 & 1;                              // 16 kb keeps the parity
    const uint16_t* a_lane = Tt + cpy * cs + (s0 - cpy);
    const int aps = a.Tb * L;                            // A plane stride (elements)
    // B: lane (n, h) reads rows 32 tr + 16 kb + 8 h + p of column 32 tc + n + j
    const uint16_t* b_lane = Gt + (32 * tc + m) * RS + 32 * tr + 8 * h;
    // the tile's Tb KR steps (tap j, K block kb), split evenly over its JS waves
    const int all = a.Tb * KR;
    const int sbeg = js * all / JS, nsteps = (js + 1) * all / JS - sbeg;

    {
        float mx = 0.f;
        if (in_regs) {
#pragma unroll
            for (int u = 0; u < kRegQuads; u++)
                if (t + NT * u < nquads) mx = fmaxf(mx, quad_max(v0[u], v1[u]));
        } else {
            for (int k = t; k < nquads; k += NT) {
                int rp, cq;
                quad_of(k, rp, cq);
                mx = fmaxf(mx, quad_max(gtile[(size_t)(2 * rp) * grow4 + cq],
                                        gtile[(size_t)(2 * rp + 1) * grow4 + cq]));
            }
        }
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
        if (lane == 0) red[wave] = mx;
        __syncthreads();   // also: the tap table is in LDS
        mx = red[0];
#pragma unroll
        for (int w2 = 1; w2 < NW; w2++) mx = fmaxf(mx, red[w2]);
        const bool zero = !(mx > 0.f);
        int e;
        frexpf(zero ? 1.f : mx, &e);                          // mx in [2^(e-1), 2^e)
        const float scale = ldexpf(1.0f, min(14 - e, 126));   // mx * scale in [2^13, 2^14)

        // ---- G planes: scaled, hi/lo split, rows (2 rp, 2 rp + 1) in one word ------
        if (!zero) {
            const int wps = gps / 2;   // plane stride in 32-bit words
            // column c of rows (2 rp, 2 rp + 1) holds cells g0, g1
            auto put_col = [&](uint32_t* base, float2 g0, float2 g1) {
                uint16_t h0, l0, h1, l1, h2, l2, h3, l3;
                split_hi_lo<FP16>(g0.x * scale, h0, l0);   // re, row 2 rp
                split_hi_lo<FP16>(g1.x * scale, h1, l1);   // re, row 2 rp + 1
                split_hi_lo<FP16>(g0.y * scale, h2, l2);   // im, row 2 rp
                split_hi_lo<FP16>(g1.y * scale, h3, l3);   // im, row 2 rp + 1
                base[0] = (uint32_t)h0 | ((uint32_t)h1 << 16);
                base[wps] = (uint32_t)l0 | ((uint32_t)l1 << 16);
                base[2 * wps] = (uint32_t)h2 | ((uint32_t)h3 << 16);
                base[3 * wps] = (uint32_t)l2 | ((uint32_t)l3 << 16);
            };
            auto put = [&](int k, const float4& g0, const float4& g1) {
                int rp, cq;
                quad_of(k, rp, cq);
                uint32_t* base = reinterpret_cast<uint32_t*>(Gt + 2 * cq * RS) + rp;
                put_col(base, make_float2(g0.x, g0.y), make_float2(g1.x, g1.y));
                put_col(base + RS / 2, make_float2(g0.z, g0.w), make_float2(g1.z, g1.w));
            };
            if (in_regs) {
#pragma unroll
                for (int u = 0; u < kRegQuads; u++)
                    if (t + NT * u < nquads) put(t + NT * u, v0[u], v1[u]);
            } else {
                for (int k = t; k < nquads; k += NT) {
                    int rp, cq;
                    quad_of(k, rp, cq);
                    put(k, gtile[(size_t)(2 * rp) * grow4 + cq],
                        gtile[(size_t)(2 * rp + 1) * grow4 + cq]);
                }
            }
        }
        if (zero) {
            for (int k = t; k < 32 * TR * 32 * TC; k += NT) {
                const int rr = r0 + k / (32 * TC), cc = c0 + k % (32 * TC);
                if (rr < a.H && cc < a.W) a.out[(size_t)rr * a.W + cc] = make_float2(0.f, 0.f);
            }
            return;
        }
        __syncthreads();

        // ---- MFMA loop over this wave's taps j = js, js + JS, ... and K blocks ------
        f32x16 acc[kChains];
#pragma unroll
        for (int c = 0; c < kChains; c++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[c][q] = 0.f;
        auto load = [&](int step, u32x4 (&A)[4], u32x4 (&B)[4]) {
            const int st = sbeg + step;
            const int j = st / KR, kb = st - j * KR;      // scalar: st is wave-uniform
            const uint16_t* ap = a_lane + (j * L + 16 * kb);
            const uint16_t* bp = b_lane + (j * RS + 16 * kb);
#pragma unroll
            for (int p = 0; p < 4; p++) {
                const uint32_t* src = reinterpret_cast<const uint32_t*>(ap + p * aps);
                A[p] = (u32x4){src[0], src[1], src[2], src[3]};
                B[p] = *reinterpret_cast<const u32x4*>(bp + p * gps);
            }
        };
        // the next step's LDS reads are interleaved with this step's MFMAs, so
        // that they complete under them (with one wave per SIMD nothing else
        // would hide their latency)
        auto interleave = []() {
#pragma unroll
            for (int k = 0; k < 12; k++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // one MFMA
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);   // one LDS read
            }
        };
        if (nsteps > 0) {
            u32x4 A0[4], B0[4], A1[4], B1[4];
            load(0, A0, B0);
            for (int step = 0; step < nsteps; step += 2) {
                load(min(step + 1, nsteps - 1), A1, B1);
                mma_step<FP16>(A0, B0, acc);
                interleave();
                if (step + 1 < nsteps) {
                    load(min(step + 2, nsteps - 1), A0, B0);
                    mma_step<FP16>(A1, B1, acc);
                    interleave();
                }
            }
        }
        f32x16 re = acc[0] + acc[1] + acc[2];
        f32x16 im = acc[3] + acc[4] + acc[5];

        // ---- partial sums of the JS waves of a tile, through LDS (over Gt) ------
        if (JS > 1) {
            __syncthreads();                                  // every wave done with Gt
            float* part = reinterpret_cast<float*>(lds);      // [wave][32 values][64 lanes]
#pragma unroll
            for (int q = 0; q < 16; q++) {
                part[(wave * 32 + q) * 64 + lane] = re[q];
                part[(wave * 32 + 16 + q) * 64 + lane] = im[q];
            }
            __syncthreads();
            // wave js of the tile sums elements [16 js / JS, 16 (js + 1) / JS)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                if (q * JS / 16 != js) continue;
                float sr = 0.f, si = 0.f;
#pragma unroll
                for (int w2 = 0; w2 < JS; w2++) {
                    sr += part[((tile * JS + w2) * 32 + q) * 64 + lane];
                    si += part[((tile * JS + w2) * 32 + 16 + q) * 64 + lane];
                }
                re[q] = sr;
                im[q] = si;
            }
        }

        // ---- epilogue: undo scales, empty-window mask, store ---------------------
        const float inv = 1.0f / (scale * a.tap_scale);
        const int cc = c0 + 32 * tc + col;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            if (JS > 1 && q * JS / 16 != js) continue;
            const int rr = r0 + 32 * tr + 8 * (q / 4) + 4 * hh + (q % 4);
            if (rr < a.H && cc < a.W) {
                float2 v2 = make_float2(re[q] * inv, im[q] * inv);
                // window size of D row 8 (q / 4) + 4 hh + q % 4 (staged above)
                if (s_win[32 * tr + 8 * (q / 4) + 4 * hh + (q % 4)] == 0) v2 = make_float2(0.f, 0.f);
                a.out[(size_t)rr * a.W + cc] = v2;
            }
        }
    }
}

}  // namespace rtsim
