// Grid (MFMA) form of the PSF convolution.
//
// Restates do_sparse_convolve_2d_f (SensorInstance2D.cu:44-83) as
//   1. a bilinear (cloud-in-cell) deposit of every in-range sample onto the
//      output grid extended by the PSF half-widths (k_deposit), and
//   2. a "valid" 2D correlation of that grid with the PSF taps,
//        out(r, b) = sum_{i < Tr, j < Tb} T[i][j] G[r + i][b + j],
//      computed as complex Toeplitz implicit GEMMs on MFMA (k_correlate),
//   3. with the empty-window rule of do_make_bins (Binner.cu:96-97) applied
//      as a row mask in the epilogue.
// In exact arithmetic this equals the gather whenever the output axes are
// uniform at the tap pitch and the Binner overlap covers the PSF support.
//
// MFMA mapping (v_mfma_f32_16x16x32_{bf16,f16}, wave64):
//   M = 16 output range rows, N = 16 output bearings, K = 32 grid columns.
//   For range tap i, A[m][k] = G[r0 + m + i][c + k] and the Toeplitz
//   B_i[k][n] = T[i][k - n] (zero outside [0, Tb)), so out += A B_i.
//   Complex products use the real embedding (re = ArBr - AiBi, im = ArBi +
//   AiBr); each real product is split hi+lo: X Y ~ XhYh + XlYh + XhYl, with
//   FP32 accumulation.  G is scaled per workgroup tile by a power of two
//   (max |G| in [2^13, 2^14)) and the taps by a host power of two, so FP16
//   operands stay in range; the epilogue divides both out exactly.
#include "device_common.h"
#include "grid_internal.h"

namespace rtsim {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// deposit
// ---------------------------------------------------------------------------
__device__ __forceinline__ float grid_row_of(const DepositArgs& a, float t)
{
    return floorf((t - a.range0) * a.inv_dr + a.off_r);
}

// G row (grid row + fr) of a sorted sample, clamped to [-2, GH + 2]; the
// travel is sorted so it is non-decreasing in the sample index (NaN travel
// sorts last and maps to the top).
__device__ __forceinline__ int g_row(const DepositArgs& a, float t)
{
    const float r = grid_row_of(a, t) + (float)a.fr;
    if (!(r >= -2.0f)) return r != r ? a.GH + 2 : -2;
    return r > (float)(a.GH + 2) ? a.GH + 2 : (int)r;
}

// row_start[rho + 1] = first sorted sample whose G row is >= rho, for rho in
// [-1, GH + 1]: sample i fills the rows in (row(i - 1), row(i)].
__global__ __launch_bounds__(256) void k_row_starts(DepositArgs a)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > a.n) return;
    const int lo = i == 0 ? -1 : g_row(a, a.sorted[i - 1].z) + 1;
    const int hi = i == a.n ? a.GH + 1 : g_row(a, a.sorted[i].z);
    for (int rho = max(lo, -1); rho <= min(hi, a.GH + 1); rho++) a.row_start[rho + 1] = i;
}

__global__ __launch_bounds__(256) void k_deposit(DepositArgs a)
{
    extern __shared__ float sacc[];                   // [kDepositRows][GW][2]
    const int t = threadIdx.x;
    const int rho0 = blockIdx.x * kDepositRows;       // first G row of this block
    const int GW = a.GW;
    for (int k = t; k < kDepositRows * GW * 2; k += blockDim.x) sacc[k] = 0.f;
    __syncthreads();
    // samples whose G row is in [rho0 - 1, rho0 + R): they touch this block
    const uint32_t s0 = a.row_start[rho0 - 1 + 1];
    const uint32_t s1 = a.row_start[min(rho0 + kDepositRows, a.GH + 1) + 1];
    for (uint32_t s = s0 + t; s < s1; s += blockDim.x) {
        if (a.skeys[s] == kOutOfRange) continue;
        const float4 rec = a.sorted[s];
        const float2 xc = a.coord[s];
        const float y = (rec.z - a.range0) * a.inv_dr + a.off_r;
        const float x = ((xc.x - a.bear0_hi) + (xc.y - a.bear0_lo)) * a.inv_db + a.off_b;
        const float fy = floorf(y), fx = floorf(x);
        const float wy1 = y - fy, wx1 = x - fx;
        const int gy = (int)fy + a.fr, gx = (int)fx + a.fb;
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            const int lr = gy + dy - rho0;
            if (lr < 0 || lr >= kDepositRows) continue;
            const float wy = dy ? wy1 : 1.0f - wy1;
#pragma unroll
            for (int dx = 0; dx < 2; dx++) {
                const int c = gx + dx;
                if (c < 0 || c >= GW) continue;
                const float w = wy * (dx ? wx1 : 1.0f - wx1);
                atomicAdd(&sacc[2 * (lr * GW + c)], w * rec.x);
                atomicAdd(&sacc[2 * (lr * GW + c) + 1], w * rec.y);
            }
        }
    }
    __syncthreads();
    for (int k = t; k < kDepositRows * GW; k += blockDim.x) {
        const int lr = k / GW, c = k - lr * GW;
        const int rho = rho0 + lr;
        if (rho < a.GH)
            a.grid[(size_t)rho * GW + c] = make_float2(sacc[2 * k], sacc[2 * k + 1]);
    }
}

hipError_t launch_deposit(const DepositArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_row_starts, dim3(a.n / 256 + 1), dim3(256), 0, s, a);
    const int blocks = (a.GH + kDepositRows - 1) / kDepositRows;
    const size_t shm = sizeof(float) * 2 * kDepositRows * a.GW;
    hipLaunchKernelGGL(k_deposit, dim3(blocks), dim3(256), shm, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// hi/lo split conversions
// ---------------------------------------------------------------------------
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

// ---------------------------------------------------------------------------
// correlation on MFMA
// ---------------------------------------------------------------------------
// LDS layout (16-bit elements):
//   A: 4 planes (re_hi, re_lo, im_hi, im_lo) x AR rows x AS stride
//   B: 2 copies (shift 0 / 1) x 4 planes x kTapChunk taps x BL
template <bool FP16>
__global__ __launch_bounds__(256) void k_correlate(CorrArgs a)
{
    extern __shared__ uint16_t lds[];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r0 = blockIdx.y * kCorrRows;
    const int c0 = blockIdx.x * kCorrCols;
    const int AR = kCorrRows + a.Tr - 1;
    const int AC = 16 + 32 * a.KB;
    const int AS = a.AS;
    const int BL = a.BL;
    uint16_t* A = lds;
    uint16_t* B = lds + 4 * AR * AS;
    __shared__ float red[4];

    // ---- tile scale: max |G| over the A region -----------------------------
    float mx = 0.f;
    for (int k = t; k < AR * AC; k += 256) {
        const int rr = k / AC, cc = k - rr * AC;
        const float2 g = a.grid[(size_t)(r0 + rr) * a.GW + (c0 + cc)];
        mx = fmaxf(mx, fmaxf(fabsf(g.x), fabsf(g.y)));
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) red[wave] = mx;
    __syncthreads();
    mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    if (!(mx > 0.f)) {
        // empty tile (or non-finite input): the ping is zero here
        for (int k = t; k < kCorrRows * kCorrCols; k += 256) {
            const int rr = r0 + k / kCorrCols, cc = c0 + k % kCorrCols;
            if (rr < a.H && cc < a.W)
                a.out[(size_t)rr * a.W + cc] = mx == mx ? make_float2(0.f, 0.f)
                                                        : make_float2(mx, mx);
        }
        return;
    }
    int e;
    frexpf(mx, &e);                          // mx in [2^(e-1), 2^e)
    const float scale = ldexpf(1.0f, min(14 - e, 126)); // mx * scale in [2^13, 2^14)

    // ---- stage A (scaled, hi/lo split) ---------------------------------------
    for (int k = t; k < AR * AC; k += 256) {
        const int rr = k / AC, cc = k - rr * AC;
        const float2 g = a.grid[(size_t)(r0 + rr) * a.GW + (c0 + cc)];
        uint16_t h, l;
        split_hi_lo<FP16>(g.x * scale, h, l);
        A[(0 * AR + rr) * AS + cc] = h;
        A[(1 * AR + rr) * AS + cc] = l;
        split_hi_lo<FP16>(g.y * scale, h, l);
        A[(2 * AR + rr) * AS + cc] = h;
        A[(3 * AR + rr) * AS + cc] = l;
    }

    f32x4 acc_re[2], acc_im[2];
#pragma unroll
    for (int n = 0; n < 2; n++) {
        acc_re[n] = (f32x4){0.f, 0.f, 0.f, 0.f};
        acc_im[n] = (f32x4){0.f, 0.f, 0.f, 0.f};
    }

    const int m = lane & 15, kg = lane >> 4;
    for (int i0 = 0; i0 < a.Tr; i0 += kTapChunk) {
        const int tc = min(kTapChunk, a.Tr - i0);
        __syncthreads();   // A staged / previous chunk consumed
        // ---- stage the Toeplitz tap rows of this chunk ------------------------
        // B[c][p][ii][x] = Tpad[p][i0 + ii][x + c], Tpad[.][x] = T[.][x - 15]
        for (int k = t; k < 2 * 4 * tc * BL; k += 256) {
            int x = k % BL;
            int rest = k / BL;
            const int ii = rest % tc;
            rest /= tc;
            const int p = rest % 4;
            const int c = rest / 4;
            const int j = x + c - 15;
            uint16_t v = 0;
            if (j >= 0 && j < a.Tb) v = a.taps[((size_t)p * a.Tr + (i0 + ii)) * a.Tb + j];
            B[((c * 4 + p) * kTapChunk + ii) * BL + x] = v;
        }
        __syncthreads();
        for (int ii = 0; ii < tc; ii++) {
            const int i = i0 + ii;
            for (int q = 0; q < a.KB; q++) {
                const int x0 = 32 * q + 8 * kg + 15 - m;
                const int cpy = x0 & 1;
                const int xb = x0 - cpy;
                u32x4 b[4];
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    const uint32_t* src = reinterpret_cast<const uint32_t*>(
                        B + ((cpy * 4 + p) * kTapChunk + ii) * BL + xb);
                    b[p] = (u32x4){src[0], src[1], src[2], src[3]};
                }
                const u32x4 sgn = (u32x4){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
                const u32x4 nbi_hi = b[2] ^ sgn, nbi_lo = b[3] ^ sgn;
#pragma unroll
                for (int n = 0; n < 2; n++) {
                    const int row = wave * 16 + m + i;
                    const int col = 16 * n + 32 * q + 8 * kg;
                    const u32x4 ar_hi = *reinterpret_cast<const u32x4*>(A + (0 * AR + row) * AS + col);
                    const u32x4 ar_lo = *reinterpret_cast<const u32x4*>(A + (1 * AR + row) * AS + col);
                    const u32x4 ai_hi = *reinterpret_cast<const u32x4*>(A + (2 * AR + row) * AS + col);
                    const u32x4 ai_lo = *reinterpret_cast<const u32x4*>(A + (3 * AR + row) * AS + col);
                    // re += Ar Br - Ai Bi
                    acc_re[n] = mfma<FP16>(ar_hi, b[0], acc_re[n]);
                    acc_re[n] = mfma<FP16>(ar_lo, b[0], acc_re[n]);
                    acc_re[n] = mfma<FP16>(ar_hi, b[1], acc_re[n]);
                    acc_re[n] = mfma<FP16>(ai_hi, nbi_hi, acc_re[n]);
                    acc_re[n] = mfma<FP16>(ai_lo, nbi_hi, acc_re[n]);
                    acc_re[n] = mfma<FP16>(ai_hi, nbi_lo, acc_re[n]);
                    // im += Ar Bi + Ai Br
                    acc_im[n] = mfma<FP16>(ar_hi, b[2], acc_im[n]);
                    acc_im[n] = mfma<FP16>(ar_lo, b[2], acc_im[n]);
                    acc_im[n] = mfma<FP16>(ar_hi, b[3], acc_im[n]);
                    acc_im[n] = mfma<FP16>(ai_hi, b[0], acc_im[n]);
                    acc_im[n] = mfma<FP16>(ai_lo, b[0], acc_im[n]);
                    acc_im[n] = mfma<FP16>(ai_hi, b[1], acc_im[n]);
                }
            }
        }
    }

    // ---- epilogue: undo scales, empty-window mask, store ---------------------
    const float inv = 1.0f / (scale * a.tap_scale);
#pragma unroll
    for (int n = 0; n < 2; n++) {
        const int cc = c0 + 16 * n + m;
#pragma unroll
        for (int e2 = 0; e2 < 4; e2++) {
            const int rr = r0 + wave * 16 + 4 * kg + e2;
            if (rr < a.H && cc < a.W) {
                float2 v = make_float2(acc_re[n][e2] * inv, acc_im[n][e2] * inv);
                if (a.win[rr].y == 0) v = make_float2(0.f, 0.f);
                a.out[(size_t)rr * a.W + cc] = v;
            }
        }
    }
}

size_t correlate_lds_bytes(const CorrArgs& a)
{
    const int AR = kCorrRows + a.Tr - 1;
    return sizeof(uint16_t) * (size_t)(4 * AR * a.AS + 2 * 4 * kTapChunk * a.BL);
}

hipError_t launch_correlate(const CorrArgs& a, bool fp16, hipStream_t s)
{
    dim3 grid((a.W + kCorrCols - 1) / kCorrCols, (a.H + kCorrRows - 1) / kCorrRows);
    const size_t shm = correlate_lds_bytes(a);
    if (fp16) hipLaunchKernelGGL(k_correlate<true>, grid, dim3(256), shm, s, a);
    else hipLaunchKernelGGL(k_correlate<false>, grid, dim3(256), shm, s, a);
    return hipGetLastError();
}

void init_grid_attributes(int max_lds)
{
    set_dynamic_lds_limit((const void*)k_correlate<true>, max_lds);
    set_dynamic_lds_limit((const void*)k_correlate<false>, max_lds);
    set_dynamic_lds_limit((const void*)k_deposit, max_lds);
}

}  // namespace rtsim
