// rtsim_measure_peaks: the running GPU's own peaks, for roofline fractions
// next to the published figures.
//   HBM: a grid-stride float4 copy between two 512 MiB buffers (reads and
//        writes counted), best of three launches.
//   MFMA: every wave runs four independent accumulator chains of
//        v_mfma_f32_32x32x16_{f16,bf16} (32 x 32 x 16 x 2 flop each), enough
//        waves for two per SIMD, best of three launches.
#include <hip/hip_runtime.h>

#include <functional>

#include "../../include/rtsim.h"
#include "sensor.h"

#define PEAK_TRY(expr)                                  \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

namespace rtsim {

namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kPeakThreads = 256;
constexpr int kPeakChains = 4;
constexpr int kPeakIters = 1024;
constexpr double kMfmaFlop = 32.0 * 32.0 * 16.0 * 2.0;

__global__ __launch_bounds__(kPeakThreads) void k_peak_copy(const float4* __restrict__ in,
                                                            float4* __restrict__ out, size_t n)
{
    const size_t stride = (size_t)gridDim.x * kPeakThreads;
    for (size_t i = (size_t)blockIdx.x * kPeakThreads + threadIdx.x; i < n; i += stride) out[i] = in[i];
}

template <bool FP16>
__global__ __launch_bounds__(kPeakThreads) void k_peak_mfma(float* out, float seed)
{
    f32x16 acc[kPeakChains];
#pragma unroll
    for (int c = 0; c < kPeakChains; c++)
        for (int e = 0; e < 16; e++) acc[c][e] = 0.0f;
    const float x = seed + (float)(threadIdx.x & 63) * 1e-3f;
    f16x8 ah, bh;
    bf16x8 ab, bb;
    for (int e = 0; e < 8; e++) {
        ah[e] = (_Float16)(x + e);
        bh[e] = (_Float16)(x - e);
        ab[e] = (__bf16)(x + e);
        bb[e] = (__bf16)(x - e);
    }
    for (int k = 0; k < kPeakIters; k++) {
#pragma unroll
        for (int c = 0; c < kPeakChains; c++) {
            if (FP16)
                acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc[c], 0, 0, 0);
            else
                acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ab, bb, acc[c], 0, 0, 0);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kPeakChains; c++)
        for (int e = 0; e < 16; e++) s += acc[c][e];
    // a data-dependent store keeps the chains live; it never happens for
    // these finite inputs
    if (s != s) out[blockIdx.x * kPeakThreads + threadIdx.x] = s;
}

int timed(hipStream_t s, hipEvent_t a, hipEvent_t b, float* ms, const std::function<void()>& launch)
{
    float best = 0.0f;
    for (int rep = 0; rep < 3; rep++) {
        PEAK_TRY(hipEventRecord(a, s));
        launch();
        PEAK_TRY(hipGetLastError());
        PEAK_TRY(hipEventRecord(b, s));
        PEAK_TRY(hipEventSynchronize(b));
        float t = 0.0f;
        PEAK_TRY(hipEventElapsedTime(&t, a, b));
        if (rep == 0 || t < best) best = t;
    }
    *ms = best;
    return RTSIM_OK;
}

}  // namespace

}  // namespace rtsim

using namespace rtsim;

extern "C" int rtsim_measure_peaks(void* stream, double* hbm_gbs, double* mfma_fp16_tflops,
                                   double* mfma_bf16_tflops)
{
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, cus = 0;
    PEAK_TRY(hipGetDevice(&dev));
    PEAK_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t bytes = (size_t)512 << 20, n = bytes / sizeof(float4);
    float4 *src = nullptr, *dst = nullptr;
    float* sink = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = RTSIM_OK;
    auto cleanup = [&]() {
        if (src) (void)hipFree(src);
        if (dst) (void)hipFree(dst);
        if (sink) (void)hipFree(sink);
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    };
    hipError_t e;
    if ((e = hipMalloc((void**)&src, bytes)) != hipSuccess || (e = hipMalloc((void**)&dst, bytes)) != hipSuccess ||
        (e = hipMalloc((void**)&sink, sizeof(float) * kPeakThreads * 8 * (size_t)cus)) != hipSuccess ||
        (e = hipEventCreate(&a)) != hipSuccess || (e = hipEventCreate(&b)) != hipSuccess ||
        (e = hipMemsetAsync(src, 0, bytes, s)) != hipSuccess) {
        cleanup();
        return hip_fail(e, "rtsim_measure_peaks: buffers");
    }
    float ms = 0.0f;
    const uint32_t copy_blocks = 8u * (uint32_t)cus;
    if (hbm_gbs && !(rc = timed(s, a, b, &ms, [&]() {
            hipLaunchKernelGGL(k_peak_copy, dim3(copy_blocks), dim3(kPeakThreads), 0, s, src, dst, n);
        })))
        *hbm_gbs = 2.0 * (double)bytes / (ms * 1e-3) / 1e9;
    // two waves per SIMD: 4 SIMDs per CU, kPeakThreads / 64 waves per workgroup
    const uint32_t mfma_blocks = 8u * (uint32_t)cus / (kPeakThreads / 64);
    const double flop = (double)mfma_blocks * (kPeakThreads / 64) * kPeakIters * kPeakChains * kMfmaFlop;
    if (!rc && mfma_fp16_tflops && !(rc = timed(s, a, b, &ms, [&]() {
            hipLaunchKernelGGL(k_peak_mfma<true>, dim3(mfma_blocks), dim3(kPeakThreads), 0, s, sink, 1.0f);
        })))
        *mfma_fp16_tflops = flop / (ms * 1e-3) / 1e12;
    if (!rc && mfma_bf16_tflops && !(rc = timed(s, a, b, &ms, [&]() {
            hipLaunchKernelGGL(k_peak_mfma<false>, dim3(mfma_blocks), dim3(kPeakThreads), 0, s, sink, 1.0f);
        })))
        *mfma_bf16_tflops = flop / (ms * 1e-3) / 1e12;
    cleanup();
    return rc;
}
