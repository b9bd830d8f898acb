// Streaming sample weighting: ReceiverView::set_sample + raygen phase.
//
// Reference semantics (rtac_simulation source tree): RayCaster.cu:62-74
// (phase, spreading) + Receiver.h:51-60 (set_sample) + Directivity.h:30-34
// (lookup) + Sample.h:144-160.  Shared by the streaming launches
// (kernels_bin.hip) and the persistent launch (kernels_persist.hip).
#pragma once

#include "device_common.h"

namespace rtsim {

// ---------------------------------------------------------------------------
// streaming kernel: 24 B in (range, dir, amplitude) -> 16 B record out
// ---------------------------------------------------------------------------
__device__ __forceinline__ float mirror_coord(float x)
{
    float f = floorf(x);
    float fr = x - f;
    int fi = (int)f;
    return (fi & 1) ? 1.0f - fr : fr;
}

__device__ __forceinline__ int mirror_index(int i, int n)
{
    if (i < 0) i = -1 - i;
    if (i >= n) i = 2 * n - 1 - i;
    return min(max(i, 0), n - 1);
}

// tex2D<float> with linear filtering, mirror addressing, normalised
// coordinates (Directivity.cpp:109-113), weights in full FP32; `az` is
// atan2(ly, lx).
__device__ __forceinline__ float directivity_lookup(const StreamArgs& p, float az, float lx,
                                                   float lz)
{
    float u = mirror_coord(p.dir_xs * az);
    float v = mirror_coord(p.dir_ys * atan2f(lz, lx));
    float xb = u * (float)p.dir_w - 0.5f, yb = v * (float)p.dir_h - 0.5f;
    float fi = floorf(xb), fj = floorf(yb);
    float a = xb - fi, b = yb - fj;
    int i0 = (int)fi, j0 = (int)fj;
    int c0 = mirror_index(i0, p.dir_w), c1 = mirror_index(i0 + 1, p.dir_w);
    int r0 = mirror_index(j0, p.dir_h), r1 = mirror_index(j0 + 1, p.dir_h);
    const float* t = p.dir;
    float t00 = t[r0 * p.dir_w + c0], t10 = t[r0 * p.dir_w + c1];
    float t01 = t[r1 * p.dir_w + c0], t11 = t[r1 * p.dir_w + c1];
    float top = t00 + a * (t10 - t00);
    float bot = t01 + a * (t11 - t01);
    return top + b * (bot - top);
}

// set_sample (Receiver.h:51-60) of one ray hit
__device__ __forceinline__ float4 stream_sample(const StreamArgs& p, float r, float d0, float d1,
                                                float d2, float2 A)
{
    if (r < 1.0e-4f) {
        // set_null_sample (Receiver.h:62-65): value 0, travel -1
        return make_float4(0.f, 0.f, -1.f, 0.f);
    }
    float lx = p.R[0] * d0 + p.R[3] * d1 + p.R[6] * d2;
    float ly = p.R[1] * d0 + p.R[4] * d1 + p.R[7] * d2;
    float lz = p.R[2] * d0 + p.R[5] * d1 + p.R[8] * d2;
    const float az = atan2f(ly, lx);
    float D = p.dir ? directivity_lookup(p, az, lx, lz) : 1.0f;
    // round-trip phase 2 k r reduced modulo one wavelength in FP64
    double q = (double)r * p.k2;
    q -= floor(q);
    float sn, cs;
    sincospif(2.0f * (float)q, &sn, &cs);
    float vr = A.x * cs - A.y * sn;
    float vi = A.x * sn + A.y * cs;
    if (p.spreading) {
        float g = 1.0f / (r * r);
        vr *= g;
        vi *= g;
    }
    float4 o;
    o.x = vr * D;
    o.y = vi * D;
    o.z = r;
    o.w = p.is2d ? -az : 0.0f;   // atan2(-ly, lx): atan2 is odd in y
    return o;
}

// One tile of TILE hits per call, NT threads, TILE / NT hits per thread with
// their loads issued together (the streaming launch: NT = TILE = 1024, one
// hit per thread; more per thread there measured 5-10 % slower: fewer
// resident waves hide the sample's dependent table lookups less well; the
// persistent launch, with fewer waves per CU: 512 threads, 8 hits each).  With
// with_hist (the grid form's bucketing) the workgroup also makes that tile's
// bucket counts in LDS `hist` ([NB] words) and its largest finite |re|, |im|
// of the binned samples, as k_bucket_hist makes them from the same records
// (counts and a maximum: the thread mapping is irrelevant).  Every launch
// (one ping, fused or not, batched, persistent) runs this one body with the
// counts switched at run time (a compile-time switch lets the compiler fuse
// multiply-adds differently, and the records would differ in their last bits
// between launches).
constexpr uint32_t kStreamTile = 1024;

template <int NT, int TILE = (int)kStreamTile>
__device__ __forceinline__ void stream_tile(const StreamArgs& p, const BucketArgs& bk, int with_hist,
                                            uint32_t blk, uint32_t* hist)
{
    constexpr int U = TILE / NT;
    constexpr int NW = NT / 64;
    static_assert(U * NT == TILE, "a tile is a whole number of hits per thread");
    const BucketArgs* b = with_hist ? &bk : nullptr;
    __shared__ float wmax[NW];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (b)
        for (int k = (int)t; k < b->NB; k += NT) hist[k] = 0u;
    float r[U], d0[U], d1[U], d2[U];
    float2 A[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const uint32_t ic = min(blk * (uint32_t)TILE + u * NT + t, p.n - 1);
        r[u] = p.range[ic];
        d0[u] = p.dx[ic];
        d1[u] = p.dy[ic];
        d2[u] = p.dz[ic];
        A[u] = p.amp[ic];
    }
    if (b) __syncthreads();   // counts cleared
    float m = 0.f;
#pragma unroll
    for (int u = 0; u < U; u++) {
        const uint32_t i = blk * (uint32_t)TILE + u * NT + t;
        if (i < p.n) {
            const float4 o = stream_sample(p, r[u], d0[u], d1[u], d2[u], A[u]);
            p.out[i] = o;
            if (b && bin_key(b->key, o.z) != kOutOfRange) {
                int b0, b1;
                bucket_pair(*b, g_row_of(b->rows, o.z), b0, b1);
                if (b0 >= 0) atomicAdd(&hist[b0], 1u);
                if (b1 >= 0) atomicAdd(&hist[b1], 1u);
                const float c = fmaxf(fabsf(o.x), fabsf(o.y));
                if (c <= 3.402823466e38f) m = fmaxf(m, c);
            }
        }
    }
    if (!b) return;
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if (lane == 0) wmax[w] = m;
    __syncthreads();
    for (int k = (int)t; k < b->NB; k += NT) b->counts[(size_t)k * b->tiles + blk] = hist[k];
    if (t == 0) {
        float mm = wmax[0];
        for (int k = 1; k < NW; k++) mm = fmaxf(mm, wmax[k]);
        b->tmax[blk] = __float_as_uint(mm);
    }
}

}  // namespace rtsim
