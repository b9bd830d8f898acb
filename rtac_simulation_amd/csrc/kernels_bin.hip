// Streaming sample weighting, sort by travel, keys/segments and windows.
//
// Reference semantics (rtac_simulation source tree):
//   k_stream   : RayCaster.cu:62-74 (phase, spreading) + Receiver.h:51-60
//                (set_sample) + Directivity.h:30-34 (lookup) + Sample.h:144-160
//   sort       : Receiver.cu:9-42 (thrust::sort by travel)
//   k_gather   : Binner.cu:24-49 (keys) + Binner.cu:51-80 (segments)
//   k_windows  : Binner.cu:82-136 (views with the empty-window rule)
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
// coordinates (Directivity.cpp:109-113), weights in full FP32.
__device__ __forceinline__ float directivity_lookup(const StreamArgs& p, float lx, float ly,
                                                   float lz)
{
    float u = mirror_coord(p.dir_xs * atan2f(ly, lx));
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

__global__ __launch_bounds__(256) void k_stream(StreamArgs p)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.n; i += stride) {
        float r = p.range[i];
        float4 o;
        if (r < 1.0e-4f) {
            // set_null_sample (Receiver.h:62-65): value 0, travel -1
            o = make_float4(0.f, 0.f, -1.f, 0.f);
        } else {
            float d0 = p.dx[i], d1 = p.dy[i], d2 = p.dz[i];
            float lx = p.R[0] * d0 + p.R[3] * d1 + p.R[6] * d2;
            float ly = p.R[1] * d0 + p.R[4] * d1 + p.R[7] * d2;
            float lz = p.R[2] * d0 + p.R[5] * d1 + p.R[8] * d2;
            float D = p.dir ? directivity_lookup(p, lx, ly, lz) : 1.0f;
            // round-trip phase 2 k r reduced modulo one wavelength in FP64
            double q = (double)r * p.k2;
            q -= floor(q);
            float sn, cs;
            sincospif(2.0f * (float)q, &sn, &cs);
            float2 A = p.amp[i];
            float vr = A.x * cs - A.y * sn;
            float vi = A.x * sn + A.y * cs;
            if (p.spreading) {
                float g = 1.0f / (r * r);
                vr *= g;
                vi *= g;
            }
            o.x = vr * D;
            o.y = vi * D;
            o.z = r;
            o.w = p.is2d ? atan2f(-ly, lx) : 0.0f;
        }
        p.out[i] = o;
    }
}

hipError_t launch_stream(const StreamArgs& a, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    uint32_t blocks = (a.n + 255) / 256;
    if (blocks > 256 * 64) blocks = 256 * 64;
    hipLaunchKernelGGL(k_stream, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// fill
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    size_t blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, s, p, value, count);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// LSD radix sort, 4 passes of 8 bits over (travel image, index).
//
// The array is cut into P contiguous stripes, one workgroup each.  Per pass:
//   k_radix_hist    : per-stripe digit histogram -> hist[digit][stripe]
//   k_radix_scan    : one workgroup per digit scans its row over stripes
//                     (exclusive) and writes the digit total
//   k_radix_scatter : each stripe walks its items in order, ranks them
//                     stably (wave ballots over the digit bits give the
//                     peers with the same digit, LDS carries counts across
//                     waves and rounds) and scatters them.
// Stable across stripes and within a stripe, so the final order is the
// stable order of the travel images.
// ---------------------------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 8;                          // items per thread per tile
constexpr int kSortTile = kSortThreads * kSortItems;   // 2048 items staged in LDS

__global__ __launch_bounds__(256) void k_radix_init(const float4* __restrict__ rec, uint32_t n,
                                                    uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        keys[i] = travel_image(rec[i].z);
        vals[i] = i;
    }
}

__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                    uint32_t n, uint32_t stripe, int shift,
                                                    uint32_t* __restrict__ hist, uint32_t P)
{
    __shared__ uint32_t h[kSortWaves][256];
    const uint32_t t = threadIdx.x, w = t >> 6;
    for (int k = 0; k < kSortWaves; k++) h[k][t] = 0;
    __syncthreads();
    const uint32_t beg = blockIdx.x * stripe;     // multiple of 256
    const uint32_t end = min(beg + stripe, n);
    const uint32_t end4 = beg + ((end - beg) & ~3u);
    const uint4* k4 = reinterpret_cast<const uint4*>(keys);
    for (uint32_t i = beg + 4 * t; i < end4; i += 4 * kSortThreads) {
        const uint4 v = k4[i >> 2];
        atomicAdd(&h[w][(v.x >> shift) & 255u], 1u);
        atomicAdd(&h[w][(v.y >> shift) & 255u], 1u);
        atomicAdd(&h[w][(v.z >> shift) & 255u], 1u);
        atomicAdd(&h[w][(v.w >> shift) & 255u], 1u);
    }
    for (uint32_t i = end4 + t; i < end; i += kSortThreads)
        atomicAdd(&h[w][(keys[i] >> shift) & 255u], 1u);
    __syncthreads();
    uint32_t c = 0;
    for (int k = 0; k < kSortWaves; k++) c += h[k][t];
    hist[t * P + blockIdx.x] = c;
}

// Inclusive wave scan (wave64) of a 32-bit value.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, uint32_t lane)
{
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    return x;
}

// One workgroup per digit: exclusive scan of hist[digit][0..P) in place,
// digit total to totals[digit].  P <= 4 * 256.
__global__ __launch_bounds__(256) void k_radix_scan(uint32_t* __restrict__ hist, uint32_t P,
                                                    uint32_t* __restrict__ totals)
{
    __shared__ uint32_t wsum[kSortWaves];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t* row = hist + blockIdx.x * P;
    uint32_t v[4];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint32_t idx = t * 4 + k;
        v[k] = idx < P ? row[idx] : 0u;
        s += v[k];
    }
    uint32_t incl = wave_incl_scan(s, lane);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t k = 0; k < w; k++) base += wsum[k];
    uint32_t run = base + incl - s;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint32_t idx = t * 4 + k;
        if (idx < P) row[idx] = run;
        run += v[k];
    }
    if (t == kSortThreads - 1) totals[blockIdx.x] = run;
}

__global__ __launch_bounds__(256) void k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, uint32_t stripe,
    int shift, const uint32_t* __restrict__ hist, uint32_t P,
    const uint32_t* __restrict__ totals)
{
    __shared__ uint32_t run[256];
    __shared__ uint32_t wc[kSortWaves][256];
    __shared__ uint32_t wsum[kSortWaves];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;

    // digit base = exclusive scan of the digit totals (block-wide)
    uint32_t tot = totals[t];
    uint32_t incl = wave_incl_scan(tot, lane);
    if (lane == 63) wsum[w] = incl;
    for (int k = 0; k < kSortWaves; k++) wc[k][t] = 0;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t k = 0; k < w; k++) base += wsum[k];
    run[t] = base + incl - tot + hist[t * P + blockIdx.x];
    __syncthreads();

    __shared__ uint32_t tcnt[256];     // digit counts of the current tile
    __shared__ uint32_t toff[256];     // exclusive prefix of tcnt
    __shared__ uint32_t skey[kSortTile];
    __shared__ uint32_t sval[kSortTile];

    const uint32_t beg = blockIdx.x * stripe;
    const uint32_t end = min(beg + stripe, n);
    for (uint32_t tb = beg; tb < end; tb += kSortTile) {
        const uint32_t tn = min((uint32_t)kSortTile, end - tb);
        // all items of the tile in registers: item j of thread t is tb + j*256 + t
        uint32_t key[kSortItems], val[kSortItems], lrank[kSortItems];
#pragma unroll
        for (int j = 0; j < kSortItems; j++) {
            const uint32_t q = j * kSortThreads + t;
            key[j] = q < tn ? kin[tb + q] : 0u;
            val[j] = q < tn ? vin[tb + q] : 0u;
        }
        tcnt[t] = 0;
        __syncthreads();
        // stable ranks within the tile, one round of 256 items at a time
#pragma unroll
        for (int j = 0; j < kSortItems; j++) {
            const bool valid = j * kSortThreads + t < tn;
            const uint32_t d = (key[j] >> shift) & 255u;
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const uint32_t bit = (d >> b) & 1u;
                const uint64_t m = __ballot(bit);
                peers &= bit ? m : ~m;
            }
            const uint32_t rank = popc_below(peers);
            if (valid && rank == 0) wc[w][d] = (uint32_t)__popcll(peers);
            __syncthreads();
            uint32_t r = tcnt[d] + rank;
            for (uint32_t k = 0; k < w; k++) r += wc[k][d];
            lrank[j] = r;
            __syncthreads();
            uint32_t add = 0;
            for (int k = 0; k < kSortWaves; k++) {
                add += wc[k][t];
                wc[k][t] = 0;
            }
            tcnt[t] += add;
            __syncthreads();
        }
        // local digit offsets, then stage the tile in LDS in digit order
        {
            const uint32_t c = tcnt[t];
            const uint32_t inc = wave_incl_scan(c, lane);
            if (lane == 63) wsum[w] = inc;
            __syncthreads();
            uint32_t base = 0;
            for (uint32_t k = 0; k < w; k++) base += wsum[k];
            toff[t] = base + inc - c;
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < kSortItems; j++) {
            if (j * kSortThreads + t < tn) {
                const uint32_t d = (key[j] >> shift) & 255u;
                const uint32_t p = toff[d] + lrank[j];
                skey[p] = key[j];
                sval[p] = val[j];
            }
        }
        __syncthreads();
        // consecutive lanes write consecutive addresses inside each digit run
        for (uint32_t q = t; q < tn; q += kSortThreads) {
            const uint32_t k = skey[q];
            const uint32_t d = (k >> shift) & 255u;
            const uint32_t pos = run[d] + (q - toff[d]);
            kout[pos] = k;
            vout[pos] = sval[q];
        }
        __syncthreads();
        run[t] += tcnt[t];
        __syncthreads();
    }
}

hipError_t launch_sort(const float4* records, uint32_t n, const SortBuffers& b,
                       const uint32_t** order_out, const uint32_t** img_out, hipStream_t s)
{
    *order_out = b.vals[0];
    *img_out = b.keys[0];
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_radix_init, dim3((n + 255) / 256), dim3(256), 0, s, records, n,
                       b.keys[0], b.vals[0]);
    // stripes of at least 2048 items, at most max_stripes (<= 1024) of them
    uint32_t P = (n + 2047) / 2048;
    if (P > b.max_stripes) P = b.max_stripes;
    if (P < 1) P = 1;
    uint32_t stripe = (n + P - 1) / P;
    stripe = (stripe + kSortTile - 1) / kSortTile * kSortTile;   // whole tiles
    P = (n + stripe - 1) / stripe;
    int cur = 0;
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 8 * pass;
        hipLaunchKernelGGL(k_radix_hist, dim3(P), dim3(256), 0, s, b.keys[cur], n, stripe,
                           shift, b.hist, P);
        hipLaunchKernelGGL(k_radix_scan, dim3(256), dim3(256), 0, s, b.hist, P, b.totals);
        hipLaunchKernelGGL(k_radix_scatter, dim3(P), dim3(256), 0, s, b.keys[cur],
                           b.vals[cur], b.keys[cur ^ 1], b.vals[cur ^ 1], n, stripe, shift,
                           b.hist, P, b.totals);
        cur ^= 1;
    }
    *order_out = b.vals[cur];
    *img_out = b.keys[cur];
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// gather sorted records, keys (Binner.cu:24-49) and segments (Binner.cu:51-69)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gather(GatherArgs a)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float4 r = a.rec[a.order[i]];
    a.sorted[i] = r;
    const uint32_t k = bin_key(a.key, r.z);
    a.skeys[i] = k;
    const double x = a.axis == 1 ? sin((double)r.w) : (double)r.w;
    const float xh = (float)x;
    a.coord[i] = make_float2(xh, (float)(x - (double)xh));
    if (k != kOutOfRange) {
        if (i == 0)
            a.seg[0].x = 0;  // Binner.cu:56-58 (equal to the zero fill)
        else if (bin_key(a.key, image_travel(a.img[i - 1])) != k)
            a.seg[k].x = i;
        if (i == a.n - 1)
            a.seg[k].y = a.n;
        else if (bin_key(a.key, image_travel(a.img[i + 1])) != k)
            a.seg[k].y = i + 1;
    }
}

hipError_t launch_gather(const GatherArgs& a, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// windows: do_make_bins (Binner.cu:82-106)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_windows(const uint2* __restrict__ seg, uint32_t N,
                                                 int ov, uint2* __restrict__ win)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (int)N) return;
    int istart = max(0, tid - ov);
    int iend = min((int)N - 1, tid + ov);
    while (istart < iend && seg[istart].y == seg[istart].x) istart++;
    while (iend > istart && seg[iend].y == seg[iend].x) iend--;
    if (istart >= iend)
        win[tid] = make_uint2(0u, 0u);
    else
        win[tid] = make_uint2(seg[istart].x, seg[iend].y - seg[istart].x);
}

hipError_t launch_windows(const uint2* seg, uint32_t N, int ov, uint2* win, hipStream_t s)
{
    if (N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_windows, dim3((N + 255) / 256), dim3(256), 0, s, seg, N, ov, win);
    return hipGetLastError();
}

}  // namespace rtsim
