// FP64 CPU oracle for the sample-to-ping path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// it.  The product path (rtac_simulation_amd/librtsim.so) never links,
// imports or calls anything here.
//
// Every function restates the reference's algorithm (file:line cites are
// relative to the rtac_simulation source tree) in plain C++:
//   * integer/bit work (Binner::Key, sort order, segments, windows) is
//     restated bit-for-bit in FP32 with `-ffp-contract=off`, exactly as the
//     reference's host and device code evaluates it;
//   * floating point work (directivity, phase, PSF, gathers) is evaluated in
//     FP64 with exact bilinear weights; the product's FP32/MFMA results are
//     compared against it with a stated relative-L2 tolerance.
//
// Pinning of each function (which independent check fixes it) is listed in
// the header comment of oracle/oracle.h.

#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>
#include <thread>
#include <vector>

namespace {

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

// cvt.rzi.s32.f32 semantics: truncation toward zero, saturating.  NaN is
// handled by the callers (it never reaches here).
inline int32_t trunc_sat(float v)
{
    if (v >= 2147483648.0f)  return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}

// Runs body(b, e) over [begin, end) in interleaved chunks on nthreads threads
// (chunks balance rows of uneven cost).
void parallel_rows(int nthreads, long begin, long end,
                   const std::function<void(long, long)>& body)
{
    long n = end - begin;
    if (n <= 0) return;
    if (nthreads <= 1 || n < 2) { body(begin, end); return; }
    if (nthreads > n) nthreads = (int)n;
    long chunk = std::max<long>(1, n / (nthreads * 8L));
    std::atomic_long next(begin);
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++) {
        pool.emplace_back([&]() {
            for (;;) {
                long b = next.fetch_add(chunk);
                if (b >= end) break;
                body(b, std::min(end, b + chunk));
            }
        });
    }
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

// --------------------------------------------------------------------------
// Binner::Key (Binner.h:34-45 host constructor, Binner.h:47-53 device call)
// --------------------------------------------------------------------------
void orc_key_make(uint32_t N, float lo, float hi, float margin, orc_key* k)
{
    // Binner.h:37-44, FP32 host arithmetic without contraction
    float resolution = (hi - lo) / (float)N;
    float m = lo - 0.5f * resolution;
    float M = hi + 0.5f * resolution;
    k->n = N;
    k->a = (float)(N - 1u) / (M - m);
    k->b = -m * k->a + 0.5f;
    k->lower = m - margin;
    k->upper = M + margin;
}

uint32_t orc_key_eval(const orc_key* k, float t)
{
    // Binner.h:48-53.  NaN travel is out of range (documented deviation: the
    // reference converts NaN to key 0 but cannot sort it, Receiver.cu:15).
    if (std::isnan(t)) return ORC_OOR;
    float v = std::fmaf(k->a, t, k->b);
    if (std::isnan(v)) return ORC_OOR;
    int32_t res = trunc_sat(v);
    if (res < 0) return ((float)res >= k->lower) ? 0u : ORC_OOR;
    if ((uint32_t)res >= k->n) return ((float)res <= k->upper) ? k->n - 1u : ORC_OOR;
    return (uint32_t)res;
}

void orc_keys(const orc_key* k, const float* travel, size_t n, size_t stride, uint32_t* keys)
{
    // Binner.cu:24-49 (thrust::transform with Key)
    for (size_t i = 0; i < n; i++) keys[i] = orc_key_eval(k, travel[i * stride]);
}

// --------------------------------------------------------------------------
// sort() (Receiver.cu:9-42): order of samples by travel.  The reference's
// comparator `<=` leaves equal-travel order unspecified; the restated order
// is the stable one (ties keep input order), -0 == +0, NaN last.
// --------------------------------------------------------------------------
uint32_t orc_travel_image(float t)
{
    uint32_t u = f2u(t);
    if (std::isnan(t)) u = 0x7fc00000u;
    if (u == 0x80000000u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

void orc_sort_order(const float* travel, size_t n, size_t stride, uint32_t* order)
{
    std::vector<uint32_t> img(n);
    for (size_t i = 0; i < n; i++) img[i] = orc_travel_image(travel[i * stride]);
    std::iota(order, order + n, 0u);
    std::stable_sort(order, order + n,
                     [&](uint32_t a, uint32_t b) { return img[a] < img[b]; });
}

// --------------------------------------------------------------------------
// Binner::make_segments (Binner.cu:51-80): memset, then per sorted sample
// --------------------------------------------------------------------------
void orc_segments(const uint32_t* keys, size_t n, uint32_t N, uint32_t* seg)
{
    std::memset(seg, 0, sizeof(uint32_t) * 2 * N);
    for (size_t i = 0; i < n; i++) {
        uint32_t k = keys[i];
        if (k == ORC_OOR) continue;
        if (i == 0)                 seg[0] = 0;          // Binner.cu:56-58
        else if (keys[i - 1] != k)  seg[2 * k] = (uint32_t)i;
        if (i == n - 1)             seg[2 * k + 1] = (uint32_t)n;
        else if (k != keys[i + 1])  seg[2 * k + 1] = (uint32_t)(i + 1);
    }
}

// --------------------------------------------------------------------------
// Binner::make_bins overlap (Binner.cu:112-113, :126-127)
// --------------------------------------------------------------------------
int orc_overlap(float margin, float lo, float hi, uint32_t N)
{
    float resolution = (hi - lo) / (float)(N - 1u);
    float q = margin / resolution;
    if (!(q == q)) return 0;
    if (q >= 2147483647.0f) return INT32_MAX;
    if (q <= -2147483648.0f) return INT32_MIN;
    return (int)q;
}

// --------------------------------------------------------------------------
// do_make_bins (Binner.cu:82-106): windows with the empty-window rule
// --------------------------------------------------------------------------
void orc_windows(const uint32_t* seg, uint32_t N, int ov, uint32_t* win)
{
    // An overlap >= N covers every bin; a negative overlap empties every
    // window exactly as ov = 0 does.  Clamping keeps `tid + ov` defined.
    if (ov < 0) ov = 0;
    if (ov > (int)N) ov = (int)N;
    for (int tid = 0; tid < (int)N; tid++) {
        int istart = std::max(0, tid - ov);
        int iend   = std::min((int)N - 1, tid + ov);
        for (; istart < iend && seg[2 * istart + 1] == seg[2 * istart]; istart++);
        for (; iend > istart && seg[2 * iend + 1] == seg[2 * iend]; iend--);
        if (istart >= iend) {
            win[2 * tid] = 0; win[2 * tid + 1] = 0;
        } else {
            win[2 * tid]     = seg[2 * istart];
            win[2 * tid + 1] = seg[2 * iend + 1] - seg[2 * istart];
        }
    }
}

// --------------------------------------------------------------------------
// Texture fetch restatement (CUDA linear filtering on normalised coordinates,
// as used by DirectivityView (Directivity.h:30-34) and KernelView1D/2D
// (ReductionKernel.h:52-73)): xB = x*n - 1/2, i = floor(xB), a = xB - i,
// value = (1-a) T[i] + a T[i+1].  Address modes: border -> 0 outside;
// mirror -> the normalised coordinate is folded (frac(x) if floor(x) is
// even, 1-frac(x) if odd) and edge texels repeat.
// --------------------------------------------------------------------------
static double mirror_coord(double x)
{
    double f = std::floor(x);
    double fr = x - f;
    long long fi = (long long)f;
    return ((fi & 1LL) == 0) ? fr : 1.0 - fr;
}

static inline long mirror_index(long i, long n)
{
    if (i < 0) i = -1 - i;
    if (i >= n) i = 2 * n - 1 - i;
    if (i < 0) i = 0;
    if (i >= n) i = n - 1;
    return i;
}

double orc_directivity(const float* table, uint32_t W, uint32_t H,
                       float xscale, float yscale, double lx, double ly, double lz)
{
    // Directivity.h:30-34: u = xs*atan2(y,x), v = ys*atan2(z,x);
    // Directivity.cpp:109-113: linear filter, mirror wrap, normalised coords.
    double u = mirror_coord((double)xscale * std::atan2(ly, lx));
    double v = mirror_coord((double)yscale * std::atan2(lz, lx));
    double xb = u * W - 0.5, yb = v * H - 0.5;
    double fi = std::floor(xb), fj = std::floor(yb);
    double a = xb - fi, b = yb - fj;
    long i0 = (long)fi, j0 = (long)fj;
    auto T = [&](long i, long j) {
        return (double)table[mirror_index(j, H) * (long)W + mirror_index(i, W)];
    };
    return (1 - a) * (1 - b) * T(i0, j0) + a * (1 - b) * T(i0 + 1, j0)
         + (1 - a) * b * T(i0, j0 + 1) + a * b * T(i0 + 1, j0 + 1);
}

// --------------------------------------------------------------------------
// Sample weighting: RayCaster.cu:62-74 (phase, spreading, set_sample call),
// Receiver.h:51-60 (set_sample: localDir = R^T dir, value *= D(localDir)),
// Sample.h:144-160 (SimSample2D::Make bearing = atan2(-y, x); Null travel -1).
// The phase is reduced in FP64 as exp(i 2 pi frac(2 r f / c)).
// --------------------------------------------------------------------------
void orc_stream(size_t n, const float* range, const float* dx, const float* dy,
                const float* dz, const float* amp, const orc_stream_params* p,
                double* out)
{
    const double k2 = 2.0 * (double)p->frequency / (double)p->sound_speed;
    for (size_t s = 0; s < n; s++) {
        double* o = out + 4 * s;
        float r = range[s];
        if (r < 1.0e-4f) {                      // RayCaster.cu:63-66
            o[0] = 0; o[1] = 0; o[2] = -1.0; o[3] = 0;
            continue;
        }
        double d[3] = {dx[s], dy[s], dz[s]};
        double l[3];
        for (int i = 0; i < 3; i++)
            l[i] = (double)p->rotation[0 * 3 + i] * d[0]
                 + (double)p->rotation[1 * 3 + i] * d[1]
                 + (double)p->rotation[2 * 3 + i] * d[2];
        double D = 1.0;
        if (p->directivity)
            D = orc_directivity(p->directivity, p->dir_width, p->dir_height,
                                p->dir_xscale, p->dir_yscale, l[0], l[1], l[2]);
        double q = (double)r * k2;
        double ph = 2.0 * M_PI * (q - std::floor(q));
        double c = std::cos(ph), sn = std::sin(ph);
        double ar = amp[2 * s], ai = amp[2 * s + 1];
        double vr = ar * c - ai * sn, vi = ar * sn + ai * c;
        if (p->spreading == 1) { double g = 1.0 / ((double)r * (double)r); vr *= g; vi *= g; }
        o[0] = vr * D; o[1] = vi * D; o[2] = r;
        o[3] = std::atan2(-l[1], l[0]);
    }
}

// --------------------------------------------------------------------------
// PSF restatement: KernelView2D (ReductionKernel.h:66-73) with the scaling of
// SensorInstance2D::kernel (SensorInstance2D.cu:35-42): tex2D at
// (dx/span_b + 1/2, dr/span_r + 1/2), linear filter, border 0
// (SensorInstance2D.cu:14-15).  Table = [range tap][bearing tap] complex.
// --------------------------------------------------------------------------
void orc_psf(const float* taps, int Tr, int Tb, double span_r, double span_b,
             double dr, double dxb, double* re, double* im)
{
    double xb = (dxb / span_b + 0.5) * Tb - 0.5;
    double yb = (dr / span_r + 0.5) * Tr - 0.5;
    double fi = std::floor(xb), fj = std::floor(yb);
    double a = xb - fi, b = yb - fj;
    long i0 = (long)fi, j0 = (long)fj;
    double sr = 0, si = 0;
    for (int dj = 0; dj < 2; dj++) {
        long j = j0 + dj;
        if (j < 0 || j >= Tr) continue;
        double wy = dj ? b : 1 - b;
        for (int di = 0; di < 2; di++) {
            long i = i0 + di;
            if (i < 0 || i >= Tb) continue;
            double w = wy * (di ? a : 1 - a);
            sr += w * taps[2 * (j * Tb + i)];
            si += w * taps[2 * (j * Tb + i) + 1];
        }
    }
    *re = sr; *im = si;
}

static inline double out_range(const orc_grid* g, long r)
{
    // Linspace<float>(lo, hi, N)[r] = lo + r (hi - lo)/(N - 1)
    if (g->range_count < 2) return (double)g->range_lo;
    return (double)g->range_lo + (double)r * ((double)g->range_hi - (double)g->range_lo)
                                 / (double)(g->range_count - 1);
}

static inline double bearing_coord(int axis, double beta)
{
    return axis == ORC_AXIS_SINE ? std::sin(beta) : beta;
}

// --------------------------------------------------------------------------
// do_sparse_convolve_2d_f (SensorInstance2D.cu:44-83) in FP64:
// out(r, b) = sum_{s in view(r)} v_s K(beta_s - beta_b, t_s - r_r), output
// row-major [range][bearing].  `sorted` holds 4 floats per sample
// {re, im, travel, bearing}; `win` holds (start, size) per range bin.
// --------------------------------------------------------------------------
void orc_conv2d(const float* sorted, const uint32_t* win, const orc_grid* g,
                const float* bearings, const float* taps, double* out, int nthreads)
{
    const long H = g->range_count, W = g->bearing_count;
    std::vector<double> bc(W);
    for (long b = 0; b < W; b++) bc[b] = bearing_coord(g->axis, bearings[b]);
    parallel_rows(nthreads, 0, H, [&](long r0, long r1) {
        std::vector<double> xs;
        for (long r = r0; r < r1; r++) {
            double rr = out_range(g, r);
            uint32_t st = win[2 * r], sz = win[2 * r + 1];
            xs.resize(sz);
            for (uint32_t k = 0; k < sz; k++)
                xs[k] = bearing_coord(g->axis, sorted[4 * (size_t)(st + k) + 3]);
            for (long b = 0; b < W; b++) {
                double ar = 0, ai = 0;
                for (uint32_t k = 0; k < sz; k++) {
                    const float* s = sorted + 4 * (size_t)(st + k);
                    double kr, ki;
                    orc_psf(taps, g->psf_range_taps, g->psf_bearing_taps,
                            g->psf_range_span, g->psf_bearing_span,
                            (double)s[2] - rr, xs[k] - bc[b], &kr, &ki);
                    ar += (double)s[0] * kr - (double)s[1] * ki;
                    ai += (double)s[0] * ki + (double)s[1] * kr;
                }
                out[2 * (r * W + b)] = ar;
                out[2 * (r * W + b) + 1] = ai;
            }
        }
    });
}

// --------------------------------------------------------------------------
// orc_conv2d on a list of output rows (out: nrows x W), for pings whose full
// gather is too long: each window sample visits only the bearings inside the
// PSF's bearing support |x_s - x_b| < (Tb + 1)/2 * span_b / Tb (the border
// address mode makes every other term exactly 0), and every cell still sums
// its terms in window order, so the rows equal orc_conv2d's.
// --------------------------------------------------------------------------
void orc_conv2d_rows(const float* sorted, const uint32_t* win, const orc_grid* g,
                     const float* bearings, const float* taps, const uint32_t* rows,
                     size_t nrows, double* out, int nthreads)
{
    const long W = g->bearing_count;
    std::vector<double> bc(W);
    for (long b = 0; b < W; b++) bc[b] = bearing_coord(g->axis, bearings[b]);
    // the support bound, widened a little: terms just inside it are evaluated
    // by orc_psf like any other
    const double supp = 0.5 * (g->psf_bearing_taps + 1) * (double)g->psf_bearing_span /
                        g->psf_bearing_taps * (1.0 + 1e-6);
    const bool ascending = std::is_sorted(bc.begin(), bc.end());
    parallel_rows(nthreads, 0, (long)nrows, [&](long i0, long i1) {
        std::vector<double> acc(2 * W);
        for (long i = i0; i < i1; i++) {
            const long r = rows[i];
            const double rr = out_range(g, r);
            std::fill(acc.begin(), acc.end(), 0.0);
            const uint32_t st = win[2 * r], sz = win[2 * r + 1];
            for (uint32_t k = 0; k < sz; k++) {
                const float* s = sorted + 4 * (size_t)(st + k);
                const double x = bearing_coord(g->axis, s[3]);
                long b0 = 0, b1 = W;
                if (ascending) {
                    b0 = std::lower_bound(bc.begin(), bc.end(), x - supp) - bc.begin();
                    b1 = std::upper_bound(bc.begin(), bc.end(), x + supp) - bc.begin();
                }
                for (long b = b0; b < b1; b++) {
                    double kr, ki;
                    orc_psf(taps, g->psf_range_taps, g->psf_bearing_taps, g->psf_range_span,
                            g->psf_bearing_span, (double)s[2] - rr, x - bc[b], &kr, &ki);
                    acc[2 * b] += (double)s[0] * kr - (double)s[1] * ki;
                    acc[2 * b + 1] += (double)s[0] * ki + (double)s[1] * kr;
                }
            }
            std::copy(acc.begin(), acc.end(), out + 2 * (size_t)i * W);
        }
    });
}

// --------------------------------------------------------------------------
// do_sparse_convolve_1d_f (SensorInstance1D.cu:35-72) in FP64 with the
// intended PSF scaling 1/(c tau) (the reference writes 1.0f/duration*c,
// SensorInstance1D.cu:30, which is c/tau: a documented deviation).
// --------------------------------------------------------------------------
void orc_conv1d(const float* sorted, const uint32_t* win, const orc_grid* g,
                const float* taps, double* out, int nthreads)
{
    const long N = g->range_count;
    parallel_rows(nthreads, 0, N, [&](long r0, long r1) {
        for (long r = r0; r < r1; r++) {
            double rr = out_range(g, r);
            uint32_t st = win[2 * r], sz = win[2 * r + 1];
            double ar = 0, ai = 0;
            for (uint32_t k = 0; k < sz; k++) {
                const float* s = sorted + 4 * (size_t)(st + k);
                double kr, ki;
                orc_psf(taps, g->psf_range_taps, 1, g->psf_range_span, 1.0,
                        (double)s[2] - rr, 0.0, &kr, &ki);
                ar += (double)s[0] * kr - (double)s[1] * ki;
                ai += (double)s[0] * ki + (double)s[1] * kr;
            }
            out[2 * r] = ar; out[2 * r + 1] = ai;
        }
    });
}

// --------------------------------------------------------------------------
// Grid form (the restatement the MFMA path computes): bilinear (cloud-in-
// cell) deposit of every in-range sample onto the output grid extended by a
// halo, discrete 2D correlation with the taps, then the empty-window mask of
// do_make_bins (Binner.cu:96-97).  Equal to orc_conv2d in exact arithmetic
// when the output axes are uniform with the PSF tap pitch and the binner
// overlap is >= (Tr + 3) / 2 bins.
// --------------------------------------------------------------------------
void orc_grid2d(const float* samples, size_t n, const orc_key* key,
                const uint32_t* win, const orc_grid* g, double bearing0,
                double bearing_pitch, const float* taps, double* out)
{
    const long H = g->range_count, W = g->bearing_count;
    const int Tr = g->psf_range_taps, Tb = g->psf_bearing_taps;
    const double mr = 0.5 * (Tr - 1), mb = 0.5 * (Tb - 1);
    const long fr = (long)std::floor(mr), fb = (long)std::floor(mb);
    const double dr_off = mr - fr, db_off = mb - fb;
    const double pitch_r = H > 1 ? ((double)g->range_hi - (double)g->range_lo) / (H - 1) : 1.0;
    const long Er = fr + 2, Eb = fb + 2;
    const long GH = H + 2 * Er, GW = W + 2 * Eb;
    std::vector<double> G(2 * GH * GW, 0.0);
    for (size_t s = 0; s < n; s++) {
        const float* p = samples + 4 * s;
        if (orc_key_eval(key, p[2]) == ORC_OOR) continue;
        double y = ((double)p[2] - (double)g->range_lo) / pitch_r + dr_off;
        double x = (bearing_coord(g->axis, p[3]) - bearing0) / bearing_pitch + db_off;
        double fy = std::floor(y), fx = std::floor(x);
        double wy = y - fy, wx = x - fx;
        for (int dy = 0; dy < 2; dy++) {
            long gy = (long)fy + dy + Er;
            if (gy < 0 || gy >= GH) continue;
            for (int dx = 0; dx < 2; dx++) {
                long gx = (long)fx + dx + Eb;
                if (gx < 0 || gx >= GW) continue;
                double w = (dy ? wy : 1 - wy) * (dx ? wx : 1 - wx);
                G[2 * (gy * GW + gx)]     += w * p[0];
                G[2 * (gy * GW + gx) + 1] += w * p[1];
            }
        }
    }
    for (long r = 0; r < H; r++) {
        bool masked = win[2 * r + 1] == 0;
        for (long b = 0; b < W; b++) {
            double ar = 0, ai = 0;
            if (!masked) {
                for (int i = 0; i < Tr; i++) {
                    long gy = r + i - fr + Er;
                    for (int j = 0; j < Tb; j++) {
                        long gx = b + j - fb + Eb;
                        double tr = taps[2 * (i * Tb + j)], ti = taps[2 * (i * Tb + j) + 1];
                        double gr = G[2 * (gy * GW + gx)], gi = G[2 * (gy * GW + gx) + 1];
                        ar += tr * gr - ti * gi;
                        ai += tr * gi + ti * gr;
                    }
                }
            }
            out[2 * (r * W + b)] = ar;
            out[2 * (r * W + b) + 1] = ai;
        }
    }
}

// --------------------------------------------------------------------------
// CPU baseline ("port"): the reference pipeline in FP32 exactly as the
// reference schedules it -- sort by travel (Receiver.cu:28-34), keys
// (Binner.cu:33-40), segments (Binner.cu:51-80), windows (Binner.cu:82-136),
// then for every output cell a gather over the whole window with a bilinear
// FP32 PSF evaluation (SensorInstance2D.cu:56-65).  Multithreaded over
// range rows.  Rows [row_begin, row_end) of the gather are computed (the
// rest are left untouched) so a bounded sample can be timed.
// --------------------------------------------------------------------------
static inline void psf_f32(const float* taps, int Tr, int Tb, float sx, float sy,
                           float dx, float dr, float* re, float* im)
{
    float xn = std::fmaf(sx, dx, 0.5f), yn = std::fmaf(sy, dr, 0.5f);
    float xb = xn * (float)Tb - 0.5f, yb = yn * (float)Tr - 0.5f;
    float fi = std::floor(xb), fj = std::floor(yb);
    float a = xb - fi, b = yb - fj;
    int i0 = (int)fi, j0 = (int)fj;
    float sr = 0.f, si = 0.f;
    if (j0 >= -1 && j0 < Tr && i0 >= -1 && i0 < Tb) {
        for (int dj = 0; dj < 2; dj++) {
            int j = j0 + dj;
            if (j < 0 || j >= Tr) continue;
            float wy = dj ? b : 1.f - b;
            for (int di = 0; di < 2; di++) {
                int i = i0 + di;
                if (i < 0 || i >= Tb) continue;
                float w = wy * (di ? a : 1.f - a);
                sr += w * taps[2 * (j * Tb + i)];
                si += w * taps[2 * (j * Tb + i) + 1];
            }
        }
    }
    *re = sr; *im = si;
}

void cpu_binning(const float* samples, size_t n, const orc_grid* g, float margin,
                 float* sorted, uint32_t* win, int nthreads)
{
    (void)nthreads;
    std::vector<uint32_t> order(n);
    orc_sort_order(samples + 2, n, 4, order.data());
    for (size_t i = 0; i < n; i++) std::memcpy(sorted + 4 * i, samples + 4 * (size_t)order[i], 16);
    orc_key key;
    orc_key_make(g->range_count, g->range_lo, g->range_hi, margin, &key);
    std::vector<uint32_t> keys(n), seg(2 * (size_t)g->range_count);
    orc_keys(&key, sorted + 2, n, 4, keys.data());
    orc_segments(keys.data(), n, g->range_count, seg.data());
    int ov = orc_overlap(margin, g->range_lo, g->range_hi, g->range_count);
    orc_windows(seg.data(), g->range_count, ov, win);
}

// 1D: do_sparse_convolve_1d_f (SensorInstance1D.cu:35-72) in FP32, every
// window sample through the linearly interpolated range PSF (scaling
// 1/(c tau), the documented deviation of the 1D kernel), border 0.
void cpu_gather1d_f32(const float* sorted, const uint32_t* win, const orc_grid* g,
                      const float* taps, float* out, int nthreads)
{
    const long N = g->range_count;
    const float sy = 1.0f / g->psf_range_span;
    const float step = N > 1 ? (g->range_hi - g->range_lo) / (float)(N - 1) : 0.f;
    parallel_rows(nthreads, 0, N, [&](long r0, long r1) {
        for (long r = r0; r < r1; r++) {
            const float rr = g->range_lo + (float)r * step;
            const uint32_t st = win[2 * r], sz = win[2 * r + 1];
            float ar = 0.f, ai = 0.f;
            for (uint32_t k = 0; k < sz; k++) {
                const float* s = sorted + 4 * (size_t)(st + k);
                float kr, ki;
                psf_f32(taps, g->psf_range_taps, 1, 0.f, sy, 0.f, s[2] - rr, &kr, &ki);
                ar += s[0] * kr - s[1] * ki;
                ai += s[0] * ki + s[1] * kr;
            }
            out[2 * r] = ar;
            out[2 * r + 1] = ai;
        }
    });
}

void cpu_gather2d_f32(const float* sorted, const uint32_t* win, const orc_grid* g,
                      const float* bearings, const float* taps, float* out,
                      long row_begin, long row_end, int nthreads)
{
    const long H = g->range_count, W = g->bearing_count;
    const float sx = 1.0f / g->psf_bearing_span, sy = 1.0f / g->psf_range_span;
    const float step = H > 1 ? (g->range_hi - g->range_lo) / (float)(H - 1) : 0.f;
    std::vector<float> bc(W);
    for (long b = 0; b < W; b++)
        bc[b] = g->axis == ORC_AXIS_SINE ? std::sin(bearings[b]) : bearings[b];
    row_begin = std::max(0L, row_begin); row_end = std::min(H, row_end);
    parallel_rows(nthreads, row_begin, row_end, [&](long r0, long r1) {
        std::vector<float> xs;
        for (long r = r0; r < r1; r++) {
            float rr = g->range_lo + (float)r * step;
            uint32_t st = win[2 * r], sz = win[2 * r + 1];
            xs.resize(sz);
            for (uint32_t k = 0; k < sz; k++) {
                float be = sorted[4 * (size_t)(st + k) + 3];
                xs[k] = g->axis == ORC_AXIS_SINE ? std::sin(be) : be;
            }
            for (long b = 0; b < W; b++) {
                float ar = 0.f, ai = 0.f;
                for (uint32_t k = 0; k < sz; k++) {
                    const float* s = sorted + 4 * (size_t)(st + k);
                    float kr, ki;
                    psf_f32(taps, g->psf_range_taps, g->psf_bearing_taps, sx, sy,
                            xs[k] - bc[b], s[2] - rr, &kr, &ki);
                    ar += s[0] * kr - s[1] * ki;
                    ai += s[0] * ki + s[1] * kr;
                }
                out[2 * (r * W + b)] = ar;
                out[2 * (r * W + b) + 1] = ai;
            }
        }
    });
}

}  // extern "C"
