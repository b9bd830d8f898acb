// Grid (MFMA) form host orchestration: validity checks, extended-grid
// geometry, tap preparation (scaled, hi/lo split in BF16 and FP16) and the
// deposit + correlation launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/rtsim.h"
#include "grid_internal.h"
#include "sensor.h"

namespace rtsim {

namespace {

uint16_t host_bf16(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float host_bf16_f32(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// IEEE binary16, round to nearest even (inputs here are finite, |x| < 2^15).
uint16_t host_f16(float x)
{
    _Float16 h = (_Float16)x;
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}

float host_f16_f32(uint16_t u)
{
    _Float16 h;
    std::memcpy(&h, &u, 2);
    return (float)h;
}

// Tuned tile shapes, cached per shape class for the whole process (a
// sensor of the same grid and PSF size starts from the tuned shape).
struct TuneKey {
    int H, W, Tr, Tb;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(H, W, Tr, Tb) < std::tie(o.H, o.W, o.Tr, o.Tb);
    }
};
std::mutex g_tune_mutex;
std::map<TuneKey, int> g_tuned;

int tuned_shape(const TuneKey& k)
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    auto it = g_tuned.find(k);
    return it == g_tuned.end() ? -1 : it->second;
}

void store_tuned(const TuneKey& k, int shape)
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tuned[k] = shape;
}

struct GridState {
    bool ready = false;
    int status = RTSIM_OK;
    std::string why;
    int Hp = 0, Wp = 0, GH = 0, GHa = 0, GW = 0, fr = 0, fb = 0;
    int KR = 0, L = 0;                   // correlation K blocks, tap column length
    int shape = 0;                       // correlation tile shape in use
    int waves8 = 0;                      // 1: 8 waves per correlation workgroup
    bool fits[kCorrShapes] = {false, false, false, false};    // 4 waves
    bool fits8[kCorrShapes] = {false, false, false, false};   // 8 waves
    float inv_dr = 0, off_r = 0, inv_db = 0, off_b = 0;
    double bear0 = 0;
    float tap_scale = 1.f;
    int max_lds = 0;                     // device: LDS per workgroup
    // bucketing: NB buckets of R grid rows, key occupancy in occ_words words
    int R = 1, NB = 0, occ_words = 0;
    float2* d_grid = nullptr;
    uint32_t* d_vmax = nullptr;          // deposit scale (set by the bucket scan)
    uint32_t* d_totals = nullptr;        // [NB] items per bucket
    // sized for `cap` samples: bucket counts [NB][tiles], key occupancy [tiles][occ_words]
    size_t cap = 0;
    uint32_t* d_counts = nullptr;
    size_t counts_words = 0;
    uint32_t* d_occ = nullptr;
    size_t occ_cap = 0;
    uint16_t* d_toep[2] = {nullptr, nullptr};   // tap columns, [0] bf16, [1] fp16
    size_t lds = 0;

    void free_sized()
    {
        if (d_counts) (void)hipFree(d_counts);
        if (d_occ) (void)hipFree(d_occ);
        d_counts = nullptr;
        d_occ = nullptr;
        cap = counts_words = occ_cap = 0;
    }
    ~GridState()
    {
        free_sized();
        if (d_grid) (void)hipFree(d_grid);
        if (d_vmax) (void)hipFree(d_vmax);
        if (d_totals) (void)hipFree(d_totals);
        for (auto p : d_toep)
            if (p) (void)hipFree(p);
    }
};

// Bucket-count words the bucketing may use for `samples` samples: NB rows of
// one count per tile of 1024 samples, at most kBucketCountBudget (larger pings
// take larger tiles), plus one tile column of slack.
size_t bucket_count_words(int NB, size_t samples)
{
    const size_t t1 = (samples + 1023) / 1024;
    return std::min((size_t)NB * std::max<size_t>(t1, 1), kBucketCountBudget) + (size_t)NB;
}

}  // namespace

void Sensor::free_grid()
{
    delete static_cast<GridState*>(grid_state);
    grid_state = nullptr;
}

// LDS-sizing arguments of a correlation shape
static CorrArgs shape_probe(const Sensor* s, const GridState* g, int sh, int waves8)
{
    CorrArgs probe{};
    probe.Tr = (int)s->Tr;
    probe.Tb = (int)s->Tb;
    probe.shape = sh;
    probe.KR = g->KR;
    probe.L = g->L;
    probe.RS = corr_row_stride(sh, g->KR);
    probe.waves8 = waves8;
    return probe;
}

// 8 waves when the shape's LDS allows one workgroup per CU only (and 8 fit)
static int default_waves8(const Sensor* s, const GridState* g, int sh)
{
    return g->fits8[sh] && corr_waves8(shape_probe(s, g, sh, 0), g->max_lds) ? 1 : 0;
}

static int prepare_grid(Sensor* s, GridState* g)
{
    // a 1D ping is the grid form with one column: every sample (bearing 0)
    // lands on column 0 with weight 1 and the taps are one column wide
    const bool one_d = s->kind == RTSIM_SENSOR_1D;
    if (!one_d && s->kind != RTSIM_SENSOR_2D)
        return fail(RTSIM_ERR_UNSUPPORTED, "MFMA form: unknown sensor kind %d", s->kind);
    if (s->N < 2 || (!one_d && s->W < 2))
        return fail(RTSIM_ERR_UNSUPPORTED, "MFMA form needs >= 2 ranges (and >= 2 bearings in 2D)");
    double x0 = 0.0, psf_pb = 1.0;
    if (!one_d) {
        // uniform bearing axis at the PSF bearing pitch
        std::vector<double> x(s->W);
        for (uint32_t b = 0; b < s->W; b++)
            x[b] = s->axis == RTSIM_AXIS_SINE ? std::sin((double)s->bearings[b]) : (double)s->bearings[b];
        const double pb = (x[s->W - 1] - x[0]) / (double)(s->W - 1);
        double dev = 0;
        for (uint32_t b = 0; b < s->W; b++) dev = std::fmax(dev, std::fabs(x[b] - (x[0] + b * pb)));
        if (dev > 5e-5 * pb)
            return fail(RTSIM_ERR_UNSUPPORTED,
                        "MFMA form needs bearings uniform in the PSF axis (deviation %.3g pitch)",
                        dev / pb);
        psf_pb = (double)s->psf_bearing_span / s->Tb;
        if (std::fabs(psf_pb - pb) > 5e-5 * pb)
            return fail(RTSIM_ERR_UNSUPPORTED,
                        "MFMA form needs the PSF bearing pitch (%.6g) equal to the bearing pitch (%.6g)",
                        psf_pb, pb);
        x0 = x[0];
    }
    const double dr = ((double)s->range_max - (double)s->range_min) / (double)(s->N - 1);
    const double psf_pr = (double)s->psf_range_span / s->Tr;
    if (std::fabs(psf_pr - dr) > 5e-5 * dr)
        return fail(RTSIM_ERR_UNSUPPORTED,
                    "MFMA form needs the PSF range pitch (%.6g) equal to the range step (%.6g)",
                    psf_pr, dr);
    const int ov_req = ((int)s->Tr + 3) / 2;
    if (s->overlap < ov_req)
        return fail(RTSIM_ERR_UNSUPPORTED,
                    "MFMA form needs a Binner overlap >= %d bins for %u range taps (have %d)",
                    ov_req, s->Tr, s->overlap);

    g->Hp = (int)((s->N + kCorrRows - 1) / kCorrRows) * kCorrRows;
    g->Wp = (int)((s->W + kCorrCols - 1) / kCorrCols) * kCorrCols;
    g->KR = (31 + (int)s->Tr + 15) / 16;
    g->L = 16 * g->KR + 32;
    g->GH = g->Hp + (int)s->Tr - 1;
    // the correlation stages 16 KR rows per 32-row tile: rows past GH exist and
    // stay zero (the taps there are zero too, but the data must be finite)
    g->GHa = std::max(g->GH, g->Hp - 32 + 16 * g->KR);
    // columns the correlation stages for its last workgroup (even, read in pairs)
    g->GW = g->Wp + (int)s->Tb - 1 + (((int)s->Tb - 1) & 1);
    g->fr = ((int)s->Tr - 1) / 2;
    g->fb = ((int)s->Tb - 1) / 2;
    g->off_r = (float)(0.5 * (s->Tr - 1) - g->fr);
    g->off_b = (float)(0.5 * (s->Tb - 1) - g->fb);
    g->inv_dr = (float)(1.0 / dr);
    g->inv_db = (float)(1.0 / psf_pb);
    g->bear0 = x0;

    int dev_id = 0, max_lds = 0;
    (void)hipGetDevice(&dev_id);
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev_id) !=
            hipSuccess || max_lds <= 0)
        max_lds = 64 * 1024;
    max_lds -= 64;   // the kernel's static LDS
    g->max_lds = max_lds;
    // tile shapes whose LDS fits; default: the first that does (32x32)
    g->shape = -1;
    for (int sh = 0; sh < kCorrShapes; sh++) {
        g->fits[sh] = correlate_lds_bytes(shape_probe(s, g, sh, 0)) <= (size_t)max_lds;
        g->fits8[sh] = correlate_lds_bytes(shape_probe(s, g, sh, 1)) <= (size_t)max_lds;
        if (g->fits[sh] && g->shape < 0) g->shape = sh;
    }
    if (g->shape < 0)
        return fail(RTSIM_ERR_UNSUPPORTED, "MFMA form: no tile shape fits %d B of LDS for %ux%u taps",
                    max_lds, s->Tr, s->Tb);
    g->waves8 = default_waves8(s, g, g->shape);
    const TuneKey key{(int)s->N, (int)s->W, (int)s->Tr, (int)s->Tb};
    const int cached = tuned_shape(key);   // shape + 4 waves8
    if (cached >= 0 && ((cached >> 2) ? g->fits8 : g->fits)[cached & 3]) {
        g->shape = cached & 3;
        g->waves8 = cached >> 2;
    }
    // bucketing: rows per bucket so that about 256 buckets cover the grid
    // (each is one deposit workgroup), within the deposit's LDS budget and
    // the counters' limit; a sample on a bucket's last row is binned twice
    const size_t row_lds = sizeof(long long) * 2 * (size_t)g->GW;
    const int r_lds = (int)std::min<size_t>(kDepositLdsBudget / row_lds, kMaxBucketRows);
    if (r_lds < 1 || row_lds > (size_t)max_lds)
        return fail(RTSIM_ERR_UNSUPPORTED, "MFMA form: deposit row of %d cells too wide", g->GW);
    g->R = std::max(1, std::min(r_lds, g->GH / 256));
    g->NB = (g->GH + g->R - 1) / g->R;
    if (g->NB > kMaxBuckets)
        return fail(RTSIM_ERR_UNSUPPORTED, "MFMA form: %d grid rows need more than %d buckets", g->GH,
                    kMaxBuckets);
    g->occ_words = (int)((s->N + 31) / 32);
    if (g->occ_words > kMaxOccWords)
        return fail(RTSIM_ERR_UNSUPPORTED, "MFMA form: at most %d range bins", 32 * kMaxOccWords);

    // taps: power-of-two scale so max |T| in [0.5, 1), split hi/lo, laid out
    // as padded tap columns [shift c][plane p][bearing tap j][x] =
    // T_p[x + c - 31][j] (zero outside [0, Tr)), planes re_hi, re_lo, im_hi, im_lo
    const size_t T = (size_t)s->Tr * s->Tb;
    float mx = 0.f;
    for (size_t k = 0; k < 2 * T; k++) mx = std::fmax(mx, std::fabs(s->psf[k]));
    int e = 0;
    if (mx > 0.f) std::frexp(mx, &e);
    g->tap_scale = std::ldexp(1.0f, -e);
    const size_t L = (size_t)g->L, Tb = s->Tb;
    const size_t cs = (size_t)corr_copy_stride_host((int)Tb, (int)L);
    const size_t tsize = 2 * cs;
    std::vector<uint16_t> toep[2];
    for (int d = 0; d < 2; d++) toep[d].assign(tsize, 0);
    for (size_t i = 0; i < s->Tr; i++) {
        for (size_t j = 0; j < Tb; j++) {
            for (int c = 0; c < 2; c++) {               // re, im
                const float v = s->psf[2 * (i * s->Tb + j) + c] * g->tap_scale;
                uint16_t planes[2][2];
                planes[0][0] = host_bf16(v);
                planes[0][1] = host_bf16(v - host_bf16_f32(planes[0][0]));
                planes[1][0] = host_f16(v);
                planes[1][1] = host_f16(v - host_f16_f32(planes[1][0]));
                for (int d = 0; d < 2; d++)
                    for (int hl = 0; hl < 2; hl++)
                        for (size_t sh = 0; sh < 2; sh++) {
                            const size_t x = i + 31 - sh;   // x + sh - 31 = i
                            const size_t p = 2 * c + hl;
                            toep[d][sh * cs + (p * Tb + j) * L + x] = planes[d][hl];
                        }
            }
        }
    }
    hipError_t err;
    if ((err = hipMalloc((void**)&g->d_grid, sizeof(float2) * (size_t)g->GHa * g->GW)) != hipSuccess)
        return hip_fail(err, "hipMalloc(grid)");
    // rows [GH, GHa) are read by the correlation and never written
    if ((err = hipMemset(g->d_grid, 0, sizeof(float2) * (size_t)g->GHa * g->GW)) != hipSuccess)
        return hip_fail(err, "hipMemset(grid)");
    if ((err = hipMalloc((void**)&g->d_vmax, sizeof(uint32_t))) != hipSuccess)
        return hip_fail(err, "hipMalloc(vmax)");
    if ((err = hipMalloc((void**)&g->d_totals, sizeof(uint32_t) * (size_t)g->NB)) != hipSuccess)
        return hip_fail(err, "hipMalloc(bucket totals)");
    for (int d = 0; d < 2; d++) {
        if ((err = hipMalloc((void**)&g->d_toep[d], sizeof(uint16_t) * tsize)) != hipSuccess)
            return hip_fail(err, "hipMalloc(toeplitz)");
        if ((err = hipMemcpy(g->d_toep[d], toep[d].data(), sizeof(uint16_t) * tsize,
                             hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(err, "hipMemcpy(toeplitz)");
    }
    g->ready = true;
    return RTSIM_OK;
}

int Sensor::ensure_grid()
{
    GridState* g = static_cast<GridState*>(grid_state);
    if (!g) {
        g = new GridState();
        grid_state = g;
        g->status = prepare_grid(this, g);
        if (g->status) g->why = fail_message();
    }
    if (!g->ready) return fail(g->status, "%s", g->why.c_str());
    return grid_reserve();
}

bool Sensor::grid_ready() const
{
    return grid_state && static_cast<const GridState*>(grid_state)->ready;
}

// The bucketing's per-sample buffers (bucket counts, key occupancy) for the
// sensor's sample capacity.
int Sensor::grid_reserve()
{
    GridState* g = static_cast<GridState*>(grid_state);
    if (!g || !g->ready || (g->d_counts && g->cap >= capacity)) return RTSIM_OK;
    g->free_sized();
    const size_t cw = bucket_count_words(g->NB, capacity);
    const size_t ow = std::max<size_t>((capacity + 1023) / 1024, 1) * (size_t)g->occ_words;
    hipError_t e;
    if ((e = hipMalloc((void**)&g->d_counts, sizeof(uint32_t) * cw)) != hipSuccess)
        return hip_fail(e, "hipMalloc(bucket counts)");
    if ((e = hipMalloc((void**)&g->d_occ, sizeof(uint32_t) * ow)) != hipSuccess) {
        g->free_sized();
        return hip_fail(e, "hipMalloc(key occupancy)");
    }
    g->counts_words = cw;
    g->occ_cap = ow;
    g->cap = capacity;
    generation++;
    return RTSIM_OK;
}

void Sensor::grid_buffers(PingBuffers* pb) const
{
    const GridState* g = static_cast<const GridState*>(grid_state);
    if (!g || !g->ready) return;
    pb->bcounts = g->d_counts;
    pb->bcounts_cap = g->counts_words;
    pb->occ = g->d_occ;
    pb->occ_cap = g->occ_cap;
}

// Plans the grid form of one ping over the buffers pb: the bucketing (counts,
// scan, scatter), the deposit with the windows, the correlation.
int Sensor::grid_plan(int form, BucketArgs* bp, DepositArgs* dp, CorrArgs* cp, const PingBuffers& pb)
{
    int rc = ensure_grid();
    if (rc) return rc;
    GridState* g = static_cast<GridState*>(grid_state);
    const uint32_t n = pb.n;
    if (!pb.items || !pb.bcounts || !pb.occ || !pb.tmax)
        return fail(RTSIM_ERR_INVALID, "MFMA form: bucketing buffers missing");

    // tiles of 1024 samples, or larger ones when the count matrix would not
    // fit its budget (the streaming kernel makes the counts of 1024-sample tiles)
    const size_t t1 = (n + 1023) / 1024;
    const size_t room = pb.bcounts_cap > (size_t)g->NB ? pb.bcounts_cap - (size_t)g->NB : 0;
    size_t m = 1;
    if ((size_t)g->NB * t1 > room) m = room ? ((size_t)g->NB * t1 + room - 1) / room : 0;
    if (!m) return fail(RTSIM_ERR_INVALID, "MFMA form: no room for the bucket counts");
    const uint32_t tile = (uint32_t)(1024 * m);
    const uint32_t tiles = (uint32_t)(((size_t)n + tile - 1) / tile);
    if ((size_t)g->NB * tiles > pb.bcounts_cap || (size_t)tiles * g->occ_words > pb.occ_cap ||
        tiles > 256u * 64u)
        return fail(RTSIM_ERR_INVALID, "MFMA form: %u samples exceed the bucketing buffers", n);

    BucketArgs& b = *bp;
    b = BucketArgs{};
    b.rec = pb.samples;
    b.n = n;
    b.tile = tile;
    b.tiles = tiles;
    b.key = key;
    b.rows = GridRows{range_min, g->inv_dr, g->off_r, g->fr, g->GH};
    b.R = g->R;
    b.NB = g->NB;
    b.counts = pb.bcounts;
    b.totals = g->d_totals;
    b.tmax = pb.tmax;
    b.vmax = g->d_vmax;
    b.items = pb.items;
    b.items_cap = (uint32_t)std::min<size_t>(pb.items_cap, 0xFFFFFFFFu);
    b.occ = pb.occ;
    b.occ_words = g->occ_words;
    b.bear0 = g->bear0;
    b.inv_db = g->inv_db;
    b.off_b = g->off_b;
    b.axis = axis;

    DepositArgs& d = *dp;
    d.items = pb.items;
    d.items_cap = b.items_cap;
    d.totals = g->d_totals;
    d.NB = g->NB;
    d.R = g->R;
    d.vmax = g->d_vmax;
    d.grid = g->d_grid;
    d.GH = g->GH;
    d.GW = g->GW;
    d.fb = g->fb;
    d.rows = b.rows;
    d.occ = pb.occ;
    d.occ_words = g->occ_words;
    d.tiles = tiles;
    d.N = N;
    d.ov = overlap;
    d.win = pb.win;

    CorrArgs& c = *cp;
    c.grid = g->d_grid;
    c.toep = g->d_toep[form == RTSIM_FORM_MFMA_FP16 ? 1 : 0];
    c.win = pb.win;
    c.out = pb.out;
    c.H = (int)N;
    c.W = (int)W;
    c.Tr = (int)Tr;
    c.Tb = (int)Tb;
    c.GH = g->GH;
    c.GW = g->GW;
    c.KR = g->KR;
    c.L = g->L;
    c.shape = g->shape;
    c.RS = corr_row_stride(g->shape, g->KR);
    c.tap_scale = g->tap_scale;
    c.waves8 = g->waves8;
    return RTSIM_OK;
}

// Times every tile shape that fits, with 4 and with 8 waves per workgroup
// (correlation kernel only, HIP events on the stream, after one full compute
// so G is current) and keeps the fastest; the result is cached for the shape
// class (H, W, Tr, Tb).
int Sensor::autotune(int form, int iterations, hipStream_t s, int* best_shape, float* best_ms)
{
    if (form != RTSIM_FORM_MFMA_BF16 && form != RTSIM_FORM_MFMA_FP16)
        return fail(RTSIM_ERR_INVALID, "autotune applies to the MFMA forms");
    if (iterations < 1) iterations = 1;
    int rc = compute(form, s);
    if (rc) return rc;
    GridState* g = static_cast<GridState*>(grid_state);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(RTSIM_ERR_HIP, "hipEventCreate failed");
    int best = -1, best8 = 0;
    float bms = 0.f;
    for (int v = 0; v < 2 * kCorrShapes; v++) {
        const int sh = v % kCorrShapes, w8 = v / kCorrShapes;
        if (!(w8 ? g->fits8 : g->fits)[sh]) continue;
        CorrArgs c = last_corr;
        c.shape = sh;
        c.RS = corr_row_stride(sh, g->KR);
        c.waves8 = w8;
        hipError_t err = launch_correlate(c, form == RTSIM_FORM_MFMA_FP16, s);   // warm
        (void)hipEventRecord(e0, s);
        for (int it = 0; it < iterations && err == hipSuccess; it++)
            err = launch_correlate(c, form == RTSIM_FORM_MFMA_FP16, s);
        (void)hipEventRecord(e1, s);
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        if (err != hipSuccess) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return hip_fail(err, "autotune");
        }
        ms /= iterations;
        if (best < 0 || ms < bms) { best = sh; best8 = w8; bms = ms; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g->shape = best;
    g->waves8 = best8;
    store_tuned(TuneKey{(int)N, (int)W, (int)Tr, (int)Tb}, best + 4 * best8);
    generation++;   // captured graphs hold the old shape
    if (best_shape) *best_shape = best;
    if (best_ms) *best_ms = bms;
    // leave the ping consistent with the chosen shape
    return compute(form, s);
}

}  // namespace rtsim

using namespace rtsim;

extern "C" int rtsim_sensor_autotune(rtsim_sensor h, int form, int iterations, void* stream,
                                     int* shape, float* ms)
{
    Sensor* s = reinterpret_cast<Sensor*>(h);
    if (!s) return fail(RTSIM_ERR_INVALID, "sensor handle is NULL");
    return s->autotune(form, iterations, (hipStream_t)stream, shape, ms);
}

extern "C" int rtsim_sensor_set_tile_shape(rtsim_sensor h, int shape)
{
    Sensor* s = reinterpret_cast<Sensor*>(h);
    if (!s) return fail(RTSIM_ERR_INVALID, "sensor handle is NULL");
    if (shape < 0 || shape >= kCorrShapes) return fail(RTSIM_ERR_INVALID, "unknown tile shape %d", shape);
    int rc = s->ensure_grid();
    if (rc) return rc;
    GridState* g = static_cast<GridState*>(s->grid_state);
    if (!g->fits[shape])
        return fail(RTSIM_ERR_UNSUPPORTED, "tile shape %d does not fit this sensor's LDS budget", shape);
    g->shape = shape;
    g->waves8 = default_waves8(s, g, shape);
    s->generation++;
    return RTSIM_OK;
}

extern "C" int rtsim_sensor_set_tile_waves(rtsim_sensor h, int waves)
{
    Sensor* s = reinterpret_cast<Sensor*>(h);
    if (!s) return fail(RTSIM_ERR_INVALID, "sensor handle is NULL");
    if (waves != 4 && waves != 8) return fail(RTSIM_ERR_INVALID, "waves must be 4 or 8 (got %d)", waves);
    int rc = s->ensure_grid();
    if (rc) return rc;
    GridState* g = static_cast<GridState*>(s->grid_state);
    if (!(waves == 8 ? g->fits8 : g->fits)[g->shape])
        return fail(RTSIM_ERR_UNSUPPORTED, "%d waves do not fit tile shape %d's LDS budget", waves, g->shape);
    g->waves8 = waves == 8 ? 1 : 0;
    s->generation++;
    return RTSIM_OK;
}

extern "C" int rtsim_sensor_tile_waves(rtsim_sensor h, int* waves)
{
    Sensor* s = reinterpret_cast<Sensor*>(h);
    if (!s) return fail(RTSIM_ERR_INVALID, "sensor handle is NULL");
    if (!waves) return fail(RTSIM_ERR_INVALID, "waves is NULL");
    int rc = s->ensure_grid();
    if (rc) return rc;
    *waves = static_cast<GridState*>(s->grid_state)->waves8 ? 8 : 4;
    return RTSIM_OK;
}

extern "C" int rtsim_sensor_tile_shape(rtsim_sensor h, int* shape, int* rows, int* cols)
{
    Sensor* s = reinterpret_cast<Sensor*>(h);
    if (!s) return fail(RTSIM_ERR_INVALID, "sensor handle is NULL");
    int rc = s->ensure_grid();
    if (rc) return rc;
    GridState* g = static_cast<GridState*>(s->grid_state);
    if (shape) *shape = g->shape;
    int r = 0, c = 0;
    corr_tile(g->shape, &r, &c);
    if (rows) *rows = r;
    if (cols) *cols = c;
    return RTSIM_OK;
}
