// Host runtime of the C ABI declared in include/rtsim.h.
//
// One rtsim_sensor owns every device buffer of one sensor instance, the way
// SensorInstance{1D,2D} own receivedSamples_, bins_, the Binner buffers, the
// PSF texture and the ping (SensorInstance.h:21-72, SensorInstance2D.h:11-93).
// All work is queued on the caller's stream; nothing here synchronises the
// device except the explicit download / debug helpers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rtsim.h"
#include "rtsim_internal.h"
#include "persist.h"
#include "sensor.h"

namespace rtsim {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

const char* fail_message() { return g_last_error.c_str(); }

int hip_fail(hipError_t e, const char* what)
{
    return fail(RTSIM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr)                                   \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

// Binner::Key host constructor (Binner.h:34-45), FP32 without contraction.
#pragma clang fp contract(off)
KeyParams make_key(uint32_t N, float lo, float hi, float margin)
{
    KeyParams k;
    float resolution = (hi - lo) / (float)N;
    float m = lo - 0.5f * resolution;
    float M = hi + 0.5f * resolution;
    k.n = N;
    k.a = (float)(N - 1u) / (M - m);
    k.b = -m * k.a + 0.5f;
    k.lower = m - margin;
    k.upper = M + margin;
    return k;
}

// binOverlap = margin / resolution truncated (Binner.cu:112-113), clamped to
// [0, N] (any overlap >= N - 1 covers every bin; <= 0 empties every window).
int make_overlap(float margin, float lo, float hi, uint32_t N)
{
    float resolution = (hi - lo) / (float)(N - 1u);
    float q = margin / resolution;
    int ov;
    if (!(q == q)) ov = 0;
    else if (q >= 2147483647.0f) ov = INT32_MAX;
    else if (q <= -2147483648.0f) ov = INT32_MIN;
    else ov = (int)q;
    if (ov < 0) ov = 0;
    if (ov > (int)N) ov = (int)N;
    return ov;
}
#pragma clang fp contract(on)

// Waveform_Sine (Waveform.h:115-134): sample_rate = 2 f os in FP32, size =
// (unsigned)(sample_rate * duration), sample idx = exp(i phase) with phase =
// idx * (2 pi f / sample_rate) evaluated in double and rounded to float; the
// PSF table is wave[h] * beam[w] in complex FP32 (SensorInstance2D.cu:20-33).
std::vector<float> sine_psf_table(float frequency, float duration, uint32_t oversampling,
                                  const float* beam, uint32_t beam_taps, uint32_t* range_taps)
{
    const float rate = 2.0f * frequency * (float)oversampling;
    const float fn = rate * duration;
    uint32_t n = 0;
    if (fn >= 1.0f && fn < (float)kMaxWaveformTaps + 1.0f) n = (uint32_t)fn;
    *range_taps = n;
    std::vector<float> t(2 * (size_t)n * beam_taps);
    for (uint32_t h = 0; h < n; h++) {
        const float phase = (float)((double)h * (2.0 * M_PI * (double)frequency / (double)rate));
        const float c = std::cos(phase), sn = std::sin(phase);
        for (uint32_t w = 0; w < beam_taps; w++) {
            const float b = beam ? beam[w] : 1.0f;
            t[2 * ((size_t)h * beam_taps + w)] = c * b;
            t[2 * ((size_t)h * beam_taps + w) + 1] = sn * b;
        }
    }
    return t;
}

void init_kernel_attributes()
{
    static std::once_flag once;
    std::call_once(once, []() {
        int dev = 0, max_lds = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                hipSuccess || max_lds <= 0)
            max_lds = 64 * 1024;
        init_plain_attributes(max_lds);
        init_grid_attributes(max_lds);
        init_persist_attributes(max_lds);
    });
}

static float2 split_df(double x)
{
    float h = (float)x;
    return make_float2(h, (float)(x - (double)h));
}

template <typename T>
static int dev_alloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    HIP_TRY(hipMalloc((void**)p, sizeof(T) * count));
    return RTSIM_OK;
}

template <typename T>
static void dev_free(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

Sensor::~Sensor()
{
    free_samples();
    free_grid();
    dev_free(d_psf);
    dev_free(d_psf_cells);
    dev_free(d_bear);
    dev_free(d_rng);
    dev_free(d_dir);
    dev_free(d_seg);
    dev_free(d_win);
    dev_free(d_work);
    dev_free(d_out);
    dev_free(d_totals);
    free_strips();
    for (auto& p : events) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    for (auto e : event_pool) (void)hipEventDestroy(e);
}

void Sensor::free_samples()
{
    dev_free(d_samples);
    dev_free(d_sorted);
    dev_free(d_items);
    dev_free(d_coord);
    dev_free(d_kv[0]);
    dev_free(d_kv[1]);
    dev_free(d_order);
    dev_free(d_skeys);
    dev_free(d_tmax);
    dev_free(d_counts);
    free_strips();
    capacity = 0;
}

void Sensor::free_strips()
{
    dev_free(d_strip_ent);
    dev_free(d_strip_cnt);
    strip_cap = strip_cnt_cap = 0;
}

// The plain 2D gather's strip lists: capacity x strip_mult entries and the
// counts of every strip and block.  An allocation failure only leaves the
// gather scanning the windows.
int Sensor::ensure_strips()
{
    if (kind != RTSIM_SENSOR_2D || strip_mult == 0 || capacity == 0) return RTSIM_OK;
    const size_t need = capacity * (size_t)strip_mult;
    const size_t words = strip_cnt_words(W, (uint32_t)capacity);
    if (d_strip_ent && strip_cap >= need && strip_cnt_cap >= words) return RTSIM_OK;
    if (strip_failed && need >= strip_failed) return RTSIM_OK;   // that size did not fit: no lists
    free_strips();
    if (dev_alloc(&d_strip_ent, need) || dev_alloc(&d_strip_cnt, words)) {
        free_strips();
        (void)hipGetLastError();
        strip_failed = need;
    } else {
        strip_cap = need;
        strip_cnt_cap = words;
    }
    generation++;   // the gather's arguments change
    return RTSIM_OK;
}

int Sensor::reserve(size_t count)
{
    if (count <= capacity && d_samples) return RTSIM_OK;   // (a sensor of 0 samples gets 1-item buffers)
    if (count > 0xFFFFFFFFu)   // sample indices are 32-bit
        return fail(RTSIM_ERR_INVALID, "sample count %zu exceeds 2^32 - 1", count);
    free_samples();
    int rc;
    if ((rc = dev_alloc(&d_samples, count))) return rc;
    if ((rc = dev_alloc(&d_sorted, count))) return rc;
    if ((rc = dev_alloc(&d_items, 2 * count))) return rc;   // a sample is binned at most twice
    if ((rc = dev_alloc(&d_coord, count))) return rc;
    if ((rc = dev_alloc(&d_kv[0], count))) return rc;
    if ((rc = dev_alloc(&d_kv[1], count))) return rc;
    if ((rc = dev_alloc(&d_order, count))) return rc;
    if ((rc = dev_alloc(&d_skeys, count))) return rc;
    if ((rc = dev_alloc(&d_counts, sort_count_words(count)))) return rc;
    if ((rc = dev_alloc(&d_tmax, sort_count_words(count) / 256))) return rc;
    capacity = count;
    return grid_reserve();   // the grid form's count and occupancy buffers, when prepared
}

int Sensor::upload_psf()
{
    dev_free(d_psf);
    dev_free(d_psf_cells);
    int rc = dev_alloc(&d_psf, (size_t)Tr * Tb);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(d_psf, psf.data(), sizeof(float) * psf.size(), hipMemcpyHostToDevice));
    // bilinear cells for j in [-1, Tr], i in [-1, Tb] (FP64, rounded once):
    // A = T[j][i], B = T[j][i+1] - A, C = T[j+1][i] - A,
    // D = T[j+1][i+1] - T[j+1][i] - T[j][i+1] + A
    const long R = (long)Tr, B = (long)Tb, stride = B + 2;
    auto tap = [&](long j, long i, int c) {
        return j >= 0 && j < R && i >= 0 && i < B ? (double)psf[2 * (j * B + i) + c] : 0.0;
    };
    // two planes of float4 cells: (A, B) of every cell, then (C, D) of every cell
    const size_t ncells = (size_t)stride * (size_t)(R + 2);
    std::vector<float> cells(8 * ncells);
    for (long j = -1; j <= R; j++)
        for (long i = -1; i <= B; i++) {
            float* ab = &cells[4 * (size_t)((j + 1) * stride + (i + 1))];
            float* cd = ab + 4 * ncells;
            for (int c = 0; c < 2; c++) {
                const double a = tap(j, i, c), b = tap(j, i + 1, c), cc = tap(j + 1, i, c),
                             d = tap(j + 1, i + 1, c);
                ab[c] = (float)a;
                ab[2 + c] = (float)(b - a);
                cd[c] = (float)(cc - a);
                cd[2 + c] = (float)(d - cc - b + a);
            }
        }
    if ((rc = dev_alloc(&d_psf_cells, cells.size() / 4))) return rc;
    HIP_TRY(hipMemcpy(d_psf_cells, cells.data(), sizeof(float) * cells.size(), hipMemcpyHostToDevice));
    // how many strips of the plain 2D gather a sample can reach (0: no lists)
    strip_mult = 0;
    if (kind == RTSIM_SENSOR_2D && W > 0 && bearings.size() >= W) {
        std::vector<double> x(W);
        for (uint32_t b = 0; b < W; b++)
            x[b] = axis == RTSIM_AXIS_SINE ? std::sin((double)bearings[b]) : (double)bearings[b];
        const double pitch = (double)psf_bearing_span / (double)Tb;
        strip_mult = strip_multiplicity(x.data(), W, 0.5 * (double)(Tb + 1) * pitch, pitch);
    }
    return RTSIM_OK;
}

// SensorInstance::set_ranges (SensorInstance.cu:23-34) for a sensor with a
// waveform recipe: an adaptive duration follows the range resolution
// (Linspace resolution in FP32), then generate_psf_data rebuilds the table
// and the PSF range span becomes c * tau (SensorInstance2D.cu:35-42).
int Sensor::apply_waveform()
{
    if (!has_wave) return RTSIM_OK;
    float duration = wave_duration;
    if (!wave_fixed) {
        const float resolution = N > 1 ? (range_max - range_min) / (float)(N - 1u) : 0.0f;
        duration = resolution / sound_speed;
    }
    uint32_t T = 0;
    std::vector<float> table = sine_psf_table(wave_frequency, duration, wave_os, beam.data(),
                                              (uint32_t)beam.size(), &T);
    if (T == 0)
        return fail(RTSIM_ERR_INVALID,
                    "the sine waveform has no samples (or more than %u) at duration %g s",
                    kMaxWaveformTaps, (double)duration);
    wave_duration = duration;
    psf.swap(table);
    Tr = T;
    Tb = (uint32_t)beam.size();
    psf_range_span = duration * sound_speed;
    return upload_psf();
}

// Ranges-dependent state: Binner::reconfigure (Binner.cu:15-22) + ping resize.
int Sensor::configure_ranges()
{
    margin = margin_request >= 0.f ? margin_request : psf_range_span;
    key = make_key(N, range_min, range_max, margin);
    overlap = make_overlap(margin, range_min, range_max, N);
    std::vector<float2> rc(N);
    for (uint32_t r = 0; r < N; r++) {
        double rr = N > 1 ? (double)range_min + (double)r * ((double)range_max - (double)range_min)
                                                    / (double)(N - 1)
                          : (double)range_min;
        rc[r] = split_df(rr);
    }
    dev_free(d_rng);
    dev_free(d_seg);
    dev_free(d_win);
    dev_free(d_work);
    dev_free(d_out);
    int rc_;
    if ((rc_ = dev_alloc(&d_rng, N))) return rc_;
    if ((rc_ = dev_alloc(&d_seg, N))) return rc_;
    if ((rc_ = dev_alloc(&d_win, N))) return rc_;
    if ((rc_ = dev_alloc(&d_work, plain2d_work_words(N)))) return rc_;
    if ((rc_ = dev_alloc(&d_out, (size_t)N * W))) return rc_;
    HIP_TRY(hipMemcpy(d_rng, rc.data(), sizeof(float2) * N, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(d_out, 0, sizeof(float2) * (size_t)N * W));
    free_grid();
    generation++;
    return RTSIM_OK;
}

int Sensor::prepare(int form)
{
    if (form == RTSIM_FORM_PLAIN_FP32) return check_plain_limits();
    if (form == RTSIM_FORM_MFMA_BF16 || form == RTSIM_FORM_MFMA_FP16) return ensure_grid();
    return fail(RTSIM_ERR_INVALID, "unknown form %d", form);
}

StreamArgs Sensor::stream_args(const rtsim_hits& hits, size_t n, size_t offset) const
{
    StreamArgs a;
    a.range = hits.range;
    a.dx = hits.dir_x;
    a.dy = hits.dir_y;
    a.dz = hits.dir_z;
    a.amp = reinterpret_cast<const float2*>(hits.amplitude);
    a.out = d_samples + offset;
    a.n = (uint32_t)n;
    std::memcpy(a.R, R, sizeof(a.R));
    a.k2 = 2.0 * (double)frequency / (double)sound_speed;
    a.spreading = spreading;
    a.is2d = kind == RTSIM_SENSOR_2D;
    a.dir = d_dir;
    a.dir_w = (int)dir_w;
    a.dir_h = (int)dir_h;
    a.dir_xs = dir_xs;
    a.dir_ys = dir_ys;
    return a;
}

hipEvent_t Sensor::take_event()
{
    if (!event_pool.empty()) {
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void Sensor::stage_begin(int stage, hipStream_t s)
{
    if (!profiling) return;
    StageEvent p;
    p.stage = stage;
    p.start = take_event();
    p.stop = take_event();
    if (!p.start || !p.stop) return;
    (void)hipEventRecord(p.start, s);
    events.push_back(p);
    open_event = (int)events.size() - 1;
}

void Sensor::stage_end(hipStream_t s)
{
    if (!profiling || open_event < 0) return;
    (void)hipEventRecord(events[open_event].stop, s);
    open_event = -1;
}

int Sensor::check_plain_limits() const
{
    if (kind == RTSIM_SENSOR_2D) {
        int dev = 0, max_lds = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                hipSuccess || max_lds <= 0)
            max_lds = 64 * 1024;
        ConvArgs c = conv_args();
        if (plain2d_lds_bytes(c) > (size_t)max_lds)
            return fail(RTSIM_ERR_UNSUPPORTED, "plain form: %zu B of LDS exceed the device's %d B",
                        plain2d_lds_bytes(c), max_lds);
    }
    return RTSIM_OK;
}

PingBuffers Sensor::own_buffers() const
{
    PingBuffers b{};
    b.samples = d_samples;
    b.n = (uint32_t)count;
    b.tiles_cap = (uint32_t)(sort_count_words(capacity) / 256);
    b.sorted = d_sorted;
    b.coord = d_coord;
    b.items = d_items;
    b.items_cap = d_items ? 2 * capacity : 0;
    grid_buffers(&b);   // bucket counts and key occupancy (when the grid form is prepared)
    b.kv[0] = d_kv[0];
    b.kv[1] = d_kv[1];
    b.order = d_order;
    b.skeys = d_skeys;
    b.counts = d_counts;
    b.totals = d_totals;
    b.tmax = d_tmax;
    b.seg = d_seg;
    b.win = d_win;
    b.work = d_work;
    if (d_strip_ent && strip_mult) {
        b.strip_ent = d_strip_ent;
        b.strip_cnt = d_strip_cnt;
        b.strip_cap = strip_cap;
    }
    b.out = d_out;
    return b;
}

// Sort by travel (Receiver.cu:28-34) whose last pass gathers the records,
// their bearing coordinates (and on request the Binner keys, Binner.cu:24-49,
// and the permutation) and bounds the segments (Binner.cu:51-80), then the
// windows (Binner::compute, Binner.h:96-102).
static int plan_sort(Sensor* s, const PingBuffers& pb, bool introspect, PingPlan* p)
{
    SortBuffers sb{};
    sb.kv[0] = pb.kv[0];
    sb.kv[1] = pb.kv[1];
    sb.order = introspect ? pb.order : nullptr;
    sb.counts = pb.counts;
    sb.totals = pb.totals;
    sb.tiles_cap = pb.tiles_cap;
    sb.seg_init = pb.seg;                             // Binner segments start at (~0, 0)
    sb.seg_count = s->N;
    sb.out.sorted = pb.sorted;
    sb.out.skeys = introspect ? pb.skeys : nullptr;
    sb.out.coord = pb.coord;
    sb.out.key = s->key;
    sb.out.axis = s->axis;
    sb.out.seg = pb.seg;
    sb.by_key = !introspect;   // the travel order only for the Binner introspection
    hipError_t e = make_sort_plan(pb.samples, pb.n, sb, &p->sort);
    if (e != hipSuccess) return hip_fail(e, "sort plan");
    BinArgs& ba = p->bins;
    ba.N = s->N;
    ba.ov = s->overlap;
    ba.seg = pb.seg;
    ba.win = pb.win;
    return RTSIM_OK;
}

int Sensor::plan(int form, PingPlan* p, const PingBuffers* pbp)
{
    const bool grid = form == RTSIM_FORM_MFMA_BF16 || form == RTSIM_FORM_MFMA_FP16;
    int rc = RTSIM_OK;
    if (form == RTSIM_FORM_PLAIN_FP32) {
        if ((rc = check_plain_limits())) return rc;
        if (!pbp && (rc = ensure_strips())) return rc;
    } else if (!grid) {
        return fail(RTSIM_ERR_INVALID, "unknown form %d", form);
    } else if ((rc = ensure_grid())) {   // also checks the grid form applies
        return rc;
    }
    const PingBuffers pb = pbp ? *pbp : own_buffers();
    p->form = form;
    p->grid = grid;
    if (grid) {
        // the deposit sums every sample exactly, so it needs them grouped by
        // grid row only (one bucketing pass) and the windows' emptiness
        p->sort = SortPlan{};
        return grid_plan(form, &p->bucket, &p->dep, &p->corr, pb);
    }
    if ((rc = plan_sort(this, pb, false, p))) return rc;
    p->conv = conv_args(pb);
    return RTSIM_OK;
}

int Sensor::plan_binning(PingPlan* p)
{
    p->form = RTSIM_FORM_PLAIN_FP32;
    p->grid = false;
    return plan_sort(this, own_buffers(), true, p);
}

int Sensor::compute(int form, hipStream_t s)
{
    (void)hipGetLastError();   // start from a clean error state: launches are checked below
    PingPlan p;
    int rc = plan(form, &p);
    if (rc) return rc;
    const bool fused = p.grid && hist_fused;
    hist_fused = false;   // the sort consumes (and overwrites) the counts
    if ((rc = launch_ping(p, form, s, fused, profiling))) return rc;
    if (p.grid) last_corr = p.corr;
    computed = true;
    computed_form = form;
    return RTSIM_OK;
}

int Sensor::launch_ping(const PingPlan& p, int form, hipStream_t s, bool skip_hist0, bool profile)
{
    const bool prof = profiling;
    profiling = profile && prof;
    auto stage = [&](int k, auto&& fn) -> hipError_t {
        stage_begin(k, s);
        const hipError_t e = fn();
        stage_end(s);
        return e;
    };
    hipError_t e;
    if (!p.grid) {
        e = stage(1, [&] { return launch_sort_plan(p.sort, s); });
        if (e == hipSuccess)
            e = stage(2, [&] {
                return kind == RTSIM_SENSOR_2D ? launch_plain_bins(p.bins, p.conv, s) : launch_windows(p.bins, s);
            });
        if (e == hipSuccess)
            e = stage(4, [&] {
                return kind == RTSIM_SENSOR_2D ? launch_conv2d_plain(p.conv, s) : launch_conv1d_plain(p.conv, s);
            });
    } else {
        e = stage(1, [&] {
            hipError_t r = skip_hist0 ? hipSuccess : launch_bucket_hist(p.bucket, s);
            if (r == hipSuccess) r = launch_bucket_scan(p.bucket, s);
            if (r == hipSuccess) r = launch_bucket_scatter(p.bucket, s);
            return r;
        });
        if (e == hipSuccess) e = stage(3, [&] { return launch_deposit(p.dep, s); });
        if (e == hipSuccess)
            e = stage(4, [&] { return launch_correlate(p.corr, form == RTSIM_FORM_MFMA_FP16, s); });
    }
    profiling = prof;
    if (e != hipSuccess) return hip_fail(e, "launching the ping");
    return RTSIM_OK;
}

ConvArgs Sensor::conv_args(const PingBuffers& pb) const
{
    ConvArgs c;
    c.sorted = pb.sorted;
    c.coord = pb.coord;
    c.win = pb.win;
    c.range_c = d_rng;
    c.bear_c = d_bear;
    c.psf = d_psf;
    c.psf_cells = d_psf_cells;
    c.work = pb.work;
    c.n = pb.n;
    const bool lists = kind == RTSIM_SENSOR_2D && pb.strip_ent && pb.strip_cap >= (size_t)pb.n * strip_mult &&
                       strip_mult > 0;
    c.strip_ent = lists ? pb.strip_ent : nullptr;
    c.strip_cnt = lists ? pb.strip_cnt : nullptr;
    c.strip_cap = lists ? pb.strip_cap : 0;
    c.out = pb.out;
    c.H = (int)N;
    c.W = (int)W;
    c.Tr = (int)Tr;
    c.Tb = (int)Tb;
    c.inv_pr = (float)((double)Tr / (double)psf_range_span);
    c.half_r = 0.5f * (float)(Tr - 1);
    c.inv_pb = (float)((double)Tb / (double)psf_bearing_span);
    c.half_b = 0.5f * (float)(Tb - 1);
    c.supp_b = (float)(0.5 * (double)(Tb + 1) * (double)psf_bearing_span / (double)Tb);
    // an (A, B) cell plane of up to kPlainLdsCells cells goes to LDS; a larger
    // one is read through L1/L2 like the (C, D) plane
    c.psf_in_lds = ((size_t)Tr + 2) * ((size_t)Tb + 2) <= kPlainLdsCells;
    // 3072-sample chunks when they and the staged plane still fit two
    // workgroups per CU (measured at 512 x 1024 with the 31 x 17 PSF: the
    // gather 7% faster than with 2048)
    c.chunk3 = plain2d_stage_bytes(3) +
                   (c.psf_in_lds ? sizeof(float4) * ((size_t)Tr + 2) * ((size_t)Tb + 2) : 0) <= kPlainLdsBudget;
    return c;
}

}  // namespace rtsim

using namespace rtsim;

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

const char* rtsim_last_error(void) { return g_last_error.c_str(); }

const char* rtsim_version(void) { return "rtsim 0.1.0 (gfx950)"; }

int rtsim_device_count(int* count)
{
    if (!count) return fail(RTSIM_ERR_INVALID, "count is NULL");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return RTSIM_OK;
}

int rtsim_stream_sync(void* stream)
{
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return RTSIM_OK;
}

static int validate_desc(const rtsim_sensor_desc* d)
{
    if (!d) return fail(RTSIM_ERR_INVALID, "desc is NULL");
    if (d->kind != RTSIM_SENSOR_1D && d->kind != RTSIM_SENSOR_2D)
        return fail(RTSIM_ERR_INVALID, "unknown sensor kind %d", d->kind);
    if (d->range_count < 1) return fail(RTSIM_ERR_INVALID, "range_count must be >= 1");
    if (!(d->range_max > d->range_min))
        return fail(RTSIM_ERR_INVALID, "range_max must exceed range_min");
    if (d->kind == RTSIM_SENSOR_2D) {
        if (!d->bearings || d->bearing_count < 1)
            return fail(RTSIM_ERR_INVALID, "2D sensor needs a bearing table");
        for (uint32_t b = 1; b < d->bearing_count; b++)
            if (!(d->bearings[b] > d->bearings[b - 1]))
                return fail(RTSIM_ERR_INVALID, "bearings must be strictly ascending");
        if (d->bearing_axis != RTSIM_AXIS_ANGLE && d->bearing_axis != RTSIM_AXIS_SINE)
            return fail(RTSIM_ERR_INVALID, "unknown bearing axis %d", d->bearing_axis);
    }
    if (const rtsim_waveform_sine* w = d->waveform) {
        if (!(w->frequency > 0.f) || w->oversampling < 1)
            return fail(RTSIM_ERR_INVALID, "waveform: frequency and oversampling must be > 0");
        if (w->fixed && !(w->duration > 0.f))
            return fail(RTSIM_ERR_INVALID, "waveform: a fixed waveform needs a duration > 0");
        if (w->beam_taps < 1 || (w->beam_taps > 1 && !w->beam))
            return fail(RTSIM_ERR_INVALID, "waveform: beam taps missing");
        if (d->kind == RTSIM_SENSOR_1D && w->beam_taps != 1)
            return fail(RTSIM_ERR_INVALID, "1D sensor needs beam_taps == 1");
    } else {
        if (!d->psf || d->psf_range_taps < 1 || d->psf_bearing_taps < 1)
            return fail(RTSIM_ERR_INVALID, "PSF taps missing");
        if (d->kind == RTSIM_SENSOR_1D && d->psf_bearing_taps != 1)
            return fail(RTSIM_ERR_INVALID, "1D sensor needs psf_bearing_taps == 1");
        if (!(d->psf_range_span > 0.f)) return fail(RTSIM_ERR_INVALID, "psf_range_span must be > 0");
    }
    if (d->kind == RTSIM_SENSOR_2D && !(d->psf_bearing_span > 0.f))
        return fail(RTSIM_ERR_INVALID, "psf_bearing_span must be > 0");
    if (d->directivity && (d->dir_width < 1 || d->dir_height < 1))
        return fail(RTSIM_ERR_INVALID, "directivity table size");
    if (!(d->sound_speed > 0.f)) return fail(RTSIM_ERR_INVALID, "sound_speed must be > 0");
    return RTSIM_OK;
}

int rtsim_sensor_create(const rtsim_sensor_desc* d, rtsim_sensor* out)
{
    if (!out) return fail(RTSIM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate_desc(d);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(RTSIM_ERR_HIP, "no HIP device available");
    init_kernel_attributes();
    Sensor* s = new Sensor();
    s->kind = d->kind;
    s->range_min = d->range_min;
    s->range_max = d->range_max;
    s->N = d->range_count;
    s->W = d->kind == RTSIM_SENSOR_2D ? d->bearing_count : 1u;
    s->axis = d->kind == RTSIM_SENSOR_2D ? d->bearing_axis : RTSIM_AXIS_ANGLE;
    if (d->kind == RTSIM_SENSOR_2D) s->bearings.assign(d->bearings, d->bearings + d->bearing_count);
    else s->bearings.assign(1, 0.f);
    if (const rtsim_waveform_sine* w = d->waveform) {
        s->has_wave = true;
        s->wave_frequency = w->frequency;
        s->wave_duration = w->duration;
        s->wave_fixed = w->fixed != 0;
        s->wave_os = w->oversampling;
        if (w->beam) s->beam.assign(w->beam, w->beam + w->beam_taps);
        else s->beam.assign(1, 1.0f);
    } else {
        s->Tr = d->psf_range_taps;
        s->Tb = d->psf_bearing_taps;
        s->psf.assign(d->psf, d->psf + 2 * (size_t)s->Tr * s->Tb);
        s->psf_range_span = d->psf_range_span;
    }
    s->psf_bearing_span = d->kind == RTSIM_SENSOR_2D ? d->psf_bearing_span : 1.0f;
    s->margin_request = d->margin;
    s->frequency = d->frequency;
    s->sound_speed = d->sound_speed;
    s->spreading = d->spreading;
    std::memcpy(s->R, d->rotation, sizeof(s->R));
    if (d->directivity) {
        s->dir.assign(d->directivity, d->directivity + (size_t)d->dir_width * d->dir_height);
        s->dir_w = d->dir_width;
        s->dir_h = d->dir_height;
        s->dir_xs = d->dir_xscale;
        s->dir_ys = d->dir_yscale;
    }
    // device-side constant tables
    std::vector<float2> bc(s->W);
    for (uint32_t b = 0; b < s->W; b++) {
        double x = s->axis == RTSIM_AXIS_SINE ? std::sin((double)s->bearings[b])
                                               : (double)s->bearings[b];
        bc[b] = split_df(x);
    }
    if ((rc = dev_alloc(&s->d_bear, s->W)) || (rc = dev_alloc(&s->d_totals, 256))) {
        delete s;
        return rc;
    }
    if (hipMemcpy(s->d_bear, bc.data(), sizeof(float2) * s->W, hipMemcpyHostToDevice) != hipSuccess) {
        delete s;
        return fail(RTSIM_ERR_HIP, "uploading sensor tables failed");
    }
    if ((rc = s->has_wave ? s->apply_waveform() : s->upload_psf())) {
        delete s;
        return rc;
    }
    if (!s->dir.empty()) {
        if ((rc = dev_alloc(&s->d_dir, s->dir.size()))) { delete s; return rc; }
        if (hipMemcpy(s->d_dir, s->dir.data(), sizeof(float) * s->dir.size(),
                      hipMemcpyHostToDevice) != hipSuccess) {
            delete s;
            return fail(RTSIM_ERR_HIP, "uploading the directivity table failed");
        }
    }
    if ((rc = s->configure_ranges())) { delete s; return rc; }
    *out = reinterpret_cast<rtsim_sensor>(s);
    return RTSIM_OK;
}

static Sensor* get(rtsim_sensor h) { return reinterpret_cast<Sensor*>(h); }

#define GET_SENSOR(h)                                                    \
    Sensor* s = get(h);                                                  \
    if (!s) return fail(RTSIM_ERR_INVALID, "sensor handle is NULL");

int rtsim_sensor_destroy(rtsim_sensor h)
{
    delete get(h);
    return RTSIM_OK;
}

int rtsim_sensor_set_ranges(rtsim_sensor h, float range_min, float range_max,
                            uint32_t range_count, float margin)
{
    GET_SENSOR(h);
    if (range_count < 1) return fail(RTSIM_ERR_INVALID, "range_count must be >= 1");
    if (!(range_max > range_min)) return fail(RTSIM_ERR_INVALID, "range_max must exceed range_min");
    const float old_min = s->range_min, old_max = s->range_max;
    const uint32_t old_n = s->N;
    s->range_min = range_min;
    s->range_max = range_max;
    s->N = range_count;
    s->computed = false;
    s->hist_fused = false;
    int rc = s->apply_waveform();
    if (rc) {   // the sensor keeps its previous ranges and table
        s->range_min = old_min;
        s->range_max = old_max;
        s->N = old_n;
        return rc;
    }
    s->margin_request = margin;
    return s->configure_ranges();
}

int rtsim_sensor_set_pose(rtsim_sensor h, const float* rotation)
{
    GET_SENSOR(h);
    if (!rotation) return fail(RTSIM_ERR_INVALID, "rotation is NULL");
    for (int k = 0; k < 9; k++)
        if (!(rotation[k] == rotation[k])) return fail(RTSIM_ERR_INVALID, "rotation has a NaN");
    std::memcpy(s->R, rotation, sizeof(s->R));
    s->pose_version++;   // batched stream launches hold the rotation
    return RTSIM_OK;
}

int rtsim_sensor_set_sample_count(rtsim_sensor h, size_t count)
{
    GET_SENSOR(h);
    if (count > RTSIM_MAX_SAMPLES)
        return fail(RTSIM_ERR_INVALID, "%zu samples exceed the sort's limit of %u per sensor", count,
                    RTSIM_MAX_SAMPLES);
    int rc = s->reserve(count);
    if (rc) return rc;
    if (s->count != count) s->generation++;
    s->count = count;
    s->computed = false;
    s->hist_fused = false;
    return RTSIM_OK;
}

int rtsim_sensor_sample_count(rtsim_sensor h, size_t* count)
{
    GET_SENSOR(h);
    if (!count) return fail(RTSIM_ERR_INVALID, "count is NULL");
    *count = s->count;
    return RTSIM_OK;
}

int rtsim_sensor_samples(rtsim_sensor h, void** dev_ptr)
{
    GET_SENSOR(h);
    if (!dev_ptr) return fail(RTSIM_ERR_INVALID, "dev_ptr is NULL");
    *dev_ptr = s->d_samples;
    s->hist_fused = false;   // the caller may write the samples
    return RTSIM_OK;
}

int rtsim_receiver_write_samples(rtsim_sensor h, const rtsim_hits* hits, size_t n,
                                 size_t offset, void* stream)
{
    GET_SENSOR(h);
    if (!hits || !hits->range || !hits->dir_x || !hits->dir_y || !hits->dir_z || !hits->amplitude)
        return fail(RTSIM_ERR_INVALID, "hits has a NULL array");
    if (offset + n > s->count)
        return fail(RTSIM_ERR_INVALID, "samples [%zu, %zu) exceed sample_count %zu", offset,
                    offset + n, s->count);
    hipStream_t st = (hipStream_t)stream;
    StreamArgs a = s->stream_args(*hits, n, offset);
    (void)hipGetLastError();
    // a write of every sample into a sensor whose grid form is prepared also
    // makes that form's bucket counts (one launch fewer per ping), when its
    // tiles are the streaming kernel's (1024 samples)
    PingPlan p;
    const bool fuse = offset == 0 && n == s->count && n > 0 && s->grid_ready() &&
                      s->plan(RTSIM_FORM_MFMA_FP16, &p) == RTSIM_OK && p.bucket.tile == 1024 &&
                      stream_blocks(a) == p.bucket.tiles;
    s->stage_begin(0, st);
    if (fuse) HIP_TRY(launch_stream_bucket(a, p.bucket, st));
    else HIP_TRY(launch_stream(a, st));
    s->stage_end(st);
    s->hist_fused = fuse;
    s->computed = false;
    return RTSIM_OK;
}

int rtsim_sensor_compute(rtsim_sensor h, int form, void* stream)
{
    GET_SENSOR(h);
    return s->compute(form, (hipStream_t)stream);
}

int rtsim_sensor_output(rtsim_sensor h, const void** dev_ptr, uint32_t* height, uint32_t* width)
{
    GET_SENSOR(h);
    if (dev_ptr) *dev_ptr = s->d_out;
    if (height) *height = s->N;
    if (width) *width = s->W;
    return RTSIM_OK;
}

int rtsim_sensor_sorted_samples(rtsim_sensor h, const void** dev_ptr)
{
    GET_SENSOR(h);
    if (!dev_ptr) return fail(RTSIM_ERR_INVALID, "dev_ptr is NULL");
    *dev_ptr = s->d_sorted;
    return RTSIM_OK;
}

int rtsim_sensor_binner_debug(rtsim_sensor h, uint32_t* keys, uint32_t* segments,
                              uint32_t* windows, uint32_t* order, int* overlap, void* stream)
{
    GET_SENSOR(h);
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipStreamSynchronize(st));
    if (!s->computed) return fail(RTSIM_ERR_INVALID, "compute has not run since the last change");
    size_t n = s->count;
    const bool grid = s->computed_form == RTSIM_FORM_MFMA_BF16 || s->computed_form == RTSIM_FORM_MFMA_FP16;
    if (((keys || order) && n) || (grid && (segments || windows))) {
        // compute writes neither the keys nor the permutation (nothing
        // downstream reads them), and the grid form does not sort at all (it
        // bins by grid row and keeps only the windows' emptiness): sort the
        // same samples by travel with all of them kept, then the windows
        // (which also write empty segments back as (0, 0)).  Both are
        // deterministic, so every array they rewrite (sorted records,
        // coordinates, segments, windows) gets the contents a plain-form
        // compute leaves.
        PingPlan p;
        int rc = s->plan_binning(&p);
        if (rc) return rc;
        (void)hipGetLastError();
        HIP_TRY(launch_sort_plan(p.sort, st));
        HIP_TRY(launch_windows(p.bins, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    if (keys && n) HIP_TRY(hipMemcpy(keys, s->d_skeys, 4 * n, hipMemcpyDeviceToHost));
    if (segments) HIP_TRY(hipMemcpy(segments, s->d_seg, 8 * (size_t)s->N, hipMemcpyDeviceToHost));
    if (windows) HIP_TRY(hipMemcpy(windows, s->d_win, 8 * (size_t)s->N, hipMemcpyDeviceToHost));
    if (order && n) HIP_TRY(hipMemcpy(order, s->d_order, 4 * n, hipMemcpyDeviceToHost));
    if (overlap) *overlap = s->overlap;
    return RTSIM_OK;
}

int rtsim_sensor_upload_samples(rtsim_sensor h, const void* host, size_t n, void* stream)
{
    GET_SENSOR(h);
    if (n > s->count) return fail(RTSIM_ERR_INVALID, "n exceeds sample_count");
    hipStream_t st = (hipStream_t)stream;
    if (n) HIP_TRY(hipMemcpyAsync(s->d_samples, host, 16 * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    s->computed = false;
    s->hist_fused = false;
    return RTSIM_OK;
}

int rtsim_sensor_download_samples(rtsim_sensor h, void* host, size_t n, int sorted, void* stream)
{
    GET_SENSOR(h);
    if (n > s->count) return fail(RTSIM_ERR_INVALID, "n exceeds sample_count");
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipStreamSynchronize(st));
    if (n)
        HIP_TRY(hipMemcpy(host, sorted ? (const void*)s->d_sorted : (const void*)s->d_samples,
                          16 * n, hipMemcpyDeviceToHost));
    return RTSIM_OK;
}

int rtsim_sensor_download_output(rtsim_sensor h, float* host, void* stream)
{
    GET_SENSOR(h);
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipMemcpy(host, s->d_out, sizeof(float2) * (size_t)s->N * s->W, hipMemcpyDeviceToHost));
    return RTSIM_OK;
}

int rtsim_sensor_set_profiling(rtsim_sensor h, int on)
{
    GET_SENSOR(h);
    HIP_TRY(hipDeviceSynchronize());
    for (auto& p : s->events) {
        s->event_pool.push_back(p.start);
        s->event_pool.push_back(p.stop);
    }
    s->events.clear();
    s->open_event = -1;
    s->profiling = on != 0;
    return RTSIM_OK;
}

int rtsim_sensor_stage_times(rtsim_sensor h, float* ms, int nstages, int* ncomputes)
{
    GET_SENSOR(h);
    if (!ms || nstages < 1) return fail(RTSIM_ERR_INVALID, "ms buffer");
    std::vector<double> sum(nstages, 0.0);
    std::vector<int> cnt(nstages, 0);
    for (auto& p : s->events) {
        HIP_TRY(hipEventSynchronize(p.stop));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, p.start, p.stop));
        if (p.stage < nstages) {
            sum[p.stage] += t;
            cnt[p.stage]++;
        }
    }
    int maxc = 0;
    for (int i = 0; i < nstages; i++) {
        ms[i] = cnt[i] ? (float)(sum[i] / cnt[i]) : 0.f;
        maxc = std::max(maxc, cnt[i]);
    }
    if (ncomputes) *ncomputes = maxc;
    return RTSIM_OK;
}

int rtsim_stream_create(void** stream)
{
    if (!stream) return fail(RTSIM_ERR_INVALID, "stream is NULL");
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *stream = st;
    return RTSIM_OK;
}

int rtsim_stream_destroy(void* stream)
{
    if (stream) HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return RTSIM_OK;
}

struct Timer {
    hipEvent_t start = nullptr, stop = nullptr;
};

int rtsim_timer_create(void** timer)
{
    if (!timer) return fail(RTSIM_ERR_INVALID, "timer is NULL");
    Timer* t = new Timer();
    if (hipEventCreate(&t->start) != hipSuccess || hipEventCreate(&t->stop) != hipSuccess) {
        delete t;
        return fail(RTSIM_ERR_HIP, "hipEventCreate failed");
    }
    *timer = t;
    return RTSIM_OK;
}

int rtsim_timer_destroy(void* timer)
{
    Timer* t = (Timer*)timer;
    if (!t) return RTSIM_OK;
    (void)hipEventDestroy(t->start);
    (void)hipEventDestroy(t->stop);
    delete t;
    return RTSIM_OK;
}

int rtsim_timer_start(void* timer, void* stream)
{
    if (!timer) return fail(RTSIM_ERR_INVALID, "timer is NULL");
    HIP_TRY(hipEventRecord(((Timer*)timer)->start, (hipStream_t)stream));
    return RTSIM_OK;
}

int rtsim_timer_stop(void* timer, void* stream)
{
    if (!timer) return fail(RTSIM_ERR_INVALID, "timer is NULL");
    HIP_TRY(hipEventRecord(((Timer*)timer)->stop, (hipStream_t)stream));
    return RTSIM_OK;
}

int rtsim_timer_elapsed(void* timer, float* ms)
{
    if (!timer || !ms) return fail(RTSIM_ERR_INVALID, "timer/ms is NULL");
    Timer* t = (Timer*)timer;
    HIP_TRY(hipEventSynchronize(t->stop));
    HIP_TRY(hipEventElapsedTime(ms, t->start, t->stop));
    return RTSIM_OK;
}

int rtsim_device_alloc(void** ptr, size_t bytes)
{
    if (!ptr) return fail(RTSIM_ERR_INVALID, "ptr is NULL");
    HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
    return RTSIM_OK;
}

int rtsim_device_free(void* ptr)
{
    if (ptr) HIP_TRY(hipFree(ptr));
    return RTSIM_OK;
}

int rtsim_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return RTSIM_OK;
}

int rtsim_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return RTSIM_OK;
}

}  // extern "C"
