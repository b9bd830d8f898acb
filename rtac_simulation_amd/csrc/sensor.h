// Host-side sensor instance state (one per rtsim_sensor handle).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/rtsim.h"
#include "grid_internal.h"
#include "rtsim_internal.h"

namespace rtsim {

// one-time kernel attribute setup (dynamic LDS limits); never under capture
void init_kernel_attributes();

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
const char* fail_message();
KeyParams make_key(uint32_t N, float lo, float hi, float margin);
int make_overlap(float margin, float lo, float hi, uint32_t N);
// Waveform_Sine samples times the beam taps (complex interleaved
// [range tap][bearing tap]); *range_taps = floor((2 f os) duration), 0 when
// the duration is too short (or out of the table limit).
std::vector<float> sine_psf_table(float frequency, float duration, uint32_t oversampling,
                                  const float* beam, uint32_t beam_taps, uint32_t* range_taps);
constexpr uint32_t kMaxWaveformTaps = 1u << 16;

// Everything one ping launches, without launching it (a batch gathers the
// plans of many pings into per-stage launches).
struct PingPlan {
    int form = 0;
    bool grid = false;
    SortPlan sort{};          // plain form (and Binner introspection)
    BinArgs bins{};
    ConvArgs conv{};
    BucketArgs bucket{};      // grid form
    DepositArgs dep{};
    CorrArgs corr{};
};

// Device buffers one ping runs through: the sensor's own (Sensor::own_buffers)
// or ones a caller provides (the scratch of rtsim_sensor_compute_noalloc).
struct PingBuffers {
    const float4* samples;   // unsorted records (read only)
    uint32_t n;              // samples
    uint32_t tiles_cap;      // tile-count columns of `counts`
    float4* sorted;
    float2* coord;           // bearing-axis coordinates (hi, lo)
    float4* items;           // grid form: 2 x n bucketed items
    size_t items_cap;        // grid form: items the buffer holds
    uint32_t* bcounts;       // grid form: bucket counts
    size_t bcounts_cap;      // words
    uint32_t* occ;           // grid form: key occupancy per tile
    size_t occ_cap;          // words
    uint2* kv[2];
    uint32_t* order;         // introspection only (may be null otherwise)
    uint32_t* skeys;         // introspection only (may be null otherwise)
    uint32_t* counts;
    uint32_t* totals;        // 256 words
    uint32_t* tmax;
    uint2* seg;              // [N]
    uint2* win;              // [N]
    uint32_t* work;          // plain 2D gather: band lengths and order
    uint2* strip_ent;        // plain 2D gather: strip lists (null: none)
    uint32_t* strip_cnt;
    size_t strip_cap;        // entries
    float2* out;             // [N][W]
};

struct StageEvent {
    int stage;
    hipEvent_t start, stop;
};

struct Sensor {
    // ---- configuration (SensorInfo / SensorInstance members) -------------
    int kind = 0;
    float range_min = 0.f, range_max = 1.f;
    uint32_t N = 1;              // range bins
    uint32_t W = 1;              // bearings (1 for 1D)
    int axis = 0;
    std::vector<float> bearings;
    uint32_t Tr = 1, Tb = 1;
    std::vector<float> psf;      // complex interleaved [Tr][Tb]
    float psf_range_span = 1.f, psf_bearing_span = 1.f;
    float margin_request = -1.f;
    float margin = 0.f;
    std::vector<float> dir;
    uint32_t dir_w = 0, dir_h = 0;
    float dir_xs = 0.f, dir_ys = 0.f;
    float frequency = 0.f, sound_speed = 1500.f;
    int spreading = 0;
    float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    // sine waveform recipe (rtsim_waveform_sine): the PSF follows the ranges
    bool has_wave = false;
    float wave_frequency = 0.f, wave_duration = 0.f;
    bool wave_fixed = false;
    uint32_t wave_os = 8;
    std::vector<float> beam;
    KeyParams key{};
    int overlap = 0;

    // ---- samples -----------------------------------------------------------
    size_t count = 0, capacity = 0;
    bool computed = false;
    // the last write_samples also made the grid form's bucket counts (and
    // tile maxima) for exactly the resident samples: compute skips them
    bool hist_fused = false;
    int computed_form = -1;       // form of the last compute (batched or not)
    float4* d_samples = nullptr;  // receivedSamples_ (unsorted)
    float4* d_sorted = nullptr;   // receivedSamples_ after sort()
    float4* d_items = nullptr;    // grid form: bucketed items (2 per sample at most)
    float2* d_coord = nullptr;    // bearing-axis coordinate (hi, lo) per sorted sample
    uint2* d_kv[2] = {nullptr, nullptr};  // sort: (key, index) between passes
    uint32_t* d_order = nullptr;  // sort permutation (written on request)
    uint32_t* d_skeys = nullptr;  // Binner::keys_ of the sorted samples (written on request)
    uint32_t* d_totals = nullptr; // sort: digit totals of a pass
    uint32_t* d_counts = nullptr; // sort: per-tile digit counts (capacity-sized)
    uint32_t* d_tmax = nullptr;   // grid form: per-tile largest |re|, |im|

    // ---- per-configuration device tables -----------------------------------
    uint2* d_seg = nullptr;       // Binner::binIndexes_
    uint2* d_win = nullptr;       // bins_ as (start, size)
    uint32_t* d_work = nullptr;   // plain 2D gather: band lengths and order
    // plain 2D gather: strip lists (ConvArgs::strip_*), sized for `capacity`
    // samples reaching at most strip_mult strips each (0: no lists)
    uint32_t strip_mult = 0;
    uint2* d_strip_ent = nullptr;
    uint32_t* d_strip_cnt = nullptr;
    size_t strip_cap = 0, strip_cnt_cap = 0;
    size_t strip_failed = 0;      // entries of a list allocation that failed (0: none)
    float2* d_out = nullptr;      // ping [N][W]
    float2* d_psf = nullptr;
    float4* d_psf_cells = nullptr;  // plain 2D gather: bilinear cells of the taps (ConvArgs::psf_cells)
    float2* d_bear = nullptr;     // bearing coordinate (hi, lo)
    float2* d_rng = nullptr;      // range coordinate (hi, lo)
    float* d_dir = nullptr;

    // ---- grid (MFMA) form state --------------------------------------------
    void* grid_state = nullptr;
    CorrArgs last_corr{};

    // ---- profiling -----------------------------------------------------------
    bool profiling = false;
    std::vector<StageEvent> events;
    std::vector<hipEvent_t> event_pool;
    int open_event = -1;

    // bumped whenever buffers or kernel arguments change (captured graphs
    // built from this sensor must then be re-captured)
    uint64_t generation = 0;
    // bumped by set_pose: a batch rewrites that ping's streaming arguments
    // in place (the rotation is a kernel argument) instead of re-capturing
    uint64_t pose_version = 0;

    ~Sensor();
    void free_samples();
    void free_grid();
    int reserve(size_t count);
    int configure_ranges();
    int apply_waveform();           // recipe -> taps, spans (no-op without a recipe)
    int upload_psf();
    int check_plain_limits() const;
    void free_strips();
    int ensure_strips();             // strip lists for `capacity` samples (when they apply)
    int prepare(int form);          // form-specific state, no device work queued
    int ensure_grid();
    bool grid_ready() const;         // grid form prepared (no preparation is started)
    int grid_reserve();              // grid form buffers sized for `capacity`
    void grid_buffers(PingBuffers* pb) const;   // the grid form's own buffers
    int compute(int form, hipStream_t s);
    // queues a planned ping (profiling events only when `profile`); skip_hist0:
    // the grid form's bucket counts are already made (the streaming kernel)
    int launch_ping(const PingPlan& p, int form, hipStream_t s, bool skip_hist0, bool profile);
    int grid_plan(int form, BucketArgs* b, DepositArgs* d, CorrArgs* c, const PingBuffers& pb);
    // validates the form; queues nothing.  pb: the buffers (default: own)
    int plan(int form, PingPlan* p, const PingBuffers* pb = nullptr);
    // Binner introspection: the sort by travel with the keys and the
    // permutation kept, then the windows (any form)
    int plan_binning(PingPlan* p);
    PingBuffers own_buffers() const;
    int autotune(int form, int iterations, hipStream_t s, int* best_shape, float* best_ms);
    StreamArgs stream_args(const rtsim_hits& hits, size_t n, size_t offset) const;
    ConvArgs conv_args() const { return conv_args(own_buffers()); }
    ConvArgs conv_args(const PingBuffers& pb) const;
    hipEvent_t take_event();
    void stage_begin(int stage, hipStream_t s);
    void stage_end(hipStream_t s);
};

}  // namespace rtsim
