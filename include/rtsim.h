/*
 * rtsim.h -- C ABI of the MI355X (gfx950) sample-to-ping library.
 *
 * The library replaces the reference's sensor-instance pipeline (rtac_simulation
 * L3): per-sample directivity x round-trip phase, sort by travel, Binner
 * (keys, segments, windows) and the sparse complex-PSF convolution into 1D /
 * 2D complex pings.  Every entry point below cites the reference interface it
 * replaces (paths relative to the rtac_simulation source tree).
 *
 * Conventions
 *   - Every function returns an int status (RTSIM_OK = 0).  On error,
 *     rtsim_last_error() returns a thread-local message.
 *   - Device pointers returned by the library are borrowed: valid until the
 *     next set_* / compute call on the same handle or its destruction.
 *   - Work is queued on the caller's HIP stream (hipStream_t passed as
 *     void*; NULL = the default stream) and is asynchronous until
 *     rtsim_stream_sync().  A handle is not thread-safe; distinct handles are.
 *   - Sample record layout (16 B, both sensor kinds):
 *       {float re, float im, float travel, float bearing}
 *     SimSample2D is exactly this (Sample.h:101-161); a 1D sensor leaves
 *     `bearing` at 0.  A null sample has travel = -1 and value 0.
 *   - Ping layout: complex64, row-major [range][bearing] (1D: [range]),
 *     as SensorInstance2D.cu:81 writes out(blockIdx.y, blockIdx.x).
 */
#ifndef RTSIM_H
#define RTSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTSIM_OK 0
#define RTSIM_ERR_INVALID 1     /* bad argument or handle              */
#define RTSIM_ERR_HIP 2         /* HIP runtime error                   */
#define RTSIM_ERR_UNSUPPORTED 3 /* valid request this form cannot do   */
#define RTSIM_ERR_CONFIG 4      /* configuration file error            */
#define RTSIM_ERR_IO 5          /* file not found / unreadable         */

#define RTSIM_SENSOR_1D 1       /* SensorInstance1D_Complex            */
#define RTSIM_SENSOR_2D 2       /* SensorInstance2D_2<Complex<float>>  */

/* PSF convolution forms */
#define RTSIM_FORM_PLAIN_FP32 0 /* reference sparse gather, FP32       */
#define RTSIM_FORM_MFMA_BF16 1  /* grid form on MFMA, BF16 hi+lo split */
#define RTSIM_FORM_MFMA_FP16 2  /* grid form on MFMA, FP16 hi+lo split */

/* Coordinate in which the PSF bearing axis is uniform */
#define RTSIM_AXIS_ANGLE 0      /* radians, as KernelView2D (reference) */
#define RTSIM_AXIS_SINE 1       /* sin(bearing) (phased-array beams)    */

#define RTSIM_SPREAD_NONE 0
#define RTSIM_SPREAD_INV_R2 1   /* value /= r^2 (RayCaster.cu:70-71)    */

#define RTSIM_KEY_OUT_OF_RANGE 0xFFFFFFFFu /* Binner::Key::OutOfRange */

/* Largest sample count of one sensor: the sort scans at most 16384 tiles of
 * 1024 samples per pass (2^24 samples); set_sample_count refuses more.       */
#define RTSIM_MAX_SAMPLES 16777216u

typedef struct rtsim_sensor_s* rtsim_sensor;

/*
 * Sine waveform recipe: Waveform_Sine (Waveform.h:96-138) times the bearing
 * taps of the beam directivity, as SensorInstance{1D,2D}::generate_psf_data
 * (SensorInstance2D.cu:20-33, SensorInstance1D.cu:15-25) tabulates them.
 * A sensor created with a recipe keeps it: at creation and at every
 * rtsim_sensor_set_ranges it rebuilds what SensorInstance::set_ranges
 * (SensorInstance.cu:23-34) rebuilds -- an adaptive waveform's duration
 * tau = resolution / c (resolution = (max - min) / (count - 1) in FP32), the
 * range taps exp(i pi n / oversampling) for n < floor((2 f oversampling) tau),
 * the PSF table wave[h] * beam[w], its range span c tau and (unless a margin
 * is given) the Binner margin c tau.  A fixed waveform keeps its duration.
 */
typedef struct {
    float frequency;            /* Hz                                              */
    float duration;             /* s; replaced by resolution / c when not fixed    */
    int fixed;                  /* Waveform::fixed()                               */
    uint32_t oversampling;      /* samples per half period (reference default 8)   */
    const float* beam;          /* bearing taps (real); NULL with beam_taps = 1 = 1 */
    uint32_t beam_taps;         /* 1 for a 1D sensor                               */
} rtsim_waveform_sine;

/*
 * Sensor description: what SensorInfo / SensorInfo2D (SensorInfo.h:18-105),
 * Directivity (Directivity.h:43-83) and the PSF table of
 * SensorInstance{1D,2D}::generate_psf_data (SensorInstance1D.cu:15-25,
 * SensorInstance2D.cu:20-33) carry, with the PSF taps given explicitly.
 */
typedef struct {
    int kind;                   /* RTSIM_SENSOR_1D / RTSIM_SENSOR_2D               */
    /* ranges: Linspace<float>(range_min, range_max, range_count)                  */
    float range_min, range_max;
    uint32_t range_count;
    /* bearings (2D): ascending table in radians (SensorInfo2D::bearings_)         */
    const float* bearings;
    uint32_t bearing_count;
    int bearing_axis;           /* RTSIM_AXIS_*                                    */
    /* PSF taps: complex64 interleaved, row-major [range tap][bearing tap]         */
    const float* psf;
    uint32_t psf_range_taps, psf_bearing_taps; /* 1D: psf_bearing_taps = 1         */
    float psf_range_span;       /* metres covered by the range taps (c * tau)      */
    float psf_bearing_span;     /* bearing-axis units covered by the bearing taps  */
    float margin;               /* Binner margin (m); < 0 selects c * tau          */
    /* receiver directivity table [dir_height][dir_width] (Directivity texture);   */
    /* NULL selects a uniform directivity                                          */
    const float* directivity;
    uint32_t dir_width, dir_height;
    float dir_xscale, dir_yscale; /* DirectivityView scalings (1/pi by default)   */
    float frequency;            /* Hz, for the round-trip phase                    */
    float sound_speed;          /* m/s                                             */
    int spreading;              /* RTSIM_SPREAD_*                                  */
    float rotation[9];          /* receiver pose R, row-major; local = R^T dir     */
    /* NULL: the PSF taps above are fixed.  Else the taps, their range span and a */
    /* negative margin's c * tau follow the recipe, and psf, psf_range_taps,       */
    /* psf_bearing_taps and psf_range_span are ignored (psf_bearing_span is not).  */
    const rtsim_waveform_sine* waveform;
} rtsim_sensor_desc;

/* Ray hits handed to the streaming kernel (SoA, device pointers). */
typedef struct {
    const float* range;         /* one-way travel r (m); r < 1e-4 -> null sample   */
    const float* dir_x;         /* direction toward the hit, world frame           */
    const float* dir_y;
    const float* dir_z;
    const float* amplitude;     /* complex64 interleaved                           */
} rtsim_hits;

/* ---- library ---------------------------------------------------------------- */
const char* rtsim_last_error(void);
const char* rtsim_version(void);
/* Number of visible HIP devices (0 when no GPU). */
int rtsim_device_count(int* count);
int rtsim_stream_sync(void* stream);

/* ---- sensor lifetime: SensorInstance2D_2<Complex<float>>::Create
 *      (SensorInstance2D.h:122-139), SensorInstance1D_Complex::Create
 *      (SensorInstance1D.h:93-98) --------------------------------------------- */
int rtsim_sensor_create(const rtsim_sensor_desc* desc, rtsim_sensor* out);
int rtsim_sensor_destroy(rtsim_sensor h);

/* SensorInstance::set_ranges (SensorInstance.h:56-58, SensorInstance.cu:23-34):
 * new output ranges and Binner margin (< 0 selects c * tau).  A sensor with a
 * waveform recipe also re-derives the duration (adaptive), the PSF taps and
 * their range span; explicit taps are kept as given.                            */
int rtsim_sensor_set_ranges(rtsim_sensor h, float range_min, float range_max,
                            uint32_t range_count, float margin);

/* Receiver pose (Receiver.h:56, SensorInstance pose set by the simulation loop,
 * OptixSimulation.cpp:76-81): row-major rotation R of the receiver; the
 * streaming kernel maps hit directions to the receiver frame as R^T dir.     */
int rtsim_sensor_set_pose(rtsim_sensor h, const float* rotation);

/* SensorInstance{1D,2D}::set_sample_count / sample_count
 * (SensorInstance1D.h:47-48, SensorInstance2D.h:59-60).  At most
 * RTSIM_MAX_SAMPLES (RTSIM_ERR_INVALID beyond).                                   */
int rtsim_sensor_set_sample_count(rtsim_sensor h, size_t count);
int rtsim_sensor_sample_count(rtsim_sensor h, size_t* count);

/* receiver_view() samples pointer (SensorInstance.cu:41-50, Receiver.h:18-35):
 * the device array of 16-B records a producer may fill directly.  Records
 * written through it must be written after this call (it discards what a
 * previous rtsim_receiver_write_samples prepared for compute).                  */
int rtsim_sensor_samples(rtsim_sensor h, void** dev_ptr);

/* ReceiverView<T>::set_sample / set_null_sample (Receiver.h:51-65) applied to
 * n hits with the phase/spreading of RayCaster.cu:62-74: writes records
 * [offset, offset + n).  A write of every sample (offset 0, n = sample count)
 * into a sensor whose MFMA form is prepared (by an earlier MFMA compute or a
 * tile-shape call) also makes that form's bucket counts, which the next MFMA
 * compute then skips; any other write, upload, count or range change and any
 * compute discard them.                                                           */
int rtsim_receiver_write_samples(rtsim_sensor h, const rtsim_hits* hits, size_t n,
                                 size_t offset, void* stream);

/* SensorInstance::compute_output (SensorInstance.h:67; 2D: fill_ping
 * SensorInstance2D.h:84-93; 1D: reduce_samples SensorInstance1D.h:63-69).
 * Plain form: sort -> Binner::compute -> the PSF gather in FP32 (the sort
 * groups the samples by Binner key, input order inside a bin: the segments
 * and windows of the sort by travel; rtsim_sensor_binner_debug sorts by
 * travel).
 * MFMA forms (1D and 2D): the samples with a Binner key are bucketed by grid
 * row, deposited on the output grid (bilinear, exact fixed point) and
 * correlated with the taps on MFMA; rows whose Binner window is empty are 0.
 * They equal the gather in exact arithmetic and are refused
 * (RTSIM_ERR_UNSUPPORTED) unless the bearing coordinates are uniform in the
 * PSF axis (within 5e-5 of a pitch), the PSF pitch equals the grid pitch in
 * range and bearing (within 5e-5), and the Binner overlap covers the PSF,
 * >= (range taps + 3) / 2 bins.                                                   */
int rtsim_sensor_compute(rtsim_sensor h, int form, void* stream);

/* No counterpart in the reference: rtsim_sensor_compute for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Computes the ping of `n`
 * caller-owned device records (16-B records as above, unsorted) into the
 * caller-owned device ping `ping` ([range][bearing] complex64, fully
 * written), in `form`, on `stream`, with the sensor's configuration (ranges,
 * bearings, PSF, pose-independent: the records are already weighted).  Calls
 * no hipMalloc / hipFree or allocating library routine: every per-ping buffer
 * is a fixed device array of RTSIM_NOALLOC_SCRATCH_BYTES that each GPU holds
 * from its first call until the process ends.  Limits: n <=
 * RTSIM_NOALLOC_MAX_SAMPLES and range_count <= RTSIM_NOALLOC_MAX_BINS
 * (RTSIM_ERR_INVALID beyond).  The MFMA forms use the sensor's prepared grid
 * form (an earlier MFMA compute or rtsim_sensor_set_tile_shape prepares it;
 * RTSIM_ERR_INVALID otherwise).  All calls on one GPU share its array and no
 * call reads what an earlier one left there, but calls that overlap on one
 * GPU (other streams, other threads) race; an MFMA-form call also rewrites
 * the sensor's grid, so it must not overlap that sensor's own computes.  The
 * sensor's own samples, sorted records, Binner arrays and ping are untouched. */
#define RTSIM_NOALLOC_MAX_SAMPLES 1048576u
#define RTSIM_NOALLOC_MAX_BINS 16384u
#define RTSIM_NOALLOC_SCRATCH_BYTES 44347392u   /* 42.3 MiB per GPU */
int rtsim_sensor_compute_noalloc(rtsim_sensor h, int form, const void* records, size_t n,
                                 void* ping, void* stream);

/* Tile-shape autotuner of the MFMA forms (no reference counterpart: the
 * reference has one fixed launch shape, SensorInstance2D.cu:93).  Runs one
 * compute, times every correlation tile shape that fits over `iterations`
 * launches with HIP events on `stream`, keeps the fastest for this sensor and
 * caches it for every sensor of the same (ranges, bearings, taps) class.
 * Synchronises the stream; shape / ms may be NULL.                             */
int rtsim_sensor_autotune(rtsim_sensor h, int form, int iterations, void* stream, int* shape,
                          float* ms);
/* Sets / reads the MFMA correlation tile shape (0: 32x32, 1: 32x64, 2: 64x32,
 * 3: 64x64 range rows x bearings).  A shape whose LDS does not fit is
 * RTSIM_ERR_UNSUPPORTED.  No reference counterpart (see autotune).            */
int rtsim_sensor_set_tile_shape(rtsim_sensor h, int shape);
int rtsim_sensor_tile_shape(rtsim_sensor h, int* shape, int* rows, int* cols);
/* Sets / reads the waves per MFMA correlation workgroup, 4 or 8 (8 share each
 * output tile's tap steps among twice as many waves).  Setting a tile shape
 * resets it to the shape's default: 8 when its LDS leaves room for one
 * workgroup per CU only, else 4.  A count whose LDS does not fit is
 * RTSIM_ERR_UNSUPPORTED.  No reference counterpart (see autotune).            */
int rtsim_sensor_set_tile_waves(rtsim_sensor h, int waves);
int rtsim_sensor_tile_waves(rtsim_sensor h, int* waves);

/* get_ping()/output() (SensorInstance2D.h:144-150, SensorInstance1D.h:103)         */
int rtsim_sensor_output(rtsim_sensor h, const void** dev_ptr, uint32_t* height,
                        uint32_t* width);

/* Sorted samples (receivedSamples_ after sort(), Receiver.cu:28) as the last
 * sort left them: rtsim_sensor_binner_debug (by travel, the reference's order)
 * or a plain-form compute (grouped by Binner key, input order inside a bin,
 * out-of-range samples last; the reference's order whenever the samples of a
 * bin arrive in travel order).  The MFMA forms do not sort the samples.          */
int rtsim_sensor_sorted_samples(rtsim_sensor h, const void** dev_ptr);

/* Binner internals (Binner.h:56-77) copied to host for bit-exact checks.  Any
 * pointer may be NULL.  keys: sorted keys (n); segments: (start, end) per bin
 * (2N); windows: (start, size) per bin (2N); order: sort permutation (n);
 * overlap: binOverlap.  Synchronises the stream.  A plain-form compute keeps
 * neither the keys nor the permutation (nothing downstream reads them), and
 * an MFMA-form compute does not sort at all (it keeps only which bins hold
 * samples): asking for the keys or the permutation, or for the segments or
 * windows after an MFMA-form compute, sorts the resident samples by travel on
 * the stream with all of them kept (the sort is deterministic, so every array
 * it rewrites keeps its contents).  The samples must be those of the last
 * compute (RTSIM_ERR_INVALID after a change).                                  */
int rtsim_sensor_binner_debug(rtsim_sensor h, uint32_t* keys, uint32_t* segments,
                              uint32_t* windows, uint32_t* order, int* overlap,
                              void* stream);

/* Host <-> device helpers (synchronous on the given stream).                      */
int rtsim_sensor_upload_samples(rtsim_sensor h, const void* host, size_t n, void* stream);
int rtsim_sensor_download_samples(rtsim_sensor h, void* host, size_t n, int sorted,
                                  void* stream);
int rtsim_sensor_download_output(rtsim_sensor h, float* host, void* stream);

/* Per-stage timing with HIP events recorded on the compute stream.  When on,
 * each compute records events around its stages; stage_times returns the
 * mean milliseconds per stage over the computes since the last reset.
 * Stages: 0 stream, 1 binning (plain form: the sort by Binner key, whose last
 * pass also gathers the records and bearing coordinates and bounds the
 * segments; MFMA forms: the bucket counts' scan and the scatter), 2 windows
 * (plain form; 2D: with the strip lists), 3 grid deposit with the windows'
 * emptiness (MFMA forms), 4 convolution (plain: band ordering + gather; MFMA:
 * correlation).                                                               */
int rtsim_sensor_set_profiling(rtsim_sensor h, int on);
int rtsim_sensor_stage_times(rtsim_sensor h, float* ms, int nstages, int* ncomputes);

/* No reference counterpart: this GPU's own peaks for roofline fractions (a
 * float4 copy between two 512 MiB buffers, reads + writes per second; four
 * independent chains of v_mfma_f32_32x32x16_{f16,bf16} per wave, two waves per
 * SIMD), best of three launches each, on `stream` (synchronised).  Allocates
 * and frees 1 GiB.  Any output pointer may be NULL.                            */
int rtsim_measure_peaks(void* stream, double* hbm_gbs, double* mfma_fp16_tflops,
                        double* mfma_bf16_tflops);

/* Streams and event timers for callers without their own HIP binding
 * (hipStreamCreate at the default priority, hipEvent pairs).                     */
int rtsim_stream_create(void** stream);
int rtsim_stream_destroy(void* stream);
int rtsim_timer_create(void** timer);
int rtsim_timer_destroy(void* timer);
int rtsim_timer_start(void* timer, void* stream);
int rtsim_timer_stop(void* timer, void* stream);
/* Waits for the stop event and returns start -> stop in milliseconds.            */
int rtsim_timer_elapsed(void* timer, float* ms);

/* Device memory helpers so callers without a GPU runtime binding can stage
 * inputs (hipMalloc / hipMemcpy / hipFree).                                       */
int rtsim_device_alloc(void** ptr, size_t bytes);
int rtsim_device_free(void* ptr);
int rtsim_memcpy_h2d(void* dst, const void* src, size_t bytes);
int rtsim_memcpy_d2h(void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* RTSIM_H */
