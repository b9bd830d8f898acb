// FP64 CPU oracle for the sample-to-ping path (TEST INFRASTRUCTURE ONLY: the
// product never includes, links or calls this).
//
// Pinning (the independent check that fixes each function; tests/ names):
//   orc_key_make / orc_key_eval   golden FP32 constants for the 512x1024
//                                 [0.1, 30] m grid (a = 34.180668f,
//                                 b = -2.4190426f) and the invariants
//                                 key(r_i) = i, monotone keys, both clamp
//                                 branches, NaN/inf; test_oracle_binner.py
//   orc_sort_order                std::stable_sort ties; brute-force
//                                 insertion sort on small inputs
//   orc_segments / orc_windows    brute force over N <= 64 from a direct
//                                 reading of the window rule; partition and
//                                 count invariants
//   orc_overlap                   closed form trunc(margin / resolution)
//   orc_directivity               closed forms at texel centres, mirror
//                                 symmetry, scipy j1 piston pattern
//   orc_stream                    closed-form phase exp(i 2 pi frac(2rf/c)),
//                                 |v| = |A| D, bearing = atan2(-y, x)
//   orc_psf                       taps reproduced at tap centres, zero beyond
//                                 the support, separable outer product
//   orc_conv2d / orc_conv1d       single sample = shifted PSF copy,
//                                 linearity, numpy brute force O(S W H)
//   orc_grid2d                    numpy correlate2d of the deposited grid
//                                 and equality with orc_conv2d
//   orc_conv2d_rows               equality with orc_conv2d on the rows it
//                                 computes (it skips only bearings outside
//                                 the PSF support, whose terms are exact 0)
//
// Not pinnable here (documented deviations): NVIDIA's 8-bit texture filter
// weights, thrust's tie order, the reference 1D PSF scaling c/tau and the
// rtac_base Texture2D default address mode of the 1D PSF.

#ifndef RTSIM_ORACLE_H
#define RTSIM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OOR 0xFFFFFFFFu
#define ORC_AXIS_ANGLE 0
#define ORC_AXIS_SINE 1

typedef struct {
    uint32_t n;
    float a, b, lower, upper;
} orc_key;

typedef struct {
    float rotation[9];        // row-major R; local = R^T dir
    float frequency, sound_speed;
    int spreading;            // 0 none, 1 divide by r^2
    const float* directivity; // [H][W] or NULL (D = 1)
    uint32_t dir_width, dir_height;
    float dir_xscale, dir_yscale;
} orc_stream_params;

typedef struct {
    float range_lo, range_hi;
    uint32_t range_count;
    uint32_t bearing_count;   // 1 for 1D
    int axis;                 // ORC_AXIS_*
    int psf_range_taps, psf_bearing_taps;
    float psf_range_span, psf_bearing_span;
} orc_grid;

void orc_key_make(uint32_t N, float lo, float hi, float margin, orc_key* k);
uint32_t orc_key_eval(const orc_key* k, float t);
void orc_keys(const orc_key* k, const float* travel, size_t n, size_t stride, uint32_t* keys);
uint32_t orc_travel_image(float t);
void orc_sort_order(const float* travel, size_t n, size_t stride, uint32_t* order);
void orc_segments(const uint32_t* keys, size_t n, uint32_t N, uint32_t* seg);
int orc_overlap(float margin, float lo, float hi, uint32_t N);
void orc_windows(const uint32_t* seg, uint32_t N, int ov, uint32_t* win);
double orc_directivity(const float* table, uint32_t W, uint32_t H, float xscale,
                       float yscale, double lx, double ly, double lz);
void orc_stream(size_t n, const float* range, const float* dx, const float* dy,
                const float* dz, const float* amp, const orc_stream_params* p,
                double* out);
void orc_psf(const float* taps, int Tr, int Tb, double span_r, double span_b,
             double dr, double dxb, double* re, double* im);
void orc_conv2d(const float* sorted, const uint32_t* win, const orc_grid* g,
                const float* bearings, const float* taps, double* out, int nthreads);
void orc_conv1d(const float* sorted, const uint32_t* win, const orc_grid* g,
                const float* taps, double* out, int nthreads);
void orc_conv2d_rows(const float* sorted, const uint32_t* win, const orc_grid* g,
                     const float* bearings, const float* taps, const uint32_t* rows,
                     size_t nrows, double* out, int nthreads);
void orc_grid2d(const float* samples, size_t n, const orc_key* key,
                const uint32_t* win, const orc_grid* g, double bearing0,
                double bearing_pitch, const float* taps, double* out);
void cpu_binning(const float* samples, size_t n, const orc_grid* g, float margin,
                 float* sorted, uint32_t* win, int nthreads);
void cpu_gather1d_f32(const float* sorted, const uint32_t* win, const orc_grid* g,
                      const float* taps, float* out, int nthreads);
void cpu_gather2d_f32(const float* sorted, const uint32_t* win, const orc_grid* g,
                      const float* bearings, const float* taps, float* out,
                      long row_begin, long row_end, int nthreads);

#ifdef __cplusplus
}
#endif

#endif
