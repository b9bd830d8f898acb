/*
 * rtsim_scene.h -- sample producers and sensor-description helpers of the
 * rtsim C ABI (host-side, no device work).
 *
 *   rtsim_generate_hits      seeded synthetic ray hits.  It stands in for the
 *                            OptiX / OpenGL ray casters that feed the path in
 *                            the reference (RayCaster.cu:17-81 raygen,
 *                            SimulationGL.cu:82-132), which are out of scope.
 *   rtsim_directivity_*      the table builders of Directivity
 *                            (Directivity.cpp:59-166), so a caller can build
 *                            the same receiver directivity tables.
 *   rtsim_emitter_*,         the Emitter's ray fan and initial values
 *   rtsim_generate_emitter_hits  (Emitter.cu:38-80) and hits of its rays.
 */
#ifndef RTSIM_SCENE_H
#define RTSIM_SCENE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTSIM_SCENE_CONE 1   /* single beam: uniform range, uniform off-axis cone */
#define RTSIM_SCENE_SEABED 2 /* bearing x elevation fan hitting planes / spheres  */

typedef struct {
    int kind;
    /* CONE */
    float range_min, range_max;   /* uniform one-way range (m)                     */
    float off_axis_max;           /* off-axis angle uniform in +-this (rad)         */
    /* SEABED: dir = (cos e cos b, -cos e sin b, sin e), b, e uniform             */
    float bearing_max, elevation_max;   /* half apertures (rad)                     */
    uint32_t plane_count;         /* <= 4                                           */
    float planes[4][4];           /* (nx, ny, nz, h): points p with n.p = -h; n is
                                     the unit normal toward the sensor             */
    float height_noise;           /* N(0, sigma^2) added to h per ray (m)           */
    uint32_t sphere_count;        /* random spheres (range/bearing clusters)        */
    float sphere_radius_min, sphere_radius_max;
    float sphere_range_min, sphere_range_max;
    float sphere_fraction;        /* fraction of rays drawn on the spheres          */
    float dynamic_range_db;       /* sphere gains 10^(-U * dB / 20)                 */
    /* amplitude = Rayleigh(sigma) e^{i U[0, 2pi)} * cos(incidence)^p (* gain)     */
    float rayleigh_sigma;
    int incidence_power;
} rtsim_scene_desc;

/* Writes hits [first, first + n) of scene `seed` into host SoA arrays
 * (amplitude is complex64 interleaved).  Misses have range 0 (the streaming
 * kernel turns r < 1e-4 into a null sample).  Deterministic for any split.   */
int rtsim_generate_hits(const rtsim_scene_desc* scene, uint64_t seed, uint64_t first, size_t n,
                        float* range, float* dir_x, float* dir_y, float* dir_z,
                        float* amplitude, int nthreads);

/* ---- emitter (Emitter.cu, Emitter.h) --------------------------------------------
 * Emitter::generate_polar_directions (Emitter.cu:59-80), restated in FP32:
 * angles in degrees (the reference's unit), H = (int)(elevation / resolution)
 * + 1 rows at phi_h = elevation (h / (H - 1) - 1/2), row h holding
 * W_h = (int)(|cos phi_h| bearing / resolution) + 1 rays at
 * theta_w = bearing (w / (W_h - 1) - 1/2), direction
 * (cos theta cos phi, sin theta cos phi, sin phi).  As in the reference, a
 * row of one ray divides 0 by 0 and that ray's direction is NaN (a tracer
 * reports it as a miss), and a fan of one row (elevation < resolution) has no
 * rays (its row width is (int) of NaN, INT_MIN on x86).  dirs holds 3 floats
 * per ray; with dirs == NULL only *count is set.                                */
int rtsim_emitter_polar_directions(float resolution_deg, float bearing_aperture_deg,
                                   float elevation_aperture_deg, float* dirs, size_t capacity,
                                   size_t* count);
/* Emitter::load_initial_values (Emitter.cu:38-55): value_i = D(dir_i) with the
 * DirectivityView lookup (Directivity.h:30-34): u = xscale atan2(y, x),
 * v = yscale atan2(z, x), linear filter, mirror addressing, normalised
 * coordinates (Directivity.cpp:109-113).  values is complex64 interleaved
 * (imaginary part 0); table == NULL gives 1 for every ray.                      */
int rtsim_emitter_initial_values(const float* dirs, size_t n, const float* table, uint32_t width,
                                 uint32_t height, float xscale, float yscale, float* values);
/* Hits of an emitter's rays on a SEABED scene (the OptiX trace of
 * RayCaster.cu:56-58 replaced by ray / plane and ray / sphere intersection):
 * ray i (i in [first, first + n)) leaves along rotation * dirs[i mod ndirs]
 * (EmitterView::ray_direction, Emitter.h:58-60; rotation row-major 3x3, NULL
 * for identity) with value values[i mod ndirs] (sample_value, Emitter.h:61-63)
 * and carries value * reflection (Rayleigh(sigma) e^{i U[0, 2pi)} *
 * cos(incidence)^p, times the sphere gain on a sphere).  Misses and NaN
 * directions have range 0.  Deterministic for any split of [first, first + n). */
int rtsim_generate_emitter_hits(const rtsim_scene_desc* scene, uint64_t seed, const float* dirs,
                                const float* values, size_t ndirs, const float* rotation,
                                uint64_t first, size_t n, float* range, float* dir_x,
                                float* dir_y, float* dir_z, float* amplitude, int nthreads);

/* Directivity::disk_antenna (Directivity.cpp:109-166).  With out == NULL only
 * the table size is returned.                                                  */
int rtsim_directivity_disk(float diameter, float wavelength, int single_sided,
                           uint32_t oversampling, float* out, size_t capacity,
                           uint32_t* width, uint32_t* height);
/* Directivity::rectangle_antenna (Directivity.cpp:59-101).                      */
int rtsim_directivity_rectangle(float width_m, float height_m, float wavelength,
                                int single_sided, uint32_t oversampling, float* out,
                                size_t capacity, uint32_t* width, uint32_t* height);
/* Directivity::make_uniform (Directivity.cpp:103-107).                          */
int rtsim_directivity_uniform(uint32_t oversampling, float* out, size_t capacity,
                              uint32_t* width, uint32_t* height);

#ifdef __cplusplus
}
#endif

#endif /* RTSIM_SCENE_H */
