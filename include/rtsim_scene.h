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
