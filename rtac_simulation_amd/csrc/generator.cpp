// Seeded synthetic ray-hit generator (replaces the OptiX/GL ray casters,
// RayCaster.cu:17-81 / SimulationGL.cu:82-132, which are out of scope), the
// Emitter's ray fan and initial values (Emitter.cu:38-80) with hits of its
// rays, and the host-side directivity table builders of Directivity
// (Directivity.cpp:35-166).
//
// The generator is counter based: sample i of a scene depends only on
// (seed, i), so any sub-range [first, first + n) can be produced on its own,
// in parallel, and always gives the same values.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rtsim.h"
#include "../../include/rtsim_scene.h"
#include "sensor.h"

namespace rtsim {
namespace {

struct Rng {
    uint64_t state;
    uint64_t next()
    {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
    double uniform(double a, double b) { return a + (b - a) * unit(); }
    double normal()
    {
        double u1 = 1.0 - unit();  // (0, 1]
        double u2 = unit();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

Rng stream_for(uint64_t seed, uint64_t index, uint64_t salt)
{
    Rng r{seed * 0xD1B54A32D192ED03ull + salt};
    r.state ^= index * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    r.next();
    return r;
}

struct Sphere {
    double c[3];
    double radius;
    double gain;
};

std::vector<Sphere> make_spheres(const rtsim_scene_desc* sc, uint64_t seed)
{
    std::vector<Sphere> out(sc->sphere_count);
    for (uint32_t k = 0; k < sc->sphere_count; k++) {
        Rng g = stream_for(seed, k, 0x5eed5be7e5ull);
        double b = g.uniform(-0.9 * sc->bearing_max, 0.9 * sc->bearing_max);
        double e = g.uniform(-0.9 * sc->elevation_max, 0.9 * sc->elevation_max);
        double rho = g.uniform(sc->sphere_range_min, sc->sphere_range_max);
        out[k].c[0] = rho * std::cos(e) * std::cos(b);
        out[k].c[1] = -rho * std::cos(e) * std::sin(b);
        out[k].c[2] = rho * std::sin(e);
        out[k].radius = g.uniform(sc->sphere_radius_min, sc->sphere_radius_max);
        out[k].gain = std::pow(10.0, -sc->dynamic_range_db / 20.0 * g.unit());
    }
    return out;
}

void gen_range(const rtsim_scene_desc* sc, uint64_t seed, uint64_t first, size_t begin,
               size_t end, const std::vector<Sphere>& spheres, float* range, float* dx,
               float* dy, float* dz, float* amp)
{
    for (size_t k = begin; k < end; k++) {
        Rng g = stream_for(seed, first + k, 0x9a7e5ull);
        double mag = sc->rayleigh_sigma * std::sqrt(-2.0 * std::log(1.0 - g.unit()));
        double ph = 2.0 * M_PI * g.unit();
        double d[3] = {1, 0, 0};
        double r = 0.0, gain = 0.0;
        if (sc->kind == RTSIM_SCENE_CONE) {
            r = g.uniform(sc->range_min, sc->range_max);
            double th = g.uniform(-sc->off_axis_max, sc->off_axis_max);
            double psi = g.uniform(0.0, 2.0 * M_PI);
            d[0] = std::cos(th);
            d[1] = std::sin(th) * std::cos(psi);
            d[2] = std::sin(th) * std::sin(psi);
            gain = 1.0;
        } else {
            double pick = g.unit();
            if (!spheres.empty() && pick < sc->sphere_fraction) {
                const Sphere& S = spheres[std::min<size_t>(spheres.size() - 1,
                                                           (size_t)(g.unit() * spheres.size()))];
                // random point on the hemisphere facing the sensor
                double w[3];
                double z = g.uniform(-1.0, 1.0), a = g.uniform(0.0, 2.0 * M_PI);
                double q = std::sqrt(std::max(0.0, 1.0 - z * z));
                w[0] = q * std::cos(a); w[1] = q * std::sin(a); w[2] = z;
                double cn = std::sqrt(S.c[0] * S.c[0] + S.c[1] * S.c[1] + S.c[2] * S.c[2]);
                double facing = -(w[0] * S.c[0] + w[1] * S.c[1] + w[2] * S.c[2]) / cn;
                if (facing < 0) { w[0] = -w[0]; w[1] = -w[1]; w[2] = -w[2]; }
                double p[3] = {S.c[0] + S.radius * w[0], S.c[1] + S.radius * w[1],
                               S.c[2] + S.radius * w[2]};
                r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
                d[0] = p[0] / r; d[1] = p[1] / r; d[2] = p[2] / r;
                double ci = -(w[0] * d[0] + w[1] * d[1] + w[2] * d[2]);
                ci = std::max(0.0, ci);
                gain = S.gain * std::pow(ci, sc->incidence_power);
            } else {
                double b = g.uniform(-sc->bearing_max, sc->bearing_max);
                double e = g.uniform(-sc->elevation_max, sc->elevation_max);
                d[0] = std::cos(e) * std::cos(b);
                d[1] = -std::cos(e) * std::sin(b);
                d[2] = std::sin(e);
                double eta = sc->height_noise > 0 ? sc->height_noise * g.normal() : 0.0;
                double best = 0.0, bestcos = 0.0;
                for (uint32_t pl = 0; pl < sc->plane_count && pl < 4; pl++) {
                    const float* P = sc->planes[pl];
                    double nd = P[0] * d[0] + P[1] * d[1] + P[2] * d[2];
                    if (nd >= -1e-9) continue;
                    double t = -((double)P[3] + eta) / nd;
                    if (t > 0 && (best == 0.0 || t < best)) { best = t; bestcos = -nd; }
                }
                r = best;
                gain = best > 0 ? std::pow(bestcos, sc->incidence_power) : 0.0;
            }
        }
        range[k] = (float)r;
        dx[k] = (float)d[0];
        dy[k] = (float)d[1];
        dz[k] = (float)d[2];
        amp[2 * k] = (float)(gain * mag * std::cos(ph));
        amp[2 * k + 1] = (float)(gain * mag * std::sin(ph));
    }
}

// Hits of emitter rays: ray i along d = rot * dirs[i mod ndirs] against the
// scene's planes (with the per-ray height noise) and spheres, closest hit.
void gen_emitter_range(const rtsim_scene_desc* sc, uint64_t seed, const float* dirs,
                       const float* values, size_t ndirs, const double* rot, uint64_t first,
                       size_t begin, size_t end, const std::vector<Sphere>& spheres, float* range,
                       float* dx, float* dy, float* dz, float* amp)
{
    for (size_t k = begin; k < end; k++) {
        const uint64_t ray = first + k;
        const size_t q = (size_t)(ray % ndirs);
        Rng g = stream_for(seed, ray, 0xe3177e5ull);
        const double mag = sc->rayleigh_sigma * std::sqrt(-2.0 * std::log(1.0 - g.unit()));
        const double ph = 2.0 * M_PI * g.unit();
        const double eta = sc->height_noise > 0 ? sc->height_noise * g.normal() : 0.0;
        const double e[3] = {dirs[3 * q], dirs[3 * q + 1], dirs[3 * q + 2]};
        double d[3];
        for (int r = 0; r < 3; r++) d[r] = rot[3 * r] * e[0] + rot[3 * r + 1] * e[1] + rot[3 * r + 2] * e[2];
        double best = 0.0, gain = 0.0;
        if (d[0] == d[0] && d[1] == d[1] && d[2] == d[2]) {
            for (uint32_t pl = 0; pl < sc->plane_count && pl < 4; pl++) {
                const float* P = sc->planes[pl];
                const double nd = P[0] * d[0] + P[1] * d[1] + P[2] * d[2];
                if (nd >= -1e-9) continue;
                const double t = -((double)P[3] + eta) / nd;
                if (t > 0 && (best == 0.0 || t < best)) {
                    best = t;
                    gain = std::pow(-nd, sc->incidence_power);
                }
            }
            for (const Sphere& S : spheres) {
                // |t d - c|^2 = R^2, nearest positive root
                const double dc = d[0] * S.c[0] + d[1] * S.c[1] + d[2] * S.c[2];
                const double cc = S.c[0] * S.c[0] + S.c[1] * S.c[1] + S.c[2] * S.c[2];
                const double disc = dc * dc - (cc - S.radius * S.radius);
                if (disc < 0) continue;
                const double t = dc - std::sqrt(disc);
                if (!(t > 0) || (best != 0.0 && t >= best)) continue;
                double nrm[3];
                for (int r = 0; r < 3; r++) nrm[r] = (t * d[r] - S.c[r]) / S.radius;
                const double ci = std::max(0.0, -(nrm[0] * d[0] + nrm[1] * d[1] + nrm[2] * d[2]));
                best = t;
                gain = S.gain * std::pow(ci, sc->incidence_power);
            }
        }
        const double vr = values ? values[2 * q] : 1.0, vi = values ? values[2 * q + 1] : 0.0;
        const double ar = gain * mag * std::cos(ph), ai = gain * mag * std::sin(ph);
        range[k] = (float)best;
        dx[k] = (float)d[0];
        dy[k] = (float)d[1];
        dz[k] = (float)d[2];
        amp[2 * k] = best > 0 ? (float)(vr * ar - vi * ai) : 0.0f;
        amp[2 * k + 1] = best > 0 ? (float)(vr * ai + vi * ar) : 0.0f;
    }
}

// Host restatement of the directivity lookup of the streaming kernel
// (DirectivityView, Directivity.h:30-34; tex2D linear filter, mirror
// addressing, normalised coordinates), in the same FP32 operations.
float mirror_coord_host(float x)
{
    const float f = std::floor(x), fr = x - f;
    return ((int)f & 1) ? 1.0f - fr : fr;
}

int mirror_index_host(int i, int n)
{
    if (i < 0) i = -1 - i;
    if (i >= n) i = 2 * n - 1 - i;
    return std::min(std::max(i, 0), n - 1);
}

float directivity_host(const float* t, int W, int H, float xs, float ys, float x, float y, float z)
{
    const float u = mirror_coord_host(xs * std::atan2(y, x));
    const float v = mirror_coord_host(ys * std::atan2(z, x));
    const float xb = u * (float)W - 0.5f, yb = v * (float)H - 0.5f;
    const float fi = std::floor(xb), fj = std::floor(yb);
    const float a = xb - fi, b = yb - fj;
    const int i0 = (int)fi, j0 = (int)fj;
    const int c0 = mirror_index_host(i0, W), c1 = mirror_index_host(i0 + 1, W);
    const int r0 = mirror_index_host(j0, H), r1 = mirror_index_host(j0 + 1, H);
    const float t00 = t[r0 * W + c0], t10 = t[r0 * W + c1];
    const float t01 = t[r1 * W + c0], t11 = t[r1 * W + c1];
    const float top = t00 + a * (t10 - t00);
    const float bot = t01 + a * (t11 - t01);
    return top + b * (bot - top);
}

}  // namespace
}  // namespace rtsim

using namespace rtsim;

extern "C" {

int rtsim_generate_hits(const rtsim_scene_desc* scene, uint64_t seed, uint64_t first, size_t n,
                        float* range, float* dir_x, float* dir_y, float* dir_z,
                        float* amplitude, int nthreads)
{
    if (!scene) return fail(RTSIM_ERR_INVALID, "scene is NULL");
    if (scene->kind != RTSIM_SCENE_CONE && scene->kind != RTSIM_SCENE_SEABED)
        return fail(RTSIM_ERR_INVALID, "unknown scene kind %d", scene->kind);
    if (n && (!range || !dir_x || !dir_y || !dir_z || !amplitude))
        return fail(RTSIM_ERR_INVALID, "NULL output array");
    std::vector<Sphere> spheres;
    if (scene->kind == RTSIM_SCENE_SEABED) spheres = make_spheres(scene, seed);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    size_t chunk = (n + nthreads - 1) / nthreads;
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++) {
        size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back(gen_range, scene, seed, first, b, e, std::cref(spheres), range, dir_x,
                          dir_y, dir_z, amplitude);
    }
    for (auto& th : pool) th.join();
    return RTSIM_OK;
}

// Emitter::generate_polar_directions (Emitter.cu:59-80) in the reference's
// FP32 arithmetic (degrees converted as `x *= M_PI / 180.0f`).
int rtsim_emitter_polar_directions(float resolution_deg, float bearing_aperture_deg,
                                   float elevation_aperture_deg, float* dirs, size_t capacity,
                                   size_t* count)
{
    if (!count) return fail(RTSIM_ERR_INVALID, "count is NULL");
    if (!(resolution_deg > 0.f) || !(bearing_aperture_deg >= 0.f) || !(elevation_aperture_deg >= 0.f))
        return fail(RTSIM_ERR_INVALID, "emitter fan: resolution must be > 0, apertures >= 0");
    float resolution = resolution_deg, bearing = bearing_aperture_deg, elevation = elevation_aperture_deg;
    resolution *= M_PI / 180.0f;
    bearing *= M_PI / 180.0f;
    elevation *= M_PI / 180.0f;
    const float hrows = elevation / resolution;
    if (!(hrows < 1.0e6f)) return fail(RTSIM_ERR_INVALID, "emitter fan too large");
    const int H = (int)hrows + 1;
    size_t total = 0;
    for (int h = 0; h < H; h++) {
        const float phi = elevation * (((float)h) / (H - 1) - 0.5f);
        const float wr = std::fabs(std::cos(phi)) * bearing / resolution;
        // H == 1 makes phi NaN (0 / 0); the reference's (int) of NaN is INT_MIN
        // on x86, so its row loop runs zero times: no rays
        const int W = wr == wr ? (int)wr + 1 : 0;
        if ((double)total + W > 1.0e9) return fail(RTSIM_ERR_INVALID, "emitter fan too large");
        if (dirs) {
            for (int w = 0; w < W; w++) {
                if (total + w >= capacity)
                    return fail(RTSIM_ERR_INVALID, "capacity below the fan's ray count");
                const float theta = bearing * (((float)w) / (W - 1) - 0.5f);
                float* d = dirs + 3 * (total + w);
                d[0] = std::cos(theta) * std::cos(phi);
                d[1] = std::sin(theta) * std::cos(phi);
                d[2] = std::sin(phi);
            }
        }
        total += (size_t)W;
    }
    *count = total;
    return RTSIM_OK;
}

int rtsim_emitter_initial_values(const float* dirs, size_t n, const float* table, uint32_t width,
                                 uint32_t height, float xscale, float yscale, float* values)
{
    if (n && (!dirs || !values)) return fail(RTSIM_ERR_INVALID, "NULL array");
    if (table && (width < 1 || height < 1)) return fail(RTSIM_ERR_INVALID, "directivity table size");
    for (size_t i = 0; i < n; i++) {
        const float* d = dirs + 3 * i;
        values[2 * i] = table ? directivity_host(table, (int)width, (int)height, xscale, yscale,
                                                 d[0], d[1], d[2])
                              : 1.0f;
        values[2 * i + 1] = 0.0f;
    }
    return RTSIM_OK;
}

int rtsim_generate_emitter_hits(const rtsim_scene_desc* scene, uint64_t seed, const float* dirs,
                                const float* values, size_t ndirs, const float* rotation,
                                uint64_t first, size_t n, float* range, float* dir_x,
                                float* dir_y, float* dir_z, float* amplitude, int nthreads)
{
    if (!scene) return fail(RTSIM_ERR_INVALID, "scene is NULL");
    if (scene->kind != RTSIM_SCENE_SEABED)
        return fail(RTSIM_ERR_INVALID, "emitter rays need a SEABED scene (planes / spheres)");
    if (!dirs || ndirs == 0) return fail(RTSIM_ERR_INVALID, "emitter has no rays");
    if (n && (!range || !dir_x || !dir_y || !dir_z || !amplitude))
        return fail(RTSIM_ERR_INVALID, "NULL output array");
    double rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (rotation)
        for (int k = 0; k < 9; k++) rot[k] = rotation[k];
    std::vector<Sphere> spheres = make_spheres(scene, seed);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    size_t chunk = (n + nthreads - 1) / nthreads;
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++) {
        size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back(gen_emitter_range, scene, seed, dirs, values, ndirs, rot, first, b, e,
                          std::cref(spheres), range, dir_x, dir_y, dir_z, amplitude);
    }
    for (auto& th : pool) th.join();
    return RTSIM_OK;
}

// Directivity::disk_antenna (Directivity.cpp:109-166): circular piston
// 2 J1(a)/a, a = (pi D / lambda) sin(off-axis), single-sided baffle zeroes
// the back half-space.  Table texel (w, h) is computed at (pi w/W, pi h/H).
int rtsim_directivity_disk(float diameter, float wavelength, int single_sided,
                           uint32_t oversampling, float* out, size_t capacity,
                           uint32_t* width, uint32_t* height)
{
    if (!(diameter > 0) || !(wavelength > 0)) return fail(RTSIM_ERR_INVALID, "disk dimensions");
    float period = 2.0f * wavelength / diameter;
    uint32_t N = std::max(64u, 8u * ((uint32_t)(oversampling * M_PI / period)));
    if (width) *width = N;
    if (height) *height = N;
    if (!out) return RTSIM_OK;
    if (capacity < (size_t)N * N) return fail(RTSIM_ERR_INVALID, "capacity below %u^2", N);
    for (uint32_t h = 0; h < N; h++) {
        double elevation = M_PI * h / N;
        double cose2 = std::cos(elevation); cose2 *= cose2;
        double sine2 = std::sin(elevation); sine2 *= sine2;
        for (uint32_t w = 0; w < N; w++) {
            double bearing = M_PI * w / N;
            double cosb2 = std::cos(bearing); cosb2 *= cosb2;
            double sinb2 = std::sin(bearing); sinb2 *= sinb2;
            double denom = 1 - sine2 * sinb2;
            double y2 = cose2 * sinb2 / denom;
            double z2 = sine2 * cosb2 / denom;
            double x = std::sqrt(1 - y2 - z2);
            float v;
            if (std::fabs(x) < 1.0e-2) {
                v = 0.0f;
            } else {
                if (elevation > 0.5 * M_PI || bearing > 0.5f * M_PI) x = -x;
                double a = (M_PI * diameter / wavelength) * std::sin(std::atan2(std::sqrt(y2 + z2), x));
                v = std::fabs(a) < 1.0e-3 ? 1.0f : (float)(2.0 * j1(a) / a);
            }
            out[(size_t)h * N + w] = v;
        }
    }
    if (single_sided)
        for (uint32_t h = 0; h < N; h++)
            for (uint32_t w = 0; w < N; w++)
                if (h > N / 2 || w > N / 2) out[(size_t)h * N + w] = 0.0f;
    return RTSIM_OK;
}

// Directivity::rectangle_antenna (Directivity.cpp:59-101): sinc x sinc.
int rtsim_directivity_rectangle(float width_m, float height_m, float wavelength,
                                int single_sided, uint32_t oversampling, float* out,
                                size_t capacity, uint32_t* width, uint32_t* height)
{
    if (!(width_m > 0) || !(height_m > 0) || !(wavelength > 0))
        return fail(RTSIM_ERR_INVALID, "rectangle dimensions");
    uint32_t Nw = 2 * (uint32_t)(0.5f * oversampling * 2 * M_PI * width_m / wavelength + 1);
    uint32_t Nh = 2 * (uint32_t)(0.5f * oversampling * 2 * M_PI * height_m / wavelength + 1);
    if (width) *width = Nw;
    if (height) *height = Nh;
    if (!out) return RTSIM_OK;
    if (capacity < (size_t)Nw * Nh) return fail(RTSIM_ERR_INVALID, "capacity below %ux%u", Nw, Nh);
    for (uint32_t h = 0; h < Nh; h++) {
        double elevation = M_PI * h / Nh;
        float eValue = std::sin(M_PI * height_m / wavelength * std::sin(elevation)) /
                       (M_PI * height_m / wavelength * std::sin(elevation));
        if (h == 0) eValue = 1;
        for (uint32_t w = 0; w < Nw; w++) {
            double bearing = M_PI * w / Nw;
            float bValue = std::sin(M_PI * width_m / wavelength * std::sin(bearing)) /
                           (M_PI * width_m / wavelength * std::sin(bearing));
            if (w == 0) bValue = 1;
            out[(size_t)h * Nw + w] = eValue * bValue;
        }
    }
    if (single_sided)
        for (uint32_t h = 0; h < Nh; h++)
            for (uint32_t w = 0; w < Nw; w++)
                if (h > Nh / 2 || w > Nw / 2) out[(size_t)h * Nw + w] = 0.0f;
    return RTSIM_OK;
}

// Directivity::make_uniform (Directivity.cpp:103-107): oversampling^2 ones
// (the reference ignores its amplitude argument).
int rtsim_directivity_uniform(uint32_t oversampling, float* out, size_t capacity,
                              uint32_t* width, uint32_t* height)
{
    if (oversampling < 1) return fail(RTSIM_ERR_INVALID, "oversampling must be >= 1");
    if (width) *width = oversampling;
    if (height) *height = oversampling;
    if (!out) return RTSIM_OK;
    if (capacity < (size_t)oversampling * oversampling) return fail(RTSIM_ERR_INVALID, "capacity");
    std::fill(out, out + (size_t)oversampling * oversampling, 1.0f);
    return RTSIM_OK;
}

}  // extern "C"
