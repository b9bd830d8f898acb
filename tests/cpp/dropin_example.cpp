// Uses the C++ drop-in (include/rtac_simulation_amd/rtac_simulation.h) the way
// a reference user would: SensorFactory::Make(yaml) -> set_sample_count ->
// set_samples (receiver view) -> compute_output -> output.
//
//   dropin_example info <config.yaml>
//       prints the parsed description (host only, no GPU)
//   dropin_example ping <config.yaml> <samples> <seed> <form> <out.bin> [sink.rtpg]
//       computes one ping from seeded synthetic hits and writes it as
//       complex64 [range][bearing] (and through FileSink)
//   dropin_example poses <poses.csv>
//       parses a pose CSV (host only) and prints each rotation
//   dropin_example drive <config.yaml> <poses.csv> <samples> <seed> <form> <in_flight>
//                        <out.rtpg> [emitter]
//       runs the simulation loop over every pose, in batches of in_flight
//       pings, into one ping file; with `emitter`, the rays are those of an
//       Emitter (0.5 deg polar fan, 130 x 20 deg, 20 mm disk directivity)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rtac_simulation_amd/rtac_simulation.h"

using namespace rtac::simulation;

static rtsim_scene_desc scene_for(const SensorInfo& info)
{
    rtsim_scene_desc s;
    std::memset(&s, 0, sizeof(s));
    s.rayleigh_sigma = 1.0f;
    const double deg = M_PI / 180.0;
    if (!info.is_2d()) {
        s.kind = RTSIM_SCENE_CONE;
        s.range_min = 0.5f;
        s.range_max = info.desc().range_max;
        s.off_axis_max = (float)(10.0 * deg);
    } else {
        // the seabed of the 512 x 1024 workload: 5 m below, rising 20 deg ahead
        const double t = 20.0 * deg;
        s.kind = RTSIM_SCENE_SEABED;
        s.bearing_max = (float)(65.0 * deg);
        s.elevation_max = (float)(10.0 * deg);
        s.plane_count = 1;
        s.planes[0][0] = (float)(-std::sin(t));
        s.planes[0][1] = 0.0f;
        s.planes[0][2] = (float)std::cos(t);
        s.planes[0][3] = 5.0f;
        s.height_noise = 0.05f;
        s.incidence_power = 2;
    }
    return s;
}

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s info|ping <config.yaml> ...\n", argv[0]);
        return 2;
    }
    std::string mode = argv[1];
    try {
        if (mode == "info") {
            SensorInfo::Ptr info = SensorInfoFactory::Make(argv[2]);
            const rtsim_sensor_desc& d = info->desc();
            std::printf("kind=%d ranges=%u [%g, %g] bearings=%u taps=%ux%u span_r=%g span_b=%g "
                        "margin=%g dir=%ux%u f=%g\n",
                        d.kind, d.range_count, d.range_min, d.range_max, d.bearing_count,
                        d.psf_range_taps, d.psf_bearing_taps, d.psf_range_span,
                        d.psf_bearing_span, d.margin, d.dir_width, d.dir_height, d.frequency);
            return 0;
        }
        if (mode == "poses") {
            PoseSourceList::Ptr poses = PoseSourceList::CreateFromCSV(argv[2]);
            std::printf("poses=%d\n", poses->size());
            while (poses->remaining() > 0) {
                const Pose p = poses->next_pose();
                const std::array<float, 9> R = p.rotation_matrix();
                std::printf("%.6f %.6f %.6f", p.x, p.y, p.z);
                for (float v : R) std::printf(" %.7f", v);
                std::printf("\n");
            }
            return 0;
        }
        if (mode == "drive" && argc >= 9) {
            SensorInfo::Ptr info = SensorInfoFactory::Make(argv[2]);
            const rtsim_scene_desc scene = scene_for(*info);
            SyntheticSimulation sim(argv[2], scene, std::strtoull(argv[4], nullptr, 10),
                                    static_cast<Form>(std::atoi(argv[6])),
                                    (unsigned)std::atoi(argv[7]), std::strtoull(argv[5], nullptr, 10));
            sim.set_receiver_poses(PoseSourceList::CreateFromCSV(argv[3]));
            if (argc >= 10 && std::string(argv[9]) == "emitter") {
                const float f = info->desc().frequency, lambda = 1500.0f / f;
                sim.set_emitter(Emitter::Create(0.5f, 130.0f, 20.0f,
                                                Directivity::disk_antenna(0.02f, lambda), f));
            }
            FileSink::Ptr sink = FileSink::Create(argv[8], true);
            sim.add_sink(sink);
            size_t rounds = 0;
            while (sim.run_batch() > 0) rounds++;
            std::printf("pings=%llu rounds=%zu file=%llu\n", (unsigned long long)sim.pings(), rounds,
                        (unsigned long long)sink->count());
            return 0;
        }
        if (mode == "ping" && argc >= 7) {
            const size_t n = std::strtoull(argv[3], nullptr, 10);
            const uint64_t seed = std::strtoull(argv[4], nullptr, 10);
            const int form = std::atoi(argv[5]);
            SensorInstance::Ptr sensor = SensorFactory::Make(argv[2]);
            rtsim_scene_desc scene = scene_for(sensor->info());
            std::vector<float> r(n), dx(n), dy(n), dz(n), a(2 * n);
            check(rtsim_generate_hits(&scene, seed, 0, n, r.data(), dx.data(), dy.data(),
                                      dz.data(), a.data(), 8));
            void* dev[5];
            const std::vector<float>* host[5] = {&r, &dx, &dy, &dz, &a};
            for (int k = 0; k < 5; k++) {
                size_t bytes = sizeof(float) * host[k]->size();
                check(rtsim_device_alloc(&dev[k], bytes));
                check(rtsim_memcpy_h2d(dev[k], host[k]->data(), bytes));
            }
            rtsim_hits hits{(const float*)dev[0], (const float*)dev[1], (const float*)dev[2],
                            (const float*)dev[3], (const float*)dev[4]};
            sensor->set_sample_count((unsigned)n);
            sensor->set_samples(hits, n);
            sensor->compute_output(static_cast<Form>(form));
            std::vector<Complex<float>> ping = sensor->output_host();
            BinnerView b = sensor->binner();
            FILE* f = std::fopen(argv[6], "wb");
            if (!f) throw FileError(argv[6]);
            std::fwrite(ping.data(), sizeof(Complex<float>), ping.size(), f);
            std::fclose(f);
            if (argc >= 8) {   // FileSink::Create + set_output (FileSink.cpp:15-42)
                FileSink::Ptr sink = FileSink::Create(argv[7], true);
                sink->set_output(sensor);
                std::printf("sink %s pings=%llu\n", sink->filename().c_str(),
                            (unsigned long long)sink->count());
            }
            double e = 0;
            for (auto& v : ping) e += std::norm(v);
            std::printf("ping %zu cells energy=%.6e overlap=%d\n", ping.size(), e, b.overlap);
            for (int k = 0; k < 5; k++) check(rtsim_device_free(dev[k]));
            return 0;
        }
    } catch (const ConfigError& e) {
        std::fprintf(stderr, "ConfigError: %s\n", e.what());
        return 3;
    } catch (const FileError& e) {
        std::fprintf(stderr, "FileError: %s\n", e.what());
        return 4;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "bad arguments\n");
    return 2;
}
