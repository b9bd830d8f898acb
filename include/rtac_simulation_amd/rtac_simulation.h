// C++ drop-in for the sample-to-ping path of rtac_simulation, over the rtsim
// C ABI (include/rtsim*.h).  Header only, host code only: any C++17 compiler
// can use it and link librtsim.so.
//
// Names and argument meaning follow the reference's sensor-instance API:
//   SimSample1D / SimSample2D            Sample.h:101-161 (16-B records here)
//   SensorInfo / SensorInfo2D            SensorInfo.h:18-105
//   SensorInstance, SensorInstance1D_Complex, SensorInstance2D_Complex,
//   SensorInstance2D_2<Complex<float>>   SensorInstance{,1D,2D}.h
//   Binner (read-only view)              Binner.h:15-141
//   Directivity builders                 Directivity.h:73-82
//   SensorFactory / SensorInfoFactory    factories/SensorInfoFactory.h
//   FileFinder                           factories/utilities.h:21-53
//   ConfigError / TypeError / FileError  reference exception types
// Errors are thrown as these exceptions (RTSIM_ERR_CONFIG -> ConfigError,
// RTSIM_ERR_IO -> FileError, others -> std::runtime_error).
#pragma once

#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <deque>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../rtsim.h"
#include "../rtsim_batch.h"
#include "../rtsim_config.h"
#include "../rtsim_scene.h"
#include "../rtsim_sink.h"

namespace rtac {
namespace simulation {

template <typename T>
using Complex = std::complex<T>;

struct ConfigError : std::runtime_error {
    explicit ConfigError(const std::string& m) : std::runtime_error("RTAC_CONFIG_ERROR : " + m) {}
};
struct FileError : std::runtime_error {
    explicit FileError(const std::string& m) : std::runtime_error("RTAC_FILE_ERROR : " + m) {}
};
struct TypeError : std::runtime_error {
    explicit TypeError(const std::string& m) : std::runtime_error("RTAC_TYPE_ERROR : " + m) {}
};

inline void check(int rc)
{
    if (rc == RTSIM_OK) return;
    std::string m = rtsim_last_error();
    if (rc == RTSIM_ERR_CONFIG) throw ConfigError(m);
    if (rc == RTSIM_ERR_IO) throw FileError(m);
    throw std::runtime_error("rtsim error " + std::to_string(rc) + " : " + m);
}

// Sample records as the library stores them (16 B each).
struct SimSample1D {
    float re, im, travel, pad;
    Complex<float> value() const { return {re, im}; }
};
struct SimSample2D {
    float re, im, travel, bearing;
    Complex<float> value() const { return {re, im}; }
};
static_assert(sizeof(SimSample2D) == 16, "SimSample2D is a float4");

// ---------------------------------------------------------------------------
// SensorInfo: a parsed / assembled sensor description (host only)
// ---------------------------------------------------------------------------
class SensorInfo {
   public:
    using Ptr = std::shared_ptr<SensorInfo>;
    using ConstPtr = std::shared_ptr<const SensorInfo>;

    virtual ~SensorInfo()
    {
        if (info_) rtsim_info_destroy(info_);
    }

    // explicit description (copies every table)
    static Ptr Create(const rtsim_sensor_desc& d) { return Ptr(new SensorInfo(d)); }
    // the description points into this object: not copyable
    SensorInfo(const SensorInfo&) = delete;
    SensorInfo& operator=(const SensorInfo&) = delete;

    const rtsim_sensor_desc& desc() const { return desc_; }
    bool is_2d() const { return desc_.kind == RTSIM_SENSOR_2D; }
    unsigned int width() const { return is_2d() ? desc_.bearing_count : 1u; }
    unsigned int height() const { return desc_.range_count; }

   protected:
    friend struct SensorInfoFactory;
    SensorInfo() = default;
    explicit SensorInfo(const rtsim_sensor_desc& d) : desc_(d)
    {
        psf_.assign(d.psf, d.psf + 2 * (size_t)d.psf_range_taps * d.psf_bearing_taps);
        desc_.psf = psf_.data();
        if (d.bearings) {
            bearings_.assign(d.bearings, d.bearings + d.bearing_count);
            desc_.bearings = bearings_.data();
        }
        if (d.directivity) {
            dir_.assign(d.directivity, d.directivity + (size_t)d.dir_width * d.dir_height);
            desc_.directivity = dir_.data();
        }
        if (d.waveform) {
            wave_ = *d.waveform;
            if (wave_.beam) {
                beam_.assign(wave_.beam, wave_.beam + wave_.beam_taps);
                wave_.beam = beam_.data();
            }
            desc_.waveform = &wave_;
        }
    }
    rtsim_info info_ = nullptr;
    rtsim_sensor_desc desc_{};
    rtsim_waveform_sine wave_{};
    std::vector<float> psf_, bearings_, dir_, beam_;
};
using SensorInfo2D = SensorInfo;

struct SensorInfoFactory {
    // SensorInfoFactory::Make(path) (SensorInfoFactory.cpp:76-80)
    static SensorInfo::Ptr Make(const std::string& path)
    {
        rtsim_info h = nullptr;
        check(rtsim_info_from_yaml(path.c_str(), &h));
        SensorInfo::Ptr p(new SensorInfo());
        p->info_ = h;
        const rtsim_sensor_desc* d = nullptr;
        check(rtsim_info_desc(h, &d));
        p->desc_ = *d;
        return p;
    }
};

struct FileFinder {
    static std::string FindOne(const std::string& name)
    {
        std::vector<char> buf(4096);
        if (rtsim_config_find(name.c_str(), buf.data(), buf.size()) != RTSIM_OK) return "";
        return std::string(buf.data());
    }
    static bool add_search_path(const std::string& p)
    {
        return rtsim_config_add_search_path(p.c_str()) == RTSIM_OK;
    }
};

// ---------------------------------------------------------------------------
// Directivity tables (Directivity.cpp:59-166)
// ---------------------------------------------------------------------------
struct DirectivityTable {
    unsigned int width = 0, height = 0;
    std::vector<float> data;   // [height][width]
};

struct Directivity {
    static DirectivityTable disk_antenna(float diameter, float wavelength,
                                         const std::string& baffleMode = "single-sided",
                                         unsigned int oversampling = 8)
    {
        DirectivityTable t;
        int single = baffleMode == "single-sided";
        check(rtsim_directivity_disk(diameter, wavelength, single, oversampling, nullptr, 0,
                                     &t.width, &t.height));
        t.data.resize((size_t)t.width * t.height);
        check(rtsim_directivity_disk(diameter, wavelength, single, oversampling, t.data.data(),
                                     t.data.size(), &t.width, &t.height));
        return t;
    }
    static DirectivityTable rectangle_antenna(float width, float height, float wavelength,
                                              const std::string& baffleMode = "single-sided",
                                              unsigned int oversampling = 8)
    {
        DirectivityTable t;
        int single = baffleMode == "single-sided";
        check(rtsim_directivity_rectangle(width, height, wavelength, single, oversampling, nullptr,
                                          0, &t.width, &t.height));
        t.data.resize((size_t)t.width * t.height);
        check(rtsim_directivity_rectangle(width, height, wavelength, single, oversampling,
                                          t.data.data(), t.data.size(), &t.width, &t.height));
        return t;
    }
    static DirectivityTable make_uniform(float /*amplitude*/ = 1.0f, unsigned int oversampling = 1)
    {
        DirectivityTable t;
        check(rtsim_directivity_uniform(oversampling, nullptr, 0, &t.width, &t.height));
        t.data.resize((size_t)t.width * t.height);
        check(rtsim_directivity_uniform(oversampling, t.data.data(), t.data.size(), &t.width,
                                        &t.height));
        return t;
    }
};

// ---------------------------------------------------------------------------
// Binner view (Binner.h:56-77 internals, read back for inspection)
// ---------------------------------------------------------------------------
struct BinnerView {
    std::vector<uint32_t> keys, order;
    std::vector<uint32_t> segments;   // (start, end) per bin
    std::vector<uint32_t> windows;    // (start, size) per bin
    int overlap = 0;
};

// ---------------------------------------------------------------------------
// Sensor instances
// ---------------------------------------------------------------------------
enum class Form { Plain = RTSIM_FORM_PLAIN_FP32, MfmaBF16 = RTSIM_FORM_MFMA_BF16,
                  MfmaFP16 = RTSIM_FORM_MFMA_FP16 };

class SensorInstance : public std::enable_shared_from_this<SensorInstance> {
   public:
    using Ptr = std::shared_ptr<SensorInstance>;
    using ConstPtr = std::shared_ptr<const SensorInstance>;

    virtual ~SensorInstance()
    {
        if (h_) rtsim_sensor_destroy(h_);
    }

    rtsim_sensor handle() const { return h_; }
    const SensorInfo& info() const { return *info_; }
    bool is_complex() const { return true; }

    // SensorInstance::set_ranges (SensorInstance.h:56-58); margin < 0 keeps c * tau
    void set_ranges(float rangeMin, float rangeMax, unsigned int count, float margin = -1.0f)
    {
        check(rtsim_sensor_set_ranges(h_, rangeMin, rangeMax, count, margin));
        range_count_ = count;
    }
    void set_sample_count(unsigned int count) { check(rtsim_sensor_set_sample_count(h_, count)); }
    unsigned int sample_count() const
    {
        size_t n = 0;
        check(rtsim_sensor_sample_count(h_, &n));
        return (unsigned int)n;
    }
    // receiver_view().samples: device array of 16-B records for producers
    void* samples_device() const
    {
        void* p = nullptr;
        check(rtsim_sensor_samples(h_, &p));
        return p;
    }
    // receiver pose (Receiver.h:56): row-major rotation, directions map as R^T dir
    void set_pose(const std::array<float, 9>& rotation) { check(rtsim_sensor_set_pose(h_, rotation.data())); }
    // ReceiverView::set_sample over n device-resident hits (Receiver.h:51-60)
    void set_samples(const rtsim_hits& hits, size_t n, size_t offset = 0, void* stream = nullptr)
    {
        check(rtsim_receiver_write_samples(h_, &hits, n, offset, stream));
    }
    // compute_output() (SensorInstance.h:67): sort -> Binner -> PSF convolution
    void compute_output(Form form = Form::Plain, void* stream = nullptr)
    {
        check(rtsim_sensor_compute(h_, (int)form, stream));
    }
    // MFMA forms: time every correlation tile shape, keep the fastest; returns
    // the shape index (no reference counterpart; synchronises the stream)
    int autotune(Form form, int iterations = 10, void* stream = nullptr)
    {
        int shape = -1;
        check(rtsim_sensor_autotune(h_, (int)form, iterations, stream, &shape, nullptr));
        return shape;
    }
    // ping as host data, row-major [range][bearing]
    std::vector<Complex<float>> output_host(void* stream = nullptr) const
    {
        uint32_t H = 0, W = 0;
        check(rtsim_sensor_output(h_, nullptr, &H, &W));
        std::vector<Complex<float>> out((size_t)H * W);
        check(rtsim_sensor_download_output(h_, reinterpret_cast<float*>(out.data()), stream));
        return out;
    }
    // device pointer to the ping (borrowed until the next compute)
    const Complex<float>* output_device() const
    {
        const void* p = nullptr;
        check(rtsim_sensor_output(h_, &p, nullptr, nullptr));
        return static_cast<const Complex<float>*>(p);
    }
    BinnerView binner(void* stream = nullptr) const
    {
        BinnerView v;
        size_t n = sample_count();
        uint32_t H = 0;
        check(rtsim_sensor_output(h_, nullptr, &H, nullptr));
        v.keys.resize(n);
        v.order.resize(n);
        v.segments.resize(2 * (size_t)H);
        v.windows.resize(2 * (size_t)H);
        check(rtsim_sensor_binner_debug(h_, v.keys.data(), v.segments.data(), v.windows.data(),
                                        v.order.data(), &v.overlap, stream));
        return v;
    }

    template <typename T>
    bool is_type() const { return dynamic_cast<const T*>(this) != nullptr; }
    template <typename T>
    std::shared_ptr<T> safe_cast()
    {
        auto p = std::dynamic_pointer_cast<T>(this->shared_from_this());
        if (!p) throw TypeError("wrong pointer conversion");
        return p;
    }

   protected:
    explicit SensorInstance(const SensorInfo::ConstPtr& info) : info_(info)
    {
        check(rtsim_sensor_create(&info->desc(), &h_));
        range_count_ = info->desc().range_count;
    }
    SensorInfo::ConstPtr info_;
    rtsim_sensor h_ = nullptr;
    unsigned int range_count_ = 0;
};

class SensorInstance1D_Complex : public SensorInstance {
   public:
    using Ptr = std::shared_ptr<SensorInstance1D_Complex>;
    static Ptr Create(const SensorInfo::ConstPtr& info)
    {
        if (info->is_2d()) throw std::runtime_error("SensorInstance1D needs a single-beam SensorInfo");
        return Ptr(new SensorInstance1D_Complex(info));
    }
    unsigned int size() const { return range_count_; }

   protected:
    explicit SensorInstance1D_Complex(const SensorInfo::ConstPtr& info) : SensorInstance(info) {}
};

class SensorInstance2D_Complex : public SensorInstance {
   public:
    using Ptr = std::shared_ptr<SensorInstance2D_Complex>;
    static Ptr Create(const SensorInfo::ConstPtr& info)
    {
        if (!info->is_2d())
            throw std::runtime_error("A SensorInstance2D must be created with a SensorInfo2D instance");
        return Ptr(new SensorInstance2D_Complex(info));
    }
    unsigned int width() const { return info_->width(); }
    unsigned int height() const { return range_count_; }
    unsigned int size() const { return width() * height(); }

   protected:
    explicit SensorInstance2D_Complex(const SensorInfo::ConstPtr& info) : SensorInstance(info) {}
};

// SensorInstance2D_2<Complex<float>> is the factory default (SensorInstance2D.h:96-151)
template <typename T>
using SensorInstance2D_2 = SensorInstance2D_Complex;

struct SensorFactory {
    // SensorFactory::Make(path) (SensorInfoFactory.cpp:7-57)
    static SensorInstance::Ptr Make(const std::string& path)
    {
        SensorInfo::ConstPtr info = SensorInfoFactory::Make(path);
        if (info->is_2d()) return SensorInstance2D_Complex::Create(info);
        return SensorInstance1D_Complex::Create(info);
    }
};

// ---------------------------------------------------------------------------
// Batched pings (one hipGraph launch for many sensor instances)
// ---------------------------------------------------------------------------
class Batch {
   public:
    Batch(const std::vector<SensorInstance::Ptr>& sensors, const std::vector<Form>& forms)
        : sensors_(sensors)
    {
        std::vector<rtsim_sensor> hs;
        std::vector<int> fs;
        for (size_t i = 0; i < sensors.size(); i++) {
            hs.push_back(sensors[i]->handle());
            fs.push_back((int)(i < forms.size() ? forms[i] : Form::Plain));
        }
        check(rtsim_batch_create(hs.data(), fs.data(), (uint32_t)hs.size(), &h_));
    }
    ~Batch()
    {
        if (h_) rtsim_batch_destroy(h_);
    }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
    void bind_hits(uint32_t index, const rtsim_hits* hits, size_t n)
    {
        check(rtsim_batch_bind_hits(h_, index, hits, n));
    }
    void compute(void* stream = nullptr) { check(rtsim_batch_compute(h_, stream)); }
    // tile shape and wave count per class of MFMA sensors and the launch
    // mode, tuned on the whole batch; returns its mean replay time in ms
    float autotune(int iterations = 5, void* stream = nullptr)
    {
        float ms = 0.f;
        check(rtsim_batch_autotune(h_, iterations, stream, &ms));
        return ms;
    }
    // MFMA-form pings in one persistent launch (rtsim_batch_set_persistent)
    void set_persistent(bool on) { check(rtsim_batch_set_persistent(h_, on ? 1 : 0)); }
    bool persistent() const
    {
        int on = 0;
        check(rtsim_batch_persistent(h_, &on));
        return on != 0;
    }
    // waits for the last launch; throws if its persistent launch was abandoned
    void check_launch() { check(rtsim_batch_check(h_)); }

   private:
    std::vector<SensorInstance::Ptr> sensors_;
    rtsim_batch h_ = nullptr;
};

// ---------------------------------------------------------------------------
// FileSink (sinks/FileSink.h): appends pings to a ping file (rtsim_sink.h)
// ---------------------------------------------------------------------------
class FileSink {
   public:
    using Ptr = std::shared_ptr<FileSink>;
    using ConstPtr = std::shared_ptr<const FileSink>;

    // FileSink::Create(filename, overwrite, sinkName) (FileSink.cpp:15-24)
    static Ptr Create(const std::string& filename, bool overwrite = false,
                      const std::string& sinkName = "file-sink")
    {
        rtsim_sink h = nullptr;
        check(rtsim_sink_open(filename.c_str(), overwrite ? 1 : 0, &h));
        return Ptr(new FileSink(h, filename, sinkName));
    }
    ~FileSink()
    {
        if (h_) rtsim_sink_close(h_);
    }
    FileSink(const FileSink&) = delete;
    FileSink& operator=(const FileSink&) = delete;

    const std::string& filename() const { return filename_; }
    const std::string& name() const { return name_; }
    // FileSink::set_output(sensor) (FileSink.cpp:26-42): 1D and 2D pings
    void set_output(const SensorInstance::Ptr& sensor, void* stream = nullptr)
    {
        check(rtsim_sink_write(h_, sensor->handle(), stream));
    }
    uint64_t count() const
    {
        uint64_t n = 0;
        check(rtsim_sink_count(h_, &n));
        return n;
    }

   private:
    FileSink(rtsim_sink h, const std::string& f, const std::string& n) : h_(h), filename_(f), name_(n) {}
    rtsim_sink h_ = nullptr;
    std::string filename_, name_;
};

// ---------------------------------------------------------------------------
// Poses and pose sources (PoseSource.h:11-87, PoseSource.cpp:7-30)
// ---------------------------------------------------------------------------
struct Pose {
    float x = 0.f, y = 0.f, z = 0.f;              // translation (m)
    float qw = 1.f, qx = 0.f, qy = 0.f, qz = 0.f;  // orientation, unit quaternion

    // row-major rotation of the (normalised) Hamilton quaternion
    std::array<float, 9> rotation_matrix() const
    {
        double w = qw, a = qx, b = qy, c = qz;
        const double n = std::sqrt(w * w + a * a + b * b + c * c);
        if (n > 0) { w /= n; a /= n; b /= n; c /= n; }
        return {(float)(1 - 2 * (b * b + c * c)), (float)(2 * (a * b - w * c)), (float)(2 * (a * c + w * b)),
                (float)(2 * (a * b + w * c)), (float)(1 - 2 * (a * a + c * c)), (float)(2 * (b * c - w * a)),
                (float)(2 * (a * c - w * b)), (float)(2 * (b * c + w * a)), (float)(1 - 2 * (a * a + b * b))};
    }
    // one row of a pose CSV: "quat,tx,ty,tz,qw,qx,qy,qz" (examples/pose_source.csv)
    static Pose decode_string(const std::string& line, char delimiter = ',')
    {
        std::vector<std::string> f;
        std::stringstream ss(line);
        std::string item;
        while (std::getline(ss, item, delimiter)) f.push_back(item);
        if (f.size() != 8 || f[0] != "quat")
            throw FileError("pose '" + line + "' is not quat,tx,ty,tz,qw,qx,qy,qz");
        Pose p;
        float* v[7] = {&p.x, &p.y, &p.z, &p.qw, &p.qx, &p.qy, &p.qz};
        for (int k = 0; k < 7; k++) {
            char* end = nullptr;
            *v[k] = std::strtof(f[k + 1].c_str(), &end);
            if (end == f[k + 1].c_str()) throw FileError("bad number in pose '" + line + "'");
        }
        return p;
    }
};

class PoseSource {
   public:
    using Ptr = std::shared_ptr<PoseSource>;
    using ConstPtr = std::shared_ptr<const PoseSource>;
    virtual ~PoseSource() = default;
    virtual int size() const = 0;        // may be -1 when undefined
    virtual int remaining() const = 0;
    virtual Pose next_pose() = 0;
};

class PoseSourceStatic : public PoseSource {
   public:
    using Ptr = std::shared_ptr<PoseSourceStatic>;
    static Ptr Create(const Pose& pose = Pose()) { return Ptr(new PoseSourceStatic(pose)); }
    void set_pose(const Pose& pose) { pose_ = pose; }
    int size() const override { return -1; }
    int remaining() const override { return -1; }
    Pose next_pose() override { return pose_; }

   protected:
    explicit PoseSourceStatic(const Pose& pose) : pose_(pose) {}
    Pose pose_;
};

class PoseSourceList : public PoseSource {
   public:
    using Ptr = std::shared_ptr<PoseSourceList>;
    static Ptr Create(const std::deque<Pose>& poses) { return Ptr(new PoseSourceList(poses)); }
    // PoseSourceList::CreateFromCSV (PoseSource.cpp:7-30): '#' lines are skipped
    static Ptr CreateFromCSV(const std::string& path, unsigned int skipLines = 0, char delimiter = ',')
    {
        std::ifstream file(path);
        if (!file.is_open()) throw FileError(" : could not open '" + path + "'");
        std::string line;
        for (unsigned int i = 0; i < skipLines && std::getline(file, line); i++) {}
        std::deque<Pose> poses;
        while (std::getline(file, line)) {
            if (line.empty() || line[0] == '#') continue;
            poses.push_back(Pose::decode_string(line, delimiter));
        }
        return Create(poses);
    }
    int size() const override { return (int)poses_.size(); }
    int remaining() const override { return (int)poses_.size() - current_; }
    Pose next_pose() override { return poses_[current_++]; }

   protected:
    explicit PoseSourceList(const std::deque<Pose>& poses) : poses_(poses) {}
    std::deque<Pose> poses_;
    int current_ = 0;
};

// ---------------------------------------------------------------------------
// Emitter (Emitter.h:67-115, Emitter.cu:21-80): ray directions (a polar fan or
// a caller's list, 3 floats per ray), their initial values D(dir) from a
// directivity table (normalised coordinates, scale 1/pi), a frequency and a
// pose; rays leave along pose.rotation * dir.
// ---------------------------------------------------------------------------
class Emitter {
   public:
    using Ptr = std::shared_ptr<Emitter>;
    using ConstPtr = std::shared_ptr<const Emitter>;

    static std::vector<float> generate_polar_directions(float resolution, float bearingAperture,
                                                        float elevationAperture)
    {
        size_t n = 0;
        check(rtsim_emitter_polar_directions(resolution, bearingAperture, elevationAperture, nullptr,
                                             0, &n));
        std::vector<float> d(3 * n);
        check(rtsim_emitter_polar_directions(resolution, bearingAperture, elevationAperture, d.data(),
                                             n, &n));
        return d;
    }
    static Ptr Create(const std::vector<float>& rayDirs, const DirectivityTable& directivity,
                      float frequency, const Pose& pose = Pose())
    {
        return Ptr(new Emitter(rayDirs, directivity, frequency, pose));
    }
    // resolution and apertures in degrees, as Emitter::Create
    static Ptr Create(float resolution, float bearingAperture, float elevationAperture,
                      const DirectivityTable& directivity, float frequency, const Pose& pose = Pose())
    {
        return Create(generate_polar_directions(resolution, bearingAperture, elevationAperture),
                      directivity, frequency, pose);
    }
    unsigned int size() const { return ray_count(); }
    unsigned int ray_count() const { return (unsigned int)(directions_.size() / 3); }
    float frequency() const { return frequency_; }
    const Pose& pose() const { return pose_; }
    Pose& pose() { return pose_; }
    const std::vector<float>& directions() const { return directions_; }
    const std::vector<float>& initial_values() const { return values_; }   // complex interleaved

   protected:
    Emitter(const std::vector<float>& dirs, const DirectivityTable& d, float frequency, const Pose& pose)
        : directions_(dirs), values_(2 * (dirs.size() / 3)), frequency_(frequency), pose_(pose)
    {
        const float inv_pi = (float)(1.0 / M_PI);
        check(rtsim_emitter_initial_values(directions_.data(), directions_.size() / 3,
                                           d.data.empty() ? nullptr : d.data.data(), d.width, d.height,
                                           inv_pi, inv_pi, values_.data()));
    }
    std::vector<float> directions_, values_;
    float frequency_;
    Pose pose_;
};

// ---------------------------------------------------------------------------
// Simulation loop: OptixSimulation1::run (OptixSimulation.cpp:56-92) with the
// ray caster (out of scope here) replaced by the seeded synthetic scene
// generator: per ping, the next receiver pose, `samples` hits of the scene
// (seed + ping index), compute_output, then every sink.  run_batch() does
// `in_flight` pings at once through a Batch (batched launches in one graph).
// ---------------------------------------------------------------------------
class SyntheticSimulation {
   public:
    SyntheticSimulation(const std::string& sensorConfig, const rtsim_scene_desc& scene, size_t samples,
                        Form form, unsigned int inFlight = 1, uint64_t seed = 1)
        : scene_(scene), samples_(samples), form_(form), seed_(seed)
    {
        if (inFlight < 1) inFlight = 1;
        for (unsigned int k = 0; k < inFlight; k++) {
            Slot sl;
            sl.sensor = SensorFactory::Make(sensorConfig);
            sl.sensor->set_sample_count((unsigned int)samples);
            for (int a = 0; a < 5; a++) {
                const size_t bytes = sizeof(float) * (a == 4 ? 2 : 1) * (samples ? samples : 1);
                check(rtsim_device_alloc(&sl.dev[a], bytes));
            }
            slots_.push_back(std::move(sl));
        }
    }
    ~SyntheticSimulation()
    {
        for (auto& sl : slots_)
            for (auto p : sl.dev)
                if (p) rtsim_device_free(p);
    }
    SyntheticSimulation(const SyntheticSimulation&) = delete;
    SyntheticSimulation& operator=(const SyntheticSimulation&) = delete;

    void set_receiver_poses(const PoseSource::Ptr& poses) { poses_ = poses; }
    // rays of an emitter instead of the scene's random fan (the scene must be
    // SEABED); ray i uses the emitter's ray i mod ray_count
    void set_emitter(const Emitter::Ptr& emitter) { emitter_ = emitter; }
    void add_sink(const FileSink::Ptr& sink) { sinks_.push_back(sink); }
    uint64_t pings() const { return ping_; }
    const SensorInstance::Ptr& receiver(unsigned int slot = 0) const { return slots_[slot].sensor; }

    // one ping (OptixSimulation1::run); false when the pose source is exhausted
    bool run() { return run_batch(1) == 1; }

    // up to `count` pings (at most in_flight) as one batch; returns pings done
    size_t run_batch(size_t count = (size_t)-1)
    {
        if (!poses_) throw std::runtime_error("Receiver pose source was not set");
        std::vector<SensorInstance::Ptr> active;
        std::vector<Form> forms;
        for (auto& sl : slots_) {
            if (active.size() >= count || poses_->remaining() == 0) break;
            sl.sensor->set_pose(poses_->next_pose().rotation_matrix());
            load_hits(sl, seed_ + ping_ + active.size());
            active.push_back(sl.sensor);
            forms.push_back(form_);
        }
        if (active.empty()) return 0;
        if (active.size() == 1) {
            Slot& sl = slots_[0];
            sl.sensor->set_samples(sl.hits(), samples_);
            sl.sensor->compute_output(form_);
        } else {
            // one batch per number of pings in flight, kept across calls: the
            // same sensors and hit buffers replay its graph (new poses are
            // rewritten in place, no re-capture)
            if (!batch_ || batch_size_ != active.size()) {
                batch_.reset();
                batch_.reset(new Batch(active, forms));
                batch_size_ = active.size();
            }
            for (size_t k = 0; k < active.size(); k++) batch_->bind_hits((uint32_t)k, &slots_[k].hits_, samples_);
            batch_->compute();
        }
        for (const auto& s : active)
            for (const auto& sink : sinks_) sink->set_output(s);
        ping_ += active.size();
        return active.size();
    }

   private:
    struct Slot {
        SensorInstance::Ptr sensor;
        void* dev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
        rtsim_hits hits_{};
        const rtsim_hits& hits() const { return hits_; }
    };
    void load_hits(Slot& sl, uint64_t seed)
    {
        std::vector<float> r(samples_), dx(samples_), dy(samples_), dz(samples_), a(2 * samples_);
        if (emitter_) {
            const auto rot = emitter_->pose().rotation_matrix();
            check(rtsim_generate_emitter_hits(&scene_, seed, emitter_->directions().data(),
                                              emitter_->initial_values().data(), emitter_->ray_count(),
                                              rot.data(), 0, samples_, r.data(), dx.data(), dy.data(),
                                              dz.data(), a.data(), 8));
        } else {
            check(rtsim_generate_hits(&scene_, seed, 0, samples_, r.data(), dx.data(), dy.data(),
                                      dz.data(), a.data(), 8));
        }
        const std::vector<float>* host[5] = {&r, &dx, &dy, &dz, &a};
        for (int k = 0; k < 5; k++)
            if (!host[k]->empty())
                check(rtsim_memcpy_h2d(sl.dev[k], host[k]->data(), sizeof(float) * host[k]->size()));
        sl.hits_ = rtsim_hits{(const float*)sl.dev[0], (const float*)sl.dev[1], (const float*)sl.dev[2],
                              (const float*)sl.dev[3], (const float*)sl.dev[4]};
    }

    rtsim_scene_desc scene_;
    size_t samples_;
    Form form_;
    uint64_t seed_, ping_ = 0;
    std::vector<Slot> slots_;
    std::unique_ptr<Batch> batch_;   // declared after the slots: destroyed first
    size_t batch_size_ = 0;
    PoseSource::Ptr poses_;
    Emitter::Ptr emitter_;
    std::vector<FileSink::Ptr> sinks_;
};

}  // namespace simulation
}  // namespace rtac
