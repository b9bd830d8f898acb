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

#include <complex>
#include <cstdint>
#include <memory>
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
    }
    rtsim_info info_ = nullptr;
    rtsim_sensor_desc desc_{};
    std::vector<float> psf_, bearings_, dir_;
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

}  // namespace simulation
}  // namespace rtac
