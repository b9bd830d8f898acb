// Sensor configuration from YAML: the SensorFactory / SensorInfoFactory
// drop-in (SensorInfoFactory.cpp:7-235, factories/utilities.cpp:22-257).
//
// Reference schema kept as is: type (front-scan | single-beam | file),
// output-node, globals through anchors/aliases, sampling.bearings (csv),
// sampling.ranges (size, range-max, unit, globals.range-min), directivity
// (rectangle | disk | uniform), waveform (sine, adaptive | fixed duration),
// beamsteering (sinc).  Extensions for explicit PSFs and binning:
//   waveform.type        hann-baseband {taps, carrier} | lfm {taps, kappa}
//   beamsteering.type    sinc2 {taps, width}
//   psf                  {type: csv, path, range-taps, bearing-taps}
//   sampling.bearings    sine-uniform | linear {size, aperture, unit}
//   directivity.type     fan {size, bearing-width, elevation-sigma, unit}
//   bearing-axis         angle (default) | sine
//   binner               {overlap: bins} | {margin: metres}
//   spreading            inverse-square (default, RayCaster.cu:69-71) | none
//   frequency            Hz of the round-trip phase (default: globals or waveform)
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rtsim.h"
#include "../../include/rtsim_config.h"
#include "../../include/rtsim_scene.h"
#include "sensor.h"
#include "yaml_lite.h"

namespace fs = std::filesystem;

namespace rtsim {

namespace {

struct ConfigError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] void cfg_error(const std::string& m) { throw ConfigError(m); }

using yaml::NodePtr;

// ---------------------------------------------------------------------------
// FileFinder (factories/utilities.cpp:22-109): search paths from
// RTAC_CONFIG_PATH (':'-separated), paths added at run time, then the
// directory of the YAML file being read.  A file matches when its path ends
// with the requested name; directories are searched recursively.
// ---------------------------------------------------------------------------
std::vector<std::string>& search_paths()
{
    static std::vector<std::string> paths;
    static bool init = false;
    if (!init) {
        init = true;
        if (const char* env = std::getenv("RTAC_CONFIG_PATH")) {
            std::istringstream iss(env);
            std::string tok;
            while (std::getline(iss, tok, ':'))
                if (!tok.empty() && fs::is_directory(tok)) paths.push_back(fs::absolute(tok).string());
        }
    }
    return paths;
}

std::string find_one(const std::string& name, const std::string& extra_dir)
{
    if (fs::is_regular_file(name)) return name;
    std::vector<std::string> dirs = search_paths();
    if (!extra_dir.empty()) dirs.insert(dirs.begin(), extra_dir);
    for (const auto& d : dirs) {
        fs::path direct = fs::path(d) / name;
        if (fs::is_regular_file(direct)) return direct.string();
        std::error_code ec;
        for (auto it = fs::recursive_directory_iterator(d, ec);
             !ec && it != fs::recursive_directory_iterator(); it.increment(ec)) {
            if (!it->is_regular_file()) continue;
            std::string p = it->path().string();
            if (p.size() >= name.size() && p.compare(p.size() - name.size(), name.size(), name) == 0)
                return p;
        }
    }
    return std::string();
}

// utilities.cpp:114-153: first token 'deg' | 'rad' | a value, comma separated
std::vector<float> bearings_from_csv(const std::string& path)
{
    std::ifstream f(path);
    if (!f.is_open()) cfg_error("cannot open bearings file '" + path + "'");
    std::vector<float> out;
    std::string tok;
    if (!std::getline(f, tok, ',')) cfg_error("'" + path + "': invalid format");
    auto trim = [](std::string s) {
        size_t b = s.find_first_not_of(" \t\r\n"), e = s.find_last_not_of(" \t\r\n");
        return b == std::string::npos ? std::string() : s.substr(b, e - b + 1);
    };
    tok = trim(tok);
    float scaling = 1.0f;
    if (tok == "deg") scaling = (float)(M_PI / 180.0f);
    else if (tok != "rad") {
        try { out.push_back(std::stof(tok)); }
        catch (...) { cfg_error("'" + path + "': first item must be deg, rad or a number"); }
    }
    while (std::getline(f, tok, ',')) {
        tok = trim(tok);
        if (tok.empty()) continue;
        try { out.push_back(scaling * std::stof(tok)); }
        catch (...) { cfg_error("'" + path + "': could not parse '" + tok + "'"); }
    }
    return out;
}

float distance_unit(const NodePtr& n)
{
    if (!n) return 1.0f;
    const std::string& u = n->as_string("unit");
    if (u == "cm") return 0.01f;
    if (u == "mm") return 0.001f;
    if (u != "m") cfg_error("unsupported distance unit '" + u + "'");
    return 1.0f;
}

float angle_unit(const NodePtr& n)
{
    if (!n) return 1.0f;
    const std::string& u = n->as_string("unit");
    if (u == "deg") return (float)(M_PI / 180.0f);
    if (u != "rad") cfg_error("unsupported angle unit '" + u + "'");
    return 1.0f;
}

NodePtr need(const NodePtr& n, const std::string& key, const std::string& where)
{
    NodePtr v = n ? n->get(key) : nullptr;
    if (!v) cfg_error(where + ": missing '" + key + "'");
    return v;
}

double num(const NodePtr& n, const std::string& key, const std::string& where)
{
    return need(n, key, where)->as_double(where + "." + key);
}

double num_or(const NodePtr& n, const std::string& key, double dflt)
{
    NodePtr v = n ? n->get(key) : nullptr;
    return v ? v->as_double(key) : dflt;
}

std::string str_or(const NodePtr& n, const std::string& key, const std::string& dflt)
{
    NodePtr v = n ? n->get(key) : nullptr;
    return v ? v->as_string(key) : dflt;
}

double global(const NodePtr& n, const std::string& key, double dflt)
{
    NodePtr g = n ? n->get("globals") : nullptr;
    return num_or(g, key, dflt);
}

std::vector<double> hann(int n)
{
    std::vector<double> w(n);
    for (int j = 0; j < n; j++) w[j] = 0.5 - 0.5 * std::cos(2.0 * M_PI * (j + 1) / (n + 1));
    return w;
}

double sinc(double x) { return x == 0.0 ? 1.0 : std::sin(M_PI * x) / (M_PI * x); }

}  // namespace

// Everything a sensor description points to.
struct Info {
    rtsim_sensor_desc desc{};
    rtsim_waveform_sine wave{};     // desc.waveform points here for a sine waveform
    std::string name;
    std::vector<float> bearings, psf, directivity, beam;
};

namespace {

// BeamDirectivity_Sinc (BeamDirectivity.h:94-99) built on rtac_base's
// signal::SincFunction(span / resolution, oversampling), which is not part of
// this tree: restated as sin(pi x)/(pi x) sampled at `oversampling` points per
// unit of x over x in [-L/2, L/2], L = span / resolution.
std::vector<float> sinc_beam(float span, float resolution, unsigned os)
{
    double L = (double)span / resolution;
    unsigned n = (unsigned)(os * L) + 1;
    std::vector<float> b(n);
    for (unsigned k = 0; k < n; k++) b[k] = (float)sinc(-0.5 * L + (double)k / os);
    return b;
}

void read_csv_taps(const std::string& path, int Tr, int Tb, std::vector<float>& out)
{
    std::ifstream f(path);
    if (!f.is_open()) cfg_error("cannot open PSF file '" + path + "'");
    std::vector<float> v;
    std::string tok;
    while (f >> tok) {
        std::stringstream ss(tok);
        std::string item;
        while (std::getline(ss, item, ',')) {
            if (item.empty()) continue;
            try { v.push_back(std::stof(item)); }
            catch (...) { cfg_error("'" + path + "': bad number '" + item + "'"); }
        }
    }
    if (v.size() != 2 * (size_t)Tr * Tb)
        cfg_error("'" + path + "': expected " + std::to_string(2 * Tr * Tb) + " values (re, im)");
    out = v;
}

void build_info(const NodePtr& root, const std::string& base_dir, Info& I, int depth)
{
    if (!root || root->kind != yaml::Node::Map) cfg_error("sensor config is not a mapping");
    NodePtr tnode = root->get("type");
    if (!tnode) cfg_error("sensor config has no type");
    std::string type = tnode->as_string("type");
    if (type == "file") {
        if (depth > 4) cfg_error("too many nested 'file' configs");
        std::string p = find_one(need(root, "path", "file")->as_string("path"), base_dir);
        if (p.empty()) cfg_error("config file not found");
        build_info(yaml::parse_file(p), fs::path(p).parent_path().string(), I, depth + 1);
        return;
    }
    std::string output = str_or(root, "output-node", "complex");
    if (output == "real") cfg_error("real sensor output mode not supported yet");
    if (output != "complex") cfg_error("unsupported output mode (" + output + ")");
    bool twod;
    if (type == "front-scan") twod = true;
    else if (type == "single-beam") twod = false;
    else cfg_error("unsupported sensor type '" + type + "'");
    I.name = str_or(root, "name", "unnamed");

    rtsim_sensor_desc& d = I.desc;
    std::memset(&d, 0, sizeof(d));
    d.kind = twod ? RTSIM_SENSOR_2D : RTSIM_SENSOR_1D;
    d.rotation[0] = d.rotation[4] = d.rotation[8] = 1.0f;
    d.dir_xscale = d.dir_yscale = (float)(1.0f / M_PI);

    // ---- sampling (SensorInfoFactory.cpp:83-148) --------------------------
    NodePtr sampling = root->get("sampling");
    if (!sampling) cfg_error("no 'sampling' node");
    NodePtr rn = sampling->get("ranges");
    if (!rn) cfg_error("invalid ranges node");
    float range_min = (float)global(rn, "range-min", num_or(rn, "range-min", 0.0));
    long size = need(rn, "size", "ranges")->as_long("ranges.size");
    if (size < 1) cfg_error("ranges.size must be >= 1");
    float range_max = distance_unit(rn->get("unit")) * (float)num(rn, "range-max", "ranges");
    d.range_min = range_min;
    d.range_max = range_max;
    d.range_count = (uint32_t)size;
    const float c = 1500.0f;   // SensorInfo.h:38: the YAML sound-celerity is not used
    d.sound_speed = c;
    // Linspace<float>::resolution() in FP32 (the adaptive waveform duration);
    // explicit-tap spans and overlap margins use the exact step
    const float resolution = size > 1 ? (range_max - range_min) / (float)(size - 1) : 0.f;
    const double step = size > 1 ? ((double)range_max - (double)range_min) / (double)(size - 1) : 0.0;

    if (twod) {
        NodePtr bn = sampling->get("bearings");
        if (!bn) cfg_error("invalid bearings node");
        std::string bt = need(bn, "type", "bearings")->as_string("bearings.type");
        if (bt == "csv") {
            std::string p = find_one(need(bn, "path", "bearings")->as_string("path"), base_dir);
            if (p.empty()) cfg_error("bearings file not found");
            I.bearings = bearings_from_csv(p);
        } else if (bt == "sine-uniform" || bt == "linear") {
            long n = need(bn, "size", "bearings")->as_long("bearings.size");
            double ap = angle_unit(bn->get("unit")) * num(bn, "aperture", "bearings");
            if (n < 2) cfg_error("bearings.size must be >= 2");
            I.bearings.resize(n);
            double s = std::sin(0.5 * ap);
            for (long b = 0; b < n; b++) {
                double t = (double)b / (n - 1);
                I.bearings[b] = bt == "linear" ? (float)(-0.5 * ap + t * ap)
                                               : (float)std::asin(-s + t * 2 * s);
            }
        } else {
            cfg_error("unsupported bearing config type: " + bt);
        }
        std::string axis = str_or(root, "bearing-axis", "angle");
        if (axis == "sine") d.bearing_axis = RTSIM_AXIS_SINE;
        else if (axis == "angle") d.bearing_axis = RTSIM_AXIS_ANGLE;
        else cfg_error("bearing-axis must be 'angle' or 'sine'");
    }

    // ---- waveform (SensorInfoFactory.cpp:168-200) -> range taps ------------
    NodePtr wn = root->get("waveform");
    if (!wn) cfg_error("invalid 'waveform' node");
    std::string wt = need(wn, "type", "waveform")->as_string("waveform.type");
    unsigned os = (unsigned)num_or(wn, "oversampling", 8);
    if (os < 1) cfg_error("waveform.oversampling must be >= 1");
    std::vector<std::complex<double>> wave;
    float span_r = 0.f;
    double wave_freq = 0;
    bool sine = false;
    if (wt == "sine") {
        float f = (float)num(wn, "frequency", "waveform");
        if (!(f > 0.f)) cfg_error("waveform.frequency must be > 0");
        wave_freq = f;
        std::string mode = str_or(wn, "duration-mode", "adaptive");
        float duration;
        if (mode == "fixed") duration = (float)num(wn, "duration", "waveform");
        else if (mode == "adaptive") duration = resolution / c;   // set_ranges (SensorInstance.cu:23-34)
        else cfg_error("waveform.duration-mode must be 'adaptive' or 'fixed'");
        // the taps come from the recipe the sensor keeps (the table below is
        // the same one, for the description's readers)
        sine = true;
        I.wave.frequency = f;
        I.wave.duration = duration;
        I.wave.fixed = mode == "fixed";
        I.wave.oversampling = os;
        uint32_t T = 0;
        std::vector<float> t = sine_psf_table(f, duration, os, nullptr, 1, &T);
        for (uint32_t h = 0; h < T; h++) wave.emplace_back(t[2 * h], t[2 * h + 1]);
        span_r = duration * c;
    } else if (wt == "hann-baseband" || wt == "lfm") {
        long T = need(wn, "taps", "waveform")->as_long("waveform.taps");
        if (T < 1) cfg_error("waveform.taps must be >= 1");
        auto w = hann((int)T);
        for (long j = 0; j < T; j++) {
            double x = j - 0.5 * (T - 1);
            double ph = wt == "lfm" ? M_PI * num_or(wn, "kappa", 0.02) * x * x
                                    : M_PI * num_or(wn, "carrier", 0.25) * x;
            wave.push_back(w[j] * std::exp(std::complex<double>(0, ph)));
        }
        span_r = (float)(T * step);
        wave_freq = num_or(wn, "frequency", 0.0);
    } else {
        cfg_error("unsupported waveform type: '" + wt + "'");
    }
    if (wave.empty()) cfg_error("the waveform has no samples (duration too short)");
    if (wave.size() > kMaxWaveformTaps) cfg_error("the waveform has more than 65536 samples");

    // ---- beamsteering (SensorInfoFactory.cpp:202-235) -> bearing taps -----
    std::vector<double> beam(1, 1.0);
    float span_b = 1.0f;
    if (twod) {
        NodePtr bs = root->get("beamsteering");
        if (!bs) cfg_error("invalid 'beamsteering' node");
        std::string bt = need(bs, "type", "beamsteering")->as_string("beamsteering.type");
        if (bt == "sinc") {
            float sc = angle_unit(bs->get("unit"));
            float span = sc * (float)num(bs, "span", "beamsteering");
            float res = sc * (float)num(bs, "resolution", "beamsteering");
            unsigned bos = (unsigned)num_or(bs, "oversampling", 8);
            beam.clear();
            for (float v : sinc_beam(span, res, bos)) beam.push_back(v);
            span_b = span;
        } else if (bt == "sinc2") {
            long T = need(bs, "taps", "beamsteering")->as_long("beamsteering.taps");
            double width = num(bs, "width", "beamsteering");
            beam.resize(T);
            for (long k = 0; k < T; k++) beam[k] = std::pow(sinc((k - 0.5 * (T - 1)) / width), 2);
            // tap pitch = the bearing pitch of the output grid (in the PSF axis)
            const auto& B = I.bearings;
            double x0 = d.bearing_axis == RTSIM_AXIS_SINE ? std::sin((double)B.front()) : B.front();
            double x1 = d.bearing_axis == RTSIM_AXIS_SINE ? std::sin((double)B.back()) : B.back();
            span_b = (float)(T * (x1 - x0) / (B.size() - 1));
        } else {
            cfg_error("unsupported beamsteering type: '" + bt + "'");
        }
    }

    // ---- PSF table (generate_psf_data, SensorInstance2D.cu:20-33) ---------
    const int Tr = (int)wave.size(), Tb = (int)beam.size();
    I.psf.assign(2 * (size_t)Tr * Tb, 0.f);
    for (int h = 0; h < Tr; h++)
        for (int w = 0; w < Tb; w++) {
            std::complex<float> v = std::complex<float>((float)wave[h].real(), (float)wave[h].imag()) *
                                    (float)beam[w];
            I.psf[2 * ((size_t)h * Tb + w)] = v.real();
            I.psf[2 * ((size_t)h * Tb + w) + 1] = v.imag();
        }
    int psfTr = Tr, psfTb = Tb;
    if (sine) {
        I.beam.assign(beam.begin(), beam.end());
        I.wave.beam = I.beam.data();
        I.wave.beam_taps = (uint32_t)I.beam.size();
        d.waveform = &I.wave;
    }
    if (NodePtr pn = root->get("psf")) {
        d.waveform = nullptr;   // explicit taps are fixed
        std::string pt = need(pn, "type", "psf")->as_string("psf.type");
        if (pt != "csv") cfg_error("psf.type must be 'csv'");
        psfTr = (int)need(pn, "range-taps", "psf")->as_long("psf.range-taps");
        psfTb = (int)need(pn, "bearing-taps", "psf")->as_long("psf.bearing-taps");
        if (psfTr != Tr || psfTb != Tb)
            cfg_error("psf file shape differs from the waveform x beamsteering taps");
        std::string p = find_one(need(pn, "path", "psf")->as_string("path"), base_dir);
        if (p.empty()) cfg_error("psf file not found");
        read_csv_taps(p, psfTr, psfTb, I.psf);
    }
    d.psf_range_taps = (uint32_t)psfTr;
    d.psf_bearing_taps = (uint32_t)psfTb;
    d.psf_range_span = span_r;
    d.psf_bearing_span = span_b;

    // ---- binner margin (SensorInstance.cu:31: margin = c * tau) -------------
    d.margin = -1.0f;
    if (NodePtr bn = root->get("binner")) {
        if (NodePtr ov = bn->get("overlap")) {
            long o = ov->as_long("binner.overlap");
            d.margin = (float)(o * step * (1.0 + std::ldexp(1.0, -10)));
        } else if (NodePtr m = bn->get("margin")) {
            d.margin = (float)m->as_double("binner.margin");
        }
    }

    // ---- directivity (utilities.cpp:155-211) ---------------------------------
    NodePtr dn = root->get("directivity");
    if (!dn) cfg_error("invalid directivity node");
    std::string dt = need(dn, "type", "directivity")->as_string("directivity.type");
    float scaling = distance_unit(dn->get("unit"));
    bool single = str_or(dn, "baffle-mode", "") == "single-sided";
    if (str_or(dn, "baffle-mode", "") == "cardioid") cfg_error("baffleMode 'cardioid' unsupported");
    float wavelength = -1.0f;
    if (NodePtr wl = dn->get("wavelength")) wavelength = scaling * (float)wl->as_double("wavelength");
    uint32_t W = 0, H = 0;
    int rc = RTSIM_OK;
    if (dt == "rectangle" || dt == "disk") {
        if (wavelength < 0) cfg_error("no wavelength info to create directivity function");
        if (dt == "rectangle") {
            float w = scaling * (float)num(dn, "width", "directivity");
            float h = scaling * (float)num(dn, "height", "directivity");
            rc = rtsim_directivity_rectangle(w, h, wavelength, single, 8, nullptr, 0, &W, &H);
            I.directivity.resize((size_t)W * H);
            if (!rc) rc = rtsim_directivity_rectangle(w, h, wavelength, single, 8, I.directivity.data(),
                                                      I.directivity.size(), &W, &H);
        } else {
            float D = scaling * (float)num(dn, "diameter", "directivity");
            rc = rtsim_directivity_disk(D, wavelength, single, 8, nullptr, 0, &W, &H);
            I.directivity.resize((size_t)W * H);
            if (!rc) rc = rtsim_directivity_disk(D, wavelength, single, 8, I.directivity.data(),
                                                 I.directivity.size(), &W, &H);
        }
        if (rc) cfg_error(fail_message());
    } else if (dt == "uniform") {
        W = H = 1;
        I.directivity.assign(1, 1.0f);
    } else if (dt == "fan") {
        long n = (long)num_or(dn, "size", 256);
        if (n < 1) cfg_error("directivity.size must be >= 1");
        float au = dn->get("angle-unit") ? angle_unit(dn->get("angle-unit")) : (float)(M_PI / 180.0);
        double bw = au * num(dn, "bearing-width", "directivity");
        double es = au * num(dn, "elevation-sigma", "directivity");
        W = H = (uint32_t)n;
        I.directivity.resize((size_t)n * n);
        for (long h = 0; h < n; h++) {
            double a_e = M_PI * (h + 0.5) / n;
            double e = a_e > M_PI / 2 ? 0.0 : std::exp(-0.5 * (a_e / es) * (a_e / es));
            for (long w = 0; w < n; w++) {
                double a_b = M_PI * (w + 0.5) / n;
                double b = a_b > M_PI / 2 ? 0.0 : std::pow(sinc(a_b / bw), 2);
                I.directivity[(size_t)h * n + w] = (float)(e * b);
            }
        }
    } else {
        cfg_error("unsupported directivity type: " + dt);
    }
    d.dir_width = W;
    d.dir_height = H;

    // ---- acoustics ---------------------------------------------------------
    double f = num_or(root, "frequency", 0.0);
    if (f == 0.0) f = global(sampling->get("ranges"), "frequency", 0.0);
    if (f == 0.0) f = global(dn, "frequency", 0.0);
    if (f == 0.0) f = wave_freq;
    d.frequency = (float)f;
    // RayCaster.cu:69-71 always divides by the squared travel
    std::string sp = str_or(root, "spreading", "inverse-square");
    if (sp == "none") d.spreading = RTSIM_SPREAD_NONE;
    else if (sp == "inverse-square") d.spreading = RTSIM_SPREAD_INV_R2;
    else cfg_error("spreading must be 'none' or 'inverse-square'");

    d.bearings = I.bearings.empty() ? nullptr : I.bearings.data();
    d.bearing_count = (uint32_t)I.bearings.size();
    d.psf = I.psf.data();
    d.directivity = I.directivity.data();
}

}  // namespace
}  // namespace rtsim

using namespace rtsim;

extern "C" {

int rtsim_info_from_yaml(const char* path, rtsim_info* out)
{
    if (!path || !out) return fail(RTSIM_ERR_INVALID, "path/out is NULL");
    *out = nullptr;
    std::string p = find_one(path, "");
    if (p.empty()) return fail(RTSIM_ERR_IO, "config file '%s' not found", path);
    Info* I = new Info();
    try {
        build_info(yaml::parse_file(p), fs::path(p).parent_path().string(), *I, 0);
    } catch (const ConfigError& e) {
        delete I;
        return fail(RTSIM_ERR_CONFIG, "%s: %s", p.c_str(), e.what());
    } catch (const std::exception& e) {
        delete I;
        return fail(RTSIM_ERR_CONFIG, "%s: %s", p.c_str(), e.what());
    }
    *out = reinterpret_cast<rtsim_info>(I);
    return RTSIM_OK;
}

int rtsim_info_desc(rtsim_info info, const rtsim_sensor_desc** desc)
{
    if (!info || !desc) return fail(RTSIM_ERR_INVALID, "info/desc is NULL");
    *desc = &reinterpret_cast<Info*>(info)->desc;
    return RTSIM_OK;
}

int rtsim_info_destroy(rtsim_info info)
{
    delete reinterpret_cast<Info*>(info);
    return RTSIM_OK;
}

int rtsim_sensor_create_from_yaml(const char* path, rtsim_sensor* out)
{
    if (!out) return fail(RTSIM_ERR_INVALID, "out is NULL");
    rtsim_info info = nullptr;
    int rc = rtsim_info_from_yaml(path, &info);
    if (rc) return rc;
    const rtsim_sensor_desc* d = nullptr;
    rtsim_info_desc(info, &d);
    rc = rtsim_sensor_create(d, out);
    rtsim_info_destroy(info);
    return rc;
}

int rtsim_config_add_search_path(const char* path)
{
    if (!path || !fs::is_directory(path)) return fail(RTSIM_ERR_IO, "'%s' is not a directory", path ? path : "");
    search_paths().push_back(fs::absolute(path).string());
    return RTSIM_OK;
}

int rtsim_config_find(const char* name, char* out, size_t capacity)
{
    if (!name || !out || capacity == 0) return fail(RTSIM_ERR_INVALID, "bad arguments");
    std::string p = find_one(name, "");
    if (p.empty()) return fail(RTSIM_ERR_IO, "'%s' not found", name);
    if (p.size() + 1 > capacity) return fail(RTSIM_ERR_INVALID, "buffer too small");
    std::memcpy(out, p.c_str(), p.size() + 1);
    return RTSIM_OK;
}

}  // extern "C"
