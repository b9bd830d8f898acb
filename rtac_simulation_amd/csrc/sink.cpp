// Ping files: the record stream declared in include/rtsim_sink.h, written by
// the FileSink drop-in (FileSink.cpp:8-42) and read back by the readers.
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/rtsim.h"
#include "../../include/rtsim_sink.h"
#include "sensor.h"

namespace rtsim {

namespace {

constexpr char kMagic[4] = {'R', 'T', 'P', 'G'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeaderBytes = 48;

uint64_t fnv1a(uint64_t h, const void* p, size_t n)
{
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    return h;
}
constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ull;

void put32(unsigned char* p, uint32_t v) { std::memcpy(p, &v, 4); }
void put64(unsigned char* p, uint64_t v) { std::memcpy(p, &v, 8); }
uint32_t get32(const unsigned char* p) { uint32_t v; std::memcpy(&v, p, 4); return v; }
uint64_t get64(const unsigned char* p) { uint64_t v; std::memcpy(&v, p, 8); return v; }

// files::prepare_path: create the parent directories
bool make_parents(const std::string& path)
{
    for (size_t i = 1; i < path.size(); i++) {
        if (path[i] != '/') continue;
        const std::string dir = path.substr(0, i);
        if (mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST) return false;
    }
    return true;
}

struct Sink {
    std::FILE* f = nullptr;
    uint64_t count = 0;
    std::vector<float> host;
    ~Sink()
    {
        if (f) std::fclose(f);
    }
};

int write_record(Sink* s, int kind, uint32_t H, uint32_t W, float lo, float hi, const float* bearings,
                 const float* data)
{
    if (kind != RTSIM_SENSOR_1D && kind != RTSIM_SENSOR_2D)
        return fail(RTSIM_ERR_INVALID, "ping kind %d", kind);
    if (kind == RTSIM_SENSOR_1D) W = 1;
    const uint32_t nb = kind == RTSIM_SENSOR_2D ? W : 0;
    if ((nb && !bearings) || (!data && H && W)) return fail(RTSIM_ERR_INVALID, "missing ping data");
    uint64_t h = fnv1a(kFnvBasis, bearings, sizeof(float) * nb);
    h = fnv1a(h, data, sizeof(float) * 2 * (size_t)H * W);
    unsigned char hd[kHeaderBytes] = {};
    std::memcpy(hd, kMagic, 4);
    put32(hd + 4, kVersion);
    put32(hd + 8, (uint32_t)kind);
    put32(hd + 12, 1u);
    put32(hd + 16, H);
    put32(hd + 20, W);
    std::memcpy(hd + 24, &lo, 4);
    std::memcpy(hd + 28, &hi, 4);
    put64(hd + 32, s->count);
    put64(hd + 40, h);
    if (std::fwrite(hd, 1, kHeaderBytes, s->f) != kHeaderBytes ||
        (nb && std::fwrite(bearings, sizeof(float), nb, s->f) != nb) ||
        std::fwrite(data, sizeof(float), 2 * (size_t)H * W, s->f) != 2 * (size_t)H * W ||
        std::fflush(s->f) != 0)
        return fail(RTSIM_ERR_IO, "writing a ping record failed");
    s->count++;
    return RTSIM_OK;
}

// Walks the records of `path`; calls back with (index, header, payload offset).
template <typename F>
int walk(const char* path, F&& fn)
{
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return fail(RTSIM_ERR_IO, "cannot open '%s'", path);
    // the file size bounds every record the headers announce (before anything
    // is sized from them)
    long fsize = -1;
    if (std::fseek(f, 0, SEEK_END) == 0) fsize = std::ftell(f);
    if (fsize < 0 || std::fseek(f, 0, SEEK_SET) != 0) {
        std::fclose(f);
        return fail(RTSIM_ERR_IO, "cannot size '%s'", path);
    }
    uint64_t index = 0;
    int rc = RTSIM_OK;
    for (;;) {
        unsigned char hd[kHeaderBytes];
        const size_t got = std::fread(hd, 1, kHeaderBytes, f);
        if (got == 0) break;
        if (got != kHeaderBytes || std::memcmp(hd, kMagic, 4) != 0 || get32(hd + 4) != kVersion) {
            rc = fail(RTSIM_ERR_IO, "'%s': record %llu is not a version-%u ping record", path,
                      (unsigned long long)index, kVersion);
            break;
        }
        rtsim_ping_header h;
        h.kind = get32(hd + 8);
        h.scalar = get32(hd + 12);
        h.height = get32(hd + 16);
        h.width = get32(hd + 20);
        std::memcpy(&h.range_min, hd + 24, 4);
        std::memcpy(&h.range_max, hd + 28, 4);
        h.index = get64(hd + 32);
        h.checksum = get64(hd + 40);
        const long payload = std::ftell(f);
        if ((h.kind != RTSIM_SENSOR_1D && h.kind != RTSIM_SENSOR_2D) ||
            (h.kind == RTSIM_SENSOR_1D && h.width != 1)) {
            rc = fail(RTSIM_ERR_IO, "'%s': record %llu has an invalid kind / width", path,
                      (unsigned long long)index);
            break;
        }
        // record bytes = 4 (bearings) + 8 H W, checked against what is left of
        // the file without overflow (H, W < 2^32, so H W < 2^64)
        const uint64_t left = (uint64_t)(fsize - payload);
        const uint64_t nb = h.kind == RTSIM_SENSOR_2D ? h.width : 0;
        const uint64_t cells = (uint64_t)h.height * h.width;
        if (cells > left / 8 || 8 * cells + 4 * nb > left) {
            rc = fail(RTSIM_ERR_IO, "'%s': truncated record %llu", path, (unsigned long long)index);
            break;
        }
        const size_t bytes = (size_t)(8 * cells + 4 * nb);
        int stop = fn(index, h, f, payload);
        if (stop < 0) {
            rc = RTSIM_ERR_IO;
            break;
        }
        if (stop > 0) break;
        if (std::fseek(f, payload + (long)bytes, SEEK_SET) != 0) {
            rc = fail(RTSIM_ERR_IO, "'%s': truncated record %llu", path, (unsigned long long)index);
            break;
        }
        index++;
    }
    std::fclose(f);
    return rc;
}

}  // namespace

}  // namespace rtsim

using namespace rtsim;

extern "C" {

int rtsim_sink_open(const char* path, int overwrite, rtsim_sink* out)
{
    if (!path || !out) return fail(RTSIM_ERR_INVALID, "path/out is NULL");
    *out = nullptr;
    struct stat st;
    if (stat(path, &st) == 0 && !overwrite)
        return fail(RTSIM_ERR_IO, "could not prepare path for file '%s': it exists", path);
    if (!make_parents(path)) return fail(RTSIM_ERR_IO, "could not prepare path for file '%s'", path);
    Sink* s = new Sink();
    s->f = std::fopen(path, "wb");
    if (!s->f) {
        delete s;
        return fail(RTSIM_ERR_IO, "cannot create '%s'", path);
    }
    *out = reinterpret_cast<rtsim_sink>(s);
    return RTSIM_OK;
}

int rtsim_sink_write_host(rtsim_sink sink, int kind, uint32_t height, uint32_t width,
                          float range_min, float range_max, const float* bearings,
                          const float* data)
{
    Sink* s = reinterpret_cast<Sink*>(sink);
    if (!s) return fail(RTSIM_ERR_INVALID, "sink is NULL");
    return write_record(s, kind, height, width, range_min, range_max, bearings, data);
}

int rtsim_sink_write(rtsim_sink sink, rtsim_sensor sensor, void* stream)
try {
    Sink* s = reinterpret_cast<Sink*>(sink);
    Sensor* se = reinterpret_cast<Sensor*>(sensor);
    if (!s || !se) return fail(RTSIM_ERR_INVALID, "sink/sensor is NULL");
    const size_t cells = (size_t)se->N * se->W;
    s->host.resize(2 * cells);
    int rc = rtsim_sensor_download_output(sensor, s->host.data(), stream);
    if (rc) return rc;
    return write_record(s, se->kind, se->N, se->W, se->range_min, se->range_max,
                        se->kind == RTSIM_SENSOR_2D ? se->bearings.data() : nullptr, s->host.data());
} catch (const std::exception& e) {
    return fail(RTSIM_ERR_IO, "writing a ping record: %s", e.what());
}

int rtsim_sink_count(rtsim_sink sink, uint64_t* pings)
{
    Sink* s = reinterpret_cast<Sink*>(sink);
    if (!s || !pings) return fail(RTSIM_ERR_INVALID, "sink/pings is NULL");
    *pings = s->count;
    return RTSIM_OK;
}

int rtsim_sink_close(rtsim_sink sink)
{
    delete reinterpret_cast<Sink*>(sink);
    return RTSIM_OK;
}

int rtsim_ping_file_count(const char* path, uint64_t* pings)
try {
    if (!path || !pings) return fail(RTSIM_ERR_INVALID, "path/pings is NULL");
    uint64_t n = 0;
    int rc = walk(path, [&](uint64_t, const rtsim_ping_header&, std::FILE*, long) {
        n++;
        return 0;
    });
    if (rc) return rc;
    *pings = n;
    return RTSIM_OK;
} catch (const std::exception& e) {
    return fail(RTSIM_ERR_IO, "reading '%s': %s", path, e.what());
}

int rtsim_ping_file_read(const char* path, uint64_t index, rtsim_ping_header* header,
                         float* bearings, float* data)
try {
    if (!path || !header) return fail(RTSIM_ERR_INVALID, "path/header is NULL");
    bool found = false;
    int inner = RTSIM_OK;
    int rc = walk(path, [&](uint64_t i, const rtsim_ping_header& h, std::FILE* f, long) {
        if (i != index) return 0;
        found = true;
        *header = h;
        const size_t nb = h.kind == RTSIM_SENSOR_2D ? h.width : 0;
        const size_t nd = 2 * (size_t)h.height * h.width;
        if (!bearings && !data) return 1;
        std::vector<float> b(nb), d(nd);
        if (std::fread(b.data(), sizeof(float), nb, f) != nb ||
            std::fread(d.data(), sizeof(float), nd, f) != nd) {
            inner = fail(RTSIM_ERR_IO, "'%s': truncated record %llu", path, (unsigned long long)i);
            return -1;
        }
        uint64_t c = fnv1a(kFnvBasis, b.data(), sizeof(float) * nb);
        c = fnv1a(c, d.data(), sizeof(float) * nd);
        if (c != h.checksum) {
            inner = fail(RTSIM_ERR_IO, "'%s': record %llu fails its checksum", path,
                         (unsigned long long)i);
            return -1;
        }
        if (bearings && nb) std::memcpy(bearings, b.data(), sizeof(float) * nb);
        if (data) std::memcpy(data, d.data(), sizeof(float) * nd);
        return 1;
    });
    if (inner) return inner;
    if (rc) return rc;
    if (!found)
        return fail(RTSIM_ERR_INVALID, "'%s' has no record %llu", path, (unsigned long long)index);
    return RTSIM_OK;
} catch (const std::exception& e) {
    return fail(RTSIM_ERR_IO, "reading '%s': %s", path, e.what());
}

}  // extern "C"
