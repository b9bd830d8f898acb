// rtsim_sensor_compute_noalloc: rtsim_sensor_compute for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  The records come from the caller and the
// ping goes to the caller; every per-ping buffer the kernels need beyond them
// lives in one file-scope device array of fixed size: for the plain form the
// sort's key / index pairs, tile counts and digit totals, the sorted records
// and their coordinates; for the grid form the bucketed items, the bucket
// counts, the key occupancy and the tile maxima (the two sets overlap: a call
// uses one of them); for both the Binner segments and windows, and the plain
// 2D gather's band order.  Each GPU holds
// its own copy of that array until the process ends; its address on a device
// is looked up once, on that device's first call, and every later call on the
// device reuses it.  The sensor's own configuration tables (PSF, coordinates)
// and, for the MFMA forms, its prepared grid state are read as set up
// beforehand; nothing here allocates.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/rtsim.h"
#include "rtsim_internal.h"
#include "sensor.h"

namespace rtsim {

namespace {

constexpr size_t kMaxS = RTSIM_NOALLOC_MAX_SAMPLES;
constexpr size_t kMaxN = RTSIM_NOALLOC_MAX_BINS;
constexpr size_t kTilesCap = (kMaxS + kSortTileSmall - 1) / kSortTileSmall;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// scratch layout (bytes), every buffer 256-byte aligned; the plain-form and
// grid-form sets share the space from offset 0
constexpr size_t kOffSorted = 0;
constexpr size_t kOffCoord = kOffSorted + align256(sizeof(float4) * kMaxS);
constexpr size_t kOffKv0 = kOffCoord + align256(sizeof(float2) * kMaxS);
constexpr size_t kOffKv1 = kOffKv0 + align256(sizeof(uint2) * kMaxS);
constexpr size_t kOffCounts = kOffKv1 + align256(sizeof(uint2) * kMaxS);
constexpr size_t kOffTotals = kOffCounts + align256(sizeof(uint32_t) * 256 * kTilesCap);
constexpr size_t kEndPlain = kOffTotals + align256(sizeof(uint32_t) * 256);

constexpr size_t kOffItems = 0;
constexpr size_t kBcountWords = kBucketCountBudget + (size_t)kMaxBuckets;
constexpr size_t kOffBcounts = kOffItems + align256(sizeof(float4) * 2 * kMaxS);
constexpr size_t kOccWords = kTilesCap * (kMaxN / 32);
constexpr size_t kOffOcc = kOffBcounts + align256(sizeof(uint32_t) * kBcountWords);
constexpr size_t kOffTmax = kOffOcc + align256(sizeof(uint32_t) * kOccWords);
constexpr size_t kEndGrid = kOffTmax + align256(sizeof(uint32_t) * kTilesCap);

constexpr size_t kOffSeg = kEndPlain > kEndGrid ? kEndPlain : kEndGrid;
constexpr size_t kOffWin = kOffSeg + align256(sizeof(uint2) * kMaxN);
constexpr size_t kOffWork = kOffWin + align256(sizeof(uint2) * kMaxN);
constexpr size_t kScratchBytes = kOffWork + align256(sizeof(uint32_t) * plain2d_work_words(kMaxN));
static_assert(kScratchBytes == RTSIM_NOALLOC_SCRATCH_BYTES, "scratch size in rtsim.h");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch[kMaxDevices] = {};

}  // namespace

// The array itself.  In host code this name is only a handle for
// hipGetSymbolAddress: kernels receive the per-device address found there.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// This device's address of the scratch array (looked up on its first call).
static int scratch_address(unsigned char** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(RTSIM_ERR_INVALID, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(RTSIM_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return RTSIM_OK;
}

}  // namespace rtsim

using namespace rtsim;

extern "C" int rtsim_sensor_compute_noalloc(rtsim_sensor h, int form, const void* records, size_t n,
                                            void* ping, void* stream)
{
    Sensor* s = reinterpret_cast<Sensor*>(h);
    if (!s) return fail(RTSIM_ERR_INVALID, "sensor handle is NULL");
    const bool grid = form == RTSIM_FORM_MFMA_BF16 || form == RTSIM_FORM_MFMA_FP16;
    if (form != RTSIM_FORM_PLAIN_FP32 && !grid) return fail(RTSIM_ERR_INVALID, "unknown form %d", form);
    if (n > kMaxS)
        return fail(RTSIM_ERR_INVALID, "%zu samples exceed the no-allocation limit of %zu", n, kMaxS);
    if (s->N > kMaxN)
        return fail(RTSIM_ERR_INVALID, "%u range bins exceed the no-allocation limit of %zu", s->N, kMaxN);
    if ((n && !records) || !ping) return fail(RTSIM_ERR_INVALID, "records/ping is NULL");
    // preparing the grid form allocates its grid: it must have happened before
    if (grid && !s->grid_ready())
        return fail(RTSIM_ERR_INVALID,
                    "the MFMA form is not prepared on this sensor (compute it once, or set a tile "
                    "shape, before calling without allocation)");
    unsigned char* base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    PingBuffers pb{};
    pb.samples = static_cast<const float4*>(records);
    pb.n = (uint32_t)n;
    pb.tiles_cap = (uint32_t)kTilesCap;
    pb.sorted = reinterpret_cast<float4*>(base + kOffSorted);
    pb.coord = reinterpret_cast<float2*>(base + kOffCoord);
    pb.kv[0] = reinterpret_cast<uint2*>(base + kOffKv0);
    pb.kv[1] = reinterpret_cast<uint2*>(base + kOffKv1);
    pb.counts = reinterpret_cast<uint32_t*>(base + kOffCounts);
    pb.totals = reinterpret_cast<uint32_t*>(base + kOffTotals);
    pb.items = reinterpret_cast<float4*>(base + kOffItems);
    pb.items_cap = 2 * kMaxS;
    pb.bcounts = reinterpret_cast<uint32_t*>(base + kOffBcounts);
    pb.bcounts_cap = kBcountWords;
    pb.occ = reinterpret_cast<uint32_t*>(base + kOffOcc);
    pb.occ_cap = kOccWords;
    pb.tmax = reinterpret_cast<uint32_t*>(base + kOffTmax);
    pb.seg = reinterpret_cast<uint2*>(base + kOffSeg);
    pb.win = reinterpret_cast<uint2*>(base + kOffWin);
    pb.work = reinterpret_cast<uint32_t*>(base + kOffWork);
    pb.out = static_cast<float2*>(ping);
    (void)hipGetLastError();
    PingPlan p;
    if ((rc = s->plan(form, &p, &pb))) return rc;
    // every array the ping reads is written first by its own kernels (the
    // sort's first histogram clears the segments; the bucketing writes every
    // count, occupancy word and item it reads; the band ordering writes the
    // plain gather's work words): nothing a previous call left in the
    // scratch is read
    return s->launch_ping(p, form, (hipStream_t)stream, false, false);
}
