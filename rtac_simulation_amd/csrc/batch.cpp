// Batched pings: many sensor instances (mixed 1D / 2D, uneven sizes, either
// form) computed from one hipGraph launch.  The reference computes one ping
// per call with a device synchronisation after every kernel
// (SensorInstance2D.cu:100,117, Binner.cu:78,118,134); here the whole batch
// (optionally including each ping's streaming kernel) is captured once into a
// graph and replayed, and re-captured only when a sensor's configuration or
// sample count changes.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/rtsim.h"
#include "../../include/rtsim_batch.h"
#include "rtsim_internal.h"
#include "sensor.h"

namespace rtsim {

struct BatchEntry {
    Sensor* sensor;
    int form;
    bool has_hits = false;
    rtsim_hits hits{};
    size_t nhits = 0;
    uint64_t generation = 0;
};

struct Batch {
    std::vector<BatchEntry> entries;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;

    ~Batch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
    }
};

static int capture(Batch* b)
{
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    hipError_t e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    int rc = RTSIM_OK;
    for (auto& en : b->entries) {
        Sensor* s = en.sensor;
        if (en.has_hits && en.nhits) {
            StreamArgs a = s->stream_args(en.hits, en.nhits, 0);
            if ((e = launch_stream(a, b->capture_stream)) != hipSuccess) {
                rc = hip_fail(e, "stream kernel");
                break;
            }
        }
        if ((rc = s->compute(en.form, b->capture_stream))) break;
        en.generation = s->generation;
    }
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->capture_stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    b->graph = g;
    if ((e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0)) != hipSuccess)
        return hip_fail(e, "hipGraphInstantiate");
    b->dirty = false;
    return RTSIM_OK;
}

}  // namespace rtsim

using namespace rtsim;

extern "C" {

int rtsim_batch_create(const rtsim_sensor* sensors, const int* forms, uint32_t count,
                       rtsim_batch* out)
{
    if (!out || (!sensors && count)) return fail(RTSIM_ERR_INVALID, "sensors/out is NULL");
    *out = nullptr;
    Batch* b = new Batch();
    for (uint32_t i = 0; i < count; i++) {
        Sensor* s = reinterpret_cast<Sensor*>(sensors[i]);
        if (!s) { delete b; return fail(RTSIM_ERR_INVALID, "sensor %u is NULL", i); }
        int form = forms ? forms[i] : RTSIM_FORM_PLAIN_FP32;
        // resolve form-specific state now: no allocation may happen under capture
        int rc = s->prepare(form);
        if (rc) { delete b; return rc; }
        BatchEntry en;
        en.sensor = s;
        en.form = form;
        b->entries.push_back(en);
    }
    hipError_t e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete b; return hip_fail(e, "hipStreamCreate"); }
    *out = reinterpret_cast<rtsim_batch>(b);
    return RTSIM_OK;
}

int rtsim_batch_bind_hits(rtsim_batch batch, uint32_t index, const rtsim_hits* hits, size_t n)
{
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b || index >= b->entries.size()) return fail(RTSIM_ERR_INVALID, "bad batch/index");
    BatchEntry& en = b->entries[index];
    if (!hits) {
        en.has_hits = false;
    } else {
        if (!hits->range || !hits->dir_x || !hits->dir_y || !hits->dir_z || !hits->amplitude)
            return fail(RTSIM_ERR_INVALID, "hits has a NULL array");
        if (n != en.sensor->count)
            return fail(RTSIM_ERR_INVALID, "bound hits (%zu) must equal the sample count (%zu)", n,
                        en.sensor->count);
        en.has_hits = true;
        en.hits = *hits;
        en.nhits = n;
    }
    b->dirty = true;
    return RTSIM_OK;
}

int rtsim_batch_compute(rtsim_batch batch, void* stream)
{
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return fail(RTSIM_ERR_INVALID, "batch is NULL");
    for (auto& en : b->entries)
        if (en.generation != en.sensor->generation || en.sensor->profiling) b->dirty = true;
    for (auto& en : b->entries)
        if (en.sensor->profiling)
            return fail(RTSIM_ERR_INVALID, "turn sensor profiling off before batching");
    if (b->dirty) {
        // the capture stream must not run ahead of work already queued by the caller
        hipError_t e = hipStreamSynchronize((hipStream_t)stream);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        int rc = capture(b);
        if (rc) return rc;
    }
    hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    for (auto& en : b->entries) en.sensor->computed = true;
    return RTSIM_OK;
}

int rtsim_batch_destroy(rtsim_batch batch)
{
    delete reinterpret_cast<Batch*>(batch);
    return RTSIM_OK;
}

}  // extern "C"
