// Batched pings: many sensor instances (mixed 1D / 2D, uneven sizes, either
// form) computed together.  The reference computes one ping per call with a
// device synchronisation after every kernel (SensorInstance2D.cu:100,117,
// Binner.cu:78,118,134).  Here every stage of the path runs ONCE for the whole
// batch: one launch covers the workgroups of all pings (each workgroup finds
// its ping in a prefix table of workgroup counts and runs the single-ping
// kernel body with that ping's arguments), so a batch of 64 pings costs the
// ~20 launches of one ping instead of ~1000 small ones.  The launches are
// captured once into a hipGraph and replayed; the per-stage descriptor
// arrays and the graph are rebuilt only when a sensor's configuration,
// sample count or tile shape changes.
//
// Persistent mode (rtsim_batch_set_persistent): the grid-form pings run in
// ONE persistent launch instead (persist.h): every stage of every grid-form
// ping from one ticket-ordered task list, a ping's stages overlapping the
// other pings' stages; plain-form pings keep their per-stage launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/rtsim.h"
#include "../../include/rtsim_batch.h"
#include "grid_internal.h"
#include "persist.h"
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
    uint64_t pose_version = 0;
    int stream_index = -1;   // its entry in the streaming stage (-1: no bound hits)
    int persist_index = -1;  // its entry among the persistent launch's pings (-1: none)
};

// Host image of one batched launch: the arguments of the pings it covers and
// their workgroup prefix, placed in the batch's device buffer.
template <typename A>
struct Stage {
    std::vector<A> args;
    std::vector<uint32_t> first{0};
    size_t args_off = 0, first_off = 0;
    size_t shm = 0;
    int variant = 0;   // kernel instantiation (correlation: form and shape)

    void add(const A& a, uint32_t blocks)
    {
        if (!blocks) return;
        args.push_back(a);
        first.push_back(first.back() + blocks);
    }
    uint32_t blocks() const { return first.back(); }
    uint32_t nping() const { return (uint32_t)args.size(); }
};

struct Batch {
    std::vector<BatchEntry> entries;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;
    void* d_desc = nullptr;
    size_t desc_cap = 0;
    // recorded after every graph launch: the graph, its descriptors and the
    // pinned pose staging are only rewritten or freed once it has completed
    hipEvent_t done = nullptr;
    bool launched = false;
    StreamArgs* h_pose = nullptr;   // pinned staging of rewritten streaming arguments

    Stage<StreamArgs> stream;
    Stage<BucketArgs> stream_bucket;     // parallel to stream: the bucket counts made on the way
    // plain pings: the sort by travel, per pass
    Stage<HistArgs> hist[4];
    Stage<ScanArgs> scan[4];
    Stage<ScatterArgs> scatter[4], scatter_last[4];   // a ping's last pass: scatter_last
    // grid pings: the bucketing (counts, their scan, the scatter)
    Stage<BucketArgs> bucket_hist, bucket_scatter;
    Stage<ScanArgs> bucket_scan;
    Stage<BinArgs> windows;                // plain pings (the grid form's are in its deposit)
    Stage<ConvArgs> strip_count, strip_scan, strip_scatter;   // plain 2D pings' strip lists
    Stage<ConvArgs> plain_rank, plain2d, plain1d;
    Stage<DepositArgs> deposit;
    std::vector<Stage<CorrArgs>> corr;   // one per (form, tile shape, waves per workgroup)

    // persistent mode: the grid-form pings, their task segments, control words
    bool persistent = false;
    std::vector<PersistPing> pp;
    std::vector<uint32_t> seg_first, seg_ping;
    size_t pp_off = 0, segf_off = 0, segp_off = 0;
    uint32_t ntasks = 0, persist_grid = 0;
    size_t persist_shm = 0;
    uint64_t spin_limit = 0;
    uint32_t* d_ctl = nullptr;
    size_t ctl_cap = 0;   // words

    // waits for the last launch (a no-op when none is in flight)
    hipError_t wait()
    {
        if (!launched) return hipSuccess;
        launched = false;
        return hipEventSynchronize(done);
    }

    ~Batch()
    {
        (void)wait();
        if (done) (void)hipEventDestroy(done);
        if (h_pose) (void)hipHostFree(h_pose);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
        if (d_desc) (void)hipFree(d_desc);
        if (d_ctl) (void)hipFree(d_ctl);
    }
};

template <typename A>
static void layout(Stage<A>& st, size_t& off)
{
    off = (off + 15) & ~(size_t)15;
    st.args_off = off;
    off += sizeof(A) * st.args.size();
    off = (off + 15) & ~(size_t)15;
    st.first_off = off;
    off += sizeof(uint32_t) * st.first.size();
}

template <typename A>
static void pack(const Stage<A>& st, std::vector<char>& host)
{
    if (!st.args.empty()) std::memcpy(host.data() + st.args_off, st.args.data(), sizeof(A) * st.args.size());
    std::memcpy(host.data() + st.first_off, st.first.data(), sizeof(uint32_t) * st.first.size());
}

template <typename A>
static const A* dargs(const Batch* b, const Stage<A>& st)
{
    return reinterpret_cast<const A*>(static_cast<const char*>(b->d_desc) + st.args_off);
}

template <typename A>
static BatchMap dmap(const Batch* b, const Stage<A>& st)
{
    return BatchMap{reinterpret_cast<const uint32_t*>(static_cast<const char*>(b->d_desc) + st.first_off),
                    st.nping()};
}

// A grid-form ping of the persistent launch: its stage arguments and the
// task count of every level (persist.h); *nflags advances past its
// deposit-task flags.
static int persist_ping(Batch* b, BatchEntry& en, const PingPlan& p, uint32_t* nflags)
{
    Sensor* s = en.sensor;
    if (!p.corr.waves8)
        return fail(RTSIM_ERR_UNSUPPORTED,
                    "the persistent launch runs 8-wave correlation tiles (rtsim_sensor_set_tile_waves(8))");
    PersistPing q{};
    q.bk = p.bucket;
    q.dep = p.dep;
    // tiles of kPersistTile samples (fewer count columns; the deposit sums
    // exactly, so the grid does not depend on the tiling) unless the ping's
    // bucket counts already need larger ones
    if (q.bk.tile < kPersistTile) {
        q.bk.tile = kPersistTile;
        q.bk.tiles = (uint32_t)(((size_t)q.bk.n + kPersistTile - 1) / kPersistTile);
        q.dep.tiles = q.bk.tiles;
    }
    if (en.has_hits && en.nhits) {
        q.st = s->stream_args(en.hits, en.nhits, 0);
        q.ntask[kLvStream] = (uint32_t)((en.nhits + kPersistTile - 1) / kPersistTile);
        q.fuse = q.bk.tile == kPersistTile && q.bk.rec == q.st.out && q.bk.n == q.st.n ? 1 : 0;
    }
    q.ntask[kLvHist] = q.fuse ? 0u : q.bk.tiles;
    q.ntask[kLvScan] = (uint32_t)q.bk.NB;
    q.ntask[kLvScatter] = q.bk.tiles;
    q.staged = q.bk.tiles && bucket_scatter_staged(q.bk) ? 1 : 0;
    q.ntask[kLvDeposit] = deposit_blocks(q.dep);
    q.corr = p.corr;
    q.ntask[kLvCorr] = correlate_blocks(q.corr);
    q.variant = (en.form == RTSIM_FORM_MFMA_FP16 ? 4 : 0) + q.corr.shape;
    // a task waits for its ping's nearest earlier level that has tasks
    int last = -1;
    for (int lv = 0; lv < kLevels; lv++) {
        q.wait[lv] = last;
        if (q.ntask[lv]) last = lv;
    }
    q.flags = *nflags;
    *nflags += q.ntask[kLvDeposit];
    b->persist_shm = std::max(b->persist_shm, persist_lds_bytes(q));
    en.persist_index = (int)b->pp.size();
    if (en.has_hits && en.nhits) en.stream_index = en.persist_index;
    b->pp.push_back(q);
    s->last_corr = p.corr;
    return RTSIM_OK;
}

// The persistent launch's task list (level-major segments) and its place in
// the descriptor buffer.
static void persist_layout(Batch* b, size_t& off)
{
    b->seg_first.clear();
    b->seg_ping.clear();
    uint32_t total = 0;
    for (int lv = 0; lv < kLevels; lv++)
        for (size_t q = 0; q < b->pp.size(); q++) {
            const uint32_t n = b->pp[q].ntask[lv];
            if (!n) continue;
            b->seg_first.push_back(total);
            b->seg_ping.push_back((uint32_t)lv << 24 | (uint32_t)q);
            total += n;
        }
    b->seg_first.push_back(total);
    b->ntasks = total;
    off = (off + 255) & ~(size_t)255;
    b->pp_off = off;
    off += sizeof(PersistPing) * b->pp.size();
    off = (off + 15) & ~(size_t)15;
    b->segf_off = off;
    off += sizeof(uint32_t) * b->seg_first.size();
    off = (off + 15) & ~(size_t)15;
    b->segp_off = off;
    off += sizeof(uint32_t) * b->seg_ping.size();
}

// Per-stage descriptors of every ping (no device work is queued).
static int build(Batch* b)
{
    b->stream = {};
    b->stream_bucket = {};
    for (int k = 0; k < 4; k++) {
        b->hist[k] = {};
        b->scan[k] = {};
        b->scatter[k] = {};
        b->scatter_last[k] = {};
    }
    b->bucket_hist = {};
    b->bucket_scan = {};
    b->bucket_scatter = {};
    b->bucket_scatter.variant = 1;   // staged unless a ping has too many buckets for it
    b->windows = {};
    b->strip_count = {};
    b->strip_scan = {};
    b->strip_scatter = {};
    b->plain_rank = {};
    b->plain2d = {};
    b->plain2d.variant = 3;   // bit 0: taps in LDS, bit 1: 3072-sample chunks (unless a ping's do not fit)
    b->plain1d = {};
    b->deposit = {};
    b->corr.clear();
    b->pp.clear();
    b->seg_first.clear();
    b->seg_ping.clear();
    b->ntasks = 0;
    b->persist_shm = 0;
    std::map<std::tuple<int, int, int>, size_t> corr_index;
    uint32_t nflags = 0;
    size_t plain_taps = 0;   // the largest (A, B) tap plane of the batch's plain 2D pings
    for (auto& en : b->entries) {
        Sensor* s = en.sensor;
        PingPlan p;
        int rc = s->plan(en.form, &p);
        if (rc) return rc;
        en.stream_index = -1;
        en.persist_index = -1;
        if (p.grid && b->persistent) {
            int prc = persist_ping(b, en, p, &nflags);
            if (prc) return prc;
            continue;
        }
        bool counted = false;   // the grid form's bucket counts made by the streaming kernel
        if (en.has_hits && en.nhits) {
            StreamArgs a = s->stream_args(en.hits, en.nhits, 0);
            const uint32_t blocks = stream_blocks(a);
            if (blocks) en.stream_index = (int)b->stream.args.size();
            b->stream.add(a, blocks);
            BucketArgs fused{};
            if (p.grid && stream_bucket_fusable(a, p.bucket)) {
                fused = p.bucket;
                counted = true;
                b->stream_bucket.shm = std::max(b->stream_bucket.shm, sizeof(uint32_t) * (size_t)fused.NB);
            }
            b->stream_bucket.add(fused, blocks);
        }
        if (!p.grid) {
            for (int k = 0; k < p.sort.passes; k++) {
                b->hist[k].add(p.sort.hist[k], p.sort.hist_blocks[k]);
                b->scan[k].add(p.sort.scan[k], p.sort.pass_blocks[k] ? p.sort.hist[k].mask + 1 : 0u);
                (k == p.sort.passes - 1 ? b->scatter_last[k] : b->scatter[k])
                    .add(p.sort.scatter[k], p.sort.pass_blocks[k]);
            }
            b->windows.add(p.bins, windows_blocks(p.bins));
            if (s->kind == RTSIM_SENSOR_2D) {
                const bool lists = p.conv.strip_ent && p.conv.n;
                b->strip_count.add(p.conv, lists ? strip_blocks(p.conv.n) : 0u);
                b->strip_scan.add(p.conv, lists ? plain_strips((uint32_t)p.conv.W) : 0u);
                b->strip_scatter.add(p.conv, lists ? strip_blocks(p.conv.n) : 0u);
                const uint32_t blocks = conv2d_plain_blocks(p.conv);
                b->plain_rank.add(p.conv, blocks ? 1u : 0u);
                b->plain2d.add(p.conv, blocks);
                if (blocks) {
                    if (!p.conv.psf_in_lds) b->plain2d.variant &= ~1;
                    if (!p.conv.chunk3) b->plain2d.variant &= ~2;
                    plain_taps = std::max(plain_taps, sizeof(float4) * (size_t)(p.conv.Tr + 2) * (size_t)(p.conv.Tb + 2));
                }
            } else {
                b->plain1d.add(p.conv, conv1d_plain_blocks(p.conv));
            }
        } else {
            const BucketArgs& bk = p.bucket;
            if (!counted) {
                b->bucket_hist.add(bk, bk.tiles);
                b->bucket_hist.shm = std::max(b->bucket_hist.shm, sizeof(uint32_t) * (size_t)bk.NB);
            }
            b->bucket_scan.add(ScanArgs{bk.counts, bk.tiles, bk.totals, bk.tmax, bk.vmax}, (uint32_t)bk.NB);
            b->bucket_scatter.add(bk, bk.tiles);
            b->bucket_scatter.shm = std::max(b->bucket_scatter.shm, bucket_lds_bytes(bk));
            if (bk.tiles && !bucket_scatter_staged(bk)) b->bucket_scatter.variant = 0;
            b->deposit.add(p.dep, deposit_blocks(p.dep));
            b->deposit.shm = std::max(b->deposit.shm, deposit_lds_bytes(p.dep));
            const std::tuple<int, int, int> key{en.form, p.corr.shape, p.corr.waves8};
            auto it = corr_index.find(key);
            if (it == corr_index.end()) {
                it = corr_index.emplace(key, b->corr.size()).first;
                b->corr.emplace_back();
                b->corr.back().variant = en.form;
            }
            Stage<CorrArgs>& cs = b->corr[it->second];
            cs.add(p.corr, correlate_blocks(p.corr));
            cs.shm = std::max(cs.shm, correlate_lds_bytes(p.corr));
            s->last_corr = p.corr;
        }
        en.generation = s->generation;
        en.pose_version = s->pose_version;
    }
    b->plain2d.shm = plain2d_stage_bytes((b->plain2d.variant & 2) ? 3 : 2) +
                     ((b->plain2d.variant & 1) ? plain_taps : 0);
    // one device buffer for every descriptor array and prefix table
    size_t off = 0;
    layout(b->stream, off);
    layout(b->stream_bucket, off);
    for (int k = 0; k < 4; k++) {
        layout(b->hist[k], off);
        layout(b->scan[k], off);
        layout(b->scatter[k], off);
        layout(b->scatter_last[k], off);
    }
    layout(b->bucket_hist, off);
    layout(b->bucket_scan, off);
    layout(b->bucket_scatter, off);
    layout(b->windows, off);
    layout(b->strip_count, off);
    layout(b->strip_scan, off);
    layout(b->strip_scatter, off);
    layout(b->plain_rank, off);
    layout(b->plain2d, off);
    layout(b->plain1d, off);
    layout(b->deposit, off);
    for (auto& c : b->corr) layout(c, off);
    if (!b->pp.empty()) persist_layout(b, off);
    if (off > b->desc_cap) {
        if (b->d_desc) (void)hipFree(b->d_desc);
        b->d_desc = nullptr;
        b->desc_cap = 0;
        hipError_t e = hipMalloc(&b->d_desc, off);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(batch descriptors)");
        b->desc_cap = off;
    }
    std::vector<char> host(off, 0);
    pack(b->stream, host);
    pack(b->stream_bucket, host);
    for (int k = 0; k < 4; k++) {
        pack(b->hist[k], host);
        pack(b->scan[k], host);
        pack(b->scatter[k], host);
        pack(b->scatter_last[k], host);
    }
    pack(b->bucket_hist, host);
    pack(b->bucket_scan, host);
    pack(b->bucket_scatter, host);
    pack(b->windows, host);
    pack(b->strip_count, host);
    pack(b->strip_scan, host);
    pack(b->strip_scatter, host);
    pack(b->plain_rank, host);
    pack(b->plain2d, host);
    pack(b->plain1d, host);
    pack(b->deposit, host);
    for (auto& c : b->corr) pack(c, host);
    if (!b->pp.empty()) {
        std::memcpy(host.data() + b->pp_off, b->pp.data(), sizeof(PersistPing) * b->pp.size());
        std::memcpy(host.data() + b->segf_off, b->seg_first.data(), sizeof(uint32_t) * b->seg_first.size());
        std::memcpy(host.data() + b->segp_off, b->seg_ping.data(), sizeof(uint32_t) * b->seg_ping.size());
    }
    hipError_t e = hipMemcpy(b->d_desc, host.data(), off, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(batch descriptors)");
    if (!b->pp.empty()) {
        // control words: zero, sized for this build's counters and flags
        const size_t words = (size_t)kCtlWords + b->pp.size() * kLevels * kCtrWords + nflags;
        if (words > b->ctl_cap) {
            if (b->d_ctl) (void)hipFree(b->d_ctl);
            b->d_ctl = nullptr;
            b->ctl_cap = 0;
            if ((e = hipMalloc((void**)&b->d_ctl, sizeof(uint32_t) * words)) != hipSuccess)
                return hip_fail(e, "hipMalloc(persistent control words)");
            b->ctl_cap = words;
        }
        if ((e = hipMemset(b->d_ctl, 0, sizeof(uint32_t) * b->ctl_cap)) != hipSuccess)
            return hip_fail(e, "hipMemset(persistent control words)");
        b->persist_grid = std::min(persist_blocks(b->persist_shm), b->ntasks);
        if (!b->persist_grid) return fail(RTSIM_ERR_HIP, "persistent launch: no workgroup fits");
        b->spin_limit = 2 * persist_clock_hz();   // two seconds
    }
    return RTSIM_OK;
}

// The batch's launches, stage after stage (each covers every ping).
static hipError_t enqueue(const Batch* b, hipStream_t s)
{
    hipError_t e;
#define RT_TRY(x) if ((e = (x)) != hipSuccess) return e
    RT_TRY(launch_stream_batch(dargs(b, b->stream), dmap(b, b->stream), b->stream.blocks(),
                               dargs(b, b->stream_bucket), b->stream_bucket.shm, s));
    for (int k = 0; k < 4; k++) {
        RT_TRY(launch_sort_hist_batch(k == 0, dargs(b, b->hist[k]), dmap(b, b->hist[k]), b->hist[k].blocks(),
                                      s));
        RT_TRY(launch_sort_scan_batch(dargs(b, b->scan[k]), dmap(b, b->scan[k]), b->scan[k].blocks(), s));
        RT_TRY(launch_sort_scatter_batch(k == 0, false, dargs(b, b->scatter[k]), dmap(b, b->scatter[k]),
                                         b->scatter[k].blocks(), s));
        RT_TRY(launch_sort_scatter_batch(k == 0, true, dargs(b, b->scatter_last[k]), dmap(b, b->scatter_last[k]),
                                         b->scatter_last[k].blocks(), s));
    }
    RT_TRY(launch_bucket_hist_batch(dargs(b, b->bucket_hist), dmap(b, b->bucket_hist), b->bucket_hist.blocks(),
                                    b->bucket_hist.shm, s));
    RT_TRY(launch_sort_scan_batch(dargs(b, b->bucket_scan), dmap(b, b->bucket_scan), b->bucket_scan.blocks(), s));
    RT_TRY(launch_bucket_scatter_batch(dargs(b, b->bucket_scatter), dmap(b, b->bucket_scatter),
                                       b->bucket_scatter.blocks(), b->bucket_scatter.shm,
                                       b->bucket_scatter.variant != 0, s));
    RT_TRY(launch_windows_batch(dargs(b, b->windows), dmap(b, b->windows), b->windows.blocks(), s));
    RT_TRY(launch_strip_count_batch(dargs(b, b->strip_count), dmap(b, b->strip_count), b->strip_count.blocks(), s));
    RT_TRY(launch_strip_scan_batch(dargs(b, b->strip_scan), dmap(b, b->strip_scan), b->strip_scan.blocks(), s));
    RT_TRY(launch_strip_scatter_batch(dargs(b, b->strip_scatter), dmap(b, b->strip_scatter),
                                      b->strip_scatter.blocks(), s));
    RT_TRY(launch_plain_rank_batch(dargs(b, b->plain_rank), dmap(b, b->plain_rank), b->plain_rank.blocks(), s));
    RT_TRY(launch_conv2d_plain_batch(dargs(b, b->plain2d), dmap(b, b->plain2d), b->plain2d.blocks(),
                                     b->plain2d.shm, b->plain2d.variant, s));
    RT_TRY(launch_conv1d_plain_batch(dargs(b, b->plain1d), dmap(b, b->plain1d), b->plain1d.blocks(), s));
    RT_TRY(launch_deposit_batch(dargs(b, b->deposit), dmap(b, b->deposit), b->deposit.blocks(),
                                b->deposit.shm, s));
    for (const auto& c : b->corr)
        RT_TRY(launch_correlate_batch(c.variant == RTSIM_FORM_MFMA_FP16, c.args.empty() ? 0 : c.args[0].shape,
                                      !c.args.empty() && c.args[0].waves8, dargs(b, c), dmap(b, c),
                                      c.blocks(), c.shm, s));
    if (b->ntasks) {
        const char* base = static_cast<const char*>(b->d_desc);
        PersistArgs a{};
        a.pings = reinterpret_cast<const PersistPing*>(base + b->pp_off);
        a.seg_first = reinterpret_cast<const uint32_t*>(base + b->segf_off);
        a.seg_ping = reinterpret_cast<const uint32_t*>(base + b->segp_off);
        a.nseg = (uint32_t)b->seg_ping.size();
        a.ntasks = b->ntasks;
        a.nping = (uint32_t)b->pp.size();
        a.ctl = b->d_ctl;
        a.spin_limit = b->spin_limit;
        RT_TRY(launch_persist(a, b->persist_grid, b->persist_shm, s));
    }
#undef RT_TRY
    return hipSuccess;
}

static int capture(Batch* b)
{
    hipError_t we = b->wait();   // the old graph and descriptors may be in flight
    if (we != hipSuccess) return hip_fail(we, "waiting for the previous batch launch");
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    int rc = build(b);
    if (rc) return rc;
    (void)hipGetLastError();
    hipError_t e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    const hipError_t le = enqueue(b, b->capture_stream);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->capture_stream, &g);
    if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(le, "batched launch");
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
        for (const auto& en : b->entries)
            if (en.sensor == s) {
                delete b;
                return fail(RTSIM_ERR_INVALID, "sensor %u appears twice in the batch", i);
            }
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
    e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
    if (e != hipSuccess) { delete b; return hip_fail(e, "hipEventCreate"); }
    if (count) {
        e = hipHostMalloc((void**)&b->h_pose, sizeof(StreamArgs) * count, hipHostMallocDefault);
        if (e != hipSuccess) { delete b; return hip_fail(e, "hipHostMalloc"); }
    }
    *out = reinterpret_cast<rtsim_batch>(b);
    return RTSIM_OK;
}

int rtsim_batch_bind_hits(rtsim_batch batch, uint32_t index, const rtsim_hits* hits, size_t n)
{
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b || index >= b->entries.size()) return fail(RTSIM_ERR_INVALID, "bad batch/index");
    BatchEntry& en = b->entries[index];
    if (!hits) {
        if (en.has_hits) b->dirty = true;
        en.has_hits = false;
    } else {
        if (!hits->range || !hits->dir_x || !hits->dir_y || !hits->dir_z || !hits->amplitude)
            return fail(RTSIM_ERR_INVALID, "hits has a NULL array");
        if (n != en.sensor->count)
            return fail(RTSIM_ERR_INVALID, "bound hits (%zu) must equal the sample count (%zu)", n,
                        en.sensor->count);
        // the same arrays again: the captured graph already reads them
        if (!en.has_hits || en.nhits != n || std::memcmp(&en.hits, hits, sizeof(rtsim_hits)) != 0)
            b->dirty = true;
        en.has_hits = true;
        en.hits = *hits;
        en.nhits = n;
    }
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
    hipStream_t st = (hipStream_t)stream;
    if (b->dirty) {
        // the capture stream must not run ahead of work already queued by the caller
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        int rc = capture(b);
        if (rc) return rc;
    } else {
        // new receiver poses: rewrite those pings' streaming arguments in the
        // descriptor buffer, ordered before this launch on its stream
        bool waited = false;
        for (auto& en : b->entries) {
            if (en.pose_version == en.sensor->pose_version) continue;
            en.pose_version = en.sensor->pose_version;
            if (en.stream_index < 0) continue;
            if (!waited) {   // the staging may still feed the previous launch's copies
                hipError_t e = b->wait();
                if (e != hipSuccess) return hip_fail(e, "waiting for the previous batch launch");
                waited = true;
            }
            const bool pers = en.persist_index >= 0;
            StreamArgs& a = pers ? b->pp[en.persist_index].st : b->stream.args[en.stream_index];
            std::memcpy(a.R, en.sensor->R, sizeof(a.R));
            // staging slot: one per entry (stream indices and persistent
            // indices can coincide, entry indices cannot)
            const size_t slot = (size_t)(&en - b->entries.data());
            b->h_pose[slot] = a;
            char* dst = static_cast<char*>(b->d_desc) +
                        (pers ? b->pp_off + sizeof(PersistPing) * (size_t)en.persist_index + offsetof(PersistPing, st)
                              : b->stream.args_off + sizeof(StreamArgs) * (size_t)en.stream_index);
            hipError_t e = hipMemcpyAsync(dst, &b->h_pose[slot], sizeof(StreamArgs),
                                          hipMemcpyHostToDevice, st);
            if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(pose)");
        }
    }
    hipError_t e = hipGraphLaunch(b->exec, st);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    e = hipEventRecord(b->done, st);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    b->launched = true;
    for (auto& en : b->entries) {
        en.sensor->computed = true;
        en.sensor->hist_fused = false;   // the batch's own sort rewrote the counts
        en.sensor->computed_form = en.form;
    }
    return RTSIM_OK;
}

// Mean milliseconds of one replay of the batch's graph (captured first when
// out of date).
int rtsim_batch_set_persistent(rtsim_batch batch, int on);
static int time_batch(Batch* b, int iterations, hipStream_t st, hipEvent_t e0, hipEvent_t e1, float* ms)
{
    int rc = rtsim_batch_compute(reinterpret_cast<rtsim_batch>(b), st);   // capture + warm
    if (rc) return rc;
    hipError_t e = hipEventRecord(e0, st);
    for (int it = 0; it < iterations && e == hipSuccess; it++) e = hipGraphLaunch(b->exec, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventRecord(b->done, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    if (e != hipSuccess) return hip_fail(e, "batch autotune");
    *ms /= (float)iterations;
    return RTSIM_OK;
}

// MFMA-form sensors of the batch by class (grid and PSF size)
using SensorClasses = std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t>, std::vector<Sensor*>>;

// Per class in turn (the others held at their current choice), times every
// tile shape and wave count the launch mode allows and keeps the fastest;
// *bms holds the batch's time with the current choice on entry and the best
// one on return.
static int tune_tiles(Batch* b, SensorClasses& classes, int iterations, hipStream_t st, hipEvent_t e0,
                      hipEvent_t e1, float* bms)
{
    int rc = RTSIM_OK;
    for (auto it = classes.begin(); rc == RTSIM_OK && it != classes.end(); ++it) {
        std::vector<Sensor*>& ss = it->second;
        int best_shape = 0, best_waves = 4;
        (void)rtsim_sensor_tile_waves(reinterpret_cast<rtsim_sensor>(ss[0]), &best_waves);
        int r0 = 0, c0 = 0;
        (void)rtsim_sensor_tile_shape(reinterpret_cast<rtsim_sensor>(ss[0]), &best_shape, &r0, &c0);
        auto apply = [&](int shape, int waves) {
            for (Sensor* s : ss) {
                const rtsim_sensor h = reinterpret_cast<rtsim_sensor>(s);
                if (rtsim_sensor_set_tile_shape(h, shape) || rtsim_sensor_set_tile_waves(h, waves))
                    return false;
            }
            return true;
        };
        for (int v = 0; v < 2 * kCorrShapes && rc == RTSIM_OK; v++) {
            const int shape = v % kCorrShapes, waves = v < kCorrShapes ? 4 : 8;
            if (shape == best_shape && waves == best_waves) continue;
            if (b->persistent && waves != 8) continue;   // the persistent launch runs 8-wave tiles
            float ms = 0.f;
            if (!apply(shape, waves)) continue;   // does not fit some sensor's LDS budget
            if ((rc = time_batch(b, iterations, st, e0, e1, &ms))) break;
            if (ms < *bms) {
                *bms = ms;
                best_shape = shape;
                best_waves = waves;
            }
        }
        if (!apply(best_shape, best_waves) && rc == RTSIM_OK)
            rc = fail(RTSIM_ERR_HIP, "batch autotune could not restore a tile shape");
    }
    return rc;
}

int rtsim_batch_autotune(rtsim_batch batch, int iterations, void* stream, float* best_ms)
{
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return fail(RTSIM_ERR_INVALID, "batch is NULL");
    if (iterations < 1) iterations = 1;
    hipStream_t st = (hipStream_t)stream;
    SensorClasses classes;
    std::vector<std::pair<int, int>> tiles;   // every MFMA-form entry's (shape, waves), in entry order
    for (auto& en : b->entries)
        if (en.form == RTSIM_FORM_MFMA_BF16 || en.form == RTSIM_FORM_MFMA_FP16) {
            Sensor* s = en.sensor;
            classes[std::make_tuple((uint32_t)s->N, (uint32_t)s->W, (uint32_t)s->Tr, (uint32_t)s->Tb)]
                .push_back(s);
        }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return fail(RTSIM_ERR_HIP, "hipEventCreate failed");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(RTSIM_ERR_HIP, "hipEventCreate failed");
    }
    float bms = 0.f;
    int rc = time_batch(b, iterations, st, e0, e1, &bms);
    if (rc == RTSIM_OK) rc = tune_tiles(b, classes, iterations, st, e0, e1, &bms);
    // the other launch mode of the MFMA-form pings, tuned the same way; the
    // faster mode and its tiles are kept
    if (rc == RTSIM_OK && !classes.empty()) {
        const bool was = b->persistent;
        auto save = [&]() {
            tiles.clear();
            for (auto& en : b->entries) {
                if (en.form != RTSIM_FORM_MFMA_BF16 && en.form != RTSIM_FORM_MFMA_FP16) continue;
                const rtsim_sensor h = reinterpret_cast<rtsim_sensor>(en.sensor);
                int shape = 0, waves = 4, r = 0, c = 0;
                (void)rtsim_sensor_tile_shape(h, &shape, &r, &c);
                (void)rtsim_sensor_tile_waves(h, &waves);
                tiles.emplace_back(shape, waves);
            }
        };
        auto restore = [&]() {
            size_t k = 0;
            for (auto& en : b->entries) {
                if (en.form != RTSIM_FORM_MFMA_BF16 && en.form != RTSIM_FORM_MFMA_FP16) continue;
                const rtsim_sensor h = reinterpret_cast<rtsim_sensor>(en.sensor);
                if (rtsim_sensor_set_tile_shape(h, tiles[k].first) || rtsim_sensor_set_tile_waves(h, tiles[k].second))
                    return fail(RTSIM_ERR_HIP, "batch autotune could not restore a tile shape");
                k++;
            }
            return (int)RTSIM_OK;
        };
        save();
        float oms = 0.f;
        if (rtsim_batch_set_persistent(batch, was ? 0 : 1) == RTSIM_OK) {
            rc = time_batch(b, iterations, st, e0, e1, &oms);
            if (rc == RTSIM_OK) rc = tune_tiles(b, classes, iterations, st, e0, e1, &oms);
            if (rc == RTSIM_OK) {
                if (oms < bms) {
                    bms = oms;
                } else {
                    rc = rtsim_batch_set_persistent(batch, was ? 1 : 0);
                    if (rc == RTSIM_OK) rc = restore();
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (best_ms) *best_ms = bms;
    // leave every ping consistent with the chosen mode and shapes
    return rtsim_batch_compute(batch, stream);
}

int rtsim_batch_set_persistent(rtsim_batch batch, int on)
{
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return fail(RTSIM_ERR_INVALID, "batch is NULL");
    if (on) {
        // 8-wave correlation tiles for every grid-form ping: the current
        // shape when it allows 8 waves, else the first shape that does
        for (auto& en : b->entries) {
            if (en.form != RTSIM_FORM_MFMA_BF16 && en.form != RTSIM_FORM_MFMA_FP16) continue;
            const rtsim_sensor h = reinterpret_cast<rtsim_sensor>(en.sensor);
            if (rtsim_sensor_set_tile_waves(h, 8) == RTSIM_OK) continue;
            bool ok = false;
            for (int shape = 0; shape < kCorrShapes && !ok; shape++)
                ok = rtsim_sensor_set_tile_shape(h, shape) == RTSIM_OK && rtsim_sensor_set_tile_waves(h, 8) == RTSIM_OK;
            if (!ok)
                return fail(RTSIM_ERR_UNSUPPORTED, "no 8-wave correlation tile fits sensor %zu's LDS budget",
                            (size_t)(&en - b->entries.data()));
        }
    }
    if (b->persistent != (on != 0)) b->dirty = true;
    b->persistent = on != 0;
    return RTSIM_OK;
}

int rtsim_batch_persistent(rtsim_batch batch, int* on)
{
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b || !on) return fail(RTSIM_ERR_INVALID, "batch/on is NULL");
    *on = b->persistent ? 1 : 0;
    return RTSIM_OK;
}

int rtsim_batch_check(rtsim_batch batch)
{
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return fail(RTSIM_ERR_INVALID, "batch is NULL");
    hipError_t e = b->wait();
    if (e != hipSuccess) return hip_fail(e, "waiting for the batch");
    if (!b->d_ctl) return RTSIM_OK;
    uint32_t status = 0;
    if ((e = hipMemcpy(&status, b->d_ctl + kCtlStatus, sizeof(status), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "reading the persistent launch's status");
    if (!status) return RTSIM_OK;
    // a timed-out launch left its counters as they were: start the next one clean
    if ((e = hipMemset(b->d_ctl, 0, sizeof(uint32_t) * b->ctl_cap)) != hipSuccess)
        return hip_fail(e, "clearing the persistent control words");
    return fail(RTSIM_ERR_HIP, "persistent launch: a task waited longer than its limit and the launch was abandoned");
}

int rtsim_batch_destroy(rtsim_batch batch)
{
    delete reinterpret_cast<Batch*>(batch);
    return RTSIM_OK;
}

}  // extern "C"
