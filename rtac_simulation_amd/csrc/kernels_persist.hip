// Persistent launch of grid-form pings (persist.h): one kernel runs every
// stage of every grid-form ping of a batch from an atomic ticket, the same
// stage bodies as the per-stage launches (stream_body.h, bucket_body.h,
// deposit_body.h, corr_body.h) at 512 threads per workgroup.
//
// Ordering between workgroups: a task's results are published by its
// workgroup's barrier followed by one agent-scope release (the level counter
// or the deposit-task flag); a waiting workgroup polls with relaxed
// agent-scope loads, then takes one agent-scope acquire before its barrier.
// A workgroup remembers the last dependency it acquired and does not wait or
// acquire again for the next task with the same one.
#include <algorithm>

#include "bucket_body.h"
#include "corr_body.h"
#include "deposit_body.h"
#include "device_common.h"
#include "grid_internal.h"
#include "persist.h"
#include "stream_body.h"

namespace rtsim {

namespace {

constexpr int NT = kPersistThreads;
// samples per thread with their loads in flight together: with 8 waves per CU
// (the correlation's registers) the memory-bound stages need more of them
// than their own launches (16-32 waves per CU) do
constexpr int kHistU = 8;                                     // counts: a pass of 4096
constexpr int kScatterU = 4;                                  // direct scatter: a pass of 2048
constexpr int kStagedU = kScatterPass / NT;                   // staged scatter: a pass of 1024
constexpr int kDepositPU = 8;                                 // deposit items
constexpr uint32_t kSegLds = 512;                             // task segments kept in LDS

__device__ __forceinline__ uint32_t ld_agent(uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// true when the wait exceeded the limit or another workgroup aborted (one thread)
__device__ __forceinline__ bool give_up(uint32_t* ctl, uint64_t t0, uint64_t limit)
{
    if (ld_agent(ctl + kCtlAbort)) return true;
    if (wall_clock64() - t0 > limit) {
        __hip_atomic_fetch_or(ctl + kCtlStatus, kPersistTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_or(ctl + kCtlAbort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return true;
    }
    __builtin_amdgcn_s_sleep(2);
    return false;
}

// The XCD this workgroup runs on (hardware register XCC_ID, bits 3:0): a
// level's finished tasks are counted in one slot per XCD, on separate cache
// lines, so that the increments of workgroups on different XCDs do not
// contend for one address.
__device__ __forceinline__ uint32_t xcd_slot() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & (kCtrSlots - 1); }

// one wave: waits until the slots of counter c sum to >= target; false on
// abort or time-out
__device__ __forceinline__ bool wait_count(uint32_t* c, uint32_t target, uint32_t* ctl, uint64_t limit)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t t0 = wall_clock64();
    for (;;) {
        uint32_t v = lane < kCtrSlots ? ld_agent(c + lane * kCtrStride) : 0u;
        for (int o = kCtrSlots / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (__builtin_amdgcn_readfirstlane(v) >= target) break;
        int quit = 0;
        if (lane == 0) quit = give_up(ctl, t0, limit) ? 1 : 0;
        if (__builtin_amdgcn_readfirstlane(quit)) return false;
    }
    acquire_agent();
    return true;
}

// one wave: waits until flags[f0, f1) all hold `want`; false on abort or time-out
__device__ __forceinline__ bool wait_flags(uint32_t* flags, uint32_t f0, uint32_t f1, uint32_t want, uint32_t* ctl,
                                           uint64_t limit)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t t0 = wall_clock64();
    for (;;) {
        bool mine = true;
        for (uint32_t f = f0 + lane; f < f1; f += 64) mine = mine && ld_agent(flags + f) == want;
        if (__all(mine)) break;
        int quit = 0;
        if (lane == 0) quit = give_up(ctl, t0, limit) ? 1 : 0;
        if (__builtin_amdgcn_readfirstlane(quit)) return false;
    }
    acquire_agent();
    return true;
}

// Exclusive scan of one row of bucket counts over the tiles (staged in LDS
// srow), its total, and for row 0 the largest tile maximum (the deposit's
// fixed-point scale): what k_sort_scan computes for a bucket row.
__device__ __forceinline__ void scan_row(const BucketArgs& b, uint32_t row, uint32_t* srow)
{
    __shared__ uint32_t ws[NT / 64];
    __shared__ float wm[NT / 64];
    const uint32_t t = threadIdx.x;
    const uint32_t tiles = b.tiles;
    uint32_t* grow = b.counts + (size_t)row * tiles;
    for (uint32_t i = t; i < tiles; i += NT) srow[i] = grow[i];
    __syncthreads();
    const uint32_t per = (tiles + NT - 1) / NT;
    const uint32_t i0 = min(t * per, tiles), i1 = min(i0 + per, tiles);
    uint32_t s = 0;
    for (uint32_t i = i0; i < i1; i++) s += srow[i];
    uint32_t total;
    uint32_t run = sort_block_scan<NT>(s, ws, &total);
    for (uint32_t i = i0; i < i1; i++) {
        const uint32_t v = srow[i];
        srow[i] = run;
        run += v;
    }
    __syncthreads();
    for (uint32_t i = t; i < tiles; i += NT) grow[i] = srow[i];
    if (t == 0) b.totals[row] = total;
    if (row == 0) {
        float m = 0.f;
        for (uint32_t i = t; i < tiles; i += NT) m = fmaxf(m, __uint_as_float(b.tmax[i]));
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        if ((t & 63) == 0) wm[t >> 6] = m;
        __syncthreads();
        if (t == 0) {
            float mm = wm[0];
            for (int k = 1; k < NT / 64; k++) mm = fmaxf(mm, wm[k]);
            b.vmax[0] = __float_as_uint(mm);
        }
    }
}

// correlation tile of variant v = 4 fp16 + shape (8 waves)
__device__ __forceinline__ void correlate_variant(const CorrArgs& c, int v, int tile)
{
    switch (v) {
    case 0: correlate_body<false, 1, 1, true>(c, tile, (c.W + 31) / 32); break;
    case 1: correlate_body<false, 1, 2, true>(c, tile, (c.W + 63) / 64); break;
    case 2: correlate_body<false, 2, 1, true>(c, tile, (c.W + 31) / 32); break;
    case 3: correlate_body<false, 2, 2, true>(c, tile, (c.W + 63) / 64); break;
    case 4: correlate_body<true, 1, 1, true>(c, tile, (c.W + 31) / 32); break;
    case 5: correlate_body<true, 1, 2, true>(c, tile, (c.W + 63) / 64); break;
    case 6: correlate_body<true, 2, 1, true>(c, tile, (c.W + 31) / 32); break;
    default: correlate_body<true, 2, 2, true>(c, tile, (c.W + 63) / 64); break;
    }
}

// the deposit-task flags a correlation tile reads: the buckets of its staged
// grid rows, then the window groups of its output rows
__device__ __forceinline__ void corr_needs(const PersistPing& p, int tile, uint32_t& b0, uint32_t& b1,
                                           uint32_t& g0, uint32_t& g1)
{
    const CorrArgs& c = p.corr;
    const int TR = (p.variant & 3) >= 2 ? 2 : 1, TC = (p.variant & 1) ? 2 : 1;
    const int gx = (c.W + 32 * TC - 1) / (32 * TC);
    const int r0 = (tile / gx) * 32 * TR;
    const int SR = 32 * (TR - 1) + 16 * c.KR;
    const int R = p.dep.R;
    b0 = (uint32_t)(r0 / R);
    b1 = (uint32_t)((min(r0 + SR, c.GH) - 1) / R) + 1;
    g0 = (uint32_t)(r0 / kWindowBins);
    g1 = (uint32_t)((min(r0 + 32 * TR, c.H) - 1) / kWindowBins) + 1;
}

}  // namespace

// The descriptor arrays are separate restrict arguments and the task index
// is made wave-uniform, so the pings' arguments are read with scalar loads.
__global__ __launch_bounds__(NT) void k_persist(const PersistPing* __restrict__ pings,
                                                const uint32_t* __restrict__ seg_first,
                                                const uint32_t* __restrict__ seg_ping, PersistArgs a)
{
    extern __shared__ __align__(16) uint32_t plds[];
    __shared__ uint32_t s_task[2], s_ok, s_epoch, s_last;
    __shared__ uint32_t s_segf[kSegLds + 1], s_segp[kSegLds];
    const uint32_t t = threadIdx.x;
    // the segment table in LDS (each task looks its segment up)
    const bool seg_lds = a.nseg <= kSegLds;
    if (seg_lds) {
        for (uint32_t i = t; i <= a.nseg; i += NT) s_segf[i] = seg_first[i];
        for (uint32_t i = t; i < a.nseg; i += NT) s_segp[i] = seg_ping[i];
    }
    const uint32_t* segf = seg_lds ? s_segf : seg_first;
    const uint32_t* segp = seg_lds ? s_segp : seg_ping;
    uint32_t* ctl = a.ctl;
    uint32_t* ctr = ctl + kCtlWords;                                   // counter (ping, level): ctr_of
    uint32_t* flags = ctr + (size_t)a.nping * kLevels * kCtrWords;
    auto ctr_of = [&](uint32_t q, int lv) { return ctr + ((size_t)q * kLevels + lv) * kCtrWords; };
    const uint32_t slot = xcd_slot() * kCtrStride;
    // wave 0: the last dependency this workgroup acquired (ping << 8 |
    // level, or the ping's whole deposit level for the correlation)
    uint32_t have = ~0u;
    if (t == 0) {
        s_task[0] = atomicAdd(ctl + kCtlTicket, 1u);
        s_epoch = ld_agent(ctl + kCtlEpoch);
    }
    for (uint32_t it = 0;; it++) {
        __syncthreads();
        const uint32_t task = __builtin_amdgcn_readfirstlane(s_task[it & 1]);
        if (task >= a.ntasks) break;
        // segment of the task: largest seg with seg_first[seg] <= task
        uint32_t lo = 0, hi = a.nseg;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (segf[mid] <= task) lo = mid;
            else hi = mid;
        }
        // wave-uniform in registers: the ping's arguments are then read with
        // scalar loads and the bodies keep their uniform values in SGPRs
        const uint32_t lq = __builtin_amdgcn_readfirstlane(segp[lo]);
        const int lv = (int)(lq >> 24);
        const uint32_t q = lq & 0xFFFFFFu;
        const uint32_t local = __builtin_amdgcn_readfirstlane(task - segf[lo]);
        const PersistPing& P = pings[q];
        const uint32_t want = s_epoch + 1;

        // ---- dependency (wave 0 waits, the workgroup's barrier follows) ---------
        const int wl = P.wait[lv];
        bool run = true;
        if (lv == kLvCorr || wl >= 0) {
            if (t < 64) {
                bool ok = true;
                if (lv == kLvCorr) {
                    const uint32_t key = q << 8 | kLvDeposit;
                    if (have != key) {
                        uint32_t b0, b1, g0, g1;
                        corr_needs(P, (int)local, b0, b1, g0, g1);
                        const uint32_t base = P.flags;
                        ok = wait_flags(flags, base + b0, base + b1, want, ctl, a.spin_limit) &&
                             wait_flags(flags, base + (uint32_t)P.dep.NB + g0, base + (uint32_t)P.dep.NB + g1,
                                        want, ctl, a.spin_limit);
                        // every deposit task of the ping done: its later tiles need no wait
                        uint32_t v = t < kCtrSlots ? ld_agent(ctr_of(q, kLvDeposit) + t * kCtrStride) : 0u;
                        for (int o = kCtrSlots / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
                        if (ok && __builtin_amdgcn_readfirstlane(v) >= P.ntask[kLvDeposit]) {
                            acquire_agent();
                            have = key;
                        }
                    }
                } else {
                    const uint32_t key = q << 8 | (uint32_t)wl;
                    if (have != key) {
                        ok = wait_count(ctr_of(q, wl), P.ntask[wl], ctl, a.spin_limit);
                        if (ok) have = key;
                    }
                }
                if (t == 0) s_ok = ok ? 1u : 0u;
            }
            __syncthreads();
            run = s_ok != 0;
        }

        // ---- the task --------------------------------------------------------
        if (run) switch (lv) {
        case kLvStream: {
            const StreamArgs st = P.st;
            const BucketArgs bk = P.bk;
            stream_tile<NT, (int)kPersistTile>(st, bk, P.fuse, local, plds);
            break;
        }
        case kLvHist: {
            const BucketArgs bk = P.bk;
            bucket_hist_body<NT, kHistU>(bk, local, plds);
            break;
        }
        case kLvScan: {
            const BucketArgs bk = P.bk;
            scan_row(bk, local, plds);
            break;
        }
        case kLvScatter: {
            const BucketArgs bk = P.bk;
            if (P.staged) bucket_scatter_body<true, NT, kStagedU>(bk, local, plds);
            else bucket_scatter_body<false, NT, kScatterU>(bk, local, plds);
            break;
        }
        case kLvDeposit: {
            const DepositArgs dp = P.dep;
            deposit_or_windows<NT, kDepositPU>(dp, local, plds);
            break;
        }
        default: {
            const CorrArgs c = P.corr;
            correlate_variant(c, P.variant, (int)local);
            break;
        }
        }

        // ---- the next ticket, taken now that this task is done; publish --------
        __syncthreads();
        if (t == 0) {
            s_task[(it + 1) & 1] = atomicAdd(ctl + kCtlTicket, 1u);
            if (run) {
                if (lv == kLvDeposit)
                    __hip_atomic_store(flags + P.flags + local, want, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(ctr_of(q, lv) + slot, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }

    // ---- the last workgroup resets the launch's counters ------------------------
    if (t == 0)
        s_last = __hip_atomic_fetch_add(ctl + kCtlExited, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
                         gridDim.x - 1
                     ? 1u
                     : 0u;
    __syncthreads();
    if (s_last) {
        for (uint32_t i = t; i < a.nping * kLevels * kCtrSlots; i += NT) ctr[(size_t)i * kCtrStride] = 0u;
        if (t == 0) {
            ctl[kCtlTicket] = 0u;
            ctl[kCtlExited] = 0u;
            ctl[kCtlAbort] = 0u;
            ctl[kCtlEpoch] = s_epoch + 1;
        }
    }
}

hipError_t launch_persist(const PersistArgs& a, uint32_t blocks, size_t shm, hipStream_t s)
{
    if (!a.ntasks || !blocks) return hipSuccess;
    hipLaunchKernelGGL(k_persist, dim3(blocks), dim3(NT), shm, s, a.pings, a.seg_first, a.seg_ping, a);
    return hipGetLastError();
}

uint32_t persist_blocks(size_t shm)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_persist, NT, shm) != hipSuccess) per = 1;
    (void)hipGetLastError();
    return (uint32_t)(std::max(per, 1) * std::max(cus, 1));
}

size_t persist_lds_bytes(const PersistPing& p)
{
    size_t m = 0;
    auto take = [&](size_t v) { m = v > m ? v : m; };
    if (p.ntask[kLvStream] && p.fuse) take(sizeof(uint32_t) * (size_t)p.bk.NB);
    if (p.ntask[kLvHist]) take(sizeof(uint32_t) * (size_t)p.bk.NB);
    if (p.ntask[kLvScan]) take(sizeof(uint32_t) * (size_t)p.bk.tiles);
    if (p.ntask[kLvScatter])
        take(p.staged ? bucket_staged_bytes(p.bk) : sizeof(uint32_t) * ((size_t)p.bk.NB + p.bk.occ_words));
    if (p.ntask[kLvDeposit]) take(deposit_lds_bytes(p.dep));
    if (p.ntask[kLvCorr]) take(correlate_lds_bytes(p.corr));
    return m;
}

uint64_t persist_clock_hz()
{
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) {
        (void)hipGetLastError();
        return 100000000ull;   // 100 MHz, the usual rate
    }
    return (uint64_t)khz * 1000ull;
}

void init_persist_attributes(int max_lds) { set_dynamic_lds_limit((const void*)k_persist, max_lds); }

}  // namespace rtsim
