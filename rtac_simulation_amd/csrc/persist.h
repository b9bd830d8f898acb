// Persistent launch of grid-form (MFMA) pings: shared host/device layout.
//
// One launch runs every stage of every grid-form ping of a batch: streaming
// (+ bucket counts), bucket counts from records, the bucket scan, the bucket
// scatter, the deposit (+ windows) and the correlation.  Each stage of a ping
// is a level of tasks (a tile, a bucket row, a bucket, an output tile); the
// tasks of all pings form one list in level-major order (every ping's
// streaming tiles, then every ping's bucket rows, ...) and workgroups take
// them in that order from an atomic ticket.  A task waits only for its own
// ping's previous level (a counter of finished tasks), and a correlation tile
// only for the deposit buckets and window groups its rows read (one flag per
// deposit task), so a ping's stages overlap other pings' stages and a
// correlation tile starts while other buckets are still depositing.  Every
// task depends only on tasks earlier in the list, which were taken by running
// workgroups, so the launch needs no co-residency and can have any size.
// Every wait is bounded by a wall-clock limit: a wait that exceeds it sets an
// abort flag that every other wait and task sees, so the launch always ends;
// the error is reported to the host (rtsim_batch_check).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "grid_internal.h"
#include "rtsim_internal.h"

namespace rtsim {

constexpr int kPersistThreads = 512;   // 8 waves: the 8-wave correlation tiles
// samples per streaming / bucketing tile (the per-stage launches use 1024):
// fewer, larger tasks and more loads in flight per thread
constexpr uint32_t kPersistTile = 4096;

enum PersistLevel : int {
    kLvStream = 0,   // streaming tiles (with the bucket counts when they fuse)
    kLvHist,         // bucket counts from the records (tiles)
    kLvScan,         // bucket rows
    kLvScatter,      // tiles
    kLvDeposit,      // buckets, then window groups
    kLvCorr,         // output tiles
    kLevels
};

struct PersistPing {
    StreamArgs st;
    BucketArgs bk;
    DepositArgs dep;
    CorrArgs corr;
    uint32_t ntask[kLevels];   // tasks of each level (0: the level is skipped)
    int wait[kLevels];         // level a task waits for (-1: none)
    int fuse;                  // the streaming tiles make the bucket counts
    int staged;                // the scatter stages its items in LDS
    int variant;               // correlation: 4 fp16 + tile shape
    uint32_t flags;            // first deposit-task flag of this ping
};

// Control words (device, uint32): the ticket, exited workgroups, the abort
// flag, the sticky status (kPersistTimeout), the launch epoch; then the level
// counters [ping][kLevels][kCtrWords] (zeroed by the launch's last
// workgroup), then the deposit-task flags (hold epoch + 1 once done in that
// launch; never reset).
enum PersistCtl : int { kCtlTicket = 0, kCtlExited, kCtlAbort, kCtlStatus, kCtlEpoch, kCtlWords = 64 };
// a level counter: one slot per XCD, each on its own 128-byte line
constexpr uint32_t kCtrSlots = 8, kCtrStride = 32, kCtrWords = kCtrSlots * kCtrStride;
constexpr uint32_t kPersistTimeout = 1u;

struct PersistArgs {
    const PersistPing* pings;
    const uint32_t* seg_first;   // [nseg + 1]: first task of each (level, ping) segment
    const uint32_t* seg_ping;    // [nseg]: level << 24 | ping
    uint32_t nseg, ntasks, nping;
    uint32_t* ctl;               // kCtlWords control words, counters, flags
    uint64_t spin_limit;         // wall-clock ticks a wait may last
};

// the launch (grid: `blocks` workgroups of kPersistThreads, `shm` dynamic LDS)
hipError_t launch_persist(const PersistArgs& a, uint32_t blocks, size_t shm, hipStream_t s);
// workgroups of the launch for `shm` bytes of dynamic LDS (resident ones)
uint32_t persist_blocks(size_t shm);
// dynamic LDS a ping's tasks need
size_t persist_lds_bytes(const PersistPing& p);
// wall-clock ticks per second of the device
uint64_t persist_clock_hz();
void init_persist_attributes(int max_lds);

}  // namespace rtsim
