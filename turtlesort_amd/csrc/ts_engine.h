// ts_engine.h -- host-visible definitions of the persistent sort engine:
// the device control block, queue slots, segment records and the launch
// parameters shared by ts_engine.hip and the C ABI.
#pragma once
#include <stdint.h>

namespace ts {

// error flags raised on the device (Ctrl::error)
enum : uint32_t {
    ERR_QUEUE = 1u,    // work queue full
    ERR_SEGS = 2u,     // segment table full
    ERR_TILES = 4u,    // tile-count table full
    ERR_TIMEOUT = 8u,  // watchdog expired
    ERR_BAD = 16u,     // inconsistent task (should not happen)
    ERR_RANGE = 32u    // an index failed its range check; the access was skipped
};

// task kinds; a slot's type word also carries the source depth of LDS tasks
enum : uint32_t { T_PIVOT = 1, T_COUNT = 2, T_SCATTER = 3, T_LDS = 4 };

constexpr int MAX_SPLIT = 15;
constexpr int MAX_BINS = 2 * MAX_SPLIT + 1;

// Device control block (one per handle).  Reset before every run by the
// reset kernel; gen increases by one per run so stale queue slots of earlier
// runs are never taken for new work.
struct Ctrl {
    uint32_t gen;
    uint32_t head;          // next queue slot to claim
    uint32_t tail;          // next queue slot to fill
    int32_t outstanding;    // tasks pushed and not yet finished
    uint32_t seg_alloc;
    uint32_t tile_alloc;
    uint32_t error;
    uint32_t max_depth;
    unsigned long long bytes_count;
    unsigned long long bytes_scatter;
    unsigned long long bytes_lds;
    unsigned long long tasks[5];   // indexed by task kind
    uint32_t bad_where;            // site code of the first failed range check
    uint32_t pad[3];
};

// One queue entry.  gen is written last (release); a slot is valid for the
// current run when gen matches Ctrl::gen.
struct Slot {
    uint32_t gen;
    uint32_t type;
    uint32_t a;
    uint32_t b;
};

// A segment being partitioned (length above the LDS cutoff).
struct Seg {
    uint32_t begin, len, depth, ntiles;
    uint32_t tile_base, ntasks, count_pending, scatter_pending;
    uint64_t split[MAX_SPLIT];     // splitters in order representation, ascending
    uint32_t base[MAX_BINS + 1];   // first position of each bin inside the segment
    uint32_t total[MAX_BINS + 1];  // elements per bin
};

struct Params {
    const void *kin;          // depth-0 keys (read only)
    const uint32_t *iin;      // depth-0 indices
    void *kout;               // final keys
    uint32_t *iout;           // final indices
    void *kt0, *kt1;          // ping-pong keys
    uint32_t *it0, *it1;      // ping-pong indices
    Ctrl *ctrl;
    Slot *q;
    Seg *segs;
    uint32_t *tiles;          // per-tile bin counts, then per-tile bin prefixes
    uint32_t qcap, segcap, tilecap;
    uint32_t cutoff;          // LDS sort at or below this length
    uint32_t tiles_per_task;
    uint32_t n_total;         // length of every key / index buffer
    uint64_t flip;            // sign-bit mask for signed keys, else 0
    uint64_t timeout_ticks;   // watchdog, in 100 MHz ticks
};

}  // namespace ts
