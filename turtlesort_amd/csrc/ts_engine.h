// ts_engine.h -- host-visible definitions of the sort engine: the device
// control block, per-level work lists, segment records and the launch
// parameters shared by the engine kernels and the C ABI.
#pragma once
#include <stdint.h>

namespace ts {

// error flags raised on the device (Ctrl::error)
enum : uint32_t {
    ERR_SEGS = 2u,     // segment table full
    ERR_TILES = 4u,    // tile table full
    ERR_LDSQ = 8u,     // LDS-sort task list full
    ERR_BAD = 16u,     // bin totals do not add up to the segment length
    ERR_RANGE = 32u,   // an index failed its range check; the access was skipped
    ERR_DEPTH = 64u    // segments were still left after the last level
};

// task kinds (statistics index)
enum : uint32_t { T_PIVOT = 1, T_COUNT = 2, T_SCATTER = 3, T_LDS = 4 };

// kernel kinds of one engine run (launch timing index; the C ABI's
// turtle_kernel_kind has the same values; K_LEVEL is no longer launched:
// the pivot kernel opens each level)
enum : int {
    K_RESET = 0, K_ROOTS, K_LEVEL, K_PIVOT, K_COUNT, K_SCAN, K_SCATTER, K_CHILDREN, K_LDS, K_LDSPREP, K_KINDS
};

// Key type code of the engine and MultiSort templates: the key width in
// bytes (1, 2, 4, 8), plus KT_FLOAT for IEEE floating-point keys.
constexpr int KT_FLOAT = 16;
constexpr int kt_bytes(int kt) { return kt & 15; }
constexpr bool kt_float(int kt) { return (kt & KT_FLOAT) != 0; }

constexpr int MAX_SPLIT = 31;
constexpr int MAX_BINS = 2 * MAX_SPLIT + 1;
// Depth bound: a segment whose child keeps more than half of its keys is
// split by key range next (Seg::mode), which divides the range by S+1; so a
// chain of levels holds at most log2(n / cutoff) halvings plus twice
// ceil(key bits / log2(S+1)) range steps (85 for 64-bit keys at S = 3).
constexpr int MAX_LEVELS = 128;
constexpr int LDS_CLASSES = 20;        // LDS-sort size classes
constexpr uint32_t SCAN_CHUNK = 4096;  // tile-count entries per workgroup of the level scan

// Device control block (one per handle), cleared by the reset kernel before
// every run.
struct Ctrl {
    uint32_t seg_alloc;     // segments created so far (all levels)
    uint32_t tile_alloc;    // tiles of the current level
    uint32_t lds_alloc;     // LDS-sort tasks created so far
    uint32_t error;
    uint32_t max_depth;
    uint32_t bad_where;     // site code of the first failed range check
    uint32_t pad0, pad1;
    unsigned long long bytes_count;
    unsigned long long bytes_scatter;
    unsigned long long bytes_lds;
    unsigned long long tasks[5];   // indexed by task kind
    // LDS-sort tasks by size class (filled after the last level)
    uint32_t cls_cnt[LDS_CLASSES];      // tasks per class
    uint32_t cls_cur[LDS_CLASSES];      // fill cursor per class
};
static_assert(sizeof(Ctrl) % 4 == 0, "Ctrl is cleared word by word");

// Segments of level L are the ids [sb, se) of the segment table.
struct Level {
    uint32_t sb, se, tiles, pad;
};

// A segment at or below the LDS cutoff, sorted in LDS.
struct LdsTask {
    uint32_t begin, len, depth, pad;
};

// A segment being partitioned (length above the LDS cutoff).  Its tiles are
// [tile_base, tile_base + ntiles) of the level's tile table; their bin
// counts sit at tcnt[tile_base*NB + bin*ntiles + tile] (bin-major), so the
// level-wide exclusive scan of tcnt gives, relative to the segment's first
// entry, the destination of every (bin, tile) inside the segment.
// Splitter choice of a segment (Seg::mode).
enum : uint32_t {
    SPLIT_SAMPLE = 0,   // sample quantiles
    SPLIT_RANGE = 1     // S+1 equal slices of [lo, hi]: the parent's split left
                        // this segment more than half of the parent's keys
};

struct Seg {
    uint32_t begin, len, depth, ntiles;
    uint32_t tile_base, mode, pad0, pad1;
    uint64_t lo, hi;               // bounds of the segment's order values (inclusive)
    uint64_t split[MAX_SPLIT];     // splitters in order representation, ascending
    uint64_t pad3;
};

struct Params {
    const void *kin;          // depth-0 keys (read only)
    const uint32_t *iin;      // depth-0 indices
    void *kout;               // final keys
    uint32_t *iout;           // final indices
    void *kt0, *kt1;          // ping-pong keys
    uint32_t *it0, *it1;      // ping-pong indices
    Ctrl *ctrl;
    Seg *segs;
    uint32_t *tcnt;           // per-tile bin counts (bin-major per segment)
    uint32_t *tpre;           // level-wide exclusive prefix of tcnt
    uint32_t *spart;          // per-chunk sums of the level scan
    uint32_t *tile_seg;       // segment id of every tile of the current level
    LdsTask *ldst;
    LdsTask *ldsq;            // the LDS tasks' records ordered by size class
    uint8_t *bins;            // bin of every key of the current level (count -> scatter),
                              // TILE bytes per tile of the level (see bin_slot)
    Level *levels;
    uint32_t segcap, tilecap, ldscap;
    uint32_t bins_tiles;      // tiles the bins buffer holds
    uint32_t cutoff;          // LDS sort at or below this length
    uint32_t tiles_per_wg;    // tiles a workgroup takes per step of the tile loops
    uint32_t n_total;         // length of every key / index buffer
    uint32_t pad1;
    uint64_t flip;            // sign-bit mask for signed and floating-point keys, else 0
};

}  // namespace ts
