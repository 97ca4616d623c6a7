// ts_capi.hip -- the C ABI of libturtle_hip.so (declared in include/turtle_c.h).
//
// A handle is the device state of one Turtle::TurtleSort<Type,TrackIndex>
// object (turtle-sort.cuh:31-113): key/index buffers sized for max_n, the
// engine's control block, work queue, segment and tile tables, and one HIP
// stream.  Status codes replace the reference's print-and-exit error policy
// (turtle-globals.cuh:18-30).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <dlfcn.h>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/turtle_c.h"
#include "ts_copy.h"
#include "ts_engine.h"
#include "ts_internal.h"

static_assert((int)ts::K_KINDS == (int)TURTLE_K_KINDS && (int)ts::K_SCATTER == (int)TURTLE_K_SCATTER &&
                  (int)ts::K_LDS == (int)TURTLE_K_LDS,
              "kernel kinds of the engine and the C ABI differ");

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

int fail_hip(hipError_t e, const char *what)
{
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? TURTLE_ERR_OOM : TURTLE_ERR_DEVICE;
}

#define HIPCHK(call, what)                                 \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return fail_hip(e_, what);   \
    } while (0)

int dtype_bytes(int dtype)
{
    switch (dtype) {
    case TURTLE_INT8: case TURTLE_UINT8: return 1;
    case TURTLE_INT16: case TURTLE_UINT16: return 2;
    case TURTLE_INT32: case TURTLE_UINT32: return 4;
    case TURTLE_INT64: case TURTLE_UINT64: return 8;
    case TURTLE_FLOAT32: return 4;
    case TURTLE_FLOAT64: return 8;
    default: return 0;
    }
}

bool dtype_signed(int dtype)
{
    return dtype == TURTLE_INT8 || dtype == TURTLE_INT16 || dtype == TURTLE_INT32 || dtype == TURTLE_INT64;
}

bool dtype_float(int dtype) { return dtype == TURTLE_FLOAT32 || dtype == TURTLE_FLOAT64; }

// sign-bit mask of the order transform of the device key traits
uint64_t flip_of(int dtype)
{
    return dtype_signed(dtype) || dtype_float(dtype) ? (1ull << (8 * dtype_bytes(dtype) - 1)) : 0ull;
}

// key type code of the kernel templates
int kt_of(int dtype) { return dtype_bytes(dtype) | (dtype_float(dtype) ? ts::KT_FLOAT : 0); }

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

uint64_t env_u64(const char *name, uint64_t dflt)
{
    const char *v = getenv(name);
    if (!v || !*v) return dflt;
    return strtoull(v, nullptr, 10);
}

// Launch timing: an event pair per kernel launch, taken from a pool that
// grows on demand and is reused after every turtle_get_profile.
struct EventTimer : ts::LaunchTimer {
    struct Rec {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> pool;
    std::vector<Rec> recs;
    size_t used = 0;
    hipError_t err = hipSuccess;
    uint32_t mask = ~0u;   // kernel kinds that are timed (bit per kind)

    hipEvent_t take()
    {
        if (used == pool.size()) {
            hipEvent_t e = nullptr;
            const hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) {
                err = r;
                return nullptr;
            }
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(int kind, hipStream_t st) override
    {
        if (err != hipSuccess || !((mask >> kind) & 1u)) return;
        hipEvent_t a = take(), b = take();
        if (!a || !b) return;
        const hipError_t r = hipEventRecord(a, st);
        if (r != hipSuccess) {
            err = r;
            return;
        }
        recs.push_back({kind, a, b});
    }
    void end(int kind, hipStream_t st) override
    {
        if (err != hipSuccess || !((mask >> kind) & 1u) || recs.empty() || recs.back().kind != kind) return;
        const hipError_t r = hipEventRecord(recs.back().b, st);
        if (r != hipSuccess) err = r;
    }
    void reset()
    {
        recs.clear();
        used = 0;
        err = hipSuccess;
    }
    ~EventTimer() override
    {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace

// Everything an engine run writes besides the key and index buffers: the
// control block, segment, tile and LDS-task tables, level records and the
// per-key bins.  A handle owns one set; every captured graph owns its own,
// so that no later call on the handle frees or resizes what a graph's
// kernel arguments point at.
struct EngineTables {
    ts::Ctrl *ctrl = nullptr;
    ts::Seg *segs = nullptr;
    uint32_t *tcnt = nullptr, *tpre = nullptr, *spart = nullptr;
    uint32_t *tile_seg = nullptr;
    ts::LdsTask *ldst = nullptr;
    ts::LdsTask *ldsq = nullptr;
    ts::Level *levels = nullptr;
    uint8_t *bins = nullptr;    // per-key bin of the current level (count -> scatter), per tile
    uint32_t segcap = 0, tilecap = 0, ldscap = 0;
    uint32_t bins_tiles = 0;    // tiles the bins buffer holds
};

// The engine run last queued on a handle.  An eager run is queued with the
// estimated number of partition levels and returns; its level check (are
// segments left? did a table overflow?) is resolved by turtle_synchronize or
// by the next call on the handle, which queue more levels, or grow the
// tables and run again from the untouched input, as needed.
struct EngineRun {
    bool engine = false;        // the handle's last device work was an engine run
    bool pending = false;       // its level check is not resolved yet
    bool graph = false;         // a graph replay (no resolution: its levels are fixed)
    const EngineTables *tab = nullptr;
    const void *kin = nullptr;  // depth-0 input (never written by the run)
    const uint32_t *iin = nullptr;
    void *kout = nullptr;
    uint32_t *iout = nullptr;
    uint64_t total = 0;
    const uint32_t *dev_off = nullptr;
    uint32_t nseg = 0;
    hipStream_t st = nullptr;
    uint32_t levels = 0;        // partition levels queued so far
    uint32_t grows = 0;         // table growths of this run
    turtle_tuning tune{};       // tuning the run was queued with
};

struct turtle_sorter {
    int dtype = 0, kb = 0, kt = 0;   // key type, key bytes, kernel key code
    uint64_t flip = 0;
    int track = 0;
    int64_t max_n = 0;
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    hipStream_t last_stream = nullptr;
    // key / index buffers of max_n elements; kin / iin also keep an
    // in-place device call's input for a rerun
    void *kin = nullptr, *kout = nullptr, *kt0 = nullptr, *kt1 = nullptr;
    uint32_t *iin = nullptr, *iout = nullptr, *it0 = nullptr, *it1 = nullptr;
    EngineTables tab;                // the handle's engine tables
    ts::Ctrl *ctrl_host = nullptr;   // pinned readback copy
    uint32_t *dev_off = nullptr;
    size_t dev_off_cap = 0;
    std::vector<uint32_t> host_off;
    turtle_tuning tune{};        // tuning of the current / last run
    turtle_tuning tune_base{};   // built-in default, or what turtle_set_tuning set
    bool tune_fixed = false;     // set by turtle_set_tuning: the tuning table is not consulted
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t lv_ev = nullptr;   // end of a batch of levels (the read-back waits for it on stream2)
    hipEvent_t rb_ev = nullptr;   // end of the read-back of a batch's level state
    ts::Level *levels_host = nullptr;  // pinned readback (a Level followed by a Ctrl)
    bool timed = false;
    uint64_t last_n = 0;
    EngineRun run;               // the last engine run
    int live_graphs = 0;         // graphs captured on this handle and not destroyed
    // StartSorting / Sync state
    void *host_keys = nullptr;
    int32_t *host_idx = nullptr;
    int64_t host_n = 0;
    bool started = false;
    double t_start = 0.0;
    // pinned staging for host <-> device copies: user memory is only ever
    // touched by host memcpy, never handed to the HIP runtime
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_used[2] = {false, false};
    ts::CopyPool *pool = nullptr;        // threads of the user <-> staging copies
    hipStream_t stream2 = nullptr;       // read-backs; second stream of the chunked host MultiSort
    // per-kernel launch timing (turtle_set_profiling)
    EventTimer *timer = nullptr;
};

namespace {

void free_tables(EngineTables &t)
{
    void *bufs[] = {t.ctrl, t.segs, t.tcnt, t.tpre, t.spart, t.tile_seg, t.ldst, t.ldsq, t.levels, t.bins};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    t = EngineTables{};
}

// Sizes engine tables for a run over `total` keys in `nroots` root segments
// (the control block is allocated and cleared once).  Segments longer than
// the cutoff are disjoint within a level, so a level holds fewer than
// total/cutoff of them; with sample-quantile pivots each level shrinks them
// about S+1 times, so all levels together stay near 1.4 total/cutoff (8x
// that is allocated).  Tiles are per level: at most total/TILE + (segments
// of the level).  Every segment yields at most S+1 LDS tasks.  `scale` > 1
// grows the tables after an overflow.  TURTLE_SEGCAP replaces the segment
// capacity of the first sizing (scale 1) only, so that tests can force an
// overflow and the growth that follows it.
int ensure_tables(const turtle_sorter *h, const turtle_tuning &tu, EngineTables &t, uint64_t total, uint32_t nroots,
                  uint32_t scale)
{
    const uint64_t cutoff = (uint64_t)tu.lds_cutoff;
    const uint64_t tile = ts::engine_tile(h->kb);
    const uint64_t S = (uint64_t)tu.splitters;
    uint64_t segcap = (8 * (total / cutoff + 1) + 1024) * scale + nroots;
    const uint64_t force = scale == 1 ? env_u64("TURTLE_SEGCAP", 0) : 0;
    if (force) segcap = force;
    const uint64_t tilecap = (total / tile + 1) + segcap;
    const uint64_t ldscap = segcap * (S + 1) + nroots + 1024;
    // tiles of one level: every segment of a level is longer than the
    // cutoff and the segments are disjoint, so there are at most
    // total/TILE + total/(cutoff+1) + 1 of them; the bins buffer holds TILE
    // bytes for each
    uint64_t bins_tiles = total / tile + total / (cutoff + 1) + 2;
    if (bins_tiles > tilecap) bins_tiles = tilecap;
    if (segcap > 0xffffff00ull || tilecap > 0xffffff00ull || ldscap > 0xffffff00ull)
        return fail(TURTLE_ERR_CAPACITY, "work tables would exceed 2^32 entries");
    if (!t.ctrl) {
        HIPCHK(hipMalloc((void **)&t.ctrl, sizeof(ts::Ctrl)), "hipMalloc(control block)");
        HIPCHK(hipMemset(t.ctrl, 0, sizeof(ts::Ctrl)), "hipMemset(control block)");
    }
    if (t.segs && t.segcap >= segcap && t.tilecap >= tilecap && t.ldscap >= ldscap && t.bins_tiles >= bins_tiles)
        return TURTLE_OK;
    // the old tables may still be in use by a run in flight
    if (t.segs) HIPCHK(hipDeviceSynchronize(), "hipDeviceSynchronize");
    ts::Ctrl *ctrl = t.ctrl;
    t.ctrl = nullptr;
    free_tables(t);
    t.ctrl = ctrl;
    HIPCHK(hipMalloc((void **)&t.segs, segcap * sizeof(ts::Seg)), "hipMalloc(segments)");
    const uint64_t entries = tilecap * ts::MAX_BINS;
    HIPCHK(hipMalloc((void **)&t.tcnt, entries * sizeof(uint32_t)), "hipMalloc(tile counts)");
    HIPCHK(hipMalloc((void **)&t.tpre, entries * sizeof(uint32_t)), "hipMalloc(tile prefixes)");
    HIPCHK(hipMalloc((void **)&t.spart, (entries / ts::SCAN_CHUNK + 2) * sizeof(uint32_t)), "hipMalloc(scan sums)");
    HIPCHK(hipMalloc((void **)&t.tile_seg, tilecap * sizeof(uint32_t)), "hipMalloc(tile map)");
    HIPCHK(hipMalloc((void **)&t.ldst, ldscap * sizeof(ts::LdsTask)), "hipMalloc(lds tasks)");
    HIPCHK(hipMalloc((void **)&t.ldsq, ldscap * sizeof(ts::LdsTask)), "hipMalloc(lds order)");
    HIPCHK(hipMalloc((void **)&t.levels, ts::MAX_LEVELS * sizeof(ts::Level)), "hipMalloc(levels)");
    HIPCHK(hipMalloc((void **)&t.bins, bins_tiles * tile), "hipMalloc(bins)");
    t.bins_tiles = (uint32_t)bins_tiles;
    t.segcap = (uint32_t)segcap;
    t.tilecap = (uint32_t)tilecap;
    t.ldscap = (uint32_t)ldscap;
    return TURTLE_OK;
}

ts::Params make_params(const turtle_sorter *h, const turtle_tuning &tu, const EngineTables &t, const void *kin,
                       const uint32_t *iin, void *kout, uint32_t *iout)
{
    ts::Params p{};
    p.kin = kin;
    p.iin = iin;
    p.kout = kout;
    p.iout = iout;
    p.kt0 = h->kt0;
    p.kt1 = h->kt1;
    p.it0 = h->it0;
    p.it1 = h->it1;
    p.ctrl = t.ctrl;
    p.segs = t.segs;
    p.tcnt = t.tcnt;
    p.tpre = t.tpre;
    p.spart = t.spart;
    p.tile_seg = t.tile_seg;
    p.ldst = t.ldst;
    p.ldsq = t.ldsq;
    p.bins = t.bins;
    p.levels = t.levels;
    p.segcap = t.segcap;
    p.tilecap = t.tilecap;
    p.ldscap = t.ldscap;
    p.bins_tiles = t.bins_tiles;
    p.cutoff = (uint32_t)tu.lds_cutoff;
    p.tiles_per_wg = (uint32_t)(tu.tiles_per_task > 0 ? tu.tiles_per_task : 1);
    p.n_total = 0;   // set by the caller
    p.flip = h->flip;
    return p;
}

// partition levels expected for a segment of n keys: the longest segment
// shrinks about (S+1) times per level (a quarter more is allowed for uneven
// splits) until it fits the cutoff; the host reads back after these and
// adds levels only if segments are left
uint32_t levels_estimate(uint64_t n, uint64_t cutoff, uint64_t S)
{
    uint32_t L = 0;
    while (n > cutoff && L < 64) {
        n = (n * 5 / 4 + S) / (S + 1);
        L++;
    }
    return L;
}

int ensure_readback(turtle_sorter *h)
{
    if (!h->levels_host)
        HIPCHK(hipHostMalloc((void **)&h->levels_host, sizeof(ts::Level) + sizeof(ts::Ctrl), hipHostMallocDefault),
               "hipHostMalloc");
    if (!h->ctrl_host) HIPCHK(hipHostMalloc((void **)&h->ctrl_host, sizeof(ts::Ctrl), hipHostMallocDefault), "hipHostMalloc");
    if (!h->lv_ev) HIPCHK(hipEventCreateWithFlags(&h->lv_ev, hipEventDisableTiming), "hipEventCreate");
    if (!h->rb_ev) HIPCHK(hipEventCreateWithFlags(&h->rb_ev, hipEventDisableTiming), "hipEventCreate");
    if (!h->stream2) HIPCHK(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking), "hipStreamCreate");
    return TURTLE_OK;
}

// Device forms run on the caller's stream; NULL is the null stream (as in
// every HIP call), never the handle's private non-blocking stream, which is
// not ordered after the null-stream work that produced the inputs.
hipStream_t stream_of(turtle_sorter *h, void *stream)
{
    (void)h;
    return (hipStream_t)stream;
}

turtle_tuning tuning_for(const turtle_sorter *h, bool idx, uint64_t total, bool seg);

// longest root of a run: the segments' longest (h->host_off holds the
// offsets of a segmented run), else the whole input
uint64_t longest_root(const turtle_sorter *h, bool segmented, uint64_t total)
{
    if (!segmented || h->host_off.size() < 2) return total;
    uint64_t longest = 0;
    for (size_t s = 0; s + 1 < h->host_off.size(); s++) {
        const uint64_t l = (uint64_t)h->host_off[s + 1] - h->host_off[s];
        if (l > longest) longest = l;
    }
    return longest;
}

// Queues one batch of run r on its stream: (first) reset + roots, then the
// partition levels [r.levels, r.levels + nlev), then the LDS phase over the
// whole task list (a later batch first clears its size classes: a task's
// input range is never written by deeper levels, so sorting it again writes
// the same output; its bytes and task counts are counted twice in the
// statistics).  Behind the levels, the second stream reads the last level's
// record and the control block back into pinned memory; the batch is then
// pending until resolve() has looked at them.
int queue_batch(turtle_sorter *h, EngineRun &r, const turtle_tuning &tu, uint32_t nlev, bool first)
{
    ts::Params p = make_params(h, tu, *r.tab, r.kin, r.iin, r.kout, r.iout);
    p.n_total = (uint32_t)r.total;
    const bool idx = r.iin != nullptr;
    const int S = tu.splitters, wpc = tu.workgroups_per_cu;
    if (r.levels + nlev > (uint32_t)ts::MAX_LEVELS) nlev = ts::MAX_LEVELS - r.levels;
    if (!first) HIPCHK(ts::engine_clear_classes(p, r.st), "engine launch");
    HIPCHK(ts::engine_levels(p, h->kt, idx, S, h->cus, wpc, r.dev_off, r.nseg, (uint32_t)r.total, r.levels, nlev, first,
                             false, r.st, h->timer),
           "engine launch");
    r.levels += nlev;
    if (r.levels > 0) HIPCHK(hipEventRecord(h->lv_ev, r.st), "hipEventRecord");
    HIPCHK(ts::engine_levels(p, h->kt, idx, S, h->cus, wpc, r.dev_off, r.nseg, (uint32_t)r.total, r.levels, 0, false,
                             true, r.st, h->timer),
           "engine launch");
    HIPCHK(hipEventRecord(h->ev1, r.st), "hipEventRecord");
    if (r.levels > 0) {
        ts::Ctrl *ch = (ts::Ctrl *)(h->levels_host + 1);
        HIPCHK(hipStreamWaitEvent(h->stream2, h->lv_ev, 0), "hipStreamWaitEvent");
        HIPCHK(hipMemcpyAsync(h->levels_host, r.tab->levels + (r.levels - 1), sizeof(ts::Level), hipMemcpyDeviceToHost,
                              h->stream2),
               "D2H(level)");
        HIPCHK(hipMemcpyAsync(ch, r.tab->ctrl, sizeof(ts::Ctrl), hipMemcpyDeviceToHost, h->stream2), "D2H(ctrl)");
        HIPCHK(hipEventRecord(h->rb_ev, h->stream2), "hipEventRecord");
        r.pending = true;
    }
    return TURTLE_OK;
}

// TURTLE_LEVEL_BATCH caps the levels per batch (tests force the read-back
// path with it)
uint32_t level_batch_cap() { return (uint32_t)env_u64("TURTLE_LEVEL_BATCH", 0); }

constexpr uint32_t MAX_GROWS = 3;

// Resolves the pending level check of the handle's last engine run: waits
// for the read-back of its last batch (not for its LDS phase), and queues
// 4 more levels and the LDS phase again while segments are left; after a
// table overflow it grows the tables (4x, 16x, 64x) and queues the whole
// run again from its input, which no level writes.
int resolve(turtle_sorter *h)
{
    EngineRun &r = h->run;
    while (r.pending) {
        r.pending = false;
        HIPCHK(hipEventSynchronize(h->rb_ev), "hipEventSynchronize(read-back)");
        const ts::Ctrl *c = (const ts::Ctrl *)(h->levels_host + 1);
        const uint32_t left = c->seg_alloc > h->levels_host->se ? c->seg_alloc - h->levels_host->se : 0u;
        const uint32_t err = c->error;
        constexpr uint32_t OVERFLOW = ts::ERR_SEGS | ts::ERR_TILES | ts::ERR_LDSQ;
        if ((err & OVERFLOW) && !(err & ~OVERFLOW) && r.grows < MAX_GROWS && r.tab == &h->tab) {
            r.grows++;
            HIPCHK(hipStreamSynchronize(r.st), "hipStreamSynchronize");
            int rc = ensure_tables(h, r.tune, h->tab, r.total, r.nseg, 1u << (2 * r.grows));
            if (rc) return rc;
            r.levels = 0;
            const uint32_t cap = level_batch_cap();
            uint32_t nlev = levels_estimate(longest_root(h, r.dev_off != nullptr, r.total), (uint64_t)r.tune.lds_cutoff,
                                            (uint64_t)r.tune.splitters);
            if (cap && nlev > cap) nlev = cap;
            HIPCHK(hipEventRecord(h->ev0, r.st), "hipEventRecord");
            if ((rc = queue_batch(h, r, r.tune, nlev, true))) return rc;
            continue;
        }
        if (left && !err && r.levels < (uint32_t)ts::MAX_LEVELS) {
            const uint32_t cap = level_batch_cap();
            const int rc = queue_batch(h, r, r.tune, cap && cap < 4 ? cap : 4, false);
            if (rc) return rc;
        }
    }
    return TURTLE_OK;
}

// One engine run over device buffers.  In-place requests (kin == kout)
// first copy the input into the handle's input buffer, which no level
// writes, so that a rerun after a table overflow still has it.  Eager runs
// resolve the previous run first, queue the estimated levels and the LDS
// phase, and return (see resolve).  capture != nullptr: recorded into a
// graph, with the graph's own tables and a fixed level count (returned in
// *captured_levels): the estimate plus two spare levels, or
// $TURTLE_GRAPH_LEVELS (tests force a too-shallow graph with it).
int run_engine(turtle_sorter *h, const void *kin, const uint32_t *iin, void *kout, uint32_t *iout, uint64_t total,
               const uint32_t *dev_off, uint32_t nseg, hipStream_t st, EngineTables *capture = nullptr,
               uint32_t *captured_levels = nullptr)
{
    const bool capturing = capture != nullptr;
    int rc = TURTLE_OK;
    if (!capturing && (rc = resolve(h))) return rc;
    const bool idx = iin != nullptr && iout != nullptr;
    // the longest root decides the level count, and the tuning: the segments
    // of a batch are partitioned independently, so a batch is tuned like one
    // sort of its longest segment (a capture uses the tuning its creator
    // resolved and sized the tables for)
    const uint64_t longest = longest_root(h, dev_off != nullptr, total);
    if (!capturing) h->tune = tuning_for(h, idx, longest, dev_off != nullptr);
    EngineTables &tab = capturing ? *capture : h->tab;
    if ((rc = ensure_tables(h, h->tune, tab, total, nseg, 1))) return rc;
    if ((rc = ensure_readback(h))) return rc;
    if (total > 0 && kin == kout) {
        HIPCHK(hipMemcpyAsync(h->kin, kin, total * h->kb, hipMemcpyDeviceToDevice, st), "hipMemcpyAsync(in-place keys)");
        kin = h->kin;
    }
    if (idx && total > 0 && (const void *)iin == (const void *)iout) {
        HIPCHK(hipMemcpyAsync(h->iin, iin, total * 4, hipMemcpyDeviceToDevice, st), "hipMemcpyAsync(in-place idx)");
        iin = h->iin;
    }
    EngineRun r;
    r.engine = true;
    r.tab = &tab;
    r.kin = kin;
    r.iin = idx ? iin : nullptr;
    r.kout = kout;
    r.iout = idx ? iout : nullptr;
    r.total = total;
    r.dev_off = dev_off;
    r.nseg = nseg;
    r.st = st;
    r.tune = h->tune;
    uint32_t nlev = levels_estimate(longest, (uint64_t)h->tune.lds_cutoff, (uint64_t)h->tune.splitters);
    if (capturing) {
        // no read-back inside a graph: two spare levels
        const uint64_t force = env_u64("TURTLE_GRAPH_LEVELS", 0);
        nlev = force ? (uint32_t)force : nlev + 2;
        if (nlev > (uint32_t)ts::MAX_LEVELS) nlev = ts::MAX_LEVELS;
        ts::Params p = make_params(h, h->tune, tab, kin, r.iin, kout, r.iout);
        p.n_total = (uint32_t)total;
        HIPCHK(ts::engine_levels(p, h->kt, idx, h->tune.splitters, h->cus, h->tune.workgroups_per_cu, dev_off, nseg,
                                 (uint32_t)total, 0, nlev, true, true, st),
               "engine launch");
        if (captured_levels) *captured_levels = nlev;
        return TURTLE_OK;
    }
    const uint32_t cap = level_batch_cap();
    if (cap && nlev > cap) nlev = cap;
    HIPCHK(hipEventRecord(h->ev0, st), "hipEventRecord");
    if ((rc = queue_batch(h, r, h->tune, nlev, true))) return rc;
    h->run = r;
    h->timed = true;
    h->last_stream = st;
    h->last_n = total;
    return TURTLE_OK;
}

// Reads the control block (and the last level's record) of the handle's
// last engine run after its stream is done.
int read_stats(turtle_sorter *h, turtle_stats *st, uint32_t *err_out)
{
    int rc = ensure_readback(h);
    if (rc) return rc;
    const EngineRun &r = h->run;
    const bool eng = r.engine && r.tab && r.tab->ctrl;
    ts::Ctrl c{};
    if (eng) {
        HIPCHK(hipMemcpyAsync(h->ctrl_host, r.tab->ctrl, sizeof(ts::Ctrl), hipMemcpyDeviceToHost, h->last_stream),
               "D2H(ctrl)");
        if (r.levels > 0)
            HIPCHK(hipMemcpyAsync(h->levels_host, r.tab->levels + (r.levels - 1), sizeof(ts::Level),
                                  hipMemcpyDeviceToHost, h->last_stream),
                   "D2H(level)");
        HIPCHK(hipStreamSynchronize(h->last_stream), "hipStreamSynchronize");
        c = *h->ctrl_host;
    }
    uint32_t err = eng ? c.error : 0u;
    // segments created by the last level were never partitioned
    if (eng && r.levels > 0 && c.seg_alloc > h->levels_host->se) err |= ts::ERR_DEPTH;
    float ms = 0.f;
    if (h->timed) HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1), "hipEventElapsedTime");
    if (st) {
        memset(st, 0, sizeof(*st));
        st->device_ms = ms;
        st->n = h->last_n;
        if (eng) {
            st->bytes_count = c.bytes_count;
            st->bytes_scatter = c.bytes_scatter;
            st->bytes_lds = c.bytes_lds;
            st->tasks_pivot = c.tasks[ts::T_PIVOT];
            st->tasks_count = c.tasks[ts::T_COUNT];
            st->tasks_scatter = c.tasks[ts::T_SCATTER];
            st->tasks_lds = c.tasks[ts::T_LDS];
            st->segments = c.seg_alloc;
            st->max_depth = c.max_depth;
        }
        st->error = err;
    }
    *err_out = err;
    return TURTLE_OK;
}

int error_status(uint32_t err)
{
    if (err == 0) return TURTLE_OK;
    if (err & ts::ERR_RANGE) return fail(TURTLE_ERR_DEVICE, "device range check failed (access skipped)");
    if (err & ts::ERR_BAD) return fail(TURTLE_ERR_DEVICE, "inconsistent partition detected on the device");
    if (err & (ts::ERR_SEGS | ts::ERR_TILES | ts::ERR_LDSQ))
        return fail(TURTLE_ERR_QUEUE_OVERFLOW, "device work list (segments/tiles/LDS tasks) overflow");
    if (err & ts::ERR_DEPTH) return fail(TURTLE_ERR_QUEUE_OVERFLOW, "segments left after the last partition level");
    return fail(TURTLE_ERR_DEVICE, "device error");
}

constexpr uint32_t TUNE_MIN_CUTOFF = 1024;

bool valid_tuning(const turtle_tuning &t, int kb)
{
    if (!ts::engine_supports(t.splitters)) return false;
    // every segment of a level is longer than the cutoff, and each of its
    // tiles owns a TILE-byte bins slot, so the bins buffer takes up to
    // n * (1 + TILE / cutoff) bytes: the cutoff's lower bound keeps that
    // below 9 bytes per key
    if (t.lds_cutoff < (int32_t)TUNE_MIN_CUTOFF || (uint32_t)t.lds_cutoff > ts::engine_cap(kb)) return false;
    if (t.tiles_per_task < 1 || t.tiles_per_task > 1024) return false;
    if (t.workgroups_per_cu < 0 || t.workgroups_per_cu > 64) return false;
    return true;
}

turtle_tuning default_tuning(int kb)
{
    turtle_tuning t;
    t.splitters = 31;
    t.lds_cutoff = (int32_t)ts::engine_cap(kb);
    t.tiles_per_task = 1;
    t.workgroups_per_cu = 0;
    return t;
}

// ---- tuning table -------------------------------------------------------------
// Written by the autotuner (turtlesort_amd/autotune.py over turtle_autotune):
// lines "kb idx log2n splitters lds_cutoff tiles_per_task workgroups_per_cu
// ms", '#' comments; lines that start with "seg" hold the tunings of
// segmented batches (turtle_autotune_segmented), their log2n being that of
// the batch's longest segment.  Read once from $TURTLE_TUNING_FILE, else
// from tuning_table.txt next to this library.  A run uses the entry of its
// key width and index flag whose log2n is nearest to log2 of its longest
// root (a segmented run the "seg" entries when there are any); without an
// entry the built-in default applies.
struct TuneEntry {
    int kb, idx, lg;
    turtle_tuning t;
    bool seg;
};

const std::vector<TuneEntry> &tuning_table()
{
    static std::vector<TuneEntry> tab;
    static std::once_flag once;
    std::call_once(once, [] {
        std::string path;
        const char *env = getenv("TURTLE_TUNING_FILE");
        if (env && *env) {
            path = env;
        } else {
            Dl_info info;
            if (dladdr((void *)&tuning_table, &info) && info.dli_fname) {
                path = info.dli_fname;
                const size_t slash = path.rfind('/');
                path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuning_table.txt";
            }
        }
        FILE *f = path.empty() ? nullptr : fopen(path.c_str(), "r");
        if (!f) return;
        char line[256];
        while (fgets(line, sizeof line, f)) {
            if (line[0] == '#') continue;
            TuneEntry e{};
            double ms = 0;
            const char *q = line;
            e.seg = strncmp(line, "seg ", 4) == 0;
            if (e.seg) q += 4;
            if (sscanf(q, "%d %d %d %d %d %d %d %lf", &e.kb, &e.idx, &e.lg, &e.t.splitters, &e.t.lds_cutoff,
                       &e.t.tiles_per_task, &e.t.workgroups_per_cu, &ms) >= 7)
                tab.push_back(e);
        }
        fclose(f);
    });
    return tab;
}

bool valid_tuning(const turtle_tuning &t, int kb);

// tuning of a run whose longest root holds `total` keys (the handle's own
// when it was set by hand); seg: the run sorts a segmented batch
turtle_tuning tuning_for(const turtle_sorter *h, bool idx, uint64_t total, bool seg)
{
    if (h->tune_fixed) return h->tune;
    const int lg = total > 1 ? (int)std::lround(std::log2((double)total)) : 0;
    bool have_seg = false;
    if (seg)
        for (const TuneEntry &e : tuning_table())
            have_seg |= e.seg && e.kb == h->kb && e.idx == (idx ? 1 : 0) && valid_tuning(e.t, h->kb);
    const TuneEntry *best = nullptr;
    for (const TuneEntry &e : tuning_table()) {
        if (e.kb != h->kb || e.idx != (idx ? 1 : 0) || e.seg != have_seg || !valid_tuning(e.t, h->kb)) continue;
        if (!best || std::abs(e.lg - lg) < std::abs(best->lg - lg)) best = &e;
    }
    return best ? best->t : h->tune_base;
}

constexpr size_t STAGE_BYTES = 16u << 20;

int ensure_staging(turtle_sorter *h)
{
    for (int k = 0; k < 2; k++) {
        if (!h->stage[k]) HIPCHK(hipHostMalloc(&h->stage[k], STAGE_BYTES, hipHostMallocDefault), "hipHostMalloc(staging)");
        if (!h->stage_ev[k]) HIPCHK(hipEventCreateWithFlags(&h->stage_ev[k], hipEventDisableTiming), "hipEventCreate");
    }
    if (!h->pool) h->pool = new ts::CopyPool(ts::CopyPool::default_threads());
    return TURTLE_OK;
}

// host -> device through the two pinned chunks: the memcpy into one chunk
// overlaps the DMA out of the other
int stage_h2d(turtle_sorter *h, void *dev, const void *host, size_t bytes, hipStream_t st)
{
    int rc = ensure_staging(h);
    if (rc) return rc;
    int k = 0;
    for (size_t off = 0; off < bytes; off += STAGE_BYTES, k ^= 1) {
        const size_t len = bytes - off < STAGE_BYTES ? bytes - off : STAGE_BYTES;
        if (h->stage_used[k]) HIPCHK(hipEventSynchronize(h->stage_ev[k]), "hipEventSynchronize");
        h->pool->copy(h->stage[k], (const char *)host + off, len);
        HIPCHK(hipMemcpyAsync((char *)dev + off, h->stage[k], len, hipMemcpyHostToDevice, st), "H2D");
        HIPCHK(hipEventRecord(h->stage_ev[k], st), "hipEventRecord");
        h->stage_used[k] = true;
    }
    return TURTLE_OK;
}

// device -> host, pipelined the same way; returns when host holds the data
int stage_d2h(turtle_sorter *h, void *host, const void *dev, size_t bytes, hipStream_t st)
{
    int rc = ensure_staging(h);
    if (rc) return rc;
    const size_t n = (bytes + STAGE_BYTES - 1) / STAGE_BYTES;
    auto issue = [&](size_t i) -> int {
        const int k = (int)(i & 1);
        const size_t off = i * STAGE_BYTES;
        const size_t len = bytes - off < STAGE_BYTES ? bytes - off : STAGE_BYTES;
        if (h->stage_used[k]) HIPCHK(hipEventSynchronize(h->stage_ev[k]), "hipEventSynchronize");
        HIPCHK(hipMemcpyAsync(h->stage[k], (const char *)dev + off, len, hipMemcpyDeviceToHost, st), "D2H");
        HIPCHK(hipEventRecord(h->stage_ev[k], st), "hipEventRecord");
        h->stage_used[k] = true;
        return TURTLE_OK;
    };
    if (n > 0 && (rc = issue(0))) return rc;
    for (size_t i = 0; i < n; i++) {
        if (i + 1 < n && (rc = issue(i + 1))) return rc;
        const int k = (int)(i & 1);
        const size_t off = i * STAGE_BYTES;
        const size_t len = bytes - off < STAGE_BYTES ? bytes - off : STAGE_BYTES;
        HIPCHK(hipEventSynchronize(h->stage_ev[k]), "hipEventSynchronize");
        h->pool->copy((char *)host + off, h->stage[k], len);
    }
    return TURTLE_OK;
}

// The stable ranks of the sort kernels rely on the lane order of LDS
// atomic return values (rank16_add); it is checked once per device.
int ensure_rank_order(int dev)
{
    static std::mutex mu;
    static int state[64] = {0};   // 0 unchecked, 1 holds, 2 does not hold
    if (dev < 0 || dev >= 64) return fail(TURTLE_ERR_UNSUPPORTED, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (state[dev] == 0) {
        bool ok = false;
        HIPCHK(ts::check_lds_rank_order(&ok), "LDS atomic order check");
        state[dev] = ok ? 1 : 2;
    }
    if (state[dev] == 2)
        return fail(TURTLE_ERR_UNSUPPORTED,
                    "LDS atomics on this device do not return in lane order; the stable ranking needs it");
    return TURTLE_OK;
}

int alloc_buf(void **p, size_t bytes, const char *what)
{
    HIPCHK(hipMalloc(p, bytes ? bytes : 16), what);
    return TURTLE_OK;
}

}  // namespace

extern "C" {

const char *turtle_status_string(int s)
{
    switch (s) {
    case TURTLE_OK: return "ok";
    case TURTLE_ERR_INVALID: return "invalid argument";
    case TURTLE_ERR_CAPACITY: return "capacity exceeded";
    case TURTLE_ERR_OOM: return "out of memory";
    case TURTLE_ERR_DTYPE: return "bad dtype";
    case TURTLE_ERR_UNSUPPORTED: return "unsupported";
    case TURTLE_ERR_DEVICE: return "device error";
    case TURTLE_ERR_QUEUE_OVERFLOW: return "work queue overflow";
    case TURTLE_ERR_TIMEOUT: return "watchdog timeout";
    case TURTLE_ERR_STATE: return "bad call order";
    case TURTLE_ERR_BUDGET: return "memory budget exceeded";
    default: return "unknown status";
    }
}

const char *turtle_last_error(void) { return g_last_error.c_str(); }

int turtle_dtype_size(int dtype) { return dtype_bytes(dtype); }

int turtle_check_device(void)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev), "hipGetDevice");
    return ensure_rank_order(dev);
}

int turtle_device_cus(void)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, dev) != hipSuccess) return 0;
    return pr.multiProcessorCount;
}

int turtle_create(int dtype, int64_t max_n, int track_index, int compress, turtle_handle *out)
{
    if (!out) return fail(TURTLE_ERR_INVALID, "out is NULL");
    *out = nullptr;
    const int kb = dtype_bytes(dtype);
    if (!kb) return fail(TURTLE_ERR_DTYPE, "unknown dtype");
    if (max_n < 0 || max_n >= (int64_t)1 << 31) return fail(TURTLE_ERR_INVALID, "max_n must be in [0, 2^31)");
    (void)compress;   // generic compressible memory is an NVIDIA-only feature
    turtle_sorter *h = new turtle_sorter();
    h->dtype = dtype;
    h->kb = kb;
    h->kt = kt_of(dtype);
    h->flip = flip_of(dtype);
    h->track = track_index ? 1 : 0;
    h->max_n = max_n;
    h->tune = h->tune_base = default_tuning(kb);
    int rc = TURTLE_OK;
    hipDeviceProp_t pr;
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess) e = hipGetDeviceProperties(&pr, h->device);
    if (e != hipSuccess) {
        delete h;
        return fail_hip(e, "no HIP device");
    }
    h->cus = pr.multiProcessorCount;
    if ((rc = ensure_rank_order(h->device))) {
        delete h;
        return rc;
    }
    const size_t kbytes = (size_t)max_n * kb, ibytes = (size_t)max_n * 4;
    if ((rc = alloc_buf(&h->kin, kbytes, "hipMalloc(keys in)")) || (rc = alloc_buf(&h->kout, kbytes, "hipMalloc(keys out)")) ||
        (rc = alloc_buf(&h->kt0, kbytes, "hipMalloc(keys tmp0)")) || (rc = alloc_buf(&h->kt1, kbytes, "hipMalloc(keys tmp1)"))) {
        turtle_destroy(h);
        return rc;
    }
    if (h->track) {
        if ((rc = alloc_buf((void **)&h->iin, ibytes, "hipMalloc(idx in)")) ||
            (rc = alloc_buf((void **)&h->iout, ibytes, "hipMalloc(idx out)")) ||
            (rc = alloc_buf((void **)&h->it0, ibytes, "hipMalloc(idx tmp0)")) ||
            (rc = alloc_buf((void **)&h->it1, ibytes, "hipMalloc(idx tmp1)"))) {
            turtle_destroy(h);
            return rc;
        }
    }
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&h->ev0)) != hipSuccess || (e = hipEventCreate(&h->ev1)) != hipSuccess) {
        turtle_destroy(h);
        return fail_hip(e, "turtle_create");
    }
    h->last_stream = h->stream;
    if ((rc = ensure_tables(h, h->tune, h->tab, (uint64_t)max_n, 1, 1))) {
        turtle_destroy(h);
        return rc;
    }
    if ((e = hipDeviceSynchronize()) != hipSuccess) {
        turtle_destroy(h);
        return fail_hip(e, "turtle_create sync");
    }
    *out = h;
    return TURTLE_OK;
}

int turtle_destroy(turtle_handle h)
{
    if (!h) return TURTLE_OK;
    if (h->live_graphs > 0)
        return fail(TURTLE_ERR_STATE, "graphs captured on this handle are alive: destroy them first");
    (void)resolve(h);
    if (h->last_stream != h->stream) (void)hipStreamSynchronize(h->last_stream);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->stream2) (void)hipStreamSynchronize(h->stream2);
    delete h->timer;
    delete h->pool;
    void *bufs[] = {h->kin, h->kout, h->kt0, h->kt1, h->iin, h->iout, h->it0, h->it1, h->dev_off};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    free_tables(h->tab);
    if (h->ctrl_host) (void)hipHostFree(h->ctrl_host);
    if (h->levels_host) (void)hipHostFree(h->levels_host);
    for (int k = 0; k < 2; k++) {
        if (h->stage[k]) (void)hipHostFree(h->stage[k]);
        if (h->stage_ev[k]) (void)hipEventDestroy(h->stage_ev[k]);
    }
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->lv_ev) (void)hipEventDestroy(h->lv_ev);
    if (h->rb_ev) (void)hipEventDestroy(h->rb_ev);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    if (h->stream2) (void)hipStreamDestroy(h->stream2);
    delete h;
    return TURTLE_OK;
}

int turtle_compression_enabled(turtle_handle h)
{
    (void)h;
    return 0;
}

int turtle_set_profiling(turtle_handle h, int on)
{
    if (!h) return fail(TURTLE_ERR_INVALID, "NULL handle");
    if (on && !h->timer) h->timer = new EventTimer();
    if (!on && h->timer) {
        HIPCHK(hipStreamSynchronize(h->last_stream), "hipStreamSynchronize");
        delete h->timer;
        h->timer = nullptr;
    }
    return TURTLE_OK;
}

int turtle_set_profiling_kinds(turtle_handle h, uint32_t kind_mask)
{
    if (!h) return fail(TURTLE_ERR_INVALID, "NULL handle");
    if (kind_mask == 0) return turtle_set_profiling(h, 0);
    int rc = turtle_set_profiling(h, 1);
    if (rc) return rc;
    h->timer->mask = kind_mask;
    return TURTLE_OK;
}

int turtle_get_profile(turtle_handle h, turtle_profile *out)
{
    if (!h || !out) return fail(TURTLE_ERR_INVALID, "NULL argument");
    memset(out, 0, sizeof(*out));
    if (!h->timer) return fail(TURTLE_ERR_STATE, "profiling is off");
    EventTimer *t = h->timer;
    if (t->err != hipSuccess) {
        const hipError_t e = t->err;
        t->reset();
        return fail_hip(e, "launch timing");
    }
    if (!t->recs.empty()) HIPCHK(hipEventSynchronize(t->recs.back().b), "hipEventSynchronize");
    for (const EventTimer::Rec &r : t->recs) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
        if (r.kind >= 0 && r.kind < TURTLE_K_KINDS) {
            out->ms[r.kind] += ms;
            out->launches[r.kind] += 1;
        }
    }
    t->reset();
    return TURTLE_OK;
}

int turtle_get_tuning(turtle_handle h, turtle_tuning *t)
{
    if (!h || !t) return fail(TURTLE_ERR_INVALID, "NULL argument");
    *t = h->tune;
    return TURTLE_OK;
}

int turtle_set_tuning(turtle_handle h, const turtle_tuning *t)
{
    if (!h || !t) return fail(TURTLE_ERR_INVALID, "NULL argument");
    if (!valid_tuning(*t, h->kb)) return fail(TURTLE_ERR_INVALID, "tuning out of range");
    h->tune = h->tune_base = *t;
    h->tune_fixed = true;
    return TURTLE_OK;
}

// Times every candidate tuning (splitters 7/15/31 x LDS cutoff cap/4,
// cap/2, cap x tiles per step 1/2) on synthetic keys (uniform over the key
// type, iota index) in device memory: one untimed run (tables sized for
// the tuning), then reps timed runs of run(kin, iin, kout, iout).  The
// handle's own tuning is left as it was.
extern "C++" template <typename RUN>
static int autotune_candidates(turtle_handle h, int64_t n, bool idx, int reps, RUN &&run, turtle_tuning *best,
                               double *best_ms)
{
    const size_t kbytes = (size_t)n * h->kb, ibytes = (size_t)n * 4;
    void *kin = nullptr, *kout = nullptr;
    uint32_t *iin = nullptr, *iout = nullptr;
    hipStream_t st = h->stream;
    int rc = TURTLE_OK;
    auto cleanup = [&] {
        void *b[] = {kin, kout, iin, iout};
        for (void *x : b)
            if (x) (void)hipFree(x);
    };
    if ((rc = alloc_buf(&kin, kbytes, "hipMalloc(autotune keys)")) || (rc = alloc_buf(&kout, kbytes, "hipMalloc")) ||
        (idx && ((rc = alloc_buf((void **)&iin, ibytes, "hipMalloc")) || (rc = alloc_buf((void **)&iout, ibytes, "hipMalloc"))))) {
        cleanup();
        return rc;
    }
    hipError_t e = ts::fill_keys(h->kb, kin, (uint64_t)n, 0x7e5eed01ull, 0, st);
    if (e == hipSuccess && idx) e = ts::fill_iota(iin, (uint64_t)n, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        cleanup();
        return fail_hip(e, "autotune fill");
    }
    const turtle_tuning base = h->tune_base;
    const bool fixed = h->tune_fixed;
    h->tune_fixed = true;
    std::vector<turtle_tuning> cands;
    const int cap = (int)ts::engine_cap(h->kb);
    for (int S : {7, 15, 31})
        for (int cut : {cap / 4, cap / 2, cap})
            for (int tiles : {1, 2})
                cands.push_back(turtle_tuning{S, cut, tiles, 0});
    double best_t = 1e30;
    turtle_tuning best_c = base;
    for (const turtle_tuning &c : cands) {
        h->tune = h->tune_base = c;
        rc = run(kin, iin, kout, iout, st);
        if (!rc) rc = turtle_synchronize(h, nullptr);
        if (rc) break;
        const double t0 = now_s();
        for (int r = 0; r < reps && !rc; r++) rc = run(kin, iin, kout, iout, st);
        if (!rc) rc = turtle_synchronize(h, nullptr);
        if (rc) break;
        const double t = (now_s() - t0) / reps;
        if (t < best_t) {
            best_t = t;
            best_c = c;
        }
    }
    h->tune_base = base;
    h->tune_fixed = fixed;
    h->tune = base;
    cleanup();
    if (rc) return rc;
    *best = best_c;
    if (best_ms) *best_ms = best_t * 1e3;
    return TURTLE_OK;
}

int turtle_autotune(turtle_handle h, int64_t n, int with_index, int reps, turtle_tuning *best, double *best_ms)
{
    if (!h || !best) return fail(TURTLE_ERR_INVALID, "NULL argument");
    if (n < 1 || n > h->max_n) return fail(TURTLE_ERR_CAPACITY, "n must be in [1, max_n]");
    if (with_index && !h->track) return fail(TURTLE_ERR_INVALID, "handle created without index tracking");
    if (reps < 1) reps = 3;
    const bool idx = with_index != 0;
    return autotune_candidates(h, n, idx, reps,
                               [&](void *ki, uint32_t *ii, void *ko, uint32_t *io, hipStream_t st) {
                                   return run_engine(h, ki, idx ? ii : nullptr, ko, idx ? io : nullptr, (uint64_t)n,
                                                     nullptr, 1, st);
                               },
                               best, best_ms);
}

int turtle_autotune_segmented(turtle_handle h, const int64_t *offsets, int64_t num_segments, int with_index, int reps,
                              turtle_tuning *best, double *best_ms)
{
    if (!h || !best || !offsets) return fail(TURTLE_ERR_INVALID, "NULL argument");
    if (num_segments < 1) return fail(TURTLE_ERR_INVALID, "no segments");
    const int64_t n = offsets[num_segments] - offsets[0];
    if (offsets[0] != 0) return fail(TURTLE_ERR_INVALID, "offsets must start at 0");
    if (n < 1 || n > h->max_n) return fail(TURTLE_ERR_CAPACITY, "the segments must hold [1, max_n] keys");
    if (with_index && !h->track) return fail(TURTLE_ERR_INVALID, "handle created without index tracking");
    if (reps < 1) reps = 3;
    const bool idx = with_index != 0;
    return autotune_candidates(h, n, idx, reps,
                               [&](void *ki, uint32_t *ii, void *ko, uint32_t *io, hipStream_t st) {
                                   return turtle_segmented_sort_device(h, ki, idx ? (const int32_t *)ii : nullptr, ko,
                                                                       idx ? (int32_t *)io : nullptr, offsets,
                                                                       num_segments, st);
                               },
                               best, best_ms);
}

int turtle_sort_device(turtle_handle h, const void *keys_in, const int32_t *idx_in, void *keys_out, int32_t *idx_out,
                       int64_t n, void *stream)
{
    if (!h) return fail(TURTLE_ERR_INVALID, "NULL handle");
    if (n < 0) return fail(TURTLE_ERR_INVALID, "negative n");
    if (n > h->max_n) return fail(TURTLE_ERR_CAPACITY, "n exceeds the handle's max_n");
    if (n > 0 && (!keys_in || !keys_out)) return fail(TURTLE_ERR_INVALID, "NULL key buffer");
    if ((idx_in == nullptr) != (idx_out == nullptr)) return fail(TURTLE_ERR_INVALID, "idx_in and idx_out must both be set");
    if (idx_in && !h->track) return fail(TURTLE_ERR_INVALID, "handle created without index tracking");
    return run_engine(h, keys_in, (const uint32_t *)idx_in, keys_out, (uint32_t *)idx_out, (uint64_t)n, nullptr, 1,
                      stream_of(h, stream));
}

int turtle_segmented_sort_device(turtle_handle h, const void *keys_in, const int32_t *idx_in, void *keys_out,
                                 int32_t *idx_out, const int64_t *offsets, int64_t num_segments, void *stream)
{
    if (!h) return fail(TURTLE_ERR_INVALID, "NULL handle");
    if (num_segments < 0 || (num_segments > 0 && !offsets)) return fail(TURTLE_ERR_INVALID, "bad segment list");
    if ((idx_in == nullptr) != (idx_out == nullptr)) return fail(TURTLE_ERR_INVALID, "idx_in and idx_out must both be set");
    if (idx_in && !h->track) return fail(TURTLE_ERR_INVALID, "handle created without index tracking");
    if (num_segments >= (int64_t)1 << 30) return fail(TURTLE_ERR_CAPACITY, "too many segments");
    if (num_segments == 0) return TURTLE_OK;
    if (offsets[0] < 0) return fail(TURTLE_ERR_INVALID, "negative offset");
    for (int64_t s = 0; s < num_segments; s++)
        if (offsets[s + 1] < offsets[s]) return fail(TURTLE_ERR_INVALID, "offsets must be non-decreasing");
    const int64_t total = offsets[num_segments];
    if (total > h->max_n) return fail(TURTLE_ERR_CAPACITY, "segments exceed the handle's max_n");
    if (total > 0 && (!keys_in || !keys_out)) return fail(TURTLE_ERR_INVALID, "NULL key buffer");
    hipStream_t st = stream_of(h, stream);
    // the previous run's offsets may still be in use by the device (and by
    // the levels its resolution adds)
    int rc = resolve(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->last_stream), "hipStreamSynchronize");
    if (st != h->last_stream) HIPCHK(hipStreamSynchronize(st), "hipStreamSynchronize");
    const size_t need = (size_t)num_segments + 1;
    if (h->dev_off_cap < need) {
        if (h->dev_off) (void)hipFree(h->dev_off);
        h->dev_off = nullptr;
        h->dev_off_cap = 0;
        HIPCHK(hipMalloc((void **)&h->dev_off, need * sizeof(uint32_t)), "hipMalloc(offsets)");
        h->dev_off_cap = need;
    }
    h->host_off.resize(need);
    for (size_t s = 0; s < need; s++) h->host_off[s] = (uint32_t)offsets[s];
    int src = stage_h2d(h, h->dev_off, h->host_off.data(), need * sizeof(uint32_t), st);
    if (src) return src;
    return run_engine(h, keys_in, (const uint32_t *)idx_in, keys_out, (uint32_t *)idx_out, (uint64_t)total,
                      h->dev_off, (uint32_t)num_segments, st);
}

// MultiSort on device buffers, with an optional 32-bit payload (idx_out
// non-NULL: idx_in's values, or each key's position in its array when
// idx_in is NULL).  Arrays up to multisort_max_size keys: one workgroup per
// array; longer ones: equal segments through the sort engine.
static int multisort_device(turtle_handle h, int dtype, int64_t arr_size, const void *keys_in, const int32_t *idx_in,
                            void *keys_out, int32_t *idx_out, int64_t num_arrays, void *stream)
{
    if (!h) return fail(TURTLE_ERR_INVALID, "NULL handle");
    if (dtype != h->dtype) return fail(TURTLE_ERR_DTYPE, "dtype differs from the handle's");
    if (arr_size < 0 || num_arrays < 0) return fail(TURTLE_ERR_INVALID, "negative size");
    if (arr_size == 0 || num_arrays == 0) return TURTLE_OK;
    if (!keys_in || !keys_out) return fail(TURTLE_ERR_INVALID, "NULL key buffer");
    if (arr_size >= (int64_t)1 << 31 || num_arrays > ((int64_t)1 << 62) / arr_size)
        return fail(TURTLE_ERR_CAPACITY, "batch too large");
    const bool idx = idx_out != nullptr;
    if (idx && (uint64_t)arr_size * (uint64_t)num_arrays > 0xffffffffull)
        return fail(TURTLE_ERR_CAPACITY, "an indexed batch is limited to 2^32 keys");
    hipStream_t st = stream_of(h, stream);
    if ((uint64_t)arr_size <= ts::multisort_max_size(h->kt)) {
        const int rc = resolve(h);   // the previous engine run's level check
        if (rc) return rc;
        // (the launches carry the timing events themselves)
        HIPCHK(ts::multisort_run(h->kt, h->flip, (uint32_t)arr_size, keys_in, keys_out, (const uint32_t *)idx_in,
                                 (uint32_t *)idx_out, (uint64_t)num_arrays, st, h->ev0, h->ev1),
               "multisort launch");
        h->timed = true;
        h->last_stream = st;
        h->last_n = (uint64_t)arr_size * num_arrays;
        h->run = EngineRun{};
        return TURTLE_OK;
    }
    // longer arrays: equal segments through the sort engine
    const int64_t total = arr_size * num_arrays;
    if (total > h->max_n) return fail(TURTLE_ERR_CAPACITY, "arrays exceed the handle's max_n");
    if (idx && !h->track) return fail(TURTLE_ERR_INVALID, "the handle does not track indices");
    if (idx && !idx_in) {
        // positions inside each array, sorted in place
        HIPCHK(ts::fill_iota((uint32_t *)idx_out, (uint64_t)total, st, (uint64_t)arr_size), "fill_iota");
        idx_in = idx_out;
    }
    std::vector<int64_t> off((size_t)num_arrays + 1);
    for (int64_t a = 0; a <= num_arrays; a++) off[a] = a * arr_size;
    return turtle_segmented_sort_device(h, keys_in, idx ? idx_in : nullptr, keys_out, idx_out, off.data(), num_arrays,
                                        stream);
}

int turtle_multisort_device(turtle_handle h, int dtype, int64_t arr_size, const void *keys_in, void *keys_out,
                            int64_t num_arrays, void *stream)
{
    return multisort_device(h, dtype, arr_size, keys_in, nullptr, keys_out, nullptr, num_arrays, stream);
}

int turtle_multisort_index_device(turtle_handle h, int dtype, int64_t arr_size, const void *keys_in,
                                  const int32_t *idx_in, void *keys_out, int32_t *idx_out, int64_t num_arrays,
                                  void *stream)
{
    if (!idx_out) return fail(TURTLE_ERR_INVALID, "NULL index output");
    return multisort_device(h, dtype, arr_size, keys_in, idx_in, keys_out, idx_out, num_arrays, stream);
}

int turtle_synchronize(turtle_handle h, turtle_stats *st)
{
    if (!h) return fail(TURTLE_ERR_INVALID, "NULL handle");
    int rc = resolve(h);   // more levels, or a rerun with grown tables, when needed
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->last_stream), "hipStreamSynchronize");
    uint32_t err = 0;
    rc = read_stats(h, st, &err);
    if (rc) return rc;
    return error_status(err);
}

int turtle_query(turtle_handle h)
{
    if (!h) return -fail(TURTLE_ERR_INVALID, "NULL handle");
    const hipError_t e = hipStreamQuery(h->last_stream);
    if (e == hipErrorNotReady) return 1;
    if (e != hipSuccess) return -fail_hip(e, "hipStreamQuery");
    return 0;
}

int turtle_sort_start(turtle_handle h, void *keys, int32_t *idx, int64_t n)
{
    if (!h) return fail(TURTLE_ERR_INVALID, "NULL handle");
    if (n < 0) return fail(TURTLE_ERR_INVALID, "negative n");
    if (n > h->max_n) return fail(TURTLE_ERR_CAPACITY, "n exceeds the handle's max_n");
    if (n > 0 && !keys) return fail(TURTLE_ERR_INVALID, "NULL keys");
    if (idx && !h->track) return fail(TURTLE_ERR_INVALID, "handle created without index tracking");
    if (h->started) return fail(TURTLE_ERR_STATE, "StartSorting called twice without Sync");
    h->t_start = now_s();
    // the previous run's level check, and its use of the input buffers
    int rc = resolve(h);
    if (rc) return rc;
    if (h->last_stream != h->stream) HIPCHK(hipStreamSynchronize(h->last_stream), "hipStreamSynchronize");
    h->host_keys = keys;
    h->host_idx = idx;
    h->host_n = n;
    if (n > 0) {
        if ((rc = stage_h2d(h, h->kin, keys, (size_t)n * h->kb, h->stream))) return rc;
        if (idx && (rc = stage_h2d(h, h->iin, idx, (size_t)n * 4, h->stream))) return rc;
    }
    rc = run_engine(h, h->kin, idx ? h->iin : nullptr, h->kout, idx ? h->iout : nullptr, (uint64_t)n, nullptr, 1,
                        h->stream);
    if (rc) return rc;
    h->started = true;
    return TURTLE_OK;
}

int turtle_sort_sync(turtle_handle h, double *seconds)
{
    if (!h) return fail(TURTLE_ERR_INVALID, "NULL handle");
    if (!h->started) return fail(TURTLE_ERR_STATE, "Sync without StartSorting");
    h->started = false;
    // the level check of the run (more levels, or a rerun from the device
    // input buffer with grown tables after an overflow), then the wait
    int rc = turtle_synchronize(h, nullptr);
    if (rc) return rc;
    if (h->host_n > 0) {
        if ((rc = stage_d2h(h, h->host_keys, h->kout, (size_t)h->host_n * h->kb, h->stream))) return rc;
        if (h->host_idx && (rc = stage_d2h(h, h->host_idx, h->iout, (size_t)h->host_n * 4, h->stream))) return rc;
    }
    if (seconds) *seconds = now_s() - h->t_start;
    h->host_keys = nullptr;
    h->host_idx = nullptr;
    return TURTLE_OK;
}

// Host MultiSort.  Arrays that one workgroup sorts go in chunks of whole
// arrays through the two pinned staging slots: chunk c is copied into slot
// c&1 by the copy threads, then its H2D copy, sort and D2H copy are queued
// on stream c&1, and the slot's previous chunk (c-2) is copied out to the
// user's buffer first.  So the host copies of one chunk overlap the
// transfers and the sort of the other, and the two streams' H2D and D2H
// copies overlap on the link.  Longer arrays: one staged batch through the
// sort engine.  idx != NULL: a payload per key (local: output only, each
// key's position in its array); the handle must track indices.
static int multisort_host(turtle_handle h, int64_t arr_size, int64_t num_arrays, void *keys, int32_t *idx,
                          bool local)
{
    const int64_t total = arr_size * num_arrays;
    if (total == 0) return TURTLE_OK;
    if (!keys || (local && !idx)) return fail(TURTLE_ERR_INVALID, "NULL host data");
    if (idx && !h->track) return fail(TURTLE_ERR_INVALID, "the handle does not track indices");
    const size_t kb = (size_t)h->kb, ib = idx ? 4u : 0u;
    const size_t kalign = 256;
    const size_t per_arr = (size_t)arr_size * (kb + ib) + kalign;
    int rc = resolve(h);   // the previous engine run's level check
    if (rc) return rc;
    if ((rc = ensure_staging(h))) return rc;
    if ((uint64_t)arr_size > ts::multisort_max_size(h->kt) || per_arr > STAGE_BYTES) {
        if ((rc = stage_h2d(h, h->kin, keys, (size_t)total * kb, h->stream))) return rc;
        if (idx && !local && (rc = stage_h2d(h, h->iin, idx, (size_t)total * 4, h->stream))) return rc;
        rc = multisort_device(h, h->dtype, arr_size, h->kin, idx && !local ? (const int32_t *)h->iin : nullptr,
                              h->kout, idx ? (int32_t *)h->iout : nullptr, num_arrays, h->stream);
        if (rc) return rc;
        if ((rc = turtle_synchronize(h, nullptr))) return rc;
        if ((rc = stage_d2h(h, keys, h->kout, (size_t)total * kb, h->stream))) return rc;
        if (idx && (rc = stage_d2h(h, idx, h->iout, (size_t)total * 4, h->stream))) return rc;
        return TURTLE_OK;
    }
    if (!h->stream2) HIPCHK(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking), "hipStreamCreate");
    hipStream_t sts[2] = {h->stream, h->stream2};
    // order both streams after the handle's earlier work, and the slots
    // after their last transfers
    HIPCHK(hipStreamSynchronize(h->last_stream), "hipStreamSynchronize");
    for (int k = 0; k < 2; k++)
        if (h->stage_used[k]) HIPCHK(hipEventSynchronize(h->stage_ev[k]), "hipEventSynchronize");
    const int64_t cha = (int64_t)((STAGE_BYTES - kalign) / ((size_t)arr_size * (kb + ib)));   // arrays per chunk
    const int64_t nch = (num_arrays + cha - 1) / cha;
    int64_t ch_arrays[2] = {0, 0}, ch_first[2] = {0, 0};
    auto idx_off = [&](int64_t nk) { return ((size_t)nk * kb + kalign - 1) & ~(kalign - 1); };
    auto retire = [&](int k) -> int {
        HIPCHK(hipEventSynchronize(h->stage_ev[k]), "hipEventSynchronize");
        const int64_t nk = ch_arrays[k] * arr_size, e0 = ch_first[k] * arr_size;
        const char *slot = (const char *)h->stage[k];
        h->pool->copy((char *)keys + (size_t)e0 * kb, slot, (size_t)nk * kb);
        if (idx) h->pool->copy(idx + e0, slot + idx_off(nk), (size_t)nk * 4);
        ch_arrays[k] = 0;
        return TURTLE_OK;
    };
    for (int64_t c = 0; c < nch; c++) {
        const int k = (int)(c & 1);
        if (ch_arrays[k] && (rc = retire(k))) return rc;
        const int64_t a0 = c * cha, na = num_arrays - a0 < cha ? num_arrays - a0 : cha;
        const int64_t nk = na * arr_size, e0 = a0 * arr_size;
        char *slot = (char *)h->stage[k];
        h->pool->copy(slot, (const char *)keys + (size_t)e0 * kb, (size_t)nk * kb);
        if (idx && !local) h->pool->copy(slot + idx_off(nk), idx + e0, (size_t)nk * 4);
        void *dkin = (char *)h->kin + (size_t)e0 * kb, *dkout = (char *)h->kout + (size_t)e0 * kb;
        uint32_t *diin = idx ? h->iin + e0 : nullptr, *diout = idx ? h->iout + e0 : nullptr;
        hipStream_t st = sts[k];
        HIPCHK(hipMemcpyAsync(dkin, slot, (size_t)nk * kb, hipMemcpyHostToDevice, st), "H2D");
        if (idx && !local) HIPCHK(hipMemcpyAsync(diin, slot + idx_off(nk), (size_t)nk * 4, hipMemcpyHostToDevice, st), "H2D");
        HIPCHK(ts::multisort_run(h->kt, h->flip, (uint32_t)arr_size, dkin, dkout, idx && !local ? diin : nullptr,
                                 diout, (uint64_t)na, st),
               "multisort launch");
        HIPCHK(hipMemcpyAsync(slot, dkout, (size_t)nk * kb, hipMemcpyDeviceToHost, st), "D2H");
        if (idx) HIPCHK(hipMemcpyAsync(slot + idx_off(nk), diout, (size_t)nk * 4, hipMemcpyDeviceToHost, st), "D2H");
        HIPCHK(hipEventRecord(h->stage_ev[k], st), "hipEventRecord");
        h->stage_used[k] = true;
        ch_arrays[k] = na;
        ch_first[k] = a0;
    }
    for (int64_t c = nch; c < nch + 2; c++) {
        const int k = (int)(c & 1);
        if (ch_arrays[k] && (rc = retire(k))) return rc;
    }
    h->timed = false;   // no single device interval to report
    h->last_stream = h->stream;
    h->run = EngineRun{};
    return TURTLE_OK;
}

int turtle_multisort(turtle_handle h, int dtype, int64_t arr_size, int block_size, int use_lds, int64_t num_arrays,
                     void *host_data, double *seconds)
{
    (void)block_size;
    (void)use_lds;
    if (!h) return fail(TURTLE_ERR_INVALID, "NULL handle");
    const double t0 = now_s();
    if (dtype != h->dtype) return fail(TURTLE_ERR_DTYPE, "dtype differs from the handle's");
    if (arr_size < 0 || num_arrays < 0) return fail(TURTLE_ERR_INVALID, "negative size");
    if (arr_size > 0 && num_arrays > h->max_n / (arr_size ? arr_size : 1))
        return fail(TURTLE_ERR_CAPACITY, "numArrays*arrSize exceeds the handle's max_n");
    int rc = multisort_host(h, arr_size, num_arrays, host_data, nullptr, false);
    if (rc) return rc;
    if (seconds) *seconds = now_s() - t0;
    return TURTLE_OK;
}

int turtle_multisort_index(turtle_handle h, int dtype, int64_t arr_size, int64_t num_arrays, void *host_keys,
                           int32_t *host_idx, int local_positions, double *seconds)
{
    if (!h) return fail(TURTLE_ERR_INVALID, "NULL handle");
    const double t0 = now_s();
    if (dtype != h->dtype) return fail(TURTLE_ERR_DTYPE, "dtype differs from the handle's");
    if (arr_size < 0 || num_arrays < 0) return fail(TURTLE_ERR_INVALID, "negative size");
    if (arr_size > 0 && num_arrays > h->max_n / (arr_size ? arr_size : 1))
        return fail(TURTLE_ERR_CAPACITY, "numArrays*arrSize exceeds the handle's max_n");
    if (arr_size * num_arrays > 0 && !host_idx) return fail(TURTLE_ERR_INVALID, "NULL host index");
    int rc = multisort_host(h, arr_size, num_arrays, host_keys, host_idx, local_positions != 0);
    if (rc) return rc;
    if (seconds) *seconds = now_s() - t0;
    return TURTLE_OK;
}

struct turtle_graph {
    turtle_handle h = nullptr;
    uint32_t *dev_off = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t total = 0;
    uint32_t nseg = 0;
    uint32_t nlev = 0;        // partition levels captured
    EngineTables tab;         // the graph's own engine tables
    bool counted = false;     // counted in h->live_graphs
};

int turtle_graph_destroy(turtle_graph_handle g)
{
    if (!g) return TURTLE_OK;
    turtle_handle h = g->h;
    if (h && h->run.tab == &g->tab) {
        // the handle's last run is a replay of this graph: let it finish
        (void)hipStreamSynchronize(h->last_stream);
        h->run = EngineRun{};
    }
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->dev_off) (void)hipFree(g->dev_off);
    free_tables(g->tab);
    if (h && g->counted) h->live_graphs--;
    delete g;
    return TURTLE_OK;
}

int turtle_segmented_graph_create(turtle_handle h, const void *keys_in, const int32_t *idx_in, void *keys_out,
                                  int32_t *idx_out, const int64_t *offsets, int64_t num_segments,
                                  turtle_graph_handle *out)
{
    if (!out) return fail(TURTLE_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!h) return fail(TURTLE_ERR_INVALID, "NULL handle");
    if (num_segments <= 0 || !offsets) return fail(TURTLE_ERR_INVALID, "bad segment list");
    if ((idx_in == nullptr) != (idx_out == nullptr)) return fail(TURTLE_ERR_INVALID, "idx_in and idx_out must both be set");
    if (idx_in && !h->track) return fail(TURTLE_ERR_INVALID, "handle created without index tracking");
    if (num_segments >= (int64_t)1 << 30) return fail(TURTLE_ERR_CAPACITY, "too many segments");
    if (offsets[0] < 0) return fail(TURTLE_ERR_INVALID, "negative offset");
    for (int64_t s = 0; s < num_segments; s++)
        if (offsets[s + 1] < offsets[s]) return fail(TURTLE_ERR_INVALID, "offsets must be non-decreasing");
    const int64_t total = offsets[num_segments];
    if (total > h->max_n) return fail(TURTLE_ERR_CAPACITY, "segments exceed the handle's max_n");
    if (total > 0 && (!keys_in || !keys_out)) return fail(TURTLE_ERR_INVALID, "NULL key buffer");
    if (total > 0 && keys_in == keys_out) return fail(TURTLE_ERR_INVALID, "graph form needs keys_out != keys_in");
    // the handle's previous run first (its level check may queue work)
    int rc = resolve(h);
    if (rc) return rc;
    turtle_graph *g = new turtle_graph();
    g->h = h;
    g->total = (uint64_t)total;
    g->nseg = (uint32_t)num_segments;
    h->live_graphs++;
    g->counted = true;
    std::vector<uint32_t> off((size_t)num_segments + 1);
    for (size_t s = 0; s < off.size(); s++) off[s] = (uint32_t)offsets[s];
    h->host_off = off;   // the level estimate reads the longest segment from here
    // everything the run allocates is allocated before the capture starts:
    // the graph's own tables, sized for its tuning
    h->tune = tuning_for(h, idx_in != nullptr, longest_root(h, true, (uint64_t)total), true);
    rc = ensure_tables(h, h->tune, g->tab, (uint64_t)total, (uint32_t)num_segments, 1);
    if (!rc) rc = ensure_readback(h);
    hipError_t e = hipSuccess;
    if (!rc) e = hipMalloc((void **)&g->dev_off, off.size() * sizeof(uint32_t));
    if (!rc && e == hipSuccess) {
        rc = stage_h2d(h, g->dev_off, off.data(), off.size() * sizeof(uint32_t), h->stream);
        if (!rc) e = hipStreamSynchronize(h->stream);
    }
    hipStream_t cs = nullptr;
    if (!rc && e == hipSuccess) e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (!rc && e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (!rc && e == hipSuccess) {
        rc = run_engine(h, keys_in, (const uint32_t *)idx_in, keys_out, (uint32_t *)idx_out, (uint64_t)total,
                        g->dev_off, (uint32_t)num_segments, cs, &g->tab, &g->nlev);
        hipGraph_t graph = nullptr;
        hipError_t e2 = hipStreamEndCapture(cs, &graph);
        g->graph = graph;
        if (!rc && e2 != hipSuccess) e = e2;
    }
    if (!rc && e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (cs) (void)hipStreamDestroy(cs);
    if (rc || e != hipSuccess) {
        turtle_graph_destroy(g);
        return rc ? rc : fail_hip(e, "graph capture");
    }
    *out = g;
    return TURTLE_OK;
}

int turtle_graph_launch(turtle_graph_handle g, void *stream)
{
    if (!g) return fail(TURTLE_ERR_INVALID, "NULL graph");
    turtle_handle h = g->h;
    int rc = resolve(h);   // an eager run's level check queues nothing after the replay
    if (rc) return rc;
    hipStream_t st = stream_of(h, stream);
    HIPCHK(hipGraphLaunch(g->exec, st), "hipGraphLaunch");
    EngineRun r;
    r.engine = true;
    r.graph = true;
    r.tab = &g->tab;
    r.total = g->total;
    r.nseg = g->nseg;
    r.st = st;
    r.levels = g->nlev;   // turtle_synchronize reads this level's record
    h->run = r;
    h->last_stream = st;
    h->last_n = g->total;
    h->timed = false;
    return TURTLE_OK;
}

}  // extern "C"

// ---- allocation-free sort -------------------------------------------------------
// turtle_sort_device_noalloc keeps every buffer of an engine run beyond the
// caller's inputs and outputs in `scratch`, a file-scope device array sized
// for NA_MAX_N keys of any width with an index: the ping-pong key/index
// buffers, a copy of an in-place call's input, the control block, segment,
// tile and LDS-task tables, level records and per-key bins.  The tables are
// sized as ensure_tables sizes them for NA_MAX_N keys in one root at the
// smallest cutoff and tile (8-byte keys: 4096).  Host code reaches the
// array only through the address hipGetSymbolAddress returns for the
// current device (kept per device); the array's name is not a device
// address on the host.
namespace {

constexpr uint64_t NA_MAX_N = 1ull << 20;
constexpr uint64_t NA_S = 31;
constexpr uint64_t NA_CUT8 = 4096, NA_CUT4 = 8192;    // LDS cutoff = tile: 8-byte keys, narrower keys
constexpr uint64_t NA_SEGCAP = 8 * (NA_MAX_N / NA_CUT8 + 1) + 1024 + 1;
constexpr uint64_t NA_TILECAP = NA_MAX_N / NA_CUT8 + 1 + NA_SEGCAP;
constexpr uint64_t NA_LDSCAP = NA_SEGCAP * (NA_S + 1) + 1 + 1024;
constexpr uint64_t na_bins_tiles(uint64_t tile) { return NA_MAX_N / tile + NA_MAX_N / (tile + 1) + 2; }
constexpr uint64_t NA_BINS = na_bins_tiles(NA_CUT8) * NA_CUT8 > na_bins_tiles(NA_CUT4) * NA_CUT4
                                 ? na_bins_tiles(NA_CUT8) * NA_CUT8
                                 : na_bins_tiles(NA_CUT4) * NA_CUT4;
constexpr uint64_t na_al(uint64_t x) { return (x + 255) & ~255ull; }
// byte offsets inside the scratch array
constexpr uint64_t NA_O_KT0 = 0;
constexpr uint64_t NA_O_KT1 = NA_O_KT0 + na_al(8 * NA_MAX_N);
constexpr uint64_t NA_O_KIN = NA_O_KT1 + na_al(8 * NA_MAX_N);
constexpr uint64_t NA_O_IT0 = NA_O_KIN + na_al(8 * NA_MAX_N);
constexpr uint64_t NA_O_IT1 = NA_O_IT0 + na_al(4 * NA_MAX_N);
constexpr uint64_t NA_O_IIN = NA_O_IT1 + na_al(4 * NA_MAX_N);
constexpr uint64_t NA_O_CTRL = NA_O_IIN + na_al(4 * NA_MAX_N);
constexpr uint64_t NA_O_SEGS = NA_O_CTRL + na_al(sizeof(ts::Ctrl));
constexpr uint64_t NA_O_TCNT = NA_O_SEGS + na_al(NA_SEGCAP * sizeof(ts::Seg));
constexpr uint64_t NA_O_TPRE = NA_O_TCNT + na_al(NA_TILECAP * ts::MAX_BINS * 4);
constexpr uint64_t NA_O_SPART = NA_O_TPRE + na_al(NA_TILECAP * ts::MAX_BINS * 4);
constexpr uint64_t NA_O_TSEG = NA_O_SPART + na_al((NA_TILECAP * ts::MAX_BINS / ts::SCAN_CHUNK + 2) * 4);
constexpr uint64_t NA_O_LDST = NA_O_TSEG + na_al(NA_TILECAP * 4);
constexpr uint64_t NA_O_LDSQ = NA_O_LDST + na_al(NA_LDSCAP * sizeof(ts::LdsTask));
constexpr uint64_t NA_O_LEVELS = NA_O_LDSQ + na_al(NA_LDSCAP * sizeof(ts::LdsTask));
constexpr uint64_t NA_O_BINS = NA_O_LEVELS + na_al(ts::MAX_LEVELS * sizeof(ts::Level));
constexpr uint64_t NA_BYTES = NA_O_BINS + na_al(NA_BINS);
static_assert(NA_BYTES < (64ull << 20), "the allocation-free scratch stays below 64 MiB");

constexpr int NA_MAX_DEVICES = 64;
struct NoAllocDevice {
    std::mutex mu;                      // serialises the calls on this device
    unsigned char *base = nullptr;      // device address of `scratch` on this device
    int cus = 0;
};
NoAllocDevice g_na[NA_MAX_DEVICES];
std::mutex g_na_init;

}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {

// address of this device's scratch array (looked up once per device, with
// that device current)
int na_device(int dev, NoAllocDevice **out)
{
    if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(TURTLE_ERR_UNSUPPORTED, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_na_init);
    NoAllocDevice &d = g_na[dev];
    if (!d.base) {
        int rc = ensure_rank_order(dev);
        if (rc) return rc;
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        int cus = 0;
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        d.cus = cus;
        d.base = (unsigned char *)ptr;
    }
    *out = &d;
    return TURTLE_OK;
}

}  // namespace

extern "C" {

int turtle_sort_device_noalloc(int dtype, const void *keys_in, const int32_t *idx_in, void *keys_out,
                               int32_t *idx_out, int64_t n, void *stream)
{
    const int kb = dtype_bytes(dtype);
    if (!kb) return fail(TURTLE_ERR_DTYPE, "unknown dtype");
    if (n < 0) return fail(TURTLE_ERR_INVALID, "negative n");
    if ((uint64_t)n > NA_MAX_N) return fail(TURTLE_ERR_CAPACITY, "the allocation-free sort takes at most 1,048,576 keys");
    if ((idx_in == nullptr) != (idx_out == nullptr)) return fail(TURTLE_ERR_INVALID, "idx_in and idx_out must both be set");
    if (n == 0) return TURTLE_OK;
    if (!keys_in || !keys_out) return fail(TURTLE_ERR_INVALID, "NULL key buffer");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev), "hipGetDevice");
    NoAllocDevice *d = nullptr;
    int rc = na_device(dev, &d);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(d->mu);
    hipStream_t st = (hipStream_t)stream;
    unsigned char *b = d->base;
    const bool idx = idx_in != nullptr;
    const void *kin = keys_in;
    const uint32_t *iin = (const uint32_t *)idx_in;
    // an in-place call reads its input from a copy (the levels write the output)
    if (keys_in == keys_out) {
        HIPCHK(hipMemcpyAsync(b + NA_O_KIN, keys_in, (size_t)n * kb, hipMemcpyDeviceToDevice, st), "hipMemcpyAsync");
        kin = b + NA_O_KIN;
    }
    if (idx && (const void *)idx_in == (const void *)idx_out) {
        HIPCHK(hipMemcpyAsync(b + NA_O_IIN, idx_in, (size_t)n * 4, hipMemcpyDeviceToDevice, st), "hipMemcpyAsync");
        iin = (const uint32_t *)(b + NA_O_IIN);
    }
    const uint32_t cutoff = ts::engine_cap(kb), tile = ts::engine_tile(kb);
    ts::Params p{};
    p.kin = kin;
    p.iin = idx ? iin : nullptr;
    p.kout = keys_out;
    p.iout = idx ? (uint32_t *)idx_out : nullptr;
    p.kt0 = b + NA_O_KT0;
    p.kt1 = b + NA_O_KT1;
    p.it0 = (uint32_t *)(b + NA_O_IT0);
    p.it1 = (uint32_t *)(b + NA_O_IT1);
    p.ctrl = (ts::Ctrl *)(b + NA_O_CTRL);
    p.segs = (ts::Seg *)(b + NA_O_SEGS);
    p.tcnt = (uint32_t *)(b + NA_O_TCNT);
    p.tpre = (uint32_t *)(b + NA_O_TPRE);
    p.spart = (uint32_t *)(b + NA_O_SPART);
    p.tile_seg = (uint32_t *)(b + NA_O_TSEG);
    p.ldst = (ts::LdsTask *)(b + NA_O_LDST);
    p.ldsq = (ts::LdsTask *)(b + NA_O_LDSQ);
    p.levels = (ts::Level *)(b + NA_O_LEVELS);
    p.bins = b + NA_O_BINS;
    p.segcap = (uint32_t)NA_SEGCAP;
    p.tilecap = (uint32_t)NA_TILECAP;
    p.ldscap = (uint32_t)NA_LDSCAP;
    p.bins_tiles = (uint32_t)na_bins_tiles(tile);
    p.cutoff = cutoff;
    p.tiles_per_wg = 1;
    p.n_total = (uint32_t)n;
    p.flip = flip_of(dtype);
    const int kt = kt_of(dtype);
    // the levels in batches (the control block is cleared by the first
    // batch's reset kernel), a small read-back after each, then the LDS
    // sorts once
    uint32_t nlev = levels_estimate((uint64_t)n, cutoff, NA_S), L0 = 0;
    ts::Level lv{};
    ts::Ctrl c{};
    for (bool first = true;; first = false) {
        if (L0 + nlev > (uint32_t)ts::MAX_LEVELS) nlev = ts::MAX_LEVELS - L0;
        HIPCHK(ts::engine_levels(p, kt, idx, (int)NA_S, d->cus, 0, nullptr, 1, (uint32_t)n, L0, nlev, first, false, st),
               "engine launch");
        L0 += nlev;
        if (L0 == 0) break;   // every key fits one LDS sort
        HIPCHK(hipMemcpyAsync(&lv, p.levels + (L0 - 1), sizeof lv, hipMemcpyDeviceToHost, st), "D2H(level)");
        HIPCHK(hipMemcpyAsync(&c, p.ctrl, sizeof c, hipMemcpyDeviceToHost, st), "D2H(ctrl)");
        HIPCHK(hipStreamSynchronize(st), "hipStreamSynchronize");
        const uint32_t left = c.seg_alloc > lv.se ? c.seg_alloc - lv.se : 0u;
        if (left == 0 || c.error != 0 || L0 >= (uint32_t)ts::MAX_LEVELS) break;
        nlev = 4;
    }
    HIPCHK(ts::engine_levels(p, kt, idx, (int)NA_S, d->cus, 0, nullptr, 1, (uint32_t)n, L0, 0, false, true, st),
           "engine launch");
    HIPCHK(hipMemcpyAsync(&c, p.ctrl, sizeof c, hipMemcpyDeviceToHost, st), "D2H(ctrl)");
    HIPCHK(hipStreamSynchronize(st), "hipStreamSynchronize");
    uint32_t err = c.error;
    if (L0 > 0 && c.seg_alloc > lv.se) err |= ts::ERR_DEPTH;
    return error_status(err);
}

int turtle_merge_device(int dtype, const void *keys_a, const int32_t *idx_a, int64_t na, const void *keys_b,
                        const int32_t *idx_b, int64_t nb, void *keys_out, int32_t *idx_out, void *stream)
{
    const int kb = dtype_bytes(dtype);
    if (!kb) return fail(TURTLE_ERR_DTYPE, "unknown dtype");
    if (na < 0 || nb < 0) return fail(TURTLE_ERR_INVALID, "negative length");
    if (na + nb >= ((int64_t)1 << 40)) return fail(TURTLE_ERR_CAPACITY, "runs too long");
    if (na + nb == 0) return TURTLE_OK;
    // an empty run's pointers may be NULL
    const bool idx = idx_out != nullptr;
    if (idx ? ((na && !idx_a) || (nb && !idx_b)) : (idx_a || idx_b))
        return fail(TURTLE_ERR_INVALID, "idx_a, idx_b and idx_out must all be set or all be NULL");
    if ((na && !keys_a) || (nb && !keys_b) || !keys_out) return fail(TURTLE_ERR_INVALID, "NULL key buffer");
    if (keys_out == keys_a || keys_out == keys_b || (idx && (idx_out == idx_a || idx_out == idx_b)))
        return fail(TURTLE_ERR_INVALID, "the output must not overlap the runs");
    HIPCHK(ts::merge_run(kt_of(dtype), flip_of(dtype), keys_a, (const uint32_t *)idx_a, (uint64_t)na, keys_b,
                         (const uint32_t *)idx_b, (uint64_t)nb, keys_out, (uint32_t *)idx_out, (hipStream_t)stream),
           "merge launch");
    return TURTLE_OK;
}

int turtle_fill_keys(int dtype, void *keys, int64_t n, uint64_t seed, int shift, void *stream)
{
    const int kb = dtype_bytes(dtype);
    if (!kb) return fail(TURTLE_ERR_DTYPE, "unknown dtype");
    if (n < 0 || shift < 0 || shift > 63 || (n > 0 && !keys)) return fail(TURTLE_ERR_INVALID, "bad argument");
    HIPCHK(ts::fill_keys(kb, keys, (uint64_t)n, seed, shift, (hipStream_t)stream), "fill_keys");
    return TURTLE_OK;
}

int turtle_fill_iota(int32_t *idx, int64_t n, void *stream)
{
    if (n < 0 || (n > 0 && !idx)) return fail(TURTLE_ERR_INVALID, "bad argument");
    HIPCHK(ts::fill_iota((uint32_t *)idx, (uint64_t)n, (hipStream_t)stream), "fill_iota");
    return TURTLE_OK;
}

}  // extern "C"
