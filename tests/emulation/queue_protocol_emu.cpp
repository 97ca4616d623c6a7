// queue_protocol_emu.cpp -- host-thread emulation of the persistent,
// queue-drained sort engine that the library shipped before its level-
// synchronous engine (git show 7b6f2b4:turtlesort_amd/csrc/ts_engine.hip;
// the first version, 9dc7641, is the same without the range checks).
//
// CPU only.  Each workgroup of the persistent kernel is one std::thread that
// runs the kernel's drain loop; the loops over threadIdx inside a task run
// serially in that thread, so block barriers vanish.  Device memory is
// allocated at exactly the sizes the library allocated (handle buffers of
// max_n elements, queue / segment / tile tables by the same formulas), so
// AddressSanitizer sees any store that leaves them.
//
// The device's agent-scope operations map onto the C++ memory model:
//   relaxed atomics (head, tail, outstanding, counters, error) -> relaxed
//   payload stores; fence_release; store_relaxed(slot.gen)     -> payload stores; gen.store(release)
//   load_relaxed(slot.gen) == gen; fence_acquire; payload loads -> gen.load(acquire); payload loads
//   fence_release; barrier; atomic_add(pending, -1); last: fence_acquire
//                                                               -> pending.fetch_sub(1, acq_rel)
// Each fence in the engine is directly followed (or, for acquire, preceded)
// by the one atomic access its partner observes, so the pairs have the same
// synchronizes-with edges as the release/acquire accesses used here
// ([atomics.fences]); ThreadSanitizer models the latter but not standalone
// fences.
//
// Tile sizes are scaled down (64 emulated threads x 16 items) so that small
// inputs already take several levels; the protocol does not depend on them.
//
// Usage: queue_protocol_emu <scenario> [n] [workgroups] [seed] [checks]
//   protocol   sort random inputs of many shapes, sizes and workgroup
//              counts with scheduling jitter; every run must end with
//              outstanding == 0, no error, and the keys of std::sort
//   race       as the device-form tests ran it: the input's producer (a
//              null-stream kernel, unordered with the handle's non-blocking
//              stream) writes the keys while the engine runs
//   between    the producer's write lands between the root segment's count
//              pass and its scatter pass (the interleaving that turns the
//              race into out-of-segment stores); checks=0 is the first
//              version, checks=1 the version with range checks
// Prints one line "result: ..." per run.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

namespace emu {

constexpr int S = 3;                 // splitters (the engine supported S = 3)
constexpr int NB = 2 * S + 1;        // bins LL,=p0,L,=p1,R,=p2,RR
constexpr uint32_t NT = 64;          // emulated threads per workgroup
constexpr uint32_t TI = 16;          // tile items per thread
constexpr uint32_t TILE = NT * TI;   // keys per count / scatter tile
constexpr uint32_t CAP = NT * TI;    // largest LDS-sorted segment
constexpr uint32_t SAMPLE = 256;     // pivot sample

enum : uint32_t { ERR_QUEUE = 1u, ERR_SEGS = 2u, ERR_TILES = 4u, ERR_TIMEOUT = 8u, ERR_BAD = 16u, ERR_RANGE = 32u };
enum : uint32_t { T_PIVOT = 1, T_COUNT = 2, T_SCATTER = 3, T_LDS = 4 };

struct Ctrl {
    std::atomic<uint32_t> gen{0}, head{0}, tail{0};
    std::atomic<int32_t> outstanding{0};
    std::atomic<uint32_t> seg_alloc{0}, tile_alloc{0}, error{0}, max_depth{0}, bad_where{0};
    std::atomic<uint64_t> tasks[5];
};

struct Slot {
    std::atomic<uint32_t> gen{0};
    uint32_t type = 0, a = 0, b = 0;
};

struct Seg {
    uint32_t begin = 0, len = 0, depth = 0, ntiles = 0;
    uint32_t tile_base = 0, ntasks = 0;
    std::atomic<uint32_t> count_pending{0}, scatter_pending{0};
    uint32_t split[S] = {};
    uint32_t base[NB + 1] = {};
    uint32_t total[NB + 1] = {};
};

// Hook called by the count pass's last finisher of a depth-0 segment before
// it pushes the scatter tasks (the "between" scenario writes the input there).
typedef void (*between_fn)(void *ctx, uint32_t begin, uint32_t len);

struct Params {
    const uint32_t *kin = nullptr;
    const uint32_t *iin = nullptr;
    uint32_t *kout = nullptr, *iout = nullptr;
    uint32_t *kt0 = nullptr, *kt1 = nullptr, *it0 = nullptr, *it1 = nullptr;
    Ctrl *ctrl = nullptr;
    Slot *q = nullptr;
    Seg *segs = nullptr;
    uint32_t *tiles = nullptr;
    uint32_t qcap = 0, segcap = 0, tilecap = 0, cutoff = 0, tiles_per_task = 1, n_total = 0;
    bool checks = true;
    double timeout_s = 20.0;
    uint32_t jitter = 0;   // 1 in `jitter` protocol points yields (0: never)
    between_fn between = nullptr;
    void *between_ctx = nullptr;
};

thread_local std::minstd_rand tl_rng;

inline void jitter(const Params &p)
{
    if (p.jitter && tl_rng() % p.jitter == 0) std::this_thread::yield();
}

inline void range_fail(const Params &p, uint32_t where)
{
    if (p.ctrl->error.fetch_or(ERR_RANGE, std::memory_order_relaxed) == 0)
        p.ctrl->bad_where.store(where, std::memory_order_relaxed);
}

inline bool span_ok(const Params &p, uint32_t begin, uint32_t len)
{
    return len > 0 && begin <= p.n_total && len <= p.n_total - begin;
}

inline const uint32_t *src_keys(const Params &p, uint32_t depth)
{
    return depth == 0 ? p.kin : ((depth & 1) ? p.kt0 : p.kt1);
}
inline const uint32_t *src_idx(const Params &p, uint32_t depth)
{
    return depth == 0 ? p.iin : ((depth & 1) ? p.it0 : p.it1);
}
inline uint32_t *next_keys(const Params &p, uint32_t depth) { return (depth & 1) ? p.kt1 : p.kt0; }
inline uint32_t *next_idx(const Params &p, uint32_t depth) { return (depth & 1) ? p.it1 : p.it0; }

inline uint32_t classify(uint32_t x, const uint32_t (&spl)[S])
{
    uint32_t lt = 0, le = 0;
    for (int j = 0; j < S; j++) {
        lt += spl[j] < x ? 1u : 0u;
        le += spl[j] <= x ? 1u : 0u;
    }
    return le > lt ? 2 * lt + 1 : 2 * lt;
}

// bulk push of `count` tasks (type, a, b = i), as push_bulk
void push_bulk(const Params &p, uint32_t gen, uint32_t type, uint32_t a, uint32_t count)
{
    if (count == 0) return;
    p.ctrl->outstanding.fetch_add((int32_t)count, std::memory_order_relaxed);
    const uint32_t base = p.ctrl->tail.fetch_add(count, std::memory_order_relaxed);
    if (base + count > p.qcap || base + count < base) {
        p.ctrl->error.fetch_or(ERR_QUEUE, std::memory_order_relaxed);
        return;
    }
    for (uint32_t i = 0; i < count; i++) {
        Slot *s = &p.q[base + i];
        s->type = type;
        s->a = a;
        s->b = i;
    }
    jitter(p);
    for (uint32_t i = 0; i < count; i++) {
        p.q[base + i].gen.store(gen, std::memory_order_release);
        jitter(p);
    }
}

// push of up to S + 1 tasks from arrays, as push_few
void push_few(const Params &p, uint32_t gen, const uint32_t *ty, const uint32_t *ta, const uint32_t *tb, uint32_t count)
{
    if (count == 0) return;
    p.ctrl->outstanding.fetch_add((int32_t)count, std::memory_order_relaxed);
    const uint32_t base = p.ctrl->tail.fetch_add(count, std::memory_order_relaxed);
    if (base + count > p.qcap || base + count < base) {
        p.ctrl->error.fetch_or(ERR_QUEUE, std::memory_order_relaxed);
        return;
    }
    for (uint32_t i = 0; i < count; i++) {
        Slot *s = &p.q[base + i];
        s->type = ty[i];
        s->a = ta[i];
        s->b = tb[i];
    }
    jitter(p);
    for (uint32_t i = 0; i < count; i++) p.q[base + i].gen.store(gen, std::memory_order_release);
}

void lds_sort(const Params &p, uint32_t begin, uint32_t len, uint32_t depth)
{
    if (p.checks && (!span_ok(p, begin, len) || len > CAP)) {
        range_fail(p, 1);
        return;
    }
    const uint32_t *src = src_keys(p, depth);
    const uint32_t *isrc = src_idx(p, depth);
    std::vector<std::pair<uint32_t, uint32_t>> v(len);
    for (uint32_t i = 0; i < len; i++) v[i] = {src[begin + i], isrc[begin + i]};
    std::stable_sort(v.begin(), v.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    for (uint32_t i = 0; i < len; i++) {
        p.kout[begin + i] = v[i].first;
        p.iout[begin + i] = v[i].second;
    }
}

void pivot(const Params &p, uint32_t gen, uint32_t sid)
{
    if (p.checks && sid >= p.segcap) {
        range_fail(p, 2);
        return;
    }
    Seg *sg = &p.segs[sid];
    const uint32_t begin = sg->begin, len = sg->len, depth = sg->depth;
    if (p.checks && !span_ok(p, begin, len)) {
        range_fail(p, 3);
        return;
    }
    const uint32_t *src = src_keys(p, depth);
    const uint32_t m = len < SAMPLE ? len : SAMPLE;
    std::vector<uint32_t> sk(m);
    for (uint32_t i = 0; i < m; i++) sk[i] = src[begin + (uint32_t)(((2ull * i + 1ull) * len) / (2ull * m))];
    std::sort(sk.begin(), sk.end());
    const uint32_t ntiles = (len + TILE - 1) / TILE;
    const uint32_t g = p.tiles_per_task;
    const uint32_t ntasks = (ntiles + g - 1) / g;
    for (int j = 0; j < S; j++) sg->split[j] = sk[((uint32_t)(j + 1) * m) / (S + 1)];
    const uint32_t tb = p.ctrl->tile_alloc.fetch_add(ntiles, std::memory_order_relaxed);
    bool ok = true;
    if (tb + ntiles > p.tilecap || tb + ntiles < tb) {
        p.ctrl->error.fetch_or(ERR_TILES, std::memory_order_relaxed);
        ok = false;
    }
    sg->ntiles = ntiles;
    sg->tile_base = tb;
    sg->ntasks = ntasks;
    sg->count_pending.store(ntasks, std::memory_order_relaxed);
    if (ok) push_bulk(p, gen, T_COUNT, sid, ntasks);
}

// exclusive prefix of every bin over the segment's tiles, as scan_tiles
bool scan_tiles(const Params &p, Seg *sg, uint32_t tile_base, uint32_t ntiles)
{
    uint32_t *tc = p.tiles + (size_t)tile_base * NB;
    uint32_t run = 0;
    for (int b = 0; b < NB; b++) {
        uint32_t loc = 0;
        for (uint32_t t = 0; t < ntiles; t++) {
            const uint32_t v = tc[(size_t)t * NB + b];
            tc[(size_t)t * NB + b] = loc;
            loc += v;
        }
        sg->total[b] = loc;
        sg->base[b] = run;
        run += loc;
    }
    sg->total[NB] = 0;
    sg->base[NB] = run;
    if (run != sg->len) {
        p.ctrl->error.fetch_or(ERR_BAD, std::memory_order_relaxed);
        return false;
    }
    return true;
}

bool tile_task_ok(const Params &p, const Seg *sg, uint32_t task)
{
    const uint32_t begin = sg->begin, len = sg->len, ntiles = sg->ntiles, tile_base = sg->tile_base;
    return span_ok(p, begin, len) && ntiles == (len + TILE - 1) / TILE && tile_base <= p.tilecap &&
           ntiles <= p.tilecap - tile_base && task * p.tiles_per_task < ntiles;
}

void count(const Params &p, uint32_t gen, uint32_t sid, uint32_t task)
{
    if (p.checks && sid >= p.segcap) {
        range_fail(p, 4);
        return;
    }
    Seg *sg = &p.segs[sid];
    if (p.checks && !tile_task_ok(p, sg, task)) {
        range_fail(p, 5);
        return;
    }
    const uint32_t begin = sg->begin, len = sg->len, depth = sg->depth;
    const uint32_t ntiles = sg->ntiles, tile_base = sg->tile_base;
    uint32_t spl[S];
    for (int j = 0; j < S; j++) spl[j] = sg->split[j];
    const uint32_t *src = src_keys(p, depth) + begin;
    const uint32_t t0 = task * p.tiles_per_task;
    const uint32_t t1 = std::min(t0 + p.tiles_per_task, ntiles);
    for (uint32_t t = t0; t < t1; t++) {
        const uint32_t ts = t * TILE;
        const uint32_t tl = std::min(len - ts, TILE);
        uint32_t cnt[NB] = {};
        for (uint32_t e = 0; e < tl; e++) cnt[classify(src[ts + e], spl)]++;
        for (int b = 0; b < NB; b++) p.tiles[(size_t)(tile_base + t) * NB + b] = cnt[b];
    }
    jitter(p);
    // last finisher of the segment's count tasks scans the tile counts
    if (sg->count_pending.fetch_sub(1u, std::memory_order_acq_rel) != 1u) return;
    if (!scan_tiles(p, sg, tile_base, ntiles)) return;
    if (depth == 0 && p.between) p.between(p.between_ctx, begin, len);
    sg->scatter_pending.store(sg->ntasks, std::memory_order_relaxed);
    push_bulk(p, gen, T_SCATTER, sid, sg->ntasks);
}

void spawn_children(const Params &p, uint32_t gen, Seg *sg, uint32_t begin, uint32_t depth)
{
    uint32_t ty[S + 1], ta[S + 1], tb[S + 1];
    uint32_t n = 0;
    for (int b = 0; b < NB; b += 2) {
        const uint32_t cl = sg->total[b];
        if (cl == 0) continue;
        const uint32_t cb = begin + sg->base[b];
        if (cl <= p.cutoff) {
            ty[n] = T_LDS | ((depth + 1) << 8);
            ta[n] = cb;
            tb[n] = cl;
            n++;
        } else {
            const uint32_t cid = p.ctrl->seg_alloc.fetch_add(1u, std::memory_order_relaxed);
            if (cid >= p.segcap) {
                p.ctrl->error.fetch_or(ERR_SEGS, std::memory_order_relaxed);
                continue;
            }
            Seg *c = &p.segs[cid];
            c->begin = cb;
            c->len = cl;
            c->depth = depth + 1;
            ty[n] = T_PIVOT;
            ta[n] = cid;
            tb[n] = 0;
            n++;
        }
    }
    uint32_t md = p.ctrl->max_depth.load(std::memory_order_relaxed);
    while (md < depth + 1 && !p.ctrl->max_depth.compare_exchange_weak(md, depth + 1, std::memory_order_relaxed)) {
    }
    push_few(p, gen, ty, ta, tb, n);
}

void scatter(const Params &p, uint32_t gen, uint32_t sid, uint32_t task)
{
    if (p.checks && sid >= p.segcap) {
        range_fail(p, 6);
        return;
    }
    Seg *sg = &p.segs[sid];
    if (p.checks && !tile_task_ok(p, sg, task)) {
        range_fail(p, 7);
        return;
    }
    const uint32_t begin = sg->begin, len = sg->len, depth = sg->depth;
    const uint32_t ntiles = sg->ntiles, tile_base = sg->tile_base;
    uint32_t spl[S];
    for (int j = 0; j < S; j++) spl[j] = sg->split[j];
    const uint32_t *src = src_keys(p, depth) + begin;
    const uint32_t *isrc = src_idx(p, depth) + begin;
    uint32_t *knext = next_keys(p, depth), *inext = next_idx(p, depth);
    const uint32_t t0 = task * p.tiles_per_task;
    const uint32_t t1 = std::min(t0 + p.tiles_per_task, ntiles);
    // the LDS staging of one tile
    std::vector<uint32_t> stk(TILE), sti(TILE);
    std::vector<unsigned char> stb(TILE);
    for (uint32_t t = t0; t < t1; t++) {
        const uint32_t ts = t * TILE;
        const uint32_t tl = std::min(len - ts, TILE);
        // the tile's own bin counts (the scatter classifies its keys again)
        std::vector<uint32_t> x(tl), ix(tl);
        std::vector<unsigned char> bin(tl);
        uint32_t cnt[NB] = {};
        for (uint32_t e = 0; e < tl; e++) {
            x[e] = src[ts + e];
            ix[e] = isrc[ts + e];
            bin[e] = (unsigned char)classify(x[e], spl);
            cnt[bin[e]]++;
        }
        uint32_t start[NB], dbase[NB];
        uint32_t run = 0;
        for (int b = 0; b < NB; b++) {
            start[b] = run;
            run += cnt[b];
            const uint32_t tpre = p.tiles[(size_t)(tile_base + t) * NB + b];
            dbase[b] = begin + sg->base[b] + tpre - start[b];
        }
        // stable ranks into the staging, bin by bin
        uint32_t woff[NB];
        std::memcpy(woff, start, sizeof(woff));
        for (uint32_t e = 0; e < tl; e++) {
            const uint32_t pos = woff[bin[e]]++;
            if (p.checks && pos >= tl) {
                range_fail(p, 8);
                continue;
            }
            stk[pos] = x[e];
            sti[pos] = ix[e];
            stb[pos] = bin[e];
        }
        // write-out: pivot-equal bins are final, the others go to the next level
        for (uint32_t i = 0; i < tl; i++) {
            const uint32_t b = stb[i];
            const uint32_t d = dbase[b] + i;
            if (p.checks && d - begin >= len) {
                range_fail(p, 9);
                continue;
            }
            if (b & 1u) {
                p.kout[d] = stk[i];
                p.iout[d] = sti[i];
            } else {
                knext[d] = stk[i];
                inext[d] = sti[i];
            }
        }
    }
    jitter(p);
    if (sg->scatter_pending.fetch_sub(1u, std::memory_order_acq_rel) != 1u) return;
    spawn_children(p, gen, sg, begin, depth);
}

// one workgroup of the persistent kernel
void drain(const Params &p, uint32_t seed)
{
    tl_rng.seed(seed);
    const uint32_t gen = p.ctrl->gen.load(std::memory_order_relaxed);
    const auto t_begin = std::chrono::steady_clock::now();
    for (;;) {
        const uint32_t idx = p.ctrl->head.fetch_add(1u, std::memory_order_relaxed);
        uint32_t type = 0, a = 0, b = 0;
        bool got = false;
        for (;;) {
            if (idx < p.qcap && p.q[idx].gen.load(std::memory_order_acquire) == gen) {
                type = p.q[idx].type;
                a = p.q[idx].a;
                b = p.q[idx].b;
                got = true;
                break;
            }
            if (p.ctrl->outstanding.load(std::memory_order_relaxed) == 0) break;
            if (p.ctrl->error.load(std::memory_order_relaxed) != 0) break;
            if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count() > p.timeout_s) {
                p.ctrl->error.fetch_or(ERR_TIMEOUT, std::memory_order_relaxed);
                break;
            }
            std::this_thread::yield();
        }
        if (!got) break;
        jitter(p);
        const uint32_t kind = type & 0xffu, aux = type >> 8;
        switch (kind) {
        case T_LDS: lds_sort(p, a, b, aux); break;
        case T_PIVOT: pivot(p, gen, a); break;
        case T_COUNT: count(p, gen, a, b); break;
        case T_SCATTER: scatter(p, gen, a, b); break;
        default: p.ctrl->error.fetch_or(ERR_BAD, std::memory_order_relaxed); break;
        }
        p.ctrl->tasks[kind < 5 ? kind : 0].fetch_add(1u, std::memory_order_relaxed);
        p.ctrl->outstanding.fetch_sub(1, std::memory_order_release);
    }
}

// The handle: buffers of max_n elements and the work tables, sized by the
// library's formulas (ensure_tables at scale 1).
struct Handle {
    uint32_t max_n;
    uint32_t *kin, *iin, *kout, *iout, *kt0, *kt1, *it0, *it1;
    Ctrl ctrl;
    Slot *q = nullptr;
    Seg *segs = nullptr;
    uint32_t *tiles = nullptr;
    uint32_t qcap = 0, segcap = 0, tilecap = 0, cutoff;
    Handle(uint32_t n, uint32_t cut, uint32_t nroots) : max_n(n), cutoff(cut)
    {
        kin = new uint32_t[n];
        iin = new uint32_t[n];
        kout = new uint32_t[n];
        iout = new uint32_t[n];
        kt0 = new uint32_t[n];
        kt1 = new uint32_t[n];
        it0 = new uint32_t[n];
        it1 = new uint32_t[n];
        const uint64_t segc = (4 * ((uint64_t)n / cutoff + 1) + 1024) + nroots;
        const uint64_t tilec = 8 * ((uint64_t)n / TILE + 1) + segc;
        const uint64_t qc = segc * (S + 2) + 2 * tilec + nroots + 1024;
        segcap = (uint32_t)segc;
        tilecap = (uint32_t)tilec;
        qcap = (uint32_t)qc;
        q = new Slot[qcap];
        segs = new Seg[segcap];
        tiles = new uint32_t[(size_t)tilecap * NB];
        for (auto &t : ctrl.tasks) t.store(0);
    }
    ~Handle()
    {
        delete[] kin; delete[] iin; delete[] kout; delete[] iout;
        delete[] kt0; delete[] kt1; delete[] it0; delete[] it1;
        delete[] q; delete[] segs; delete[] tiles;
    }
};

struct RunResult {
    uint32_t error = 0, bad_where = 0, max_depth = 0;
    int32_t outstanding = 0;
    uint64_t tasks = 0;
    double seconds = 0;
};

// ts_reset + ts_roots + ts_drain over segments [off[s], off[s+1]) of the
// handle's input (run on the host before the workgroup threads start, as
// the kernels before the drain kernel)
RunResult run(Handle &h, const std::vector<uint32_t> &off, uint32_t workgroups, uint32_t seed, bool checks,
              uint32_t jit, between_fn between, void *between_ctx, double timeout_s)
{
    Ctrl &c = h.ctrl;
    c.gen.store(c.gen.load() + 1);
    c.head.store(0);
    c.tail.store(0);
    c.outstanding.store(0);
    c.seg_alloc.store(0);
    c.tile_alloc.store(0);
    c.error.store(0);
    c.max_depth.store(0);
    c.bad_where.store(0);
    for (auto &t : c.tasks) t.store(0);
    Params p;
    p.kin = h.kin;
    p.iin = h.iin;
    p.kout = h.kout;
    p.iout = h.iout;
    p.kt0 = h.kt0;
    p.kt1 = h.kt1;
    p.it0 = h.it0;
    p.it1 = h.it1;
    p.ctrl = &c;
    p.q = h.q;
    p.segs = h.segs;
    p.tiles = h.tiles;
    p.qcap = h.qcap;
    p.segcap = h.segcap;
    p.tilecap = h.tilecap;
    p.cutoff = h.cutoff;
    p.n_total = off.back();
    p.checks = checks;
    p.jitter = jit;
    p.between = between;
    p.between_ctx = between_ctx;
    p.timeout_s = timeout_s;
    const uint32_t gen = c.gen.load();
    // roots: big segments become pivot tasks, small ones LDS tasks
    for (size_t s = 0; s + 1 < off.size(); s++) {
        const uint32_t beg = off[s], len = off[s + 1] - off[s];
        if (len == 0) continue;
        c.outstanding.fetch_add(1);
        const uint32_t qi = c.tail.fetch_add(1);
        if (qi >= p.qcap) {
            c.error.fetch_or(ERR_QUEUE);
            continue;
        }
        Slot *sl = &p.q[qi];
        if (len > p.cutoff) {
            const uint32_t sid = c.seg_alloc.fetch_add(1);
            if (sid >= p.segcap) {
                c.error.fetch_or(ERR_SEGS);
                continue;
            }
            Seg *sg = &p.segs[sid];
            sg->begin = beg;
            sg->len = len;
            sg->depth = 0;
            sl->type = T_PIVOT;
            sl->a = sid;
            sl->b = 0;
        } else {
            sl->type = T_LDS;
            sl->a = beg;
            sl->b = len;
        }
        sl->gen.store(gen);
    }
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> wg;
    for (uint32_t w = 0; w < workgroups; w++) wg.emplace_back(drain, std::cref(p), seed * 7919u + w);
    for (auto &t : wg) t.join();
    RunResult r;
    r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    r.error = c.error.load();
    r.bad_where = c.bad_where.load();
    r.max_depth = c.max_depth.load();
    r.outstanding = c.outstanding.load();
    for (auto &t : c.tasks) r.tasks += t.load();
    return r;
}

void fill(uint32_t *k, uint32_t n, std::mt19937_64 &rng, int shape)
{
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t r = rng();
        switch (shape) {
        case 0: k[i] = (uint32_t)r; break;                          // uniform
        case 1: k[i] = (uint32_t)(r >> 60); break;                  // 16 distinct
        case 2: k[i] = 7; break;                                    // all equal
        case 3: k[i] = i; break;                                    // ascending
        case 4: k[i] = n - i; break;                                // descending
        default: k[i] = (r & 1) ? 0u : 0xffffffffu; break;          // extremes
        }
    }
}

bool sorted_ok(const Handle &h, const std::vector<uint32_t> &keys, const std::vector<uint32_t> &off)
{
    for (size_t s = 0; s + 1 < off.size(); s++) {
        std::vector<uint32_t> e(keys.begin() + off[s], keys.begin() + off[s + 1]);
        std::stable_sort(e.begin(), e.end());
        for (uint32_t i = off[s]; i < off[s + 1]; i++) {
            if (h.kout[i] != e[i - off[s]]) return false;
            const uint32_t src = h.iout[i];
            if (src < off[s] || src >= off[s + 1] || keys[src] != h.kout[i]) return false;
        }
    }
    return true;
}

void print(const char *scenario, uint32_t n, uint32_t wg, const RunResult &r, int ok)
{
    std::printf("result: scenario=%s n=%u workgroups=%u error=%u bad_where=%u outstanding=%d tasks=%llu depth=%u "
                "sorted=%d seconds=%.3f\n",
                scenario, n, wg, r.error, r.bad_where, r.outstanding, (unsigned long long)r.tasks, r.max_depth, ok,
                r.seconds);
    std::fflush(stdout);
}

// ---- scenarios ------------------------------------------------------------------------

int protocol(uint32_t n_max, uint32_t wg_max, uint32_t seed)
{
    std::mt19937_64 rng(seed);
    int bad = 0;
    const uint32_t sizes[] = {0, 1, 2, 255, 256, 257, 1023, 1024, 1025, 4097, 20000, n_max};
    for (uint32_t n : sizes) {
        for (int shape = 0; shape < 6; shape++) {
            for (uint32_t wg = 1; wg <= wg_max; wg *= 2) {
                const bool batched = (shape + wg) % 3 == 0 && n > 64;
                std::vector<uint32_t> off{0};
                if (batched) {
                    while (off.back() < n) off.push_back(std::min<uint32_t>(n, off.back() + 1 + (uint32_t)(rng() % 3000)));
                } else {
                    off.push_back(n);
                }
                Handle h(std::max<uint32_t>(n, 1), 256, (uint32_t)off.size());
                fill(h.kin, n, rng, shape);
                for (uint32_t i = 0; i < n; i++) h.iin[i] = i;
                std::vector<uint32_t> keys(h.kin, h.kin + n);
                // two runs on one handle: the second sees the first run's stale slots
                for (int rep = 0; rep < 2; rep++) {
                    const RunResult r = run(h, off, wg, seed + rep, true, 3, nullptr, nullptr, 20.0);
                    const int ok = sorted_ok(h, keys, off) ? 1 : 0;
                    print("protocol", n, wg, r, ok);
                    if (r.error || r.outstanding != 0 || !ok) bad++;
                }
            }
        }
    }
    std::printf("protocol: %d bad runs\n", bad);
    return bad ? 1 : 0;
}

struct Producer {
    uint32_t *keys;
    uint32_t n;
    uint32_t value;
};

// the unordered producer: one pass over the input, as the fill kernel
void produce(Producer *pr)
{
    for (uint32_t i = 0; i < pr->n; i++) pr->keys[i] = pr->value ^ (i * 2654435761u);
}

int race(uint32_t n, uint32_t wg, uint32_t seed, bool checks)
{
    std::mt19937_64 rng(seed);
    Handle h(n, 256, 1);
    fill(h.kin, n, rng, 0);
    for (uint32_t i = 0; i < n; i++) h.iin[i] = i;
    std::vector<uint32_t> off{0, n};
    Producer pr{h.kin, n, 0x9e3779b9u};
    std::thread t(produce, &pr);
    const RunResult r = run(h, off, wg, seed, checks, 5, nullptr, nullptr, 20.0);
    t.join();
    print("race", n, wg, r, -1);
    return 0;
}

void write_between(void *ctx, uint32_t begin, uint32_t len)
{
    // every key of the segment becomes larger than each splitter: the tiles'
    // counts now say about a quarter of the keys go to the last bin, the
    // scatter finds all of them there
    uint32_t *k = (uint32_t *)ctx;
    for (uint32_t i = begin; i < begin + len; i++) k[i] = 0xffffffffu;
}

int between(uint32_t n, uint32_t wg, uint32_t seed, bool checks)
{
    std::mt19937_64 rng(seed);
    Handle h(n, 256, 1);
    fill(h.kin, n, rng, 0);
    for (uint32_t i = 0; i < n; i++) h.iin[i] = i;
    std::vector<uint32_t> off{0, n};
    const RunResult r = run(h, off, wg, seed, checks, 0, write_between, h.kin, 20.0);
    print(checks ? "between_checked" : "between_unchecked", n, wg, r, -1);
    return 0;
}

}  // namespace emu

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s protocol|race|between [n] [workgroups] [seed] [checks]\n", argv[0]);
        return 2;
    }
    const std::string sc = argv[1];
    const uint32_t n = argc > 2 ? (uint32_t)std::strtoul(argv[2], nullptr, 10) : 100000u;
    const uint32_t wg = argc > 3 ? (uint32_t)std::strtoul(argv[3], nullptr, 10) : 8u;
    const uint32_t seed = argc > 4 ? (uint32_t)std::strtoul(argv[4], nullptr, 10) : 1u;
    const bool checks = argc > 5 ? std::strtoul(argv[5], nullptr, 10) != 0 : true;
    if (sc == "protocol") return emu::protocol(n, wg, seed);
    if (sc == "race") return emu::race(n, wg, seed, checks);
    if (sc == "between") return emu::between(n, wg, seed, checks);
    std::fprintf(stderr, "unknown scenario %s\n", sc.c_str());
    return 2;
}
