// ts_levels.hip -- the sort engine for gfx950: level-synchronous partition
// passes over a device-resident work list, then LDS sorts.
//
// The reference recurses with CUDA dynamic parallelism: each wave of
// quicksort blocks tail-launches the next wave (turtle-sort.cu:35-111,
// 586-587) and one 1024-thread block partitions one whole segment
// (turtle-sort.cu:573-934).  gfx950 has no device-side launch.  Here the
// recursion runs level by level; one level is six kernels over the segments
// of that level, and every segment is cut into tiles that all workgroups
// share:
//
//   k_level    opens level L: its segments are ids [sb, se) of the segment
//              table (created by the previous level), its tile count is 0;
//   k_pivot    per segment: sample 1024 keys evenly, sort them in LDS, take
//              S splitters at the sample quantiles (S = 3 is the reference's
//              3-pivot scheme), reserve the segment's tiles;
//   k_count    per tile: bin counts;
//   k_scan     per segment: exclusive prefix of every bin over its tiles,
//              bin totals and bin bases;
//   k_scatter  per tile: stable wave64 ranking into bins, staged in LDS,
//              written out coalesced; pivot-equal bins go straight to the
//              output (their keys are final), the other bins to the next
//              ping-pong buffer;
//   k_children per segment: every non-empty non-pivot bin becomes a segment
//              of level L+1 or, at or below the cutoff, an LDS-sort task.
// After the last level k_lds sorts every LDS task in one workgroup each.
//
// Bins follow the reference's order LL,=pL,L,=p,R,=pR,RR
// (turtle-sort.cu:755-761) generalised to S splitters: bin 2j+1 holds keys
// equal to splitter j (equality tested first, lowest splitter first, as
// turtle-sort.cu:688-699), bin 2j the keys strictly between splitters j-1
// and j.
//
// Workgroups of one kernel never wait for each other: whatever one kernel
// produces (counts, prefixes, segment records, list lengths) is read by a
// later kernel on the same stream.  Cross-workgroup writes inside a kernel
// are atomic appends to lists that only later kernels read.
#include "ts_device.h"
#include "ts_engine.h"
#include "ts_internal.h"

namespace ts {

template <int KB, bool IDX, int S>
struct Cfg {
    typedef typename KeyT<KB>::U U;
    static constexpr int NT = 512;
    static constexpr int NW = NT / WAVE;
    static constexpr int NB = 2 * S + 1;
    static constexpr int BITS = (NB + 1 <= 8) ? 3 : 5;   // bits of a bin id (NB marks "no element")
    static constexpr int TI = (KB >= 8) ? 8 : 16;       // tile items per thread
    static constexpr int SI = (KB >= 8) ? 8 : 16;       // LDS-sort items per thread
    static constexpr uint32_t TILE = NT * TI;
    static constexpr uint32_t CAP = NT * SI;
    static constexpr uint32_t SAMPLE = 1024;
    static constexpr uint32_t IB = IDX ? 4 : 0;
    static constexpr uint32_t SORT_KEY_BYTES = (padded_len<SI>(CAP) * KB + 15) & ~15u;
    static constexpr uint32_t SORT_BYTES = SORT_KEY_BYTES + (IDX ? padded_len<SI>(CAP) * 4 : 0);
    static constexpr uint32_t PIVOT_BYTES = (padded_len<SI>(SAMPLE) * KB + 15) & ~15u;
    static constexpr uint32_t SCAT_BYTES = TILE * KB + TILE * IB + TILE;
};

// ---- buffers by depth -----------------------------------------------------------
// depth 0 lives in kin; depth d >= 1 in kt[(d-1) & 1]; a depth-d segment
// scatters its non-pivot bins into kt[d & 1].
template <typename U>
__device__ __forceinline__ const U *src_keys(const Params &p, uint32_t depth)
{
    return depth == 0 ? (const U *)p.kin : (const U *)((depth & 1) ? p.kt0 : p.kt1);
}
__device__ __forceinline__ const uint32_t *src_idx(const Params &p, uint32_t depth)
{
    return depth == 0 ? p.iin : ((depth & 1) ? p.it0 : p.it1);
}
template <typename U>
__device__ __forceinline__ U *next_keys(const Params &p, uint32_t depth)
{
    return (U *)((depth & 1) ? p.kt1 : p.kt0);
}
__device__ __forceinline__ uint32_t *next_idx(const Params &p, uint32_t depth)
{
    return (depth & 1) ? p.it1 : p.it0;
}

// Records a failed range check (the caller skips the access); the run then
// ends with ERR_RANGE instead of touching memory outside its buffers.
__device__ __forceinline__ void range_fail(const Params &p, uint32_t where)
{
    if (atomic_or(&p.ctrl->error, ERR_RANGE) == 0) store_relaxed(&p.ctrl->bad_where, where);
}

__device__ __forceinline__ bool span_ok(const Params &p, uint32_t begin, uint32_t len)
{
    return len > 0 && begin <= p.n_total && len <= p.n_total - begin;
}

// bin of order value x against ascending splitters spl[0..S):
//   lt = #splitters < x, le = #splitters <= x
//   equal to a splitter -> 2*lt+1 (the lowest equal splitter), else 2*lt
template <typename U, int S>
__device__ __forceinline__ uint32_t classify(U x, const U (&spl)[S])
{
    uint32_t lt = 0, le = 0;
#pragma unroll
    for (int j = 0; j < S; j++) {
        lt += (spl[j] < x) ? 1u : 0u;
        le += (spl[j] <= x) ? 1u : 0u;
    }
    return (le > lt) ? 2 * lt + 1 : 2 * lt;
}

// appends n LDS-sort tasks, returns the first slot (or ~0u when full)
__device__ __forceinline__ uint32_t lds_reserve(const Params &p, uint32_t n)
{
    const uint32_t b = atomic_add(&p.ctrl->lds_alloc, n);
    if (b > p.ldscap || n > p.ldscap - b) {
        atomic_or(&p.ctrl->error, ERR_LDSQ);
        return ~0u;
    }
    return b;
}

// ---- set-up kernels ------------------------------------------------------------------
__global__ void k_reset(Ctrl *c)
{
    if (threadIdx.x == 0) {
        c->seg_alloc = 0;
        c->tile_alloc = 0;
        c->lds_alloc = 0;
        c->error = 0;
        c->max_depth = 0;
        c->bad_where = 0;
        c->bytes_count = 0;
        c->bytes_scatter = 0;
        c->bytes_lds = 0;
        for (int i = 0; i < 5; i++) c->tasks[i] = 0;
    }
}

// Roots of a batch of segments [off[s], off[s+1]) (off == nullptr: the
// single segment [0, single_len)): large ones become level-0 segments,
// small ones LDS tasks.  Appends are aggregated per wave.
__global__ void k_roots(Params p, const uint32_t *off, uint32_t nseg, uint32_t single_len)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t beg = 0, len = 0;
    if (s < nseg) {
        beg = off ? off[s] : 0u;
        len = off ? off[s + 1] - beg : single_len;
    }
    if (len > 0 && !span_ok(p, beg, len)) {
        range_fail(p, 10);
        len = 0;
    }
    const bool big = len > p.cutoff;
    const bool small = len > 0 && !big;
    const uint64_t mb = ballot(big), ms = ballot(small);
    uint32_t sbase = 0, lbase = 0;
    if (lane() == 0) {
        const uint32_t nb = (uint32_t)__popcll(mb), ns = (uint32_t)__popcll(ms);
        if (nb) sbase = atomic_add(&p.ctrl->seg_alloc, nb);
        if (ns) lbase = lds_reserve(p, ns);
    }
    sbase = __shfl(sbase, 0, WAVE);
    lbase = __shfl(lbase, 0, WAVE);
    if (big) {
        const uint32_t sid = sbase + mask_rank(mb);
        if (sid >= p.segcap) {
            atomic_or(&p.ctrl->error, ERR_SEGS);
            return;
        }
        Seg *sg = &p.segs[sid];
        sg->begin = beg;
        sg->len = len;
        sg->depth = 0;
    } else if (small && lbase != ~0u) {
        LdsTask *t = &p.ldst[lbase + mask_rank(ms)];
        t->begin = beg;
        t->len = len;
        t->depth = 0;
    }
}

// opens level L: segments [sb, se), no tiles yet
__global__ void k_level(Params p, uint32_t L)
{
    if (threadIdx.x == 0) {
        const uint32_t sb = L == 0 ? 0u : p.levels[L - 1].se;
        uint32_t se = p.ctrl->seg_alloc;
        if (se > p.segcap) se = p.segcap;   // overflow already flagged
        p.levels[L].sb = sb;
        p.levels[L].se = se;
        p.levels[L].tiles = 0;
        p.ctrl->tile_alloc = 0;
        if (se > sb && L + 1 > p.ctrl->max_depth) p.ctrl->max_depth = L + 1;
    }
}

template <int KB, bool IDX, int S>
struct LevelK {
    typedef Cfg<KB, IDX, S> C;
    typedef typename C::U U;
    static constexpr int NT = C::NT, NW = C::NW, NB = C::NB, BITS = C::BITS, TI = C::TI, SI = C::SI;
    static constexpr uint32_t TILE = C::TILE;

    struct SharedScan {
        uint32_t red[16];
        uint32_t tot[MAX_BINS + 1];
        uint32_t bcast[4];
    };
    struct SharedTile {
        uint32_t wcnt[NW][MAX_BINS + 1];   // per-wave bin counts
        uint32_t woff[NW][MAX_BINS + 1];   // per-wave bin offsets inside the tile
        uint32_t dbase[MAX_BINS + 1];      // destination base per bin (minus tile offset)
    };

    static __device__ __forceinline__ void load_split(const Seg *sg, U (&spl)[S])
    {
#pragma unroll
        for (int j = 0; j < S; j++) spl[j] = (U)sg->split[j];
    }

    // segment id + tile of global tile g, validated
    static __device__ __forceinline__ bool tile_of(const Params &p, uint32_t g, uint32_t &sid, uint32_t &t)
    {
        sid = p.tile_seg[g];
        if (sid >= p.segcap) return false;
        const Seg *sg = &p.segs[sid];
        if (g < sg->tile_base) return false;
        t = g - sg->tile_base;
        return t < sg->ntiles && span_ok(p, sg->begin, sg->len) && (uint64_t)t * TILE < sg->len;
    }

    // ---------------------------------------------------------------- pivot
    static __device__ void pivot(const Params &p, uint32_t L, unsigned char *arena, SharedScan &sh)
    {
        const Level lv = p.levels[L];
        U *sk = (U *)arena;
        const U flip = (U)p.flip;
        for (uint32_t sid = lv.sb + blockIdx.x; sid < lv.se; sid += gridDim.x) {
            Seg *sg = &p.segs[sid];
            const uint32_t begin = sg->begin, len = sg->len, depth = sg->depth;
            if (!span_ok(p, begin, len)) {
                if (threadIdx.x == 0) range_fail(p, 3);
                continue;
            }
            const U *src = src_keys<U>(p, depth);
            const uint32_t m = len < C::SAMPLE ? len : C::SAMPLE;
            uint32_t n_eff = next_pow2_dev(m);
            if (n_eff < (uint32_t)SI) n_eff = SI;
            for (uint32_t i = threadIdx.x; i < n_eff; i += NT) {
                U v = key_max<U>();
                if (i < m) {
                    const uint64_t pos = ((2ull * i + 1ull) * (uint64_t)len) / (2ull * m);
                    v = (U)(src[begin + (uint32_t)pos] ^ flip);
                }
                sk[pidx<SI>(i)] = v;
            }
            block_sync();
            block_sort<U, false, NT, SI>(sk, nullptr, n_eff);
            const uint32_t ntiles = (len + TILE - 1) / TILE;
            if (threadIdx.x == 0) {
                for (int j = 0; j < S; j++) sg->split[j] = (uint64_t)sk[pidx<SI>(((uint32_t)(j + 1) * m) / (S + 1))];
                uint32_t tb = atomic_add(&p.ctrl->tile_alloc, ntiles);
                if (tb > p.tilecap || ntiles > p.tilecap - tb) {
                    atomic_or(&p.ctrl->error, ERR_TILES);
                    tb = ~0u;
                }
                sg->ntiles = tb == ~0u ? 0u : ntiles;
                sg->tile_base = tb == ~0u ? 0u : tb;
                sh.bcast[0] = tb;
                atomic_add64((unsigned long long *)&p.ctrl->tasks[T_PIVOT], 1ull);
            }
            block_sync();
            const uint32_t tb = sh.bcast[0];
            if (tb != ~0u)
                for (uint32_t t = threadIdx.x; t < ntiles; t += NT) p.tile_seg[tb + t] = sid;
            block_sync();
        }
    }

    // ---------------------------------------------------------------- count
    static __device__ void count(const Params &p, SharedTile &sh)
    {
        const uint32_t ntot = p.ctrl->tile_alloc < p.tilecap ? p.ctrl->tile_alloc : p.tilecap;
        const uint32_t G = p.tiles_per_wg;
        const U flip = (U)p.flip;
        const int w = threadIdx.x / WAVE;
        uint64_t nbytes = 0;
        for (uint32_t c0 = blockIdx.x * G; c0 < ntot; c0 += gridDim.x * G) {
            const uint32_t c1 = c0 + G < ntot ? c0 + G : ntot;
            for (uint32_t g = c0; g < c1; g++) {
                uint32_t sid, t;
                if (!tile_of(p, g, sid, t)) {
                    if (threadIdx.x == 0) range_fail(p, 4);
                    continue;
                }
                const Seg *sg = &p.segs[sid];
                U spl[S];
                load_split(sg, spl);
                const U *src = src_keys<U>(p, sg->depth) + sg->begin;
                const uint32_t ts = t * TILE;
                const uint32_t tl = (sg->len - ts) < TILE ? (sg->len - ts) : TILE;
                nbytes += (uint64_t)tl * KB;
                if (lane() < (uint32_t)NB) sh.wcnt[w][lane()] = 0;
                wave_sync();
                U x[TI];
#pragma unroll
                for (int c = 0; c < TI; c++) {
                    const uint32_t e = (uint32_t)w * (TI * WAVE) + c * WAVE + lane();
                    x[c] = e < tl ? src[ts + e] : (U)0;
                }
#pragma unroll
                for (int c = 0; c < TI; c++) {
                    const uint32_t e = (uint32_t)w * (TI * WAVE) + c * WAVE + lane();
                    const bool valid = e < tl;
                    const uint32_t bin = valid ? classify<U, S>((U)(x[c] ^ flip), spl) : (uint32_t)NB;
                    const uint64_t peers = match_peers<BITS>(bin, valid);
                    if (valid && mask_rank(peers) == 0) atomicAdd(&sh.wcnt[w][bin], (uint32_t)__popcll(peers));
                }
                block_sync();
                if (threadIdx.x < (uint32_t)NB) {
                    uint32_t s = 0;
#pragma unroll
                    for (int i = 0; i < NW; i++) s += sh.wcnt[i][threadIdx.x];
                    p.tiles[(size_t)g * NB + threadIdx.x] = s;
                }
                block_sync();
            }
        }
        if (threadIdx.x == 0 && nbytes) {
            atomic_add64(&p.ctrl->bytes_count, nbytes);
            atomic_add64(&p.ctrl->tasks[T_COUNT], 1ull);
        }
    }

    // ---------------------------------------------------------------- scan
    // exclusive prefix of every bin over the segment's tiles, written over
    // the counts; bin totals and bases go to the segment record
    static __device__ void scan(const Params &p, uint32_t L, SharedScan &sh)
    {
        const Level lv = p.levels[L];
        for (uint32_t sid = lv.sb + blockIdx.x; sid < lv.se; sid += gridDim.x) {
            Seg *sg = &p.segs[sid];
            const uint32_t ntiles = sg->ntiles, tile_base = sg->tile_base;
            if (ntiles == 0) continue;   // tile table overflow already flagged
            uint32_t *tc = p.tiles + (size_t)tile_base * NB;
            const uint32_t chunk = (ntiles + NT - 1) / NT;
            const uint32_t a0 = threadIdx.x * chunk < ntiles ? threadIdx.x * chunk : ntiles;
            const uint32_t a1 = a0 + chunk < ntiles ? a0 + chunk : ntiles;
            for (int b = 0; b < NB; b++) {
                uint32_t loc = 0;
                for (uint32_t t = a0; t < a1; t++) loc += tc[(size_t)t * NB + b];
                uint32_t total;
                loc = block_excl_scan<NT>(loc, sh.red, &total);
                if (threadIdx.x == 0) sh.tot[b] = total;
                for (uint32_t t = a0; t < a1; t++) {
                    const uint32_t v = tc[(size_t)t * NB + b];
                    tc[(size_t)t * NB + b] = loc;
                    loc += v;
                }
            }
            block_sync();
            if (threadIdx.x == 0) {
                uint32_t run = 0;
                for (int b = 0; b < NB; b++) {
                    sg->total[b] = sh.tot[b];
                    sg->base[b] = run;
                    run += sh.tot[b];
                }
                sg->total[NB] = 0;
                sg->base[NB] = run;
                if (run != sg->len) {
                    atomic_or(&p.ctrl->error, ERR_BAD);
                    sg->ntiles = 0;   // the scatter skips an inconsistent segment
                }
            }
            block_sync();
        }
    }

    // ---------------------------------------------------------------- scatter
    static __device__ void scatter(const Params &p, unsigned char *arena, SharedTile &sh)
    {
        const uint32_t ntot = p.ctrl->tile_alloc < p.tilecap ? p.ctrl->tile_alloc : p.tilecap;
        const uint32_t G = p.tiles_per_wg;
        const U flip = (U)p.flip;
        U *stk = (U *)arena;
        uint32_t *sti = (uint32_t *)(arena + TILE * KB);
        unsigned char *stb = arena + TILE * KB + TILE * C::IB;
        U *kfin = (U *)p.kout;
        uint32_t *ifin = p.iout;
        const int w = threadIdx.x / WAVE;
        uint64_t nbytes = 0;
        for (uint32_t c0 = blockIdx.x * G; c0 < ntot; c0 += gridDim.x * G) {
            const uint32_t c1 = c0 + G < ntot ? c0 + G : ntot;
            for (uint32_t g = c0; g < c1; g++) {
                uint32_t sid, t;
                if (!tile_of(p, g, sid, t)) {
                    if (threadIdx.x == 0) range_fail(p, 6);
                    continue;
                }
                const Seg *sg = &p.segs[sid];
                const uint32_t begin = sg->begin, len = sg->len, depth = sg->depth;
                U spl[S];
                load_split(sg, spl);
                const U *src = src_keys<U>(p, depth) + begin;
                const uint32_t *isrc = src_idx(p, depth) + begin;
                U *knext = next_keys<U>(p, depth);
                uint32_t *inext = next_idx(p, depth);
                const uint32_t ts = t * TILE;
                const uint32_t tl = (len - ts) < TILE ? (len - ts) : TILE;
                nbytes += 2ull * tl * (KB + C::IB);
                if (lane() < (uint32_t)NB) sh.wcnt[w][lane()] = 0;
                wave_sync();
                U x[TI];
                uint32_t ix[TI];
                uint32_t bpack[(TI + 3) / 4];
#pragma unroll
                for (int c = 0; c < TI; c++) {
                    const uint32_t e = (uint32_t)w * (TI * WAVE) + c * WAVE + lane();
                    x[c] = e < tl ? src[ts + e] : (U)0;
                    if (IDX) ix[c] = e < tl ? isrc[ts + e] : 0u;
                }
#pragma unroll
                for (int q = 0; q < (TI + 3) / 4; q++) bpack[q] = 0;
#pragma unroll
                for (int c = 0; c < TI; c++) {
                    const uint32_t e = (uint32_t)w * (TI * WAVE) + c * WAVE + lane();
                    const bool valid = e < tl;
                    const uint32_t bin = valid ? classify<U, S>((U)(x[c] ^ flip), spl) : (uint32_t)NB;
                    bpack[c / 4] |= bin << (8 * (c % 4));
                    const uint64_t peers = match_peers<BITS>(bin, valid);
                    if (valid && mask_rank(peers) == 0) atomicAdd(&sh.wcnt[w][bin], (uint32_t)__popcll(peers));
                }
                block_sync();
                if (w == 0) {
                    const uint32_t l = lane();
                    uint32_t tot = 0;
                    if (l < (uint32_t)NB) {
#pragma unroll
                        for (int i = 0; i < NW; i++) tot += sh.wcnt[i][l];
                    }
                    const uint32_t start = wave_incl_scan(tot) - tot;
                    if (l < (uint32_t)NB) {
                        uint32_t run = start;
#pragma unroll
                        for (int i = 0; i < NW; i++) {
                            sh.woff[i][l] = run;
                            run += sh.wcnt[i][l];
                        }
                        const uint32_t tpre = p.tiles[(size_t)g * NB + l];
                        sh.dbase[l] = begin + sg->base[l] + tpre - start;
                    }
                }
                block_sync();
#pragma unroll
                for (int c = 0; c < TI; c++) {
                    const uint32_t e = (uint32_t)w * (TI * WAVE) + c * WAVE + lane();
                    const bool valid = e < tl;
                    const uint32_t bin = (bpack[c / 4] >> (8 * (c % 4))) & 0xffu;
                    const uint64_t peers = match_peers<BITS>(bin, valid);
                    const uint32_t r = mask_rank(peers);
                    const uint32_t base = valid ? sh.woff[w][bin] : 0u;
                    wave_sync();                   // every lane has read before a leader updates
                    if (valid) {
                        const uint32_t pos = base + r;
                        if (pos < tl) {
                            stk[pos] = x[c];
                            if (IDX) sti[pos] = ix[c];
                            stb[pos] = (unsigned char)bin;
                        } else {
                            range_fail(p, 8);
                        }
                        if (r == 0) sh.woff[w][bin] = base + (uint32_t)__popcll(peers);
                    }
                    wave_sync();
                }
                block_sync();
                for (uint32_t i = threadIdx.x; i < tl; i += NT) {
                    const uint32_t b = stb[i];
                    const uint32_t d = b < (uint32_t)NB ? sh.dbase[b] + i : 0xffffffffu;
                    if (d - begin >= len) {        // outside the segment (also catches b >= NB)
                        range_fail(p, 9);
                        continue;
                    }
                    const U kv = stk[i];
                    if (b & 1u) {
                        kfin[d] = kv;
                        if (IDX) ifin[d] = sti[i];
                    } else {
                        knext[d] = kv;
                        if (IDX) inext[d] = sti[i];
                    }
                }
                block_sync();
            }
        }
        if (threadIdx.x == 0 && nbytes) {
            atomic_add64(&p.ctrl->bytes_scatter, nbytes);
            atomic_add64(&p.ctrl->tasks[T_SCATTER], 1ull);
        }
    }

    // ---------------------------------------------------------------- children
    // one thread per segment of level L
    static __device__ void children(const Params &p, uint32_t L)
    {
        const Level lv = p.levels[L];
        for (uint32_t sid = lv.sb + blockIdx.x * blockDim.x + threadIdx.x; sid < lv.se;
             sid += gridDim.x * blockDim.x) {
            const Seg *sg = &p.segs[sid];
            if (sg->ntiles == 0) continue;   // overflowed or inconsistent segment
            const uint32_t begin = sg->begin, depth = sg->depth;
            for (int b = 0; b < NB; b += 2) {
                const uint32_t cl = sg->total[b];
                if (cl == 0) continue;
                const uint32_t cb = begin + sg->base[b];
                if (cl <= p.cutoff) {
                    const uint32_t k = lds_reserve(p, 1);
                    if (k == ~0u) continue;
                    LdsTask *tk = &p.ldst[k];
                    tk->begin = cb;
                    tk->len = cl;
                    tk->depth = depth + 1;
                } else {
                    const uint32_t cid = atomic_add(&p.ctrl->seg_alloc, 1u);
                    if (cid >= p.segcap) {
                        atomic_or(&p.ctrl->error, ERR_SEGS);
                        continue;
                    }
                    Seg *c = &p.segs[cid];
                    c->begin = cb;
                    c->len = cl;
                    c->depth = depth + 1;
                }
            }
        }
    }

    // ---------------------------------------------------------------- LDS sort
    static __device__ void lds(const Params &p, unsigned char *arena)
    {
        U *sk = (U *)arena;
        uint32_t *si = (uint32_t *)(arena + C::SORT_KEY_BYTES);
        const U flip = (U)p.flip;
        U *dst = (U *)p.kout;
        const uint32_t ntask = p.ctrl->lds_alloc < p.ldscap ? p.ctrl->lds_alloc : p.ldscap;
        uint64_t nbytes = 0, ntasks = 0;
        for (uint32_t k = blockIdx.x; k < ntask; k += gridDim.x) {
            const LdsTask tk = p.ldst[k];
            const uint32_t begin = tk.begin, len = tk.len;
            if (!span_ok(p, begin, len) || len > C::CAP) {
                if (threadIdx.x == 0) range_fail(p, 1);
                continue;
            }
            const U *src = src_keys<U>(p, tk.depth);
            const uint32_t *isrc = src_idx(p, tk.depth);
            uint32_t n_eff = next_pow2_dev(len);
            if (n_eff < (uint32_t)SI) n_eff = SI;
            for (uint32_t i = threadIdx.x; i < n_eff; i += NT) {
                const bool v = i < len;
                sk[pidx<SI>(i)] = v ? (U)(src[begin + i] ^ flip) : key_max<U>();
                if (IDX) si[pidx<SI>(i)] = v ? isrc[begin + i] : 0u;
            }
            block_sync();
            block_sort<U, IDX, NT, SI>(sk, si, n_eff);
            for (uint32_t i = threadIdx.x; i < len; i += NT) {
                dst[begin + i] = (U)(sk[pidx<SI>(i)] ^ flip);
                if (IDX) p.iout[begin + i] = si[pidx<SI>(i)];
            }
            block_sync();
            nbytes += 2ull * len * (KB + C::IB);
            ntasks++;
        }
        if (threadIdx.x == 0 && ntasks) {
            atomic_add64(&p.ctrl->bytes_lds, nbytes);
            atomic_add64(&p.ctrl->tasks[T_LDS], ntasks);
        }
    }
};

// ---- kernels ----------------------------------------------------------------------------
template <int KB, bool IDX, int S>
__global__ void __launch_bounds__(512) k_pivot(Params p, uint32_t L)
{
    typedef LevelK<KB, IDX, S> K;
    __shared__ __align__(16) unsigned char arena[Cfg<KB, IDX, S>::PIVOT_BYTES];
    __shared__ typename K::SharedScan sh;
    K::pivot(p, L, arena, sh);
}

template <int KB, bool IDX, int S>
__global__ void __launch_bounds__(512) k_count(Params p)
{
    typedef LevelK<KB, IDX, S> K;
    __shared__ typename K::SharedTile sh;
    K::count(p, sh);
}

template <int KB, bool IDX, int S>
__global__ void __launch_bounds__(512) k_scan(Params p, uint32_t L)
{
    typedef LevelK<KB, IDX, S> K;
    __shared__ typename K::SharedScan sh;
    K::scan(p, L, sh);
}

template <int KB, bool IDX, int S>
__global__ void __launch_bounds__(512) k_scatter(Params p)
{
    typedef LevelK<KB, IDX, S> K;
    __shared__ __align__(16) unsigned char arena[Cfg<KB, IDX, S>::SCAT_BYTES];
    __shared__ typename K::SharedTile sh;
    K::scatter(p, arena, sh);
}

template <int KB, bool IDX, int S>
__global__ void __launch_bounds__(256) k_children(Params p, uint32_t L)
{
    LevelK<KB, IDX, S>::children(p, L);
}

template <int KB, bool IDX, int S>
__global__ void __launch_bounds__(512) k_lds(Params p)
{
    __shared__ __align__(16) unsigned char arena[Cfg<KB, IDX, S>::SORT_BYTES];
    LevelK<KB, IDX, S>::lds(p, arena);
}

// ---- host side ------------------------------------------------------------------------
struct Launch {
    int grid_pivot, grid_tile, grid_scan, grid_lds;
};

template <int KB, bool IDX, int S>
static int occ(const void *k, int threads)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, threads, 0) != hipSuccess || nb < 1) return 1;
    return nb;
}

template <int KB, bool IDX, int S>
static hipError_t run_levels(const Params &p, int cus, int wg_per_cu, const uint32_t *dev_off, uint32_t nseg,
                             uint32_t single_len, uint32_t L0, uint32_t nlev, bool first, bool last,
                             hipStream_t st, LaunchTimer *tm)
{
    static int o_tile = 0, o_pivot = 0, o_lds = 0;
    if (!o_tile) {
        o_tile = occ<KB, IDX, S>((const void *)k_scatter<KB, IDX, S>, 512);
        o_pivot = occ<KB, IDX, S>((const void *)k_pivot<KB, IDX, S>, 512);
        o_lds = occ<KB, IDX, S>((const void *)k_lds<KB, IDX, S>, 512);
    }
    auto lim = [&](int o) { return cus * ((wg_per_cu > 0 && wg_per_cu < o) ? wg_per_cu : o); };
    const int g_tile = lim(o_tile), g_pivot = lim(o_pivot), g_lds = lim(o_lds);
    const int g_child = cus;
#define TS_LAUNCH(kind, ...)                  \
    do {                                      \
        if (tm) tm->begin(kind, st);          \
        hipLaunchKernelGGL(__VA_ARGS__);      \
        if (tm) tm->end(kind, st);            \
    } while (0)
    if (first) {
        TS_LAUNCH(K_RESET, k_reset, dim3(1), dim3(64), 0, st, p.ctrl);
        if (nseg > 0)
            TS_LAUNCH(K_ROOTS, k_roots, dim3((nseg + 255) / 256), dim3(256), 0, st, p, dev_off, nseg, single_len);
    }
    for (uint32_t L = L0; L < L0 + nlev && L < (uint32_t)MAX_LEVELS; L++) {
        TS_LAUNCH(K_LEVEL, k_level, dim3(1), dim3(64), 0, st, p, L);
        TS_LAUNCH(K_PIVOT, (k_pivot<KB, IDX, S>), dim3(g_pivot), dim3(512), 0, st, p, L);
        TS_LAUNCH(K_COUNT, (k_count<KB, IDX, S>), dim3(g_tile), dim3(512), 0, st, p);
        TS_LAUNCH(K_SCAN, (k_scan<KB, IDX, S>), dim3(g_pivot), dim3(512), 0, st, p, L);
        TS_LAUNCH(K_SCATTER, (k_scatter<KB, IDX, S>), dim3(g_tile), dim3(512), 0, st, p);
        TS_LAUNCH(K_CHILDREN, (k_children<KB, IDX, S>), dim3(g_child), dim3(256), 0, st, p, L);
    }
    if (last) TS_LAUNCH(K_LDS, (k_lds<KB, IDX, S>), dim3(g_lds), dim3(512), 0, st, p);
#undef TS_LAUNCH
    return hipGetLastError();
}

typedef hipError_t (*levels_fn)(const Params &, int, int, const uint32_t *, uint32_t, uint32_t, uint32_t, uint32_t,
                                bool, bool, hipStream_t, LaunchTimer *);

template <bool IDX, int S>
static levels_fn pick(int kb)
{
    switch (kb) {
    case 1: return run_levels<1, IDX, S>;
    case 2: return run_levels<2, IDX, S>;
    case 4: return run_levels<4, IDX, S>;
    default: return run_levels<8, IDX, S>;
    }
}

bool engine_supports(int splitters) { return splitters == 3; }
uint32_t engine_tile(int kb) { return kb >= 8 ? 512u * 8u : 512u * 16u; }
uint32_t engine_cap(int kb) { return kb >= 8 ? 512u * 8u : 512u * 16u; }

hipError_t engine_levels(const Params &p, int kb, bool idx, int splitters, int cus, int wg_per_cu,
                         const uint32_t *dev_off, uint32_t nseg, uint32_t single_len, uint32_t L0, uint32_t nlev,
                         bool first, bool last, hipStream_t st, LaunchTimer *timer)
{
    (void)splitters;
    levels_fn f = idx ? pick<true, 3>(kb) : pick<false, 3>(kb);
    return f(p, cus, wg_per_cu, dev_off, nseg, single_len, L0, nlev, first, last, st, timer);
}

}  // namespace ts
