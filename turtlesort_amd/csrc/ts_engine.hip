// ts_engine.hip -- the persistent, queue-drained sort engine for gfx950.
//
// The reference chains CUDA dynamic-parallelism launches: every wave of
// quicksort blocks tail-launches the next wave (turtle-sort.cu:35-111, 586-587)
// and one 1024-thread block partitions one whole segment
// (turtle-sort.cu:573-934).  gfx950 has no device-side launch, so here one
// persistent kernel (ts_drain) is launched with every workgroup resident and
// drains a device work queue.  A segment is cut into tiles that all
// workgroups share:
//
//   T_PIVOT(seg)    sample the segment, sort the sample in LDS, take S
//                   evenly spaced splitters (S = 3 is the reference's
//                   3-pivot scheme), push the count tasks;
//   T_COUNT(seg,k)  per-tile bin counts; the last finisher scans the tile
//                   counts into per-tile bin offsets and pushes the scatters;
//   T_SCATTER(seg,k) stable wave64 ballot ranking of each tile into its bins,
//                   staged in LDS, then written out coalesced.  Pivot-equal
//                   bins are final: they go straight to the output.  The last
//                   finisher pushes one child per non-empty non-pivot bin;
//   T_LDS(begin,len) segments at or below the cutoff: loaded into LDS,
//                   stable block sort, written to the output.
//
// Bins follow the reference's order LL,=pL,L,=p,R,=pR,RR
// (turtle-sort.cu:755-761) generalised to S splitters: bin 2j+1 holds keys
// equal to splitter j (equality tested first, lowest splitter first, as
// turtle-sort.cu:688-699), bin 2j the keys strictly between splitters j-1
// and j.
//
// No task ever waits for another task: dependencies resolve through
// "last finisher" counters, so workgroups only spin while waiting for a queue
// slot to be filled, and leave when the outstanding-task count reaches zero,
// the error word is set, or the watchdog expires.
#include "ts_device.h"
#include "ts_engine.h"
#include "ts_internal.h"

namespace ts {

template <int KB, bool IDX, int S>
struct Eng {
    typedef typename KeyT<KB>::U U;
    static constexpr int NT = 512;
    static constexpr int NW = NT / WAVE;
    static constexpr int NB = 2 * S + 1;
    static constexpr int TI = (KB >= 8) ? 8 : 16;     // tile items per thread
    static constexpr int SI = (KB >= 8) ? 8 : 16;     // LDS-sort items per thread
    static constexpr uint32_t TILE = NT * TI;
    static constexpr uint32_t CAP = NT * SI;
    static constexpr uint32_t SAMPLE = 1024;
    static constexpr uint32_t IB = IDX ? 4 : 0;
    // LDS arena: the larger of the scatter staging and the block sort
    static constexpr uint32_t SCAT_BYTES = TILE * KB + TILE * IB + TILE;
    static constexpr uint32_t SORT_KEY_BYTES = (padded_len<SI>(CAP) * KB + 15) & ~15u;
    static constexpr uint32_t SORT_BYTES = SORT_KEY_BYTES + (IDX ? padded_len<SI>(CAP) * 4 : 0);
    static constexpr uint32_t ARENA = ((SCAT_BYTES > SORT_BYTES ? SCAT_BYTES : SORT_BYTES) + 15) & ~15u;
};

struct Shared {
    uint32_t slot_type, slot_a, slot_b, state;
    uint32_t bcast[4];
    uint32_t red[16];
    uint32_t wcnt[8][MAX_BINS + 1];     // per-wave bin counts
    uint32_t woff[8][MAX_BINS + 1];     // per-wave bin offsets inside the tile
    uint32_t dbase[MAX_BINS + 1];       // destination base per bin (minus tile offset)
    uint32_t tot[MAX_BINS + 1];
};

// ---- buffers by depth -----------------------------------------------------------
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

// ---- queue pushes ----------------------------------------------------------------
// Whole-workgroup bulk push of `count` tasks (type, a, b = first_b + i).
// Must be called by every thread of the workgroup.
template <int NT>
__device__ void push_bulk(const Params &p, Shared &sh, uint32_t gen, uint32_t type, uint32_t a,
                          uint32_t count)
{
    if (count == 0) return;
    if (threadIdx.x == 0) {
        atomic_add_i(&p.ctrl->outstanding, (int32_t)count);
        uint32_t base = atomic_add(&p.ctrl->tail, count);
        if (base + count > p.qcap || base + count < base) {
            atomic_or(&p.ctrl->error, ERR_QUEUE);
            base = 0xffffffffu;
        }
        sh.bcast[0] = base;
    }
    __syncthreads();
    const uint32_t base = sh.bcast[0];
    __syncthreads();
    if (base == 0xffffffffu) return;
    for (uint32_t i = threadIdx.x; i < count; i += NT) {
        Slot *s = &p.q[base + i];
        s->type = type;
        s->a = a;
        s->b = i;
    }
    fence_release();
    for (uint32_t i = threadIdx.x; i < count; i += NT) store_relaxed(&p.q[base + i].gen, gen);
}

// Single-thread push of up to MAX_SPLIT+1 tasks held in arrays.
__device__ void push_few(const Params &p, uint32_t gen, const uint32_t *type, const uint32_t *a,
                         const uint32_t *b, uint32_t count)
{
    if (count == 0) return;
    atomic_add_i(&p.ctrl->outstanding, (int32_t)count);
    uint32_t base = atomic_add(&p.ctrl->tail, count);
    if (base + count > p.qcap || base + count < base) {
        atomic_or(&p.ctrl->error, ERR_QUEUE);
        return;
    }
    for (uint32_t i = 0; i < count; i++) {
        Slot *s = &p.q[base + i];
        s->type = type[i];
        s->a = a[i];
        s->b = b[i];
    }
    fence_release();
    for (uint32_t i = 0; i < count; i++) store_relaxed(&p.q[base + i].gen, gen);
}

// ---- classification --------------------------------------------------------------
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

template <int KB, bool IDX, int S>
struct Tasks {
    typedef Eng<KB, IDX, S> E;
    typedef typename E::U U;
    static constexpr int NT = E::NT, NW = E::NW, NB = E::NB, TI = E::TI, SI = E::SI;
    static constexpr uint32_t TILE = E::TILE;
    static constexpr int BITS = (NB + 1 <= 8) ? 3 : 5;   // bits of a bin id (NB itself marks "no element")

    // ---------------------------------------------------------------- LDS sort
    static __device__ void lds_sort(const Params &p, unsigned char *arena, uint32_t begin, uint32_t len,
                                    uint32_t depth)
    {
        U *sk = (U *)arena;
        uint32_t *si = (uint32_t *)(arena + E::SORT_KEY_BYTES);
        const U flip = (U)p.flip;
        const U *src = src_keys<U>(p, depth);
        const uint32_t *isrc = src_idx(p, depth);
        U *dst = (U *)p.kout;
        if (!span_ok(p, begin, len) || len > E::CAP) {
            if (threadIdx.x == 0) range_fail(p, 1);
            return;
        }
        uint32_t n_eff = next_pow2_dev(len);
        if (n_eff < (uint32_t)SI) n_eff = SI;
        for (uint32_t i = threadIdx.x; i < n_eff; i += NT) {
            bool v = i < len;
            sk[pidx<SI>(i)] = v ? (U)(src[begin + i] ^ flip) : key_max<U>();
            if (IDX) si[pidx<SI>(i)] = v ? isrc[begin + i] : 0u;
        }
        __syncthreads();
        block_sort<U, IDX, NT, SI>(sk, si, n_eff);
        for (uint32_t i = threadIdx.x; i < len; i += NT) {
            dst[begin + i] = (U)(sk[pidx<SI>(i)] ^ flip);
            if (IDX) p.iout[begin + i] = si[pidx<SI>(i)];
        }
        __syncthreads();
        if (threadIdx.x == 0) atomic_add64(&p.ctrl->bytes_lds, 2ull * len * (KB + E::IB));
    }

    // ---------------------------------------------------------------- pivots
    static __device__ void pivot(const Params &p, Shared &sh, unsigned char *arena, uint32_t gen, uint32_t sid)
    {
        if (sid >= p.segcap) {
            if (threadIdx.x == 0) range_fail(p, 2);
            return;
        }
        Seg *sg = &p.segs[sid];
        const uint32_t begin = sg->begin, len = sg->len, depth = sg->depth;
        if (!span_ok(p, begin, len)) {
            if (threadIdx.x == 0) range_fail(p, 3);
            return;
        }
        const U flip = (U)p.flip;
        const U *src = src_keys<U>(p, depth);
        U *sk = (U *)arena;
        const uint32_t m = len < E::SAMPLE ? len : E::SAMPLE;
        uint32_t n_eff = next_pow2_dev(m);
        if (n_eff < (uint32_t)SI) n_eff = SI;
        for (uint32_t i = threadIdx.x; i < n_eff; i += NT) {
            U v = key_max<U>();
            if (i < m) {
                uint64_t pos = ((2ull * i + 1ull) * (uint64_t)len) / (2ull * m);
                v = (U)(src[begin + (uint32_t)pos] ^ flip);
            }
            sk[pidx<SI>(i)] = v;
        }
        __syncthreads();
        block_sort<U, false, NT, SI>(sk, nullptr, n_eff);
        const uint32_t ntiles = (len + TILE - 1) / TILE;
        const uint32_t g = p.tiles_per_task;
        const uint32_t ntasks = (ntiles + g - 1) / g;
        if (threadIdx.x == 0) {
            for (int j = 0; j < S; j++) sg->split[j] = (uint64_t)sk[pidx<SI>(((uint32_t)(j + 1) * m) / (S + 1))];
            uint32_t tb = atomic_add(&p.ctrl->tile_alloc, ntiles);
            uint32_t ok = 1;
            if (tb + ntiles > p.tilecap || tb + ntiles < tb) {
                atomic_or(&p.ctrl->error, ERR_TILES);
                ok = 0;
            }
            sg->ntiles = ntiles;
            sg->tile_base = tb;
            sg->ntasks = ntasks;
            sg->count_pending = ntasks;
            sh.bcast[1] = ok;
        }
        __syncthreads();
        const uint32_t ok = sh.bcast[1];
        __syncthreads();
        if (ok) push_bulk<NT>(p, sh, gen, T_COUNT, sid, ntasks);
        __syncthreads();
    }

    // load the splitters of a segment (uniform across the workgroup)
    static __device__ __forceinline__ void load_split(const Seg *sg, U (&spl)[S])
    {
#pragma unroll
        for (int j = 0; j < S; j++) spl[j] = (U)sg->split[j];
    }

    // ---------------------------------------------------------------- count
    static __device__ void count(const Params &p, Shared &sh, unsigned char *arena, uint32_t gen, uint32_t sid,
                                 uint32_t task)
    {
        if (sid >= p.segcap) {
            if (threadIdx.x == 0) range_fail(p, 4);
            return;
        }
        Seg *sg = &p.segs[sid];
        const uint32_t begin = sg->begin, len = sg->len, depth = sg->depth;
        const uint32_t ntiles = sg->ntiles, tile_base = sg->tile_base;
        if (!span_ok(p, begin, len) || ntiles != (len + TILE - 1) / TILE || tile_base > p.tilecap ||
            ntiles > p.tilecap - tile_base || task * p.tiles_per_task >= ntiles) {
            if (threadIdx.x == 0) range_fail(p, 5);
            return;
        }
        U spl[S];
        load_split(sg, spl);
        const U flip = (U)p.flip;
        const U *src = src_keys<U>(p, depth) + begin;
        const int w = threadIdx.x / WAVE;
        const uint32_t t0 = task * p.tiles_per_task;
        uint32_t t1 = t0 + p.tiles_per_task;
        if (t1 > ntiles) t1 = ntiles;
        uint64_t nbytes = 0;
        for (uint32_t t = t0; t < t1; t++) {
            const uint32_t ts = t * TILE;
            const uint32_t tl = (len - ts) < TILE ? (len - ts) : TILE;
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
            __syncthreads();
            if (threadIdx.x < (uint32_t)NB) {
                uint32_t s = 0;
#pragma unroll
                for (int i = 0; i < NW; i++) s += sh.wcnt[i][threadIdx.x];
                p.tiles[(size_t)(tile_base + t) * NB + threadIdx.x] = s;
            }
            __syncthreads();
        }
        // last finisher of the segment's count tasks scans the tile counts
        fence_release();
        __syncthreads();
        if (threadIdx.x == 0) {
            atomic_add64(&p.ctrl->bytes_count, nbytes);
            uint32_t old = atomic_add(&sg->count_pending, 0xffffffffu);
            sh.bcast[2] = (old == 1) ? 1u : 0u;
        }
        __syncthreads();
        const bool last = sh.bcast[2] != 0;
        __syncthreads();
        if (!last) return;
        fence_acquire();
        if (!scan_tiles(p, sh, sg, tile_base, ntiles)) return;   // error raised; scatter would be unsafe
        if (threadIdx.x == 0) sg->scatter_pending = sg->ntasks;
        fence_release();
        __syncthreads();
        push_bulk<NT>(p, sh, gen, T_SCATTER, sid, sg->ntasks);
        __syncthreads();
    }

    // exclusive prefix of every bin over the segment's tiles, written over
    // the counts; bin totals and bin bases go to the segment record
    static __device__ bool scan_tiles(const Params &p, Shared &sh, Seg *sg, uint32_t tile_base, uint32_t ntiles)
    {
        uint32_t *tc = p.tiles + (size_t)tile_base * NB;
        const uint32_t chunk = (ntiles + NT - 1) / NT;
        const uint32_t a0 = threadIdx.x * chunk;
        const uint32_t a1 = (a0 + chunk) < ntiles ? (a0 + chunk) : ntiles;
        // one bin at a time keeps the register footprint small; the tile
        // counts are re-read from L2
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
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t run = 0;
            for (int b = 0; b < NB; b++) {
                sg->total[b] = sh.tot[b];
                sg->base[b] = run;
                run += sh.tot[b];
            }
            sg->total[NB] = 0;
            sg->base[NB] = run;
            const bool good = run == sg->len;
            if (!good) atomic_or(&p.ctrl->error, ERR_BAD);
            sh.bcast[3] = good ? 1u : 0u;
        }
        __syncthreads();
        const bool good = sh.bcast[3] != 0;
        __syncthreads();
        return good;
    }

    // ---------------------------------------------------------------- scatter
    static __device__ void scatter(const Params &p, Shared &sh, unsigned char *arena, uint32_t gen, uint32_t sid,
                                   uint32_t task)
    {
        if (sid >= p.segcap) {
            if (threadIdx.x == 0) range_fail(p, 6);
            return;
        }
        Seg *sg = &p.segs[sid];
        const uint32_t begin = sg->begin, len = sg->len, depth = sg->depth;
        const uint32_t ntiles = sg->ntiles, tile_base = sg->tile_base;
        if (!span_ok(p, begin, len) || ntiles != (len + TILE - 1) / TILE || tile_base > p.tilecap ||
            ntiles > p.tilecap - tile_base || task * p.tiles_per_task >= ntiles) {
            if (threadIdx.x == 0) range_fail(p, 7);
            return;
        }
        U spl[S];
        load_split(sg, spl);
        const U flip = (U)p.flip;
        const U *src = src_keys<U>(p, depth) + begin;
        const uint32_t *isrc = src_idx(p, depth) + begin;
        U *knext = next_keys<U>(p, depth);
        uint32_t *inext = next_idx(p, depth);
        U *kfin = (U *)p.kout;
        uint32_t *ifin = p.iout;
        U *stk = (U *)arena;
        uint32_t *sti = (uint32_t *)(arena + TILE * KB);
        unsigned char *stb = arena + TILE * KB + TILE * E::IB;
        const int w = threadIdx.x / WAVE;
        const uint32_t t0 = task * p.tiles_per_task;
        uint32_t t1 = t0 + p.tiles_per_task;
        if (t1 > ntiles) t1 = ntiles;
        uint64_t nbytes = 0;
        for (uint32_t t = t0; t < t1; t++) {
            const uint32_t ts = t * TILE;
            const uint32_t tl = (len - ts) < TILE ? (len - ts) : TILE;
            nbytes += 2ull * tl * (KB + E::IB);
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
            __syncthreads();
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
                    const uint32_t tpre = p.tiles[(size_t)(tile_base + t) * NB + l];
                    sh.dbase[l] = begin + sg->base[l] + tpre - start;
                }
            }
            __syncthreads();
#pragma unroll
            for (int c = 0; c < TI; c++) {
                const uint32_t e = (uint32_t)w * (TI * WAVE) + c * WAVE + lane();
                const bool valid = e < tl;
                const uint32_t bin = (bpack[c / 4] >> (8 * (c % 4))) & 0xffu;
                const uint64_t peers = match_peers<BITS>(bin, valid);
                const uint32_t r = mask_rank(peers);
                const uint32_t base = valid ? sh.woff[w][bin] : 0u;
                wave_sync();                       // every lane has read before a leader updates
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
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < tl; i += NT) {
                const uint32_t b = stb[i];
                const uint32_t d = b < (uint32_t)NB ? sh.dbase[b] + i : 0xffffffffu;
                if (d - begin >= len) {          // outside the segment (also catches b >= NB)
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
            __syncthreads();
        }
        fence_release();
        __syncthreads();
        if (threadIdx.x == 0) {
            atomic_add64(&p.ctrl->bytes_scatter, nbytes);
            uint32_t old = atomic_add(&sg->scatter_pending, 0xffffffffu);
            sh.bcast[2] = (old == 1) ? 1u : 0u;
        }
        __syncthreads();
        const bool last = sh.bcast[2] != 0;
        __syncthreads();
        if (!last) return;
        fence_acquire();
        if (threadIdx.x == 0) spawn_children(p, gen, sg, begin, depth);
        __syncthreads();
    }

    // one child per non-empty non-pivot bin: LDS task or new segment
    static __device__ void spawn_children(const Params &p, uint32_t gen, Seg *sg, uint32_t begin, uint32_t depth)
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
                const uint32_t cid = atomic_add(&p.ctrl->seg_alloc, 1u);
                if (cid >= p.segcap) {
                    atomic_or(&p.ctrl->error, ERR_SEGS);
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
        atomic_max(&p.ctrl->max_depth, depth + 1);
        push_few(p, gen, ty, ta, tb, n);
    }
};

template <int KB, bool IDX, int S>
__global__ void __launch_bounds__(512) ts_drain(Params p)
{
    typedef Eng<KB, IDX, S> E;
    typedef Tasks<KB, IDX, S> TK;
    __shared__ __align__(16) unsigned char arena[E::ARENA];
    __shared__ Shared sh;
    const uint32_t gen = load_relaxed(&p.ctrl->gen);
    const uint64_t t_begin = wall_clock64();
    for (;;) {
        if (threadIdx.x == 0) {
            const uint32_t idx = atomic_add(&p.ctrl->head, 1u);
            uint32_t st = 0;
            for (;;) {
                if (idx < p.qcap && load_relaxed(&p.q[idx].gen) == gen) {
                    fence_acquire();
                    sh.slot_type = p.q[idx].type;
                    sh.slot_a = p.q[idx].a;
                    sh.slot_b = p.q[idx].b;
                    st = 1;
                    break;
                }
                if (__hip_atomic_load(&p.ctrl->outstanding, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
                if (load_relaxed(&p.ctrl->error) != 0) break;
                if (wall_clock64() - t_begin > p.timeout_ticks) {
                    atomic_or(&p.ctrl->error, ERR_TIMEOUT);
                    break;
                }
                __builtin_amdgcn_s_sleep(4);
            }
            sh.state = st;
        }
        __syncthreads();
        if (!sh.state) break;
        fence_acquire();
        const uint32_t type = sh.slot_type & 0xffu;
        const uint32_t a = sh.slot_a, b = sh.slot_b;
        const uint32_t aux = sh.slot_type >> 8;
        __syncthreads();
        switch (type) {
        case T_LDS: TK::lds_sort(p, arena, a, b, aux); break;
        case T_PIVOT: TK::pivot(p, sh, arena, gen, a); break;
        case T_COUNT: TK::count(p, sh, arena, gen, a, b); break;
        case T_SCATTER: TK::scatter(p, sh, arena, gen, a, b); break;
        default:
            if (threadIdx.x == 0) atomic_or(&p.ctrl->error, ERR_BAD);
            break;
        }
        fence_release();
        __syncthreads();
        if (threadIdx.x == 0) {
            atomic_add64(&p.ctrl->tasks[type < 5 ? type : 0], 1ull);
            atomic_add_i(&p.ctrl->outstanding, -1);
        }
    }
}

// ---- run set-up kernels ------------------------------------------------------------
__global__ void ts_reset(Ctrl *c)
{
    if (threadIdx.x == 0) {
        c->gen = c->gen + 1;
        c->head = 0;
        c->tail = 0;
        c->outstanding = 0;
        c->seg_alloc = 0;
        c->tile_alloc = 0;
        c->error = 0;
        c->max_depth = 0;
        c->bad_where = 0;
        c->bytes_count = 0;
        c->bytes_scatter = 0;
        c->bytes_lds = 0;
        for (int i = 0; i < 5; i++) c->tasks[i] = 0;
    }
}

// Root tasks of a batch of segments [off[s], off[s+1]).  Runs after ts_reset
// and before ts_drain, so plain stores are enough; work is aggregated per
// wave to keep the counter atomics few.
// off == nullptr: a single segment [0, single_len).
__global__ void ts_roots(Params p, const uint32_t *off, uint32_t nseg, uint32_t single_len)
{
    const uint32_t gen = p.ctrl->gen;
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
    const bool has = len > 0;
    const bool big = len > p.cutoff;
    const uint64_t mt = ballot(has);
    const uint64_t mb = ballot(has && big);
    uint32_t qbase = 0, sbase = 0;
    if (lane() == 0) {
        const uint32_t nt = (uint32_t)__popcll(mt), nb = (uint32_t)__popcll(mb);
        if (nt) {
            atomic_add_i(&p.ctrl->outstanding, (int32_t)nt);
            qbase = atomic_add(&p.ctrl->tail, nt);
        }
        if (nb) sbase = atomic_add(&p.ctrl->seg_alloc, nb);
    }
    qbase = __shfl(qbase, 0, WAVE);
    sbase = __shfl(sbase, 0, WAVE);
    if (!has) return;
    const uint32_t qi = qbase + mask_rank(mt);
    if (qi >= p.qcap) {
        atomic_or(&p.ctrl->error, ERR_QUEUE);
        return;
    }
    Slot *sl = &p.q[qi];
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
        sl->type = T_PIVOT;
        sl->a = sid;
        sl->b = 0;
    } else {
        sl->type = T_LDS;   // depth 0
        sl->a = beg;
        sl->b = len;
    }
    sl->gen = gen;
}

// ---- host side ------------------------------------------------------------------------
template <int KB, bool IDX, int S>
static hipError_t launch_drain(const Params &p, int grid, hipStream_t st)
{
    hipLaunchKernelGGL((ts_drain<KB, IDX, S>), dim3(grid), dim3(512), 0, st, p);
    return hipGetLastError();
}

template <int KB, bool IDX, int S>
static int occupancy_of()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)ts_drain<KB, IDX, S>, 512, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

typedef hipError_t (*drain_fn)(const Params &, int, hipStream_t);
typedef int (*occ_fn)();

template <bool IDX, int S>
static drain_fn pick_drain(int kb)
{
    switch (kb) {
    case 1: return launch_drain<1, IDX, S>;
    case 2: return launch_drain<2, IDX, S>;
    case 4: return launch_drain<4, IDX, S>;
    default: return launch_drain<8, IDX, S>;
    }
}
template <bool IDX, int S>
static occ_fn pick_occ(int kb)
{
    switch (kb) {
    case 1: return occupancy_of<1, IDX, S>;
    case 2: return occupancy_of<2, IDX, S>;
    case 4: return occupancy_of<4, IDX, S>;
    default: return occupancy_of<8, IDX, S>;
    }
}

static drain_fn find_drain(int kb, bool idx, int)
{
    return idx ? pick_drain<true, 3>(kb) : pick_drain<false, 3>(kb);
}

int engine_occupancy(int kb, bool idx, int)
{
    return idx ? pick_occ<true, 3>(kb)() : pick_occ<false, 3>(kb)();
}

bool engine_supports(int splitters) { return splitters == 3; }

uint32_t engine_tile(int kb) { return kb >= 8 ? 512u * 8u : 512u * 16u; }
uint32_t engine_cap(int kb) { return kb >= 8 ? 512u * 8u : 512u * 16u; }

hipError_t engine_run(const Params &p, int kb, bool idx, int s, int grid, const uint32_t *dev_off,
                      uint32_t nseg, uint32_t single_len, hipStream_t st)
{
    hipLaunchKernelGGL(ts_reset, dim3(1), dim3(64), 0, st, p.ctrl);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (nseg > 0) {
        hipLaunchKernelGGL(ts_roots, dim3((nseg + 255) / 256), dim3(256), 0, st, p, dev_off, nseg, single_len);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return find_drain(kb, idx, s)(p, grid, st);
}

}  // namespace ts
