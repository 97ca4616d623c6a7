// ts_levels_impl.h -- the sort engine for gfx950: level-synchronous
// partition passes over a device-resident work list, then LDS sorts.
// Included once per key type by ts_levels_kb{1,2,4,8}.hip and
// ts_levels_kf{4,8}.hip, which instantiate it; ts_levels.hip dispatches.
//
// The reference recurses with CUDA dynamic parallelism: each wave of
// quicksort blocks tail-launches the next wave (turtle-sort.cu:35-111,
// 586-587) and one 1024-thread block partitions one whole segment
// (turtle-sort.cu:573-934).  gfx950 has no device-side launch.  Here the
// recursion runs level by level; one level is a short chain of kernels over
// the segments of that level, and every segment is cut into tiles that all
// workgroups share:
//
//   k_pivot    opens level L (its segments are ids [sb, se) of the segment
//              table, created by the previous level); per segment: sample
//              1024 keys (runs of 8 evenly spaced), sort them in LDS, take
//              S splitters at the sample quantiles (S = 3 is the reference's
//              3-pivot scheme; 7, 15 and 31 cut a segment into more bins per
//              pass), reserve the segment's tiles;
//   k_count    per tile: bin counts; every key's bin is also stored (one
//              byte, four to a dword in a per-tile slot) so that the scatter
//              does not search the splitters again;
//   k_scan1/2  level-wide exclusive scan of the bin-major tile counts: the
//              destination of every (segment, bin, tile) run;
//   k_scatter  per tile: stable wave64 ranking into bins, staged in LDS,
//              written out coalesced; pivot-equal bins go straight to the
//              output (their keys are final), the other bins to the next
//              ping-pong buffer;
//   k_children per segment (one wave): every non-empty non-pivot bin becomes
//              a segment of level L+1 or, at or below the cutoff, an LDS task.
// After the last level the LDS tasks are grouped by size class
// (k_ldscount, k_ldsbucket) and the k_lds kernels sort them (merge sort for
// small tasks and for 8-byte keys, LSD radix for the others).
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
#pragma once
#include <type_traits>
#include "ts_device.h"
#include "ts_engine.h"
#include "ts_internal.h"

namespace ts {

template <int KT, bool IDX, int S>
struct Cfg {
    static constexpr int KB = kt_bytes(KT);               // key bytes
    static constexpr bool FL = kt_float(KT);              // IEEE floating-point keys
    typedef typename KeyT<KB>::U U;
    static constexpr int NT = 512;
    static constexpr int NW = NT / WAVE;
    static constexpr int NB = 2 * S + 1;
    static constexpr int NBP = (int)next_pow2(NB + 1);   // bins + the "no element" marker NB
    static constexpr int BITS = ilog2(NBP);               // bits of a bin id
    static constexpr int LOG = ilog2(S + 1);              // depth of the splitter search tree
    static constexpr int TI = (KB >= 8) ? 8 : 16;         // tile items per thread
    static constexpr int SI = (KB >= 8) ? 8 : 16;         // LDS-sort items per thread
    static constexpr uint32_t TILE = NT * TI;
    static constexpr uint32_t CAP = NT * SI;
    static constexpr uint32_t SAMPLE = 1024;
    static constexpr uint32_t IB = IDX ? 4 : 0;
    // LDS-sort size classes (see LevelK::lds_class): merge classes
    // [CMIN, CRAD) of n_eff = 2^c keys, radix classes [CRAD, CLAST] of
    // c - CRAD + 1 whole waves of WS slots
    static constexpr int CMIN = 6, CMAX = ilog2(CAP);
    static constexpr uint32_t WS = (uint32_t)SI * WAVE;
    static constexpr int CRAD = ilog2(WS);
    static constexpr int CLAST = CRAD + NW - 1;
    static constexpr uint32_t SORT_KEY_BYTES = (padded_len<SI>(CAP) * KB + 15) & ~15u;
    static constexpr uint32_t SORT_BYTES = SORT_KEY_BYTES + (IDX ? padded_len<SI>(CAP) * 4 : 0);
    static constexpr uint32_t SCAT_BYTES = TILE * KB + TILE * IB + TILE;
    // count pass: bucket table over the segment's key range (S >= 15; for
    // S <= 7 the splitters are compared in registers)
    static constexpr bool USE_LUT = S >= 15;
    static constexpr int LUTB = 10;
    static constexpr uint32_t NLUT = 1u << LUTB;
    static_assert((S + 1) == (1 << LOG), "S + 1 must be a power of two");
    static_assert(S <= MAX_SPLIT && NBP <= WAVE, "too many splitters");
    static_assert(CLAST < LDS_CLASSES && (1 << CMIN) >= SI && CMIN < CRAD && WS * NW == CAP, "size classes");
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

// ---- splitters ----------------------------------------------------------------------
// Sorted splitters in LDS, plus the same values in Eytzinger order (node i
// has children 2i and 2i+1; tree[1] is the median) for a branch-free
// search: after LOG steps i - (S+1) is the number of splitters < x.
template <typename U, int S>
struct Split {
    U tree[S + 1];
    U srt[S + 1];
};

template <typename U, int S, int LOG>
__device__ __forceinline__ void split_load(Split<U, S> &sp, const Seg *sg)
{
    const uint32_t t = threadIdx.x;
    if (t < (uint32_t)S) sp.srt[t] = (U)sg->split[t];
    if (t >= 1 && t <= (uint32_t)S) {
        const int d = 31 - __clz(t);                 // depth of node t
        const uint32_t k = t - (1u << d);            // position inside its depth
        const uint32_t r = (2 * k + 1) * (1u << (LOG - 1 - d)) - 1;   // in-order rank
        sp.tree[t] = (U)sg->split[r];
    }
}

// bin of order value x: lt = #splitters < x; equal to a splitter -> 2*lt+1
// (lt then indexes the lowest equal splitter), else 2*lt
template <typename U, int S, int LOG>
__device__ __forceinline__ uint32_t classify(U x, const Split<U, S> &sp, const U (&reg)[S <= 7 ? S : 1],
                                             const U (&reg3)[3])
{
    uint32_t lt;
    bool eq;
    if constexpr (S <= 7) {
        uint32_t le = 0;
        lt = 0;
#pragma unroll
        for (int j = 0; j < S; j++) {
            lt += (reg[j] < x) ? 1u : 0u;
            le += (reg[j] <= x) ? 1u : 0u;
        }
        eq = le > lt;
    } else {
        // the top two levels of the tree come from registers (reg3: nodes
        // 1, 2, 3, the same for every lane), the rest from LDS
        const U r1 = reg3[0], r2 = reg3[1], r3 = reg3[2];
        const uint32_t c1 = r1 < x ? 1u : 0u;
        const U t = c1 ? r3 : r2;
        uint32_t i = 4 + 2 * c1 + (t < x ? 1u : 0u);
#pragma unroll
        for (int l = 2; l < LOG; l++) i = 2 * i + ((sp.tree[i] < x) ? 1u : 0u);
        lt = i - (S + 1);
        eq = lt < (uint32_t)S && sp.srt[lt] == x;
    }
    return eq ? 2 * lt + 1 : 2 * lt;
}

// classify of N keys at once, level by level: the N LDS loads of one tree
// level are independent, so they are issued together and wait once
template <typename U, int S, int LOG, int N>
__device__ __forceinline__ void classify_n(const U (&x)[N], uint32_t (&bin)[N], const Split<U, S> &sp,
                                           const U (&reg)[S <= 7 ? S : 1], const U (&reg3)[3])
{
    if constexpr (S <= 7) {
#pragma unroll
        for (int c = 0; c < N; c++) bin[c] = classify<U, S, LOG>(x[c], sp, reg, reg3);
    } else {
        const U r1 = reg3[0], r2 = reg3[1], r3 = reg3[2];
        uint32_t i[N];
#pragma unroll
        for (int c = 0; c < N; c++) {
            const uint32_t c1 = r1 < x[c] ? 1u : 0u;
            const U t = c1 ? r3 : r2;
            i[c] = 4 + 2 * c1 + (t < x[c] ? 1u : 0u);
        }
#pragma unroll
        for (int l = 2; l < LOG; l++) {
            U tv[N];
#pragma unroll
            for (int c = 0; c < N; c++) tv[c] = sp.tree[i[c]];
#pragma unroll
            for (int c = 0; c < N; c++) i[c] = 2 * i[c] + (tv[c] < x[c] ? 1u : 0u);
        }
        U sv[N];
#pragma unroll
        for (int c = 0; c < N; c++) {
            const uint32_t lt = i[c] - (S + 1);
            sv[c] = sp.srt[lt < (uint32_t)S ? lt : 0u];
        }
#pragma unroll
        for (int c = 0; c < N; c++) {
            const uint32_t lt = i[c] - (S + 1);
            const bool eq = lt < (uint32_t)S && sv[c] == x[c];
            bin[c] = eq ? 2 * lt + 1 : 2 * lt;
        }
    }
}

// ---- set-up kernels ------------------------------------------------------------------
__global__ void k_reset(Ctrl *c);
__global__ void k_roots(Params p, const uint32_t *off, uint32_t nseg, uint32_t single_len, uint32_t reset);

// LDS-sort plan from the per-class task counts: first slot of every class
// in the class-ordered task list and first workgroup step of every class
// (class c < crad: cap / 2^c tasks per step; c >= crad: tasks of
// c - crad + 1 whole waves, nw / (c - crad + 1) per step); wg[LDS_CLASSES]
// is the total.  Every workgroup that needs it computes it (20 entries).
// Mixed steps: a step of two 3-wave tasks leaves 2 of 8 waves idle, so the
// first *mix steps of the 3-wave class also take one 2-wave task each (the
// first *mix tasks of that class, whose own steps start after them).
__device__ __forceinline__ void lds_plan(const Ctrl *c, uint32_t crad, uint32_t nw, uint32_t cap, uint32_t *off,
                                         uint32_t *wg, uint32_t *mix)
{
    const uint32_t k2 = crad + 1, k3 = crad + 2;   // the 2-wave and 3-wave radix classes
    uint32_t m = 0;
    if (nw == 8 && k3 < (uint32_t)LDS_CLASSES) {
        const uint32_t n2 = c->cls_cnt[k2], n3 = c->cls_cnt[k3];
        m = (n3 + 1) / 2 < n2 ? (n3 + 1) / 2 : n2;
    }
    uint32_t o = 0, g = 0;
    for (uint32_t k = 0; k < (uint32_t)LDS_CLASSES; k++) {
        off[k] = o;
        wg[k] = g;
        const uint32_t n = c->cls_cnt[k];
        o += n;
        const uint32_t ns = k == k2 ? n - m : n;   // tasks in the class's own steps
        uint32_t P = k < crad ? cap >> k : (k - crad + 1 <= nw ? nw / (k - crad + 1) : 1u);
        if (P == 0) P = 1;
        if (ns) g += (ns + P - 1) / P;
    }
    wg[LDS_CLASSES] = g;
    if (mix) *mix = m;
}

template <int KT, bool IDX, int S>
struct LevelK {
    typedef Cfg<KT, IDX, S> C;
    typedef typename C::U U;
    static constexpr int KB = C::KB;
    static constexpr bool FL = C::FL;
    static constexpr int NT = C::NT, NW = C::NW, NB = C::NB, NBP = C::NBP, BITS = C::BITS, LOG = C::LOG;
    static constexpr int TI = C::TI, SI = C::SI;
    static constexpr uint32_t TILE = C::TILE, CAP = C::CAP;
    typedef U Reg[S <= 7 ? S : 1];

    struct SharedScan {
        uint32_t red[16];
        uint32_t bcast[4];
    };
    struct SharedTile {
        uint32_t wcnt[NW][NBP];   // per-wave bin counts
        uint32_t woff[NW][NBP];   // per-wave bin offsets inside the tile
        uint32_t dbase[NBP];      // destination of bin b = dbase[b] + staging position
        Split<U, S> sp;
    };
    struct TaskRec {
        uint32_t begin, len, depth;
    };

    // Bucket table of the count pass.  The order values [lo, hi] of a
    // segment are cut into NLUT buckets of 2^sh values each; the entry of a
    // bucket holds lt0 = the number of splitters below the bucket (bits
    // 0-5), 0, 1 or "2 or more" splitters inside it (bits 6-7) and, with
    // exactly one inside, that splitter's offset from the bucket's first
    // value (bits 8 up; all ones otherwise).  A key's bin is then one LDS
    // load and two compares; buckets with several splitters (and keys
    // outside [lo, hi]) take the tree search.
    typedef typename std::conditional<KB == 8, uint64_t, uint32_t>::type LutE;
    static constexpr uint32_t NLUT = C::USE_LUT ? C::NLUT : 1u;
    static constexpr LutE LUT_NOOFF = (LutE)(~(LutE)0 >> 8);   // offset field of "no single splitter"
    static constexpr LutE LUT_MANY = (LutE)0x80u;
    struct SharedCount {
        uint32_t wcnt[2][NW][NBP];   // per-wave bin counts (two sets, alternating between tiles)
        Split<U, S> sp;
        LutE lut[NLUT];
        U lut_lo, lut_span;
        uint32_t lut_sh;
    };

    static __device__ __forceinline__ void load_reg(const Split<U, S> &sp, Reg &reg, U (&reg3)[3])
    {
        if constexpr (S <= 7) {
#pragma unroll
            for (int j = 0; j < S; j++) reg[j] = sp.srt[j];
        } else {
#pragma unroll
            for (int j = 0; j < 3; j++) reg3[j] = sp.tree[j + 1];
        }
    }

    // Bins buffer: tile g of the level owns bytes [g*TILE, (g+1)*TILE).  The
    // tile kernels give thread (w, lane) the items e = w*TI*WAVE + c*WAVE +
    // lane (c < TI); the bins of its items c = 4q..4q+3 are the four bytes
    // (low first) of the aligned dword at this offset, so the count pass
    // stores and the scatter loads one dword per four items.
    static __device__ __forceinline__ uint32_t bin_dword(uint32_t w, uint32_t q, uint32_t ln)
    {
        return w * (TI * WAVE) + q * (4 * WAVE) + ln * 4;
    }
    static_assert(TI % 4 == 0, "bins are packed four to a dword");

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

    // ---------------------------------------------------------------- count
    // number of splitters below v: the Eytzinger search, all levels in LDS
    static __device__ __forceinline__ uint32_t tree_lt(const Split<U, S> &sp, U v)
    {
        uint32_t i = 1;
#pragma unroll
        for (int l = 0; l < LOG; l++) i = 2 * i + ((sp.tree[i] < v) ? 1u : 0u);
        return i - (S + 1);
    }

    // fills the bucket table of segment sg (its sorted splitters and search
    // tree are already in sh.sp)
    static __device__ void lut_build(SharedCount &sh, const Seg *sg)
    {
        const U kmax = key_max<U>();
        const U lo = (U)sg->lo;
        const U hi = sg->hi > (uint64_t)kmax ? kmax : (U)sg->hi;
        const U span = hi > lo ? (U)(hi - lo) : (U)0;
        const uint32_t bits = span ? (uint32_t)(64 - __clzll((uint64_t)span)) : 0u;
        const uint32_t shf = bits > (uint32_t)C::LUTB ? bits - (uint32_t)C::LUTB : 0u;
        const uint64_t wm = (1ull << shf) - 1ull;   // values per bucket - 1 (shf < 64)
        if (threadIdx.x == 0) {
            sh.lut_lo = lo;
            sh.lut_span = span;
            sh.lut_sh = shf;
        }
        for (uint32_t b = threadIdx.x; b < NLUT; b += NT) {
            const uint64_t o = (uint64_t)b << shf;   // first value of the bucket, relative to lo
            LutE e = LUT_MANY | (LUT_NOOFF << 8);   // past hi: no key maps here
            if (o <= (uint64_t)span) {
                const U bs = (U)(lo + (U)o);
                const U be = ((uint64_t)(U)(span - (U)o) < wm) ? hi : (U)(bs + (U)wm);
                const uint32_t lt0 = tree_lt(sh.sp, bs);
                const uint32_t le = be == kmax ? (uint32_t)S : tree_lt(sh.sp, (U)(be + 1));
                const uint32_t m = le - lt0;
                LutE off = LUT_NOOFF;
                if (m == 1) off = (LutE)(U)(sh.sp.srt[lt0] - bs);
                e = (LutE)lt0 | (LutE)((m < 2 ? m : 2u) << 6) | (off << 8);
            }
            sh.lut[b] = e;
        }
    }

    // bins of N keys through the bucket table; keys in buckets with several
    // splitters or outside [lo, hi] take the tree search.  Groups of 8 keys:
    // their table loads are issued together and wait once, and fewer
    // entries are live at a time.
    template <int N>
    static __device__ __forceinline__ void classify_lut_n(const U (&x)[N], uint32_t (&bin)[N], const SharedCount &sh,
                                                          U lo, U span, uint32_t shf, const Reg &reg, const U (&reg3)[3])
    {
        // lo, span and shf are the same for every lane (scalar registers);
        // shf < 8 * KB, so the shifts stay in the key's width
        const LutE wm = (LutE)(((LutE)1 << shf) - 1u);
        constexpr int GR = KB == 8 ? (N < 4 ? N : 4) : (N < 8 ? N : 8);
        static_assert(N % GR == 0, "groups of keys");
#pragma unroll
        for (int g0 = 0; g0 < N; g0 += GR) {
            LutE e[GR];
#pragma unroll
            for (int c = 0; c < GR; c++) {
                const LutE q = (LutE)(U)(x[g0 + c] - lo) >> shf;
                e[c] = sh.lut[q < (LutE)NLUT ? (uint32_t)q : NLUT - 1];
            }
            uint32_t full = 0;
#pragma unroll
            for (int c = 0; c < GR; c++) {
                const U d = (U)(x[g0 + c] - lo);
                const uint32_t lt0 = (uint32_t)e[c] & 63u;
                const LutE off = e[c] >> 8;
                const LutE low = (LutE)d & wm;
                bin[g0 + c] = 2 * lt0 + (low > off ? 2u : 0u) + (low == off ? 1u : 0u);
                full |= ((((uint32_t)e[c] & (uint32_t)LUT_MANY) != 0) || d > span) ? (1u << c) : 0u;
            }
            if (ballot(full != 0)) {   // wave-uniform
#pragma unroll
                for (int c = 0; c < GR; c++)
                    if ((full >> c) & 1u) bin[g0 + c] = classify<U, S, LOG>(x[g0 + c], sh.sp, reg, reg3);
            }
        }
    }

    // A tile of the current level: its segment, its index inside the
    // segment, the segment's keys and the tile's key count.
    struct TileRef {
        uint32_t sid, t, tl;       // segment, tile inside it, keys in the tile
        uint32_t kbeg, depth;      // first key of the tile in its level buffer, segment depth
        uint32_t tile_base, ntiles;
        uint32_t begin, len;       // the segment
    };

    // The three key buffers a level reads from (by segment depth), as
    // opaque scalar values: selecting between the kernel argument's fields
    // directly lets the compiler turn the selection into a load at a
    // computed offset inside the argument block, which moves the whole
    // block to scratch memory.
    // The key and index buffers a level reads (by segment depth).  The
    // kernels' only (first) argument is the Params block, at offset 0 of
    // the kernel argument segment: the selected field is read from there
    // with a scalar load.  (Selecting between the argument's fields in C++
    // lets the compiler copy the whole block to scratch memory and index
    // it; a buffer resource from a value it cannot prove uniform turns
    // every buffer load into a loop over the lanes.)
    struct KeyBases {
        __device__ __forceinline__ explicit KeyBases(const Params &) {}
        static __device__ __forceinline__ uint64_t arg_u64(uint32_t off)
        {
            typedef const __attribute__((address_space(4))) uint64_t CU64;
            return ((CU64 *)__builtin_amdgcn_kernarg_segment_ptr())[off / 8];
        }
        __device__ __forceinline__ const U *src(uint32_t depth) const
        {
            return (const U *)arg_u64(depth == 0 ? (uint32_t)offsetof(Params, kin)
                                                 : ((depth & 1) ? (uint32_t)offsetof(Params, kt0)
                                                                : (uint32_t)offsetof(Params, kt1)));
        }
        __device__ __forceinline__ const uint32_t *isrc(uint32_t depth) const
        {
            return (const uint32_t *)arg_u64(depth == 0 ? (uint32_t)offsetof(Params, iin)
                                                        : ((depth & 1) ? (uint32_t)offsetof(Params, it0)
                                                                       : (uint32_t)offsetof(Params, it1)));
        }
    };
    static_assert(offsetof(Params, kin) % 8 == 0 && offsetof(Params, kt0) % 8 == 0 && offsetof(Params, kt1) % 8 == 0 &&
                      offsetof(Params, iin) % 8 == 0 && offsetof(Params, it0) % 8 == 0 &&
                      offsetof(Params, it1) % 8 == 0,
                  "buffer pointers of the argument block are 8-byte aligned");

    // a value every lane holds alike, moved to scalar registers
    static __device__ __forceinline__ U uniform_val(U v)
    {
        // (readfirstlane returns int: each half goes through uint32_t, else
        // the low half would be sign-extended into the high one)
        if constexpr (KB == 8)
            return ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
                   (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
        else
            return (U)__builtin_amdgcn_readfirstlane((uint32_t)v);
    }

    // TileRef of tile g (false, after a range_fail, when it is inconsistent).
    // prev: the tile before g when it is known, so that a tile of the same
    // segment needs no table loads.  Every field is block-uniform and is
    // kept in scalar registers (readfirstlane), so that the key loads take
    // a scalar base address.
    // (the tile and segment tables are read through the constant address
    // space: scalar loads; the count and scatter passes do not write them)
    static __device__ __forceinline__ bool tile_ref(const Params &p, uint32_t g, const TileRef *prev, TileRef &r)
    {
        typedef const __attribute__((address_space(4))) Seg CSeg;
        typedef const __attribute__((address_space(4))) uint32_t CU32;
        if (prev && prev->t + 1 < prev->ntiles) {
            r = *prev;
            r.t = prev->t + 1;
        } else {
            const uint32_t sid = ((CU32 *)p.tile_seg)[g];
            if (sid >= p.segcap) {
                if (threadIdx.x == 0) range_fail(p, 4);
                return false;
            }
            CSeg *sg = (CSeg *)p.segs + sid;
            r.sid = sid;
            r.tile_base = sg->tile_base;
            r.ntiles = sg->ntiles;
            r.depth = sg->depth;
            r.begin = sg->begin;
            r.len = sg->len;
            r.t = g - r.tile_base;
            if (g < r.tile_base || r.t >= r.ntiles || !span_ok(p, r.begin, r.len) ||
                (uint64_t)r.t * TILE >= r.len) {
                if (threadIdx.x == 0) range_fail(p, 4);
                return false;
            }
        }
        if (g >= p.bins_tiles) {
            if (threadIdx.x == 0) range_fail(p, 12);
            return false;
        }
        const uint32_t ts = r.t * TILE;
        r.tl = (r.len - ts) < TILE ? (r.len - ts) : TILE;
        r.kbeg = r.begin + ts;
        return true;
    }

    // issues the loads of a tile's keys (buffer loads bounded by the tile:
    // positions past its end read 0, their bins are set to the marker later)
    static __device__ __forceinline__ void count_load(const KeyBases &kb, const TileRef &r, U (&x)[TI])
    {
        // the thread index is re-read through an empty asm statement: the
        // offsets derived from it are recomputed here instead of being
        // hoisted out of the tile loop and spilled
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const uint32_t ebase = (tid / WAVE) * (TI * WAVE) + tid % WAVE;
        const uint32_t depth = __builtin_amdgcn_readfirstlane(r.depth), kbeg = __builtin_amdgcn_readfirstlane(r.kbeg),
                       tl = __builtin_amdgcn_readfirstlane(r.tl);   // (uniform resource: see KeyBases)
        const BufRsrc rs = buf_rsrc(kb.src(depth) + kbeg, tl * (uint32_t)KB);
#pragma unroll
        for (int c = 0; c < TI; c++) x[c] = buf_load<U>(rs, (ebase + (uint32_t)c * WAVE) * (uint32_t)KB);
    }

    // Count pass, first half of a tile: the bins of the keys in x (loaded by
    // count_load), four to a dword in the layout the scatter reads; items
    // past the tile's end hold the marker NB.
    static __device__ __forceinline__ void count_classify(const Params &p, SharedCount &sh, const TileRef &r,
                                                          const Reg &reg, const U (&reg3)[3], U (&x)[TI],
                                                          uint32_t (&bw)[TI / 4])
    {
        const U flip = (U)p.flip;
        uint32_t bins[TI];
#pragma unroll
        for (int c = 0; c < TI; c++) x[c] = to_ord<U, FL>(x[c], flip);
        if constexpr (C::USE_LUT)
            classify_lut_n<TI>(x, bins, sh, uniform_val(sh.lut_lo), uniform_val(sh.lut_span),
                               __builtin_amdgcn_readfirstlane(sh.lut_sh), reg, reg3);
        else
            classify_n<U, S, LOG, TI>(x, bins, sh.sp, reg, reg3);
        if (r.tl < TILE) {   // block-uniform
            uint32_t tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            const uint32_t ebase = (tid / WAVE) * (TI * WAVE) + tid % WAVE;
#pragma unroll
            for (int c = 0; c < TI; c++)
                if (ebase + (uint32_t)c * WAVE >= r.tl) bins[c] = NB;
        }
#pragma unroll
        for (int q = 0; q < TI / 4; q++)
            bw[q] = bins[4 * q] | (bins[4 * q + 1] << 8) | (bins[4 * q + 2] << 16) | (bins[4 * q + 3] << 24);
    }

    // Count pass, second half of a tile: per-wave counters (wc: this tile's
    // set; two sets alternate between tiles, so one barrier per tile
    // suffices), the bins stored for the scatter, the tile's counts.
    // FULL: the tile holds TILE keys.
    template <bool FULL>
    static __device__ __forceinline__ void count_accum(const Params &p, uint32_t (*wc)[NBP], const TileRef &r,
                                                       uint32_t g, const uint32_t (&bw)[TI / 4])
    {
        uint32_t tid = threadIdx.x;   // opaque copy: see count_load
        asm volatile("" : "+v"(tid));
        const int w = __builtin_amdgcn_readfirstlane(tid / WAVE);
        const uint32_t ln = tid % WAVE;
        uint32_t *bout = (uint32_t *)(p.bins + (size_t)g * TILE);
        const uint32_t tl = FULL ? TILE : r.tl;
        const uint32_t ebase = (uint32_t)w * (TI * WAVE) + ln;
        auto binof = [&](int c) { return (bw[c / 4] >> (8 * (c % 4))) & 0xffu; };
        // An item whose valid lanes all hold one bin (sorted runs, heavy
        // duplicates) adds its count with one atomic: 64 same-address
        // atomics would be serialised by the LDS.  The items are tested only
        // when the wave's first or last item is such a one, so that random
        // keys pay two tests per tile.
        auto uniform = [&](int c, uint32_t &b0, uint64_t &vm) {
            const bool ok = FULL || ebase + (uint32_t)c * WAVE < tl;
            const uint32_t b = binof(c);
            b0 = __builtin_amdgcn_readfirstlane(b);
            vm = ballot(ok);
            return ballot(ok && b == b0) == vm;
        };
        uint32_t bf, bl;
        uint64_t mf, ml;
        if (uniform(0, bf, mf) || uniform(TI - 1, bl, ml)) {   // wave-uniform
#pragma unroll
            for (int c = 0; c < TI; c++) {
                uint32_t b0 = 0;
                uint64_t vm = 0;
                if (uniform(c, b0, vm)) {
                    if (ln == 0 && vm) atomicAdd(&wc[w][b0], (uint32_t)__popcll(vm));
                } else if (FULL || ebase + (uint32_t)c * WAVE < tl) {
                    atomicAdd(&wc[w][binof(c)], 1u);
                }
            }
        } else {
#pragma unroll
            for (int c = 0; c < TI; c++)
                if (FULL || ebase + (uint32_t)c * WAVE < tl) atomicAdd(&wc[w][binof(c)], 1u);
        }
        // the scatter reads the bins back
#pragma unroll
        for (int q = 0; q < TI / 4; q++) bout[bin_dword(w, q, ln) / 4] = bw[q];
        block_sync();
        if (threadIdx.x < (uint32_t)NB) {
            uint32_t s = 0;
#pragma unroll
            for (int i = 0; i < NW; i++) s += wc[i][threadIdx.x];
            p.tcnt[(size_t)r.tile_base * NB + (size_t)threadIdx.x * r.ntiles + r.t] = s;
        }
    }

    // Each workgroup takes a contiguous range of the level's tiles, so that
    // it changes segment (splitter and bucket table loads, two barriers)
    // rarely, and issues the loads of every tile's keys as soon as the tile
    // before is classified, so that they are in flight while that tile's
    // counters are added up.
    static __device__ void count(const Params &p, SharedCount &sh)
    {
        // (the tile count comes from a vector load and the division is done
        // in vector registers: readfirstlane keeps the range, and so the tile
        // loop and every address derived from it, in scalar registers)
        const uint32_t ntot = __builtin_amdgcn_readfirstlane(p.ctrl->tile_alloc < p.tilecap ? p.ctrl->tile_alloc
                                                                                            : p.tilecap);
        const uint32_t per = __builtin_amdgcn_readfirstlane(ntot / gridDim.x);
        const uint32_t rem = ntot - per * gridDim.x;
        const uint32_t g0 = blockIdx.x * per + (blockIdx.x < rem ? blockIdx.x : rem);
        const uint32_t g1 = g0 + per + (blockIdx.x < rem ? 1u : 0u);
        const int w = threadIdx.x / WAVE;
        const KeyBases kb(p);
        uint64_t nbytes = 0, ntiles = 0;
        uint32_t cur = ~0u;
        Reg reg;
        U reg3[3];
        U x[TI];
        TileRef r{}, rn{};
        bool ok = false;
        uint32_t g = g0;
        // the first consistent tile of the range
        while (g < g1 && !(ok = tile_ref(p, g, nullptr, r))) g++;
        if (ok) count_load(kb, r, x);
        uint32_t par = 0;
        while (g < g1) {
            if (r.sid != cur) {
                split_load<U, S, LOG>(sh.sp, &p.segs[r.sid]);
                block_sync();
                if constexpr (C::USE_LUT) {
                    lut_build(sh, &p.segs[r.sid]);
                    block_sync();
                }
                cur = r.sid;
            }
            load_reg(sh.sp, reg, reg3);
            uint32_t (*wc)[NBP] = sh.wcnt[par];
            par ^= 1u;
            if (lane() < (uint32_t)NBP) wc[w][lane()] = 0;
            wave_sync();
            nbytes += (uint64_t)r.tl * (KB + 1);
            ntiles++;
            uint32_t bw[TI / 4];
            count_classify(p, sh, r, reg, reg3, x, bw);
            // the next consistent tile of the range, its loads issued (the
            // scheduling barrier keeps them after the classification: both
            // tiles' keys would otherwise be live at once)
            __builtin_amdgcn_sched_barrier(0);
            uint32_t gn = g + 1;
            bool okn = false;
            const TileRef *pv = &r;
            while (gn < g1 && !(okn = tile_ref(p, gn, pv, rn))) {
                gn++;
                pv = nullptr;
            }
            if (okn) count_load(kb, rn, x);
            if (r.tl == TILE) count_accum<true>(p, wc, r, g, bw);
            else count_accum<false>(p, wc, r, g, bw);
            g = gn;
            r = rn;
        }
        block_sync();
        if (threadIdx.x == 0 && nbytes) {
            atomic_add64(&p.ctrl->bytes_count, nbytes);
            atomic_add64(&p.ctrl->tasks[T_COUNT], ntiles ? 1ull : 0ull);
        }
    }

    // ---------------------------------------------------------------- scan
    // Level-wide exclusive scan of tcnt[0, tiles*NB) in chunks of SCAN_CHUNK
    // entries: scan1 sums every chunk, scan2 adds the sums of the chunks
    // before it and scans inside.  Values are < 2^31 (keys of one level).
    static __device__ void scan1(const Params &p, SharedScan &sh)
    {
        const uint32_t nt = p.ctrl->tile_alloc < p.tilecap ? p.ctrl->tile_alloc : p.tilecap;
        const uint32_t E = nt * (uint32_t)NB;
        const uint32_t nch = (E + SCAN_CHUNK - 1) / SCAN_CHUNK;
        for (uint32_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
            const uint32_t a = ch * SCAN_CHUNK, b = (a + SCAN_CHUNK) < E ? a + SCAN_CHUNK : E;
            uint32_t s = 0;
            for (uint32_t i = a + threadIdx.x; i < b; i += NT) s += p.tcnt[i];
            uint32_t tot;
            (void)block_excl_scan<NT>(s, sh.red, &tot);
            if (threadIdx.x == 0) p.spart[ch] = tot;
        }
    }

    static __device__ void scan2(const Params &p, SharedScan &sh)
    {
        constexpr uint32_t PER = SCAN_CHUNK / NT;
        const uint32_t nt = p.ctrl->tile_alloc < p.tilecap ? p.ctrl->tile_alloc : p.tilecap;
        const uint32_t E = nt * (uint32_t)NB;
        const uint32_t nch = (E + SCAN_CHUNK - 1) / SCAN_CHUNK;
        for (uint32_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
            uint32_t s = 0;
            for (uint32_t c = threadIdx.x; c < ch; c += NT) s += p.spart[c];
            uint32_t off;
            (void)block_excl_scan<NT>(s, sh.red, &off);
            const uint32_t a = ch * SCAN_CHUNK + threadIdx.x * PER;
            uint32_t v[PER];
            uint32_t loc = 0;
#pragma unroll
            for (int k = 0; k < (int)PER; k++) {
                v[k] = (a + k < E) ? p.tcnt[a + k] : 0u;
                loc += v[k];
            }
            uint32_t tot;
            uint32_t run = off + block_excl_scan<NT>(loc, sh.red, &tot);
#pragma unroll
            for (int k = 0; k < (int)PER; k++) {
                if (a + k < E) p.tpre[a + k] = run;
                run += v[k];
            }
        }
    }

    // ---------------------------------------------------------------- scatter
    // One tile of the scatter.  FULL: the tile holds TILE keys, so no
    // per-element bounds predicates.  Consistency checks are branch-free
    // (a flag per thread, reported once per tile).
    template <bool FULL>
    static __device__ __forceinline__ void scatter_tile(const Params &p, unsigned char *arena, SharedTile &sh,
                                                        const Seg *sg, uint32_t t)
    {
        U *stk = (U *)arena;
        uint32_t *sti = (uint32_t *)(arena + TILE * KB);
        unsigned char *stb = arena + TILE * KB + TILE * C::IB;
        U *kfin = (U *)p.kout;
        uint32_t *ifin = p.iout;
        // the thread index is re-read through an empty asm statement in every
        // tile: the element offsets and LDS addresses derived from it are then
        // recomputed per tile instead of being hoisted out of the tile loop
        // and spilled
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int w = __builtin_amdgcn_readfirstlane(tid / WAVE);
        const uint32_t ln = tid % WAVE;
        const uint32_t begin = sg->begin, len = sg->len, depth = sg->depth;
        const uint32_t tb = sg->tile_base, ntl = sg->ntiles;
        const U *src = src_keys<U>(p, depth) + begin;
        const uint32_t *isrc = src_idx(p, depth) + begin;
        const uint32_t *bsrc = (const uint32_t *)(p.bins + (size_t)(tb + t) * TILE);
        U *knext = next_keys<U>(p, depth);
        uint32_t *inext = next_idx(p, depth);
        const uint32_t ts = t * TILE;
        const uint32_t tl = FULL ? TILE : ((len - ts) < TILE ? (len - ts) : TILE);
        bool bad = false;
        if (ln < (uint32_t)NBP) sh.wcnt[w][ln] = 0;
        wave_sync();
        U x[TI];
        uint32_t ix[TI];
        uint32_t pk[TI];   // bin | rank inside the wave's bin << 8
        // wave 0 needs the tile's destination prefixes after the ranking:
        // their loads are issued first, so that their latency is covered
        uint32_t pre_bin = 0, pre_seg = 0;
        if (w == 0 && ln < (uint32_t)NB) {
            const size_t e0 = (size_t)tb * NB;
            pre_bin = p.tpre[e0 + (size_t)ln * ntl + t];
            pre_seg = p.tpre[e0];
        }
        // keys and the bins the count pass stored for them
        uint32_t bw[TI / 4];
#pragma unroll
        for (int q = 0; q < TI / 4; q++) bw[q] = bsrc[bin_dword(w, q, ln) / 4];
#pragma unroll
        for (int c = 0; c < TI; c++) {
            const uint32_t e = (uint32_t)w * (TI * WAVE) + c * WAVE + ln;
            const uint32_t b = (bw[c / 4] >> (8 * (c % 4))) & 0xffu;
            if (FULL) {
                x[c] = src[ts + e];
                pk[c] = b;
            } else {
                x[c] = e < tl ? src[ts + e] : (U)0;
                pk[c] = e < tl ? b : (uint32_t)NB;
            }
        }
        // rank: one LDS atomic on the wave's own bin counter (one wave owns
        // wcnt[w]); its return value counts the wave's earlier keys of the
        // bin in (item, lane) order, i.e. in element order (see rank16_add)
        // An item whose valid lanes all hold one bin (sorted runs) takes its
        // ranks with one atomic: 64 same-address atomics are serialised by
        // the LDS.  Items are tested only when the wave's first or last item
        // is such one, so that random keys pay two tests per tile.
        auto uniform = [&](int c, uint32_t &b0, uint64_t &vm) {
            const uint32_t e = (uint32_t)w * (TI * WAVE) + c * WAVE + ln;
            const bool valid = (FULL || e < tl) && pk[c] < (uint32_t)NB;
            b0 = __builtin_amdgcn_readfirstlane(pk[c]);
            vm = ballot(valid);
            return vm != 0 && ballot(valid && pk[c] == b0) == vm;
        };
        uint32_t bu;
        uint64_t mu;
        const bool runs = uniform(0, bu, mu) || uniform(TI - 1, bu, mu);   // wave-uniform
#pragma unroll
        for (int c = 0; c < TI; c++) {
            const uint32_t e = (uint32_t)w * (TI * WAVE) + c * WAVE + ln;
            const bool valid = (FULL || e < tl) && pk[c] < (uint32_t)NB;
            bad |= (FULL || e < tl) && pk[c] >= (uint32_t)NB;
            const uint32_t bin = valid ? pk[c] : (uint32_t)NB;
            uint32_t old = 0, b0;
            uint64_t vm;
            if (runs && uniform(c, b0, vm)) {   // wave-uniform
                const uint32_t first = (uint32_t)__builtin_ctzll(vm);
                uint32_t o0 = 0;
                if (ln == first) o0 = atomicAdd(&sh.wcnt[w][b0], (uint32_t)__popcll(vm));
                old = __builtin_amdgcn_readlane(o0, first) + mask_rank(vm);
            } else if (valid) {
                old = atomicAdd(&sh.wcnt[w][bin], 1u);   // rank in (item, lane) order: see rank16_add
            }
            pk[c] = bin | (old << 8);
        }
        // the first half of the indices is loaded now (they are not needed
        // for the ranking; the barrier and the offset scan cover the
        // latency), the second half while the first is staged
        constexpr int H = TI / 2;
#pragma unroll
        for (int c = 0; c < H; c++) {
            const uint32_t e = (uint32_t)w * (TI * WAVE) + c * WAVE + ln;
            if (IDX) ix[c] = (FULL || e < tl) ? isrc[ts + e] : 0u;
        }
        block_sync();
        if (w == 0) {
            const uint32_t l = ln;
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
                sh.dbase[l] = begin + (pre_bin - pre_seg) - start;
            }
        }
        block_sync();
#pragma unroll
        for (int h = 0; h < 2; h++) {
            if (h == 1) {
#pragma unroll
                for (int c = H; c < TI; c++) {
                    const uint32_t e = (uint32_t)w * (TI * WAVE) + c * WAVE + ln;
                    if (IDX) ix[c] = (FULL || e < tl) ? isrc[ts + e] : 0u;
                }
            }
#pragma unroll
            for (int c = h * H; c < (h + 1) * H; c++) {
                const uint32_t e = (uint32_t)w * (TI * WAVE) + c * WAVE + ln;
                const uint32_t bin = pk[c] & 0xffu;
                uint32_t pos = sh.woff[w][bin] + (pk[c] >> 8);
                const bool valid = FULL || e < tl;
                bad |= valid && pos >= tl;
                pos &= TILE - 1;                   // stays inside the staging arrays
                if (valid) {
                    stk[pos] = x[c];
                    if (IDX) sti[pos] = ix[c];
                    stb[pos] = (unsigned char)bin;
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        block_sync();
        // write-out in groups of WU positions per thread: the LDS reads of a
        // group are issued together and wait once
        constexpr int WU = 4;
#pragma nounroll
        for (uint32_t i0 = tid; i0 < tl; i0 += WU * NT) {
            uint32_t bb[WU], dd[WU];
#pragma unroll
            for (int u = 0; u < WU; u++) {
                const uint32_t i = i0 + (uint32_t)u * NT;
                bb[u] = (FULL || i < tl) ? (uint32_t)stb[i & (TILE - 1)] : (uint32_t)NB;
            }
#pragma unroll
            for (int u = 0; u < WU; u++) dd[u] = sh.dbase[bb[u] < (uint32_t)NB ? bb[u] : 0u];
#pragma unroll
            for (int u = 0; u < WU; u++) {
                const uint32_t i = i0 + (uint32_t)u * NT;
                const uint32_t b = bb[u];
                const uint32_t d = dd[u] + i;
                const bool real = FULL || i < tl;
                const bool ok = b < (uint32_t)NB && d - begin < len;   // inside the segment
                bad |= real && !ok;
                if (real && ok) {
                    U *kd = (b & 1u) ? kfin : knext;
                    kd[d] = stk[i];
                    if (IDX) {
                        uint32_t *id = (b & 1u) ? ifin : inext;
                        id[d] = sti[i];
                    }
                }
            }
        }
        if (bad) range_fail(p, 8);
        block_sync();
    }

    static __device__ void scatter(const Params &p, unsigned char *arena, SharedTile &sh)
    {
        const uint32_t ntot = p.ctrl->tile_alloc < p.tilecap ? p.ctrl->tile_alloc : p.tilecap;
        const uint32_t G = p.tiles_per_wg;
        uint64_t nbytes = 0;
        // XCD-aware order: workgroups are dispatched round-robin over the 8
        // XCDs (workgroup b runs on XCD b % 8), and each XCD has its own L2.
        // The sweep's chunks are renumbered (an 8 x gridDim.x/8 transpose) so
        // that neighbouring chunks go to workgroups of one XCD: a bin's run
        // from one tile and the next tile's run of that bin share their edge
        // lines, which then fill up in one L2 instead of being written back
        // as partial lines from two
        uint32_t b = blockIdx.x;
        if (gridDim.x % 8 == 0) b = (b % 8) * (gridDim.x / 8) + b / 8;
        for (uint32_t c0 = b * G; c0 < ntot; c0 += gridDim.x * G) {
            const uint32_t c1 = c0 + G < ntot ? c0 + G : ntot;
            for (uint32_t g = c0; g < c1; g++) {
                uint32_t sid, t;
                if (!tile_of(p, g, sid, t)) {
                    if (threadIdx.x == 0) range_fail(p, 6);
                    continue;
                }
                if (g >= p.bins_tiles) {
                    if (threadIdx.x == 0) range_fail(p, 12);
                    continue;
                }
                const Seg *sg = &p.segs[sid];
                const uint32_t ts = t * TILE;
                const uint32_t tl = (sg->len - ts) < TILE ? (sg->len - ts) : TILE;
                nbytes += (uint64_t)tl * (2 * (KB + C::IB) + 1);
                if (tl == TILE) scatter_tile<true>(p, arena, sh, sg, t);
                else scatter_tile<false>(p, arena, sh, sg, t);
            }
        }
        if (threadIdx.x == 0 && nbytes) {
            atomic_add64(&p.ctrl->bytes_scatter, nbytes);
            atomic_add64(&p.ctrl->tasks[T_SCATTER], 1ull);
        }
    }

    // ---------------------------------------------------------------- children
    // one wave per segment of level L (CW waves per workgroup step); lane j
    // handles the non-pivot bin 2j.  The workgroup reserves the step's child
    // segments and LDS tasks with one atomic per counter (a level of 1024
    // segments otherwise queues 2048 atomics on two words).
    static constexpr uint32_t CW = 4;   // waves per children workgroup (256 threads)
    struct ChildShared {
        uint32_t nb[CW], ns[CW];
        uint32_t sbase, lbase;
    };

    static __device__ void children(const Params &p, uint32_t L, ChildShared &cs)
    {
        const Level lv = p.levels[L];
        // the tiles of this level are done with: the next level's pivot
        // kernel reserves its tiles from zero
        if (blockIdx.x == 0 && threadIdx.x == 0) p.ctrl->tile_alloc = 0;
        const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
        for (uint32_t s0 = lv.sb + blockIdx.x * CW; s0 < lv.se; s0 += gridDim.x * CW) {   // block-uniform
            const uint32_t sid = s0 + w;
            const uint32_t j = lane();
            uint32_t cl = 0, cb = 0, depth = 0, cmode = SPLIT_SAMPLE;
            uint64_t clo = 0, chi = 0;
            if (sid < lv.se) {   // wave-uniform
                const Seg *sg = &p.segs[sid];
                const uint32_t nt = sg->ntiles;
                if (nt != 0) {   // else an overflowed segment
                    const uint32_t begin = sg->begin, len = sg->len;
                    depth = sg->depth;
                    const size_t e0 = (size_t)sg->tile_base * NB, e1 = e0 + (size_t)NB * nt;
                    const uint32_t p0 = p.tpre[e0];
                    const uint32_t sum = p.tpre[e1 - 1] + p.tcnt[e1 - 1] - p0;
                    if (sum != len) {
                        if (j == 0) atomic_or(&p.ctrl->error, ERR_BAD);
                    } else if (j <= (uint32_t)S) {
                        const uint32_t b = 2 * j;
                        const uint32_t st = p.tpre[e0 + (size_t)b * nt] - p0;
                        const uint32_t en = (b + 1 < (uint32_t)NB) ? p.tpre[e0 + (size_t)(b + 1) * nt] - p0 : len;
                        cl = en - st;
                        cb = begin + st;
                        // bin 2j holds keys strictly between splitters j-1
                        // and j (a non-empty bin has a key there, so the
                        // bounds do not wrap); a child that kept more than
                        // half of its parent is split by range next, which
                        // bounds the depth by the key bits
                        clo = j == 0 ? sg->lo : sg->split[j - 1] + 1;
                        chi = j == (uint32_t)S ? sg->hi : sg->split[j] - 1;
                        cmode = cl > len / 2 ? SPLIT_RANGE : SPLIT_SAMPLE;
                    }
                }
            }
            const bool big = cl > p.cutoff;
            const bool small = cl > 0 && !big;
            const uint64_t mb = ballot(big), ms = ballot(small);
            if (j == 0) {
                cs.nb[w] = (uint32_t)__popcll(mb);
                cs.ns[w] = (uint32_t)__popcll(ms);
            }
            block_sync();
            if (threadIdx.x == 0) {
                uint32_t tb = 0, tsm = 0;
                for (uint32_t x = 0; x < CW; x++) {
                    tb += cs.nb[x];
                    tsm += cs.ns[x];
                }
                cs.sbase = tb ? atomic_add(&p.ctrl->seg_alloc, tb) : 0u;
                cs.lbase = tsm ? lds_reserve(p, tsm) : 0u;
            }
            block_sync();
            uint32_t sbase = cs.sbase, lbase = cs.lbase;
            for (uint32_t x = 0; x < w; x++) {
                sbase += cs.nb[x];
                lbase += lbase != ~0u ? cs.ns[x] : 0u;
            }
            if (big) {
                const uint32_t cid = sbase + mask_rank(mb);
                if (cid >= p.segcap) {
                    atomic_or(&p.ctrl->error, ERR_SEGS);
                } else {
                    Seg *c = &p.segs[cid];
                    c->begin = cb;
                    c->len = cl;
                    c->depth = depth + 1;
                    c->mode = cmode;
                    c->lo = clo;
                    c->hi = chi;
                }
            } else if (small && lbase != ~0u) {
                LdsTask *tk = &p.ldst[lbase + mask_rank(ms)];
                tk->begin = cb;
                tk->len = cl;
                tk->depth = depth + 1;
            }
            block_sync();   // the next step rewrites cs
        }
    }

    // ---------------------------------------------------------------- LDS sort
    // tasks per size class (per-workgroup LDS histogram, one atomic per class)
    static __device__ void ldscount(const Params &p, uint32_t *hist)
    {
        const uint32_t ntask = p.ctrl->lds_alloc < p.ldscap ? p.ctrl->lds_alloc : p.ldscap;
        if (threadIdx.x < (uint32_t)LDS_CLASSES) hist[threadIdx.x] = 0;
        block_sync();
        for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < ntask; k += gridDim.x * blockDim.x)
            atomicAdd(&hist[lds_class(p.ldst[k].len)], 1u);
        block_sync();
        if (threadIdx.x < (uint32_t)LDS_CLASSES && hist[threadIdx.x]) atomic_add(&p.ctrl->cls_cnt[threadIdx.x], hist[threadIdx.x]);
    }

    // task ids ordered by class (order inside a class is arbitrary; every
    // task is sorted on its own, so the output does not depend on it).  Each
    // workgroup takes a contiguous chunk of the tasks: it counts the chunk's
    // classes, reserves their slots with one atomic per class, then places
    // the chunk's tasks (reserving per class and 256 tasks made ~84K
    // atomics on 20 words for 1.66M tasks).
    static __device__ void ldsbucket(const Params &p, uint32_t *hist, uint32_t *wbase, uint32_t *coff, uint32_t *cwg)
    {
        const uint32_t ntask = p.ctrl->lds_alloc < p.ldscap ? p.ctrl->lds_alloc : p.ldscap;
        const uint32_t per = (ntask + gridDim.x - 1) / gridDim.x;
        const uint32_t a = blockIdx.x * per < ntask ? blockIdx.x * per : ntask;
        const uint32_t b = ntask - a < per ? ntask : a + per;
        if (threadIdx.x == 0) lds_plan(p.ctrl, CPLAN, NW, CAP, coff, cwg, nullptr);
        if (threadIdx.x < (uint32_t)LDS_CLASSES) hist[threadIdx.x] = 0;
        block_sync();
        for (uint32_t k = a + threadIdx.x; k < b; k += blockDim.x) atomicAdd(&hist[lds_class(p.ldst[k].len)], 1u);
        block_sync();
        if (threadIdx.x < (uint32_t)LDS_CLASSES)
            wbase[threadIdx.x] = hist[threadIdx.x] ? atomic_add(&p.ctrl->cls_cur[threadIdx.x], hist[threadIdx.x]) : 0u;
        block_sync();
        for (uint32_t k = a + threadIdx.x; k < b; k += blockDim.x) {
            const uint32_t c = lds_class(p.ldst[k].len);
            const uint32_t slot = coff[c] + atomicAdd(&wbase[c], 1u);   // LDS: a distinct slot per task
            if (slot < p.ldscap) p.ldsq[slot] = p.ldst[k];
            else range_fail(p, 11);
        }
    }

    // ---------------------------------------------------------------- LDS sorts
    // Size class of a leaf task of len keys (WS = the slots of one wave):
    // len <= WS/2: class c = max(CMIN, ceil(log2 len)), merge-sorted, CAP >> c
    // tasks per workgroup step in power-of-two regions; longer: class
    // CRAD + kw - 1 with kw = ceil(len / WS) whole waves, radix-sorted,
    // NW / kw tasks per step.
    static constexpr uint32_t WS = C::WS;
    static constexpr uint32_t CRAD = (uint32_t)C::CRAD, CLAST = (uint32_t)C::CLAST;

    // MERGE_ONLY: every task takes the merge path.  Off: 8-byte keys, whose
    // full-range leaves still span ~50 bits, take the radix with two passes
    // over the top bits of each task's span and a fix-up of the keys that
    // share them (radix_step), which beats the log2(n/SI) merge rounds (the
    // batched 20,000-segment int64 run: 8.9 -> 5.9 ms in the LDS sorts).
    static constexpr bool MERGE_ONLY = false;
    static constexpr uint32_t CPLAN = MERGE_ONLY ? (uint32_t)C::CMAX + 1 : (uint32_t)CRAD;   // plan's first radix class
    static __device__ __forceinline__ uint32_t lds_class(uint32_t len)
    {
        if (MERGE_ONLY || len <= WS / 2) {
            const uint32_t c = len <= 1 ? 0u : (uint32_t)(32 - __clz(len - 1));
            return c < (uint32_t)C::CMIN ? (uint32_t)C::CMIN : c;
        }
        const uint32_t kw = (len + WS - 1) / WS;
        return kw <= (uint32_t)NW ? CRAD + kw - 1 : CLAST;   // len <= CAP: kw <= NW
    }

    // Radix classes: a task's keys span a narrow range (they lie between two
    // splitters of their parent), so a stable LSD radix sort of key - min
    // over the range's bits takes a few passes of RW-bit digits, each pass
    // ranking the keys with one LDS atomic per key on per-wave digit
    // counters (rank16_add); padding slots take the largest digit in every
    // pass and so stay behind the task's keys.
    static constexpr int RW = 9;
    static constexpr uint32_t NBIN = 1u << RW;
    static constexpr uint32_t MAXPASS = (8 * KB + RW - 1) / RW;   // covers any key range
    static constexpr uint32_t EPT = NW * NBIN / NT;   // scan entries per thread

    struct RadixShared {
        uint32_t coff[LDS_CLASSES];       // LDS-sort plan (lds_plan)
        uint32_t cwg[LDS_CLASSES + 1];
        uint32_t mix;                     // mixed 3+3+2-wave steps (lds_plan)
        alignas(16) uint16_t cnt[NW][NBIN];   // per-wave digit counts (rank16_add), then the digits' first slots
        U wmin[NW], wmax[NW];             // per-wave key extremes of the step
        uint32_t red[16];
        uint32_t flag;                    // a group of the step exceeded the group limit
    };

    struct PivotShared {
        alignas(4) uint16_t cnt[NW][NBIN];   // per-wave digit counts (rank16_add), then the digits' first positions
        U buf[NW][C::SAMPLE];         // per-wave sample
        uint32_t ntl[NW];             // tiles of each wave's segment in this step
        uint32_t tbase, done;         // the step's first tile; segments done
    };

    // lane src's x for every lane; src is the same for every lane (a scalar
    // read of the lane, no LDS permute)
    template <typename T>
    static __device__ __forceinline__ T shfl_u(T x, int src)
    {
        if constexpr (sizeof(T) == 8)
            return (T)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)x >> 32), src) << 32) |
                       (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)x, src));
        else
            return (T)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, src);
    }

    // (narrow keys widened to 32 bits; the reductions run on DPP moves)
    typedef typename std::conditional<KB == 8, uint64_t, uint32_t>::type UW;
    static __device__ __forceinline__ U wave_min(U x) { return (U)wave_extreme<false, UW>((UW)x); }
    static __device__ __forceinline__ U wave_max(U x) { return (U)wave_extreme<true, UW>((UW)x); }

    // The radix classes' workgroup steps.  Step v of class c (kw = c - CRAD
    // + 1 waves per task, P = NW / kw tasks) holds the class's tasks
    // first..first+nkc-1 at waves t*kw .. t*kw+kw-1; a mixed step (the first
    // `mix` steps of the 3-wave class) also holds one 2-wave task at waves
    // nkc*kw, nkc*kw+1.  Every wave derives its own task from these numbers
    // and reads the task's record itself (the class-ordered list holds the
    // records), so no workgroup-wide task table is built.
    struct RadixWave {
        LdsTask tk;        // the wave's task (len 0 when the wave is idle)
        uint32_t wb, kq;   // the task's first wave and its waves
        uint32_t used;     // waves of the step in use (tasks hold waves [0, used))
        uint32_t nkc, kw;  // tasks of the class in the step, waves per such task
        bool mixed, act;
    };

    static __device__ __forceinline__ RadixWave radix_wave(const Params &p, const RadixShared &rs, uint32_t c,
                                                           uint32_t sv, uint32_t w)
    {
        RadixWave rw;
        rw.kw = c - CRAD + 1;
        const uint32_t P = (uint32_t)NW / rw.kw;
        rw.mixed = rw.kw == 3 && sv < rs.mix;
        const uint32_t first = sv * P + (rw.kw == 2 ? rs.mix : 0u);
        const uint32_t cnt = p.ctrl->cls_cnt[c];
        rw.nkc = first < cnt ? ((cnt - first) < P ? (cnt - first) : P) : 0u;
        rw.used = rw.nkc * rw.kw + (rw.mixed ? 2u : 0u);
        uint32_t slot = ~0u;
        rw.act = w < rw.used;
        if (w < rw.nkc * rw.kw) {
            const uint32_t q = w / rw.kw;
            rw.wb = q * rw.kw;
            rw.kq = rw.kw;
            slot = rs.coff[c] + first + q;
        } else {
            rw.wb = rw.nkc * rw.kw;
            rw.kq = rw.act ? 2u : 1u;
            if (rw.act) slot = rs.coff[c - 1] + sv;   // the mixed step's 2-wave task
        }
        // the record is read through the constant address space: a scalar
        // load (no kernel writes the list while the LDS sorts run), which the
        // caller issues a step ahead (radix_check waits for it)
        typedef const __attribute__((address_space(4))) LdsTask CTask;
        rw.tk = LdsTask{0, 0, 0, 0};
        if (rw.act && slot < p.ldscap) {
            CTask *t = (CTask *)p.ldsq + slot;
            rw.tk.begin = t->begin;
            rw.tk.len = t->len;
            rw.tk.depth = t->depth;
        }
        return rw;
    }

    // validates a wave's task record (a failed check is reported and the
    // task skipped)
    static __device__ __forceinline__ void radix_check(const Params &p, RadixWave &rw)
    {
        if (rw.act && (!span_ok(p, rw.tk.begin, rw.tk.len) || rw.tk.len > rw.kq * WS)) {
            if (lane() == 0) range_fail(p, 1);
            rw.tk.len = 0;
        }
    }

    // One workgroup step of the radix classes.  Wave w holds items i at task
    // position j = (w - wb)*WS + i*WAVE + lane, loaded from global memory
    // straight into registers (range-checked buffer loads: positions past
    // the task's end read 0 and become padding with the largest digit); the
    // step's LDS arrays are in slot order (task at slots [wb*WS, (wb+kq)*WS)),
    // and every pass leaves each task's keys sorted at the start of its
    // slots.  The wave then writes its own slots' share of the sorted task
    // (positions j of its slots) with range-checked buffer stores.
    static __device__ void radix_step(const Params &p, U *sk, uint32_t *si, const RadixWave &rw, uint32_t tid,
                                      RadixShared &rs)
    {
        const U flip = (U)p.flip;
        static_assert(NW <= 8, "one workgroup step holds at most 8 waves");
        const uint32_t w = __builtin_amdgcn_readfirstlane(tid / WAVE), l = tid % WAVE;
        const bool act = rw.act;
        const uint32_t wb = rw.wb, len = rw.tk.len;
        const uint32_t j0 = (w - wb) * WS + l;            // task position of item i: j0 + i*WAVE
        U *const rk = sk + w * WS + l;                    // slot of item i: rk[i*WAVE]
        uint32_t *const ri = si + w * WS + l;
        U k[SI];
        uint32_t v[SI];
        U mn = key_max<U>(), mx = 0;
        {
            const BufRsrc kr = buf_rsrc(src_keys<U>(p, rw.tk.depth) + rw.tk.begin, len * (uint32_t)KB);
            const BufRsrc ir = buf_rsrc(src_idx(p, rw.tk.depth) + rw.tk.begin, IDX ? len * 4u : 0u);
#pragma unroll
            for (int i = 0; i < SI; i++) k[i] = buf_load<U>(kr, (j0 + (uint32_t)i * WAVE) * (uint32_t)KB);
#pragma unroll
            for (int i = 0; i < SI; i++) v[i] = IDX ? buf_load<uint32_t>(ir, (j0 + (uint32_t)i * WAVE) * 4u) : 0u;
#pragma unroll
            for (int i = 0; i < SI; i++) {
                const bool real = j0 + (uint32_t)i * WAVE < len;
                k[i] = real ? to_ord<U, FL>(k[i], flip) : key_max<U>();
                mn = real && k[i] < mn ? k[i] : mn;
                mx = real && k[i] > mx ? k[i] : mx;
            }
        }
        mn = wave_min(mn);
        mx = wave_max(mx);
        if (l == 0) {
            rs.wmin[w] = mn;
            rs.wmax[w] = mx;
        }
        block_sync();
        // every wave derives the step's pass count (the largest key span of
        // its tasks) and its own task's minimum from the per-wave extremes
        uint32_t maxbits = 0, tbits = 0;   // the step's widest task span, this wave's task span
        U tmin = 0;
        {
            const uint32_t ntask = rw.nkc + (rw.mixed ? 1u : 0u);
            for (uint32_t t = 0; t < ntask; t++) {   // uniform
                const uint32_t a0 = t * rw.kw, a1 = t < rw.nkc ? a0 + rw.kw : a0 + 2u;
                U a = key_max<U>(), b = 0;
                for (uint32_t x = a0; x < a1; x++) {
                    const U xa = rs.wmin[x], xb = rs.wmax[x];
                    a = xa < a ? xa : a;
                    b = xb > b ? xb : b;
                }
                const uint64_t span = b > a ? (uint64_t)(U)(b - a) : 0ull;
                const uint32_t bits = span ? (uint32_t)(64 - __clzll(span)) : 0u;
                maxbits = bits > maxbits ? bits : maxbits;
                if (a0 == wb) {
                    tmin = a;
                    tbits = bits;
                }
            }
            maxbits = __builtin_amdgcn_readfirstlane(maxbits);
            tbits = __builtin_amdgcn_readfirstlane(tbits);
            tmin = uniform_val(tmin);
        }
        uint32_t nfull = (maxbits + RW - 1) / RW;
        nfull = nfull < MAXPASS ? nfull : MAXPASS;
        // Spans wider than two digits: the two passes cover the top TOPB
        // bits of each task's own span (key - min >> lowb, lowb per task);
        // keys that agree in those bits (a group: neighbours after the
        // passes) are then ordered by a comparison inside their group.  A
        // group longer than GMAX (or a group scan that would run past it)
        // makes the step redo the full passes from its input.  A task whose
        // span fits the two digits (lowb 0) is sorted exactly by them.
        // (8-byte keys: their leaves span ~50 bits; narrower keys keep the
        // registers of the plain passes)
        constexpr uint32_t TOPP = 2, TOPB = TOPP * RW, GMAX = 64;
        bool top = KB == 8 && nfull > TOPP;   // block-uniform
        uint32_t lowb = top && tbits > TOPB ? tbits - TOPB : 0u;   // wave-uniform
        uint32_t npass = top ? TOPP : nfull;
        // scan ownership: the step's entries in (task, digit, wave of the
        // task) order are NBIN per wave in use, so thread tid owns the EPT
        // consecutive entries from tid*EPT, all inside its own wave's task:
        // entry e1 + k2 of the task (e1 = (w - wb)*NBIN + l*EPT) is digit
        // (e1 + k2) / kq, wave wb + (e1 + k2) % kq; their counters' element
        // offsets in rs.cnt (< NW*NBIN = 2^12) are kept two to a word.
        // Threads of idle waves own none.
        static_assert(NBIN == (uint32_t)NT && EPT <= 8 && EPT % 2 == 0 && NW * NBIN <= 65536 && NBIN % EPT == 0,
                      "scan ownership");
        const uint32_t kq = rw.kq;
        const uint32_t e1 = (w - wb) * NBIN + l * EPT;
        uint32_t ea[EPT / 2];
        {
            uint32_t jj = e1 / kq, ww = e1 - jj * kq;
#pragma unroll
            for (int k2 = 0; k2 < (int)EPT; k2++) {
                const uint32_t o = (wb + ww) * NBIN + jj;
                if (k2 & 1) ea[k2 / 2] |= o << 16;
                else ea[k2 / 2] = o;
                if (++ww == kq) {
                    ww = 0;
                    jj++;
                }
            }
        }
        auto eoff = [&](int k2) { return (k2 & 1) ? ea[k2 / 2] >> 16 : ea[k2 / 2] & 0xffffu; };
        U *dst = (U *)p.kout;
        const BufRsrc kw = buf_rsrc(dst + rw.tk.begin, len * (uint32_t)KB);
        const BufRsrc iw = buf_rsrc(p.iout + rw.tk.begin, IDX ? len * 4u : 0u);
        for (uint32_t attempt = 0; attempt < 2; attempt++) {   // block-uniform
            if (attempt == 1) {
                // the full passes, from the task's input again
                const BufRsrc kr = buf_rsrc(src_keys<U>(p, rw.tk.depth) + rw.tk.begin, len * (uint32_t)KB);
                const BufRsrc ir = buf_rsrc(src_idx(p, rw.tk.depth) + rw.tk.begin, IDX ? len * 4u : 0u);
#pragma unroll
                for (int i = 0; i < SI; i++) k[i] = buf_load<U>(kr, (j0 + (uint32_t)i * WAVE) * (uint32_t)KB);
#pragma unroll
                for (int i = 0; i < SI; i++) v[i] = IDX ? buf_load<uint32_t>(ir, (j0 + (uint32_t)i * WAVE) * 4u) : 0u;
#pragma unroll
                for (int i = 0; i < SI; i++) {
                    const bool real = j0 + (uint32_t)i * WAVE < len;
                    k[i] = real ? to_ord<U, FL>(k[i], flip) : key_max<U>();
                }
            }
            for (uint32_t pass = 0; pass < npass; pass++) {
                const uint32_t sh = lowb + pass * RW;
                if (pass > 0 && act) {
#pragma unroll
                    for (int i = 0; i < SI; i++) {
                        k[i] = rk[i * WAVE];
                        if (IDX) v[i] = ri[i * WAVE];
                    }
                }
                uint32_t dr[SI];
                if (act) {
                    // the wave's 512 counters: 16 bytes per lane, one store each
                    static_assert(NBIN * sizeof(uint16_t) == WAVE * 16, "one 16-byte store per lane clears a row");
                    typedef uint32_t U32x4 __attribute__((ext_vector_type(4)));
                    reinterpret_cast<U32x4 *>(&rs.cnt[w][0])[l] = U32x4{0u, 0u, 0u, 0u};
                    wave_sync();
                    // after a pass the task's keys are at its first len positions
                    auto digit = [&](int i) {
                        const bool real = j0 + (uint32_t)i * WAVE < len;
                        return real ? (uint32_t)((U)(k[i] - tmin) >> sh) & (NBIN - 1) : NBIN - 1;
                    };
                    // an item whose lanes share one digit (runs of sorted keys)
                    // takes its ranks with one atomic (64 same-address atomics
                    // are serialised by the LDS); tested per item only when the
                    // wave's first or last item is such one
                    auto same = [&](uint32_t d) {
                        const uint32_t dd = __builtin_amdgcn_readfirstlane(d);
                        return ballot(d == dd) == ~0ull;
                    };
                    const bool runs = same(digit(0)) || same(digit(SI - 1));   // wave-uniform
                    if (runs) {
#pragma unroll
                        for (int i = 0; i < SI; i++) {
                            const uint32_t d = digit(i);
                            uint32_t r;
                            if (same(d)) {   // wave-uniform
                                uint32_t o0 = 0;
                                if (l == 0) o0 = rank16_add((uint32_t *)rs.cnt[w], d, (uint32_t)WAVE);
                                r = __builtin_amdgcn_readlane(o0, 0) + l;
                            } else {
                                r = rank16_add((uint32_t *)rs.cnt[w], d, 1u);
                            }
                            dr[i] = d | (r << 16);
                        }
                    } else {
#pragma unroll
                        for (int i = 0; i < SI; i++) {
                            const uint32_t d = digit(i);
                            dr[i] = d | (rank16_add((uint32_t *)rs.cnt[w], d, 1u) << 16);
                        }
                    }
                }
                block_sync();
                // exclusive scan over (task, digit, wave of the task): the first
                // slot of every (wave, digit) run of the step
                {
                    // For a task of 1, 2 or 4 waves the thread's entries are
                    // EPT/kq consecutive digits (aligned) in each of the
                    // task's kq counter rows: one vector load and store per
                    // row; other widths go entry by entry.
                    static_assert(EPT == 8, "vector forms of the scan entries");
                    typedef uint16_t H8 __attribute__((ext_vector_type(8)));
                    typedef uint16_t H4 __attribute__((ext_vector_type(4)));
                    typedef uint16_t H2 __attribute__((ext_vector_type(2)));
                    uint32_t cv[EPT], loc = 0;
#pragma unroll
                    for (int k2 = 0; k2 < (int)EPT; k2++) cv[k2] = 0u;
                    if (act) {   // wave-uniform, and so is kq
                        if (kq == 1) {
                            const H8 a = *reinterpret_cast<const H8 *>(&rs.cnt[wb][e1]);
#pragma unroll
                            for (int k2 = 0; k2 < 8; k2++) cv[k2] = a[k2];
                        } else if (kq == 2) {
                            const H4 a = *reinterpret_cast<const H4 *>(&rs.cnt[wb][e1 / 2]);
                            const H4 b = *reinterpret_cast<const H4 *>(&rs.cnt[wb + 1][e1 / 2]);
#pragma unroll
                            for (int k2 = 0; k2 < 8; k2++) cv[k2] = (k2 & 1) ? b[k2 / 2] : a[k2 / 2];
                        } else if (kq == 4) {
                            H2 r[4];
#pragma unroll
                            for (int q = 0; q < 4; q++) r[q] = *reinterpret_cast<const H2 *>(&rs.cnt[wb + q][e1 / 4]);
#pragma unroll
                            for (int k2 = 0; k2 < 8; k2++) cv[k2] = r[k2 % 4][k2 / 4];
                        } else {
#pragma unroll
                            for (int k2 = 0; k2 < (int)EPT; k2++) cv[k2] = (&rs.cnt[0][0])[eoff(k2)];
                        }
                    }
#pragma unroll
                    for (int k2 = 0; k2 < (int)EPT; k2++) loc += cv[k2];
                    uint32_t tot;
                    // (no trailing barrier: the one after the prefixes are
                    // stored comes before rs.red is written again)
                    uint32_t run = block_excl_scan<NT, false>(loc, rs.red, &tot);
                    if (act) {
                        uint16_t o[EPT];   // the entries' exclusive prefixes, in entry order
#pragma unroll
                        for (int k2 = 0; k2 < (int)EPT; k2++) {
                            o[k2] = (uint16_t)run;
                            run += cv[k2];
                        }
                        if (kq == 1) {
                            H8 a;
#pragma unroll
                            for (int k2 = 0; k2 < 8; k2++) a[k2] = o[k2];
                            *reinterpret_cast<H8 *>(&rs.cnt[wb][e1]) = a;
                        } else if (kq == 2) {
                            H4 a, b;
#pragma unroll
                            for (int k2 = 0; k2 < 4; k2++) {
                                a[k2] = o[2 * k2];
                                b[k2] = o[2 * k2 + 1];
                            }
                            *reinterpret_cast<H4 *>(&rs.cnt[wb][e1 / 2]) = a;
                            *reinterpret_cast<H4 *>(&rs.cnt[wb + 1][e1 / 2]) = b;
                        } else if (kq == 4) {
#pragma unroll
                            for (int q = 0; q < 4; q++) {
                                H2 r;
                                r[0] = o[q];
                                r[1] = o[q + 4];
                                *reinterpret_cast<H2 *>(&rs.cnt[wb + q][e1 / 4]) = r;
                            }
                        } else {
#pragma unroll
                            for (int k2 = 0; k2 < (int)EPT; k2++) (&rs.cnt[0][0])[eoff(k2)] = o[k2];
                        }
                    }
                }
                block_sync();
                if (act) {
#pragma unroll
                    for (int i = 0; i < SI; i++) {
                        const uint32_t pos = (rs.cnt[w][dr[i] & 0xffffu] + (dr[i] >> 16)) & (CAP - 1);
                        sk[pos] = k[i];
                        if (IDX) si[pos] = v[i];
                    }
                }
                block_sync();
            }
            if (KB != 8 || !top) break;
            // groups: runs of keys with equal top bits (neighbours after the
            // passes).  The lane holding a group's first position (its head:
            // the key before has other top bits, the key after the same)
            // orders the group in LDS by insertion sort (stable).  Top bits
            // are below 2^TOPB, so they are compared as 32-bit values; the
            // passes only permute a group's keys among its positions, so the
            // top bits seen at any position do not change during the fix-up.
            U *const tk = sk + wb * WS;   // the task's keys in their sorted slots
            uint32_t *const ti = si + wb * WS;
            auto top32 = [&](U key) { return (uint32_t)((U)(key - tmin) >> lowb); };
            uint32_t heads = 0;
            if (act && lowb != 0) {   // wave-uniform
                // eq bit i: the position before item i's (lane l-1 of item i,
                // lane 63 of item i-1, or the previous wave's last position)
                // has the same top bits.  The neighbour comes from a DPP
                // whole-wave shift by one lane; lane 0 takes the value of the
                // position before, read once for the wave.
                const uint32_t wfirst = (w - wb) * WS, wnext = wfirst + WS;   // the wave's first position, the next wave's
                uint32_t last = w > wb ? top32(tk[wfirst - 1]) : 0u;   // unused at position 0 (j > 0)
                uint32_t eq = 0;
#pragma unroll
                for (int i = 0; i < SI; i++) {
                    const uint32_t j = j0 + (uint32_t)i * WAVE;
                    const uint32_t t = top32(rk[i * WAVE]);
                    const uint32_t before =
                        (uint32_t)__builtin_amdgcn_update_dpp((int)last, (int)t, 0x138, 0xf, 0xf, false);
                    if (j > 0 && j < len && before == t) eq |= 1u << i;
                    last = (uint32_t)__builtin_amdgcn_readlane((int)t, WAVE - 1);
                }
                // the position after: lane l+1 of item i, lane 0 of item i+1,
                // or for the wave's last position the next wave's first
                const uint32_t eq_first = (uint32_t)__builtin_amdgcn_readfirstlane((int)eq);
                const bool next_wave_eq = wnext < len && top32(tk[wnext]) == last;   // last: the wave's last position
                const uint32_t old = (eq_first >> 1) | (next_wave_eq ? 1u << (SI - 1) : 0u);
                const uint32_t after = (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)eq, 0x130, 0xf, 0xf, false);
                heads = after & ~eq;
            }
            if (tid == 0) rs.flag = 0;
            block_sync();
            bool ovf = false;
            if (ballot(heads != 0)) {   // wave-uniform: a few groups per task
#pragma nounroll
                for (uint32_t g = heads; g; g &= g - 1) {
                    const uint32_t gs = j0 + (uint32_t)__builtin_ctz(g) * WAVE;
                    const uint32_t t = top32(tk[gs]);
                    uint32_t ge = gs + 1;
#pragma nounroll
                    while (ge < len && ge - gs <= GMAX && top32(tk[ge]) == t) ge++;
                    if (ge - gs > GMAX) {
                        ovf = true;
                        break;
                    }
#pragma nounroll
                    for (uint32_t q = gs + 1; q < ge; q++) {
                        const U kq = tk[q];
                        const uint32_t vq = IDX ? ti[q] : 0u;
                        uint32_t r = q;
#pragma nounroll
                        while (r > gs && tk[r - 1] > kq) {
                            tk[r] = tk[r - 1];
                            if (IDX) ti[r] = ti[r - 1];
                            r--;
                        }
                        tk[r] = kq;
                        if (IDX) ti[r] = vq;
                    }
                }
            }
            if (ovf) rs.flag = 1;   // every writer stores the same value
            block_sync();
            if (__builtin_amdgcn_readfirstlane(rs.flag) == 0) break;   // sorted in LDS: the write-out below
            // a group longer than GMAX: the full passes, from the input again
            top = false;
            lowb = 0;
            npass = nfull;
        }
        // write-out of the wave's slots: slot w*WS + i*WAVE + l holds task
        // position j0 + i*WAVE (every task's keys sit at the start of its
        // slots; with no pass, they are still in registers in load order)
        if (act) {
            if (npass > 0) {
#pragma unroll
                for (int i = 0; i < SI; i++) {
                    k[i] = rk[i * WAVE];
                    if (IDX) v[i] = ri[i * WAVE];
                }
            }
#pragma unroll
            for (int i = 0; i < SI; i++) {
                const uint32_t j = j0 + (uint32_t)i * WAVE;
                buf_store<U>(kw, j * (uint32_t)KB, from_ord<U, FL>(k[i], flip));
                if (IDX) buf_store<uint32_t>(iw, j * 4u, v[i]);
            }
        }
        // (no barrier here: before the next step writes the key slots, the
        // counters or the group flag, every wave has passed that step's
        // range barrier, i.e. has finished this step's reads; the only LDS
        // it writes before then are its own wmin/wmax entries, read here only
        // before this step's first pass)
    }

    // all workgroup steps of the radix classes [CRAD, CLAST]
    static __device__ void radix_all(const Params &p, unsigned char *arena, RadixShared &rs, unsigned long long *acc)
    {
        U *sk = (U *)arena;
        uint32_t *si = (uint32_t *)(arena + C::SORT_KEY_BYTES);
        if (threadIdx.x == 0) lds_plan(p.ctrl, CPLAN, NW, CAP, rs.coff, rs.cwg, &rs.mix);
        if (threadIdx.x == 0) acc[0] = acc[1] = 0;
        block_sync();
        const uint32_t vbeg = rs.cwg[CRAD], vend = rs.cwg[CLAST + 1];
        uint64_t nbytes = 0, ntasks = 0;
        // the task records of step v + gridDim.x are loaded while step v runs
        auto wave_of = [&](uint32_t v) {
            uint32_t c = CRAD;
            while (c < CLAST && v >= rs.cwg[c + 1]) c++;
            c = __builtin_amdgcn_readfirstlane(c);
            const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
            return radix_wave(p, rs, c, __builtin_amdgcn_readfirstlane(v - rs.cwg[c]), w);
        };
        RadixWave rn;
        if (vbeg + blockIdx.x < vend) rn = wave_of(vbeg + blockIdx.x);
        for (uint32_t v = vbeg + blockIdx.x; v < vend; v += gridDim.x) {
            // the thread index is re-read through an empty asm statement in
            // every step: otherwise the compiler hoists the loop-invariant
            // addresses derived from it out of the loop and spills them
            uint32_t tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            RadixWave rw = rn;
            if (v + gridDim.x < vend) rn = wave_of(v + gridDim.x);
            radix_check(p, rw);
            const uint32_t w = __builtin_amdgcn_readfirstlane(tid / WAVE);
            if (rw.act && w == rw.wb && lane() == 0 && rw.tk.len) {
                nbytes += 2ull * rw.tk.len * (KB + C::IB);
                ntasks++;
            }
            radix_step(p, sk, si, rw, tid, rs);
        }
        if (ntasks) {
            atomicAdd(&acc[0], (unsigned long long)nbytes);
            atomicAdd(&acc[1], (unsigned long long)ntasks);
        }
        block_sync();
        if (threadIdx.x == 0 && acc[1]) {
            atomic_add64(&p.ctrl->bytes_lds, acc[0]);
            atomic_add64(&p.ctrl->tasks[T_LDS], acc[1]);
        }
    }

    // the workgroup steps of the merge classes [CMIN, CPLAN) (the radix
    // classes are the other launch, radix_all: disjoint tasks)
    static __device__ void lds_merge(const Params &p, unsigned char *arena, TaskRec *rec, RadixShared &rs,
                                     unsigned long long *acc)
    {
        U *sk = (U *)arena;
        uint32_t *si = (uint32_t *)(arena + C::SORT_KEY_BYTES);
        const U flip = (U)p.flip;
        U *dst = (U *)p.kout;
        const Ctrl *cc = p.ctrl;
        if (threadIdx.x == 0) lds_plan(cc, CPLAN, NW, CAP, rs.coff, rs.cwg, &rs.mix);
        block_sync();
        // merge classes [CMIN, CPLAN) (they reach CMAX when MERGE_ONLY)
        const uint32_t c0 = (uint32_t)C::CMIN, clast = CPLAN - 1;
        const uint32_t vbeg = rs.cwg[c0], vend = rs.cwg[clast + 1];
        uint64_t nbytes = 0, ntasks = 0;
        if (threadIdx.x == 0) acc[0] = acc[1] = 0;
        for (uint32_t v = vbeg + blockIdx.x; v < vend; v += gridDim.x) {
            // the thread index is re-read through an empty asm statement in
            // every step: otherwise the compiler hoists the 2 x SI
            // loop-invariant addresses derived from it out of the loop and
            // spills them
            uint32_t tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            uint32_t c = c0;
            while (c < clast && v >= rs.cwg[c + 1]) c++;
            const uint32_t region = 1u << c;     // slots per task of the class
            const uint32_t P = CAP >> c;
            const uint32_t sv = v - rs.cwg[c];   // step inside the class
            const uint32_t first = sv * P;
            const uint32_t cnt = cc->cls_cnt[c];
            const uint32_t nk = first < cnt ? ((cnt - first) < P ? (cnt - first) : P) : 0u;
            // every record the step reads is set (tasks past nk have length 0)
            for (uint32_t q = threadIdx.x; q < P; q += NT) {
                TaskRec r{0, 0, 0};
                if (q < nk) {
                    const uint32_t slot = rs.coff[c] + first + q;
                    if (slot < p.ldscap) {
                        const LdsTask tk = p.ldsq[slot];
                        r = TaskRec{tk.begin, tk.len, tk.depth};
                    }
                    if (!span_ok(p, r.begin, r.len) || r.len > region) {
                        range_fail(p, 1);
                        r.len = 0;
                    }
                    nbytes += 2ull * r.len * (KB + C::IB);
                    ntasks += r.len ? 1u : 0u;
                }
                rec[q] = r;
            }
            block_sync();
            {
                const uint32_t n_eff = region;
                // the global loads of a batch are issued before its LDS
                // stores, so they are in flight together (batches of 8 bound
                // the registers they hold)
                for (int r0 = 0; r0 < SI; r0 += 8) {
                    U kv[8];
                    uint32_t iv[8];
                    uint32_t ok8 = 0;
#pragma unroll
                    for (int r = 0; r < 8; r++) {
                        const uint32_t i = tid + (uint32_t)(r0 + r) * NT;
                        const uint32_t q = i >> c, j = i & (n_eff - 1);
                        const TaskRec tr = rec[q];
                        kv[r] = key_max<U>();
                        iv[r] = 0u;
                        // the loaded value is only used after the loop, so
                        // the loads of the eight items overlap
                        if (j < tr.len) {
                            kv[r] = src_keys<U>(p, tr.depth)[tr.begin + j];
                            if (IDX) iv[r] = src_idx(p, tr.depth)[tr.begin + j];
                        }
                        ok8 |= (j < tr.len ? 1u : 0u) << r;
                    }
#pragma unroll
                    for (int r = 0; r < 8; r++) kv[r] = (ok8 >> r) & 1u ? to_ord<U, FL>(kv[r], flip) : key_max<U>();
                    // pidx(tid + r*NT) = pidx(tid) + r*(NT + NT/SI)
                    U *const mk = sk + pidx<SI>(tid) + (uint32_t)r0 * (NT + NT / SI);
                    uint32_t *const mi = si + pidx<SI>(tid) + (uint32_t)r0 * (NT + NT / SI);
#pragma unroll
                    for (int r = 0; r < 8; r++) {
                        mk[r * (NT + NT / SI)] = kv[r];
                        if (IDX) mi[r * (NT + NT / SI)] = iv[r];
                    }
                }
                block_sync();
                block_sort<U, IDX, NT, SI>(sk, si, CAP, n_eff);
                // a fresh opaque copy of the thread index: the slot numbers
                // below are recomputed here instead of being kept live (in
                // spilled registers) from the fill through the sort
                uint32_t tid2 = threadIdx.x;
                asm volatile("" : "+v"(tid2));
                const U *const mk = sk + pidx<SI>(tid2);   // + r*(NT + NT/SI) for tid + r*NT
                const uint32_t *const mi = si + pidx<SI>(tid2);
#pragma unroll
                for (int r = 0; r < SI; r++) {
                    const uint32_t i = tid2 + (uint32_t)r * NT;
                    const uint32_t q = i >> c, j = i & (n_eff - 1);
                    const TaskRec tr = rec[q];
                    if (j < tr.len) {
                        nt_store(&dst[tr.begin + j], from_ord<U, FL>(mk[r * (NT + NT / SI)], flip));
                        if (IDX) nt_store(&p.iout[tr.begin + j], mi[r * (NT + NT / SI)]);
                    }
                }
            }
            block_sync();
        }
        if (ntasks) {
            atomicAdd(&acc[0], (unsigned long long)nbytes);
            atomicAdd(&acc[1], (unsigned long long)ntasks);
        }
        block_sync();
        if (threadIdx.x == 0 && acc[1]) {
            atomic_add64(&p.ctrl->bytes_lds, acc[0]);
            atomic_add64(&p.ctrl->tasks[T_LDS], acc[1]);
        }
    }

    // ---------------------------------------------------------------- pivot
    // One wave per segment (8 per workgroup step): the wave samples 1024 keys
    // (128 evenly spaced runs of 8) into registers (16 per lane), sorts them
    // by stable LSD radix passes of RW-bit digits over the top 2*RW bits of
    // the sample's key range (ranks from one LDS atomic per key on per-wave
    // digit counters, no block barrier), takes S splitters at the sample
    // quantiles, orders them exactly and reserves the segment's tiles.
    static constexpr uint32_t SPL = C::SAMPLE / WAVE;   // sample keys per lane

    static __device__ void pivot(const Params &p, uint32_t L, PivotShared &ps)
    {
        // level L = the segments created since level L-1 opened (nothing
        // appends segments while this kernel runs); workgroup 0 records it
        // for the later kernels of the level and the host
        Level lv;
        lv.sb = L == 0 ? 0u : p.levels[L - 1].se;
        lv.se = p.ctrl->seg_alloc < p.segcap ? p.ctrl->seg_alloc : p.segcap;   // overflow already flagged
        lv.tiles = 0;
        lv.pad = 0;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            p.levels[L] = lv;
            if (lv.se > lv.sb && L + 1 > p.ctrl->max_depth) p.ctrl->max_depth = L + 1;
        }
        const U flip = (U)p.flip;
        constexpr uint32_t SAMPLE = C::SAMPLE;
        const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE), l = lane();
        U *const b = ps.buf[w];
        uint16_t *const cn = ps.cnt[w];
        uint32_t done = 0;
        if (threadIdx.x == 0) ps.done = 0;
        block_sync();
        // segment sb + w*grid + block (+ k*grid*NW): a level with fewer
        // segments than waves in the grid spreads over every workgroup, one
        // wave per SIMD, instead of filling the first workgroups' 8 waves.
        // The steps are block-uniform: the workgroup reserves the tiles of
        // its step's segments with one atomic (a level of 1024 segments
        // otherwise queues 1024 atomics on one word).
        for (uint32_t s0 = lv.sb + blockIdx.x; s0 < lv.se; s0 += gridDim.x * NW) {   // block-uniform
            const uint32_t sid = s0 + w * gridDim.x;
            Seg *sg = &p.segs[sid < lv.se ? sid : s0];
            uint32_t ntiles = 0;
            if (sid < lv.se) {   // wave-uniform
                done++;
                const uint32_t begin = sg->begin, len = sg->len, depth = sg->depth;
                if (!span_ok(p, begin, len)) {
                    if (l == 0) {
                        range_fail(p, 3);
                        sg->ntiles = 0;
                        sg->tile_base = 0;
                    }
                } else if (sg->mode == SPLIT_RANGE) {
                    // S+1 equal slices of the segment's bounds [lo, hi]: every
                    // child's range is at most a (S+1)-th of this one's (a
                    // range narrower than S+1 values makes every value above
                    // lo a splitter, so the next level ends it), whatever the
                    // keys' positions
                    const U kmax = key_max<U>();
                    const U lo = (U)sg->lo;
                    const U hi = sg->hi > (uint64_t)kmax ? kmax : (U)sg->hi;
                    const U span = hi > lo ? (U)(hi - lo) : (U)0;
                    U wd = (U)(span / (U)(S + 1));
                    if (wd == 0) wd = 1;
                    if (l < (uint32_t)S) {
                        const U off = (U)((U)(l + 1) * wd);
                        sg->split[l] = (uint64_t)(off > span ? hi : (U)(lo + off));
                    }
                    ntiles = (len + TILE - 1) / TILE;
                } else {
                    const uint32_t m = len < SAMPLE ? len : SAMPLE;
                    const U *src = src_keys<U>(p, depth) + begin;
                    U k[SPL];
                    U mn = key_max<U>(), mx = 0;
                    // len >= SAMPLE: SAMPLE/RUN runs of RUN consecutive keys, run r
                    // starting at r*st + (st - RUN)/2 with st = len / (SAMPLE/RUN), so
                    // that a wave touches one line per run rather than one per key;
                    // sample j = i*WAVE + l is key j % RUN of run j / RUN.  The last
                    // run ends before len: st >= RUN and (SAMPLE/RUN)*st <= len.
                    // Else every key.
                    constexpr uint32_t RUN = 8, NRUN = SAMPLE / RUN;
                    static_assert(SAMPLE >= 8 * RUN && (WAVE % RUN) == 0, "sample runs");
                    const uint32_t st = len / NRUN;
                    const uint32_t base = m == SAMPLE ? (l / RUN) * st + (st - RUN) / 2 + l % RUN : l;
                    const uint32_t step = m == SAMPLE ? (WAVE / RUN) * st : WAVE;
                    // all loads first (positions past the sample re-read key 0),
                    // then the predicates, so that the loads' latencies overlap
        #pragma unroll
                    for (int i = 0; i < (int)SPL; i++) {
                        const uint32_t j = (uint32_t)i * WAVE + l;
                        k[i] = src[j < m ? base + (uint32_t)i * step : 0u];
                    }
        #pragma unroll
                    for (int i = 0; i < (int)SPL; i++) {
                        const bool real = (uint32_t)i * WAVE + l < m;
                        k[i] = real ? to_ord<U, FL>(k[i], flip) : key_max<U>();
                        mn = real && k[i] < mn ? k[i] : mn;
                        mx = real && k[i] > mx ? k[i] : mx;
                    }
                    mn = wave_min(mn);
                    mx = wave_max(mx);
                    // The sample is sorted on the top 2*RW bits of its span only (at
                    // most two passes): the keys at the quantile ranks are then within
                    // 2^-18 of the span of the exact quantiles, and the picked
                    // splitters are put in exact order below.
                    const uint64_t span = mx > mn ? (uint64_t)(U)(mx - mn) : 0ull;
                    const uint32_t bits = span ? (uint32_t)(64 - __clzll(span)) : 0u;
                    const uint32_t low = bits > 2u * RW ? bits - 2u * RW : 0u;   // bits left unsorted
                    const uint32_t npass = (bits - low + RW - 1) / RW;
                    for (uint32_t pass = 0; pass < npass; pass++) {
                        const uint32_t sh = low + pass * RW;
                        for (uint32_t d = l; d < NBIN; d += WAVE) cn[d] = 0;
                        wave_sync();
                        uint32_t dr[SPL];
        #pragma unroll
                        for (int i = 0; i < (int)SPL; i++) {
                            // after a pass the sample's keys are at its first m positions
                            const bool real = (uint32_t)i * WAVE + l < m;
                            const uint32_t d = real ? (uint32_t)((U)(k[i] - mn) >> sh) & (NBIN - 1) : NBIN - 1;
                            dr[i] = d | (rank16_add((uint32_t *)cn, d, 1u) << 16);
                        }
                        wave_sync();
                        // exclusive scan of the NBIN digit counts: lane l owns NBIN/WAVE digits
                        {
                            constexpr uint32_t DPL = NBIN / WAVE;
                            uint32_t c8[DPL], loc = 0;
        #pragma unroll
                            for (uint32_t q = 0; q < DPL; q++) {
                                c8[q] = cn[l * DPL + q];
                                loc += c8[q];
                            }
                            uint32_t run = wave_incl_scan(loc) - loc;
        #pragma unroll
                            for (uint32_t q = 0; q < DPL; q++) {
                                cn[l * DPL + q] = (uint16_t)run;
                                run += c8[q];
                            }
                        }
                        wave_sync();
        #pragma unroll
                        for (int i = 0; i < (int)SPL; i++) b[(cn[dr[i] & 0xffffu] + (dr[i] >> 16)) & (SAMPLE - 1)] = k[i];
                        wave_sync();
        #pragma unroll
                        for (int i = 0; i < (int)SPL; i++) k[i] = b[(uint32_t)i * WAVE + l];
                        wave_sync();
                    }
                    if (npass == 0) {
        #pragma unroll
                        for (int i = 0; i < (int)SPL; i++) b[(uint32_t)i * WAVE + l] = k[i];
                        wave_sync();
                    }
                    // splitter l = the key at quantile rank (l+1) m / (S+1), then each
                    // lane's rank among the S picked keys (ties by lane) orders them
                    // exactly (they are already ordered when the sort was exact)
                    const U spv = l < (uint32_t)S ? b[((l + 1) * m) / (S + 1)] : key_max<U>();
                    uint32_t srank = 0;
        #pragma unroll
                    for (int t = 0; t < S; t++) {
                        const U o = shfl_u<U>(spv, t);
                        srank += (o < spv || (o == spv && (uint32_t)t < l)) ? 1u : 0u;
                    }
                    if (l < (uint32_t)S) sg->split[srank] = (uint64_t)spv;
                    ntiles = (len + TILE - 1) / TILE;
                }
            }
            if (l == 0) ps.ntl[w] = ntiles;
            block_sync();
            if (threadIdx.x == 0) {
                uint32_t tot = 0;
                for (uint32_t x = 0; x < (uint32_t)NW; x++) tot += ps.ntl[x];
                uint32_t tb = tot ? atomic_add(&p.ctrl->tile_alloc, tot) : 0u;
                if (tot && (tb > p.tilecap || tot > p.tilecap - tb)) {
                    atomic_or(&p.ctrl->error, ERR_TILES);
                    tb = ~0u;
                }
                ps.tbase = tb;
            }
            block_sync();
            if (ntiles) {   // wave-uniform
                uint32_t tb = ps.tbase;
                if (tb != ~0u)
                    for (uint32_t x = 0; x < w; x++) tb += ps.ntl[x];
                if (l == 0) {
                    sg->ntiles = tb == ~0u ? 0u : ntiles;
                    sg->tile_base = tb == ~0u ? 0u : tb;
                }
                if (tb != ~0u)
                    for (uint32_t t = l; t < ntiles; t += WAVE) p.tile_seg[tb + t] = sid;
            }
            block_sync();   // the next step rewrites ps.ntl
        }
        if (l == 0 && done) atomicAdd(&ps.done, done);
        block_sync();
        if (threadIdx.x == 0 && ps.done) atomic_add64((unsigned long long *)&p.ctrl->tasks[T_PIVOT], (unsigned long long)ps.done);
    }
};

// ---- kernels ----------------------------------------------------------------------------
template <int KT, bool IDX, int S>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) k_pivot(Params p, uint32_t L)
{
    typedef LevelK<KT, IDX, S> K;
    __shared__ typename K::PivotShared ps;
    K::pivot(p, L, ps);
}

template <int KT, bool IDX, int S>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(kt_bytes(KT) == 1 ? 4 : 6))) k_count(Params p)
{
    typedef LevelK<KT, IDX, S> K;
    __shared__ typename K::SharedCount sh;
    K::count(p, sh);
}

template <int KT, bool IDX, int S>
__global__ void __launch_bounds__(512) k_scan1(Params p)
{
    typedef LevelK<KT, IDX, S> K;
    __shared__ typename K::SharedScan sh;
    K::scan1(p, sh);
}

template <int KT, bool IDX, int S>
__global__ void __launch_bounds__(512) k_scan2(Params p)
{
    typedef LevelK<KT, IDX, S> K;
    __shared__ typename K::SharedScan sh;
    K::scan2(p, sh);
}

template <int KT, bool IDX, int S>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) k_scatter(Params p)
{
    typedef LevelK<KT, IDX, S> K;
    __shared__ __align__(16) unsigned char arena[Cfg<KT, IDX, S>::SCAT_BYTES];
    __shared__ typename K::SharedTile sh;
    K::scatter(p, arena, sh);
}

template <int KT, bool IDX, int S>
__global__ void __launch_bounds__(256) k_children(Params p, uint32_t L)
{
    typedef LevelK<KT, IDX, S> K;
    static_assert(K::CW * WAVE == 256, "k_children runs CW waves per workgroup");
    __shared__ typename K::ChildShared cs;
    K::children(p, L, cs);
}

template <int KT, bool IDX, int S>
__global__ void __launch_bounds__(256) k_ldscount(Params p)
{
    __shared__ uint32_t hist[LDS_CLASSES];
    LevelK<KT, IDX, S>::ldscount(p, hist);
}

template <int KT, bool IDX, int S>
__global__ void __launch_bounds__(256) k_ldsbucket(Params p)
{
    __shared__ uint32_t hist[LDS_CLASSES];
    __shared__ uint32_t wbase[LDS_CLASSES];
    __shared__ uint32_t coff[LDS_CLASSES], cwg[LDS_CLASSES + 1];
    LevelK<KT, IDX, S>::ldsbucket(p, hist, wbase, coff, cwg);
}

template <int KT, bool IDX, int S, bool RAD>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) k_lds(Params p)
{
    typedef LevelK<KT, IDX, S> K;
    __shared__ __align__(16) unsigned char arena[Cfg<KT, IDX, S>::SORT_BYTES];
    static_assert((Cfg<KT, IDX, S>::CAP >> Cfg<KT, IDX, S>::CMIN) > Cfg<KT, IDX, S>::NW, "task records of a step");
    __shared__ typename K::TaskRec rec[Cfg<KT, IDX, S>::CAP >> Cfg<KT, IDX, S>::CMIN];
    __shared__ typename K::RadixShared rs;
    __shared__ unsigned long long acc[2];
    if constexpr (RAD)
        K::radix_all(p, arena, rs, acc);
    else
        K::lds_merge(p, arena, rec, rs, acc);
}

// ---- host side ------------------------------------------------------------------------
template <int KT, bool IDX, int S>
static int occ(const void *k, int threads)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, threads, 0) != hipSuccess || nb < 1) return 1;
    return nb;
}

template <int KT, bool IDX, int S>
hipError_t run_levels(const Params &p, int cus, int wg_per_cu, const uint32_t *dev_off, uint32_t nseg,
                      uint32_t single_len, uint32_t L0, uint32_t nlev, bool first, bool last, hipStream_t st,
                      LaunchTimer *tm)
{
    // workgroups per CU of each kernel (computed once, thread-safe)
    struct Occ {
        int tile, cnt, lds;
    };
    static const Occ oc = [] {
        Occ o;
        o.cnt = occ<KT, IDX, S>((const void *)k_count<KT, IDX, S>, 512);
        o.lds = occ<KT, IDX, S>((const void *)k_lds<KT, IDX, S, true>, 512);
        o.tile = occ<KT, IDX, S>((const void *)k_scatter<KT, IDX, S>, 512);
        return o;
    }();
    const int o_tile = oc.tile, o_cnt = oc.cnt, o_lds = oc.lds;
    auto lim = [&](int o) { return cus * ((wg_per_cu > 0 && wg_per_cu < o) ? wg_per_cu : o); };
    // (two pivot workgroups per CU: a level of many short segments waits on
    // the sample loads)
    const int g_tile = lim(o_tile), g_cnt = lim(o_cnt), g_pivot = 2 * cus, g_lds = cus * o_lds;
    const int g_small = cus * 2;
#define TS_LAUNCH(kind, ...)                  \
    do {                                      \
        if (tm) tm->begin(kind, st);          \
        hipLaunchKernelGGL(__VA_ARGS__);      \
        if (tm) tm->end(kind, st);            \
    } while (0)
    if (first) {
        // up to 256 roots (every plain sort): one launch clears the control
        // block and appends the roots
        if (nseg > 0 && nseg <= 256) {
            TS_LAUNCH(K_ROOTS, k_roots, dim3(1), dim3(256), 0, st, p, dev_off, nseg, single_len, 1u);
        } else {
            TS_LAUNCH(K_RESET, k_reset, dim3(1), dim3(256), 0, st, p.ctrl);
            if (nseg > 0)
                TS_LAUNCH(K_ROOTS, k_roots, dim3((nseg + 255) / 256), dim3(256), 0, st, p, dev_off, nseg, single_len,
                          0u);
        }
    }
    for (uint32_t L = L0; L < L0 + nlev && L < (uint32_t)MAX_LEVELS; L++) {
        TS_LAUNCH(K_PIVOT, (k_pivot<KT, IDX, S>), dim3(g_pivot), dim3(512), 0, st, p, L);
        TS_LAUNCH(K_COUNT, (k_count<KT, IDX, S>), dim3(g_cnt), dim3(512), 0, st, p);
        TS_LAUNCH(K_SCAN, (k_scan1<KT, IDX, S>), dim3(g_small), dim3(512), 0, st, p);
        TS_LAUNCH(K_SCAN, (k_scan2<KT, IDX, S>), dim3(g_small), dim3(512), 0, st, p);
        TS_LAUNCH(K_SCATTER, (k_scatter<KT, IDX, S>), dim3(g_tile), dim3(512), 0, st, p);
        TS_LAUNCH(K_CHILDREN, (k_children<KT, IDX, S>), dim3(g_small), dim3(256), 0, st, p, L);
    }
    if (last) {
        TS_LAUNCH(K_LDSPREP, (k_ldscount<KT, IDX, S>), dim3(g_small), dim3(256), 0, st, p);
        TS_LAUNCH(K_LDSPREP, (k_ldsbucket<KT, IDX, S>), dim3(g_small), dim3(256), 0, st, p);
        TS_LAUNCH(K_LDS, (k_lds<KT, IDX, S, false>), dim3(g_lds), dim3(512), 0, st, p);
        if (!LevelK<KT, IDX, S>::MERGE_ONLY)
            TS_LAUNCH(K_LDS, (k_lds<KT, IDX, S, true>), dim3(g_lds), dim3(512), 0, st, p);
    }
#undef TS_LAUNCH
    return hipGetLastError();
}

typedef hipError_t (*levels_fn)(const Params &, int, int, const uint32_t *, uint32_t, uint32_t, uint32_t, uint32_t,
                                bool, bool, hipStream_t, LaunchTimer *);

template <int KT>
levels_fn pick_kb(bool idx, int S)
{
    switch (S) {
    case 3: return idx ? run_levels<KT, true, 3> : run_levels<KT, false, 3>;
    case 7: return idx ? run_levels<KT, true, 7> : run_levels<KT, false, 7>;
    case 15: return idx ? run_levels<KT, true, 15> : run_levels<KT, false, 15>;
    case 31: return idx ? run_levels<KT, true, 31> : run_levels<KT, false, 31>;
    default: return nullptr;
    }
}

levels_fn pick_kb1(bool idx, int S);
levels_fn pick_kb2(bool idx, int S);
levels_fn pick_kb4(bool idx, int S);
levels_fn pick_kb8(bool idx, int S);
levels_fn pick_kf4(bool idx, int S);
levels_fn pick_kf8(bool idx, int S);

}  // namespace ts
