// ts_device.h -- device-side building blocks shared by the gfx950 kernels:
// key traits, wave64 helpers, agent-scope synchronisation, block scans and
// the LDS-resident stable block sort.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ts {

constexpr int WAVE = 64;

// ---- key traits ------------------------------------------------------------
// Keys are moved in their stored representation.  Comparisons use an
// unsigned "order" value of the same width:
//  - unsigned integers: the key itself (flip = 0);
//  - signed integers: the sign bit flipped (flip = sign mask), which maps the
//    signed order onto the unsigned order;
//  - IEEE binary32/binary64 (FL): non-negative values get the sign bit set,
//    negative ones every bit inverted (flip = sign mask).  The result orders
//    as IEEE 754 totalOrder: -NaN < -inf < ... < -0 < +0 < ... < +inf < +NaN,
//    which agrees with the < of the type on every pair < orders.
template <int KB> struct KeyT;
template <> struct KeyT<1> { typedef uint8_t U; typedef int8_t I; };
template <> struct KeyT<2> { typedef uint16_t U; typedef int16_t I; };
template <> struct KeyT<4> { typedef uint32_t U; typedef int32_t I; };
template <> struct KeyT<8> { typedef uint64_t U; typedef int64_t I; };

// all-ones when the top bit of x is set, else 0
template <typename U>
__device__ __forceinline__ U top_mask(U x)
{
    typedef typename KeyT<sizeof(U)>::I I;
    return (U)((I)x >> (8 * sizeof(U) - 1));
}

// stored representation -> order value
template <typename U, bool FL>
__device__ __forceinline__ U to_ord(U x, U flip)
{
    if constexpr (FL)
        return (U)(x ^ (flip | top_mask(x)));
    else
        return (U)(x ^ flip);
}

// order value -> stored representation
template <typename U, bool FL>
__device__ __forceinline__ U from_ord(U x, U flip)
{
    if constexpr (FL)
        return (U)(x ^ (flip | (U)~top_mask(x)));
    else
        return (U)(x ^ flip);
}

template <typename U>
__device__ __forceinline__ U key_max() { return (U)~(U)0; }

// ---- wave helpers ----------------------------------------------------------
__device__ __forceinline__ uint32_t lane() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// orders LDS accesses of the lanes of one wave (read-all then leader-update
// sequences on per-wave counters)
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// Block barrier that first waits for this wave's LDS loads to return
// (s_waitcnt lgkmcnt(0); the vmcnt and expcnt fields keep their maximum, so
// global loads are not waited for).  The compiler emits s_barrier without
// that wait when a load's result is only used after the barrier, and on
// gfx950 such a load in flight has returned a value that another wave
// stored after the barrier (a write-after-read race on the LDS buffer).
__device__ __forceinline__ void block_sync()
{
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __syncthreads();
}

// number of set bits of m strictly below this lane
__device__ __forceinline__ uint32_t mask_rank(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Inclusive prefix sum over the 64 lanes (every lane active) from DPP lane
// moves, no LDS traffic: shifts by 1, 2 and 3 inside each row of 16 lanes
// give sums of four, shifts by 4 and 8 (upper banks only) row prefixes, and
// the row broadcasts of lanes 15 and 31 carry the row totals into the rows
// above.  Lanes a move has no source for, or that its row and bank masks
// leave out, receive 0.
template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ uint32_t dpp_or0(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, BANK_MASK, false);
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    uint32_t s = v + dpp_or0<0x111, 0xf, 0xf>(v);   // row_shr:1
    s += dpp_or0<0x112, 0xf, 0xf>(v);               // row_shr:2
    s += dpp_or0<0x113, 0xf, 0xf>(v);               // row_shr:3
    s += dpp_or0<0x114, 0xf, 0xe>(s);               // row_shr:4, lanes 4..15 of each row
    s += dpp_or0<0x118, 0xf, 0xc>(s);               // row_shr:8, lanes 8..15 of each row
    s += dpp_or0<0x142, 0xa, 0xf>(s);               // row_bcast:15 into rows 1 and 3
    s += dpp_or0<0x143, 0xc, 0xf>(s);               // row_bcast:31 into rows 2 and 3
    return s;
}

// Wave-wide minimum or maximum (every lane active; every lane gets it): the
// DPP moves of wave_incl_scan with the operation's identity in the lanes a
// move leaves out, then lane 63's running result.  T: 32- or 64-bit.
template <int CTRL, int ROW_MASK, int BANK_MASK, typename T>
__device__ __forceinline__ T dpp_or(T ident, T v)
{
    if constexpr (sizeof(T) == 8) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)ident, (int)(uint32_t)v, CTRL, ROW_MASK,
                                                                  BANK_MASK, false);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)(ident >> 32), (int)(uint32_t)(v >> 32),
                                                                  CTRL, ROW_MASK, BANK_MASK, false);
        return (T)(((uint64_t)hi << 32) | lo);
    } else {
        return (T)(uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)ident, (int)(uint32_t)v, CTRL, ROW_MASK,
                                                        BANK_MASK, false);
    }
}
template <bool MAX, typename T>
__device__ __forceinline__ T wave_extreme(T v)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    const T ident = MAX ? (T)0 : (T)~(T)0;
    auto op = [](T a, T b) { return MAX ? (a > b ? a : b) : (a < b ? a : b); };
    T s = op(v, dpp_or<0x111, 0xf, 0xf>(ident, v));
    s = op(s, dpp_or<0x112, 0xf, 0xf>(ident, v));
    s = op(s, dpp_or<0x113, 0xf, 0xf>(ident, v));
    s = op(s, dpp_or<0x114, 0xf, 0xe>(ident, s));
    s = op(s, dpp_or<0x118, 0xf, 0xc>(ident, s));
    s = op(s, dpp_or<0x142, 0xa, 0xf>(ident, s));
    s = op(s, dpp_or<0x143, 0xc, 0xf>(ident, s));
    if constexpr (sizeof(T) == 8)
        return (T)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(s >> 32), WAVE - 1) << 32) |
                   (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)s, WAVE - 1));
    else
        return (T)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)s, WAVE - 1);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, WAVE);
    return v;
}

// Stable rank through one LDS atomic: adds `add` to 16-bit counter c (half
// c & 1 of word c >> 1; counts stay below 2^16) and returns the counter's
// previous value.  On gfx950 the lanes of one wave's LDS atomic that hit
// the same word get their return values in ascending lane order, and a
// wave's LDS instructions complete in program order, so for a wave-private
// counter array the value returned to a key is the number of keys with the
// same counter before it in (instruction, lane) order.  The library checks
// that property on each device before it sorts (check_lds_rank_order).
__device__ __forceinline__ uint32_t rank16_add(uint32_t *words, uint32_t c, uint32_t add)
{
    const uint32_t sh = (c & 1u) << 4;
    const uint32_t old = atomicAdd(&words[c >> 1], add << sh);
    return (old >> sh) & 0xffffu;
}

// Lanes whose value v (< 2^BITS, or invalid) equals this lane's: AND of the
// per-bit ballots, restricted to valid lanes.  Per bit: s = bit b of v
// sign-extended (0 or ~0), bb = ballot of that bit; the lanes that differ
// from this one in that bit are s ^ bb, accumulated with OR, a 3-input
// bitwise function per 32-bit half (one v_bitop3 each on gfx950).
template <int BITS>
__device__ __forceinline__ uint64_t match_peers(uint32_t v, bool valid)
{
    // dl/dh collect the lanes that differ from this one in some bit
    uint32_t dl = 0, dh = 0;
#pragma unroll
    for (int b = 0; b < BITS; b++) {
        const int32_t si = __builtin_amdgcn_sbfe((int32_t)v, b, 1);   // bit b, sign-extended
        const uint64_t bb = ballot(si < 0);
        dl |= (uint32_t)si ^ (uint32_t)bb;
        dh |= (uint32_t)si ^ (uint32_t)(bb >> 32);
    }
    const uint64_t vm = ballot(valid);
    return ((uint64_t)(~dh & (uint32_t)(vm >> 32)) << 32) | (~dl & (uint32_t)vm);
}

// ---- agent-scope synchronisation ----------------------------------------------
// Work handed between workgroups (possibly on different XCDs, each with its
// own L2) is published with a release fence before the signalling atomic and
// consumed with an acquire fence after it.
__device__ __forceinline__ void fence_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

__device__ __forceinline__ uint32_t atomic_add(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long atomic_add64(unsigned long long *p, unsigned long long v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t atomic_add_i(int32_t *p, int32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_or(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_max(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t load_acquire(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t load_relaxed(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_relaxed(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- block-wide exclusive scan (NT threads) -------------------------------------
// red: LDS scratch of at least NT/64 + 1 words.  Returns the exclusive prefix
// of v over thread order; *total receives the block sum.  Contains barriers:
// every thread of the block must call it.
// TRAILING: end with a barrier, so that red may be rewritten right away;
// without it the caller passes a barrier before red is written again.
template <int NT, bool TRAILING = true>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *red, uint32_t *total)
{
    constexpr int NW = NT / WAVE;
    const int w = threadIdx.x / WAVE;
    uint32_t inc = wave_incl_scan(v);
    if (lane() == WAVE - 1) red[w] = inc;
    block_sync();
    uint32_t wpre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; i++) {
        uint32_t s = red[i];
        wpre += (i < w) ? s : 0u;
        tot += s;
    }
    if (TRAILING) block_sync();
    *total = tot;
    return wpre + inc - v;
}

// ---- range-checked loads (buffer instructions) ------------------------------------
// A raw buffer resource over [base, base + bytes): loads at byte offsets at or
// past `bytes` return 0 instead of touching memory, so a tile's last partial
// wave needs no per-item predicate or clamped address, and full and partial
// tiles share one load sequence (scalar base, one offset register, constant
// offsets).  Word 3 = 0x00020000: 32-bit data format, no swizzle, no index
// stride, the raw-buffer setting for gfx9-family parts.
typedef __amdgpu_buffer_rsrc_t BufRsrc;

__device__ __forceinline__ BufRsrc buf_rsrc(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}

template <typename T>
__device__ __forceinline__ T buf_load(BufRsrc r, uint32_t off)
{
    if constexpr (sizeof(T) == 8) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
        return (T)(((uint64_t)v[1] << 32) | v[0]);
    } else if constexpr (sizeof(T) == 4) {
        return (T)__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
    } else if constexpr (sizeof(T) == 2) {
        return (T)__builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0);
    } else {
        return (T)__builtin_amdgcn_raw_buffer_load_b8(r, (int)off, 0, 0);
    }
}

// store at a byte offset of a buffer resource; stores at or past its end are
// dropped.  The stores are non-temporal (cache policy NT): they write final
// outputs that no kernel of the run reads again, and a kernel that leaves
// less dirty data in the L2s ends sooner (65,536 x 1024 int32 MultiSort:
// 0.137 -> 0.127 ms per back-to-back launch).
constexpr int CPOL_NT = 2;

// the same for a plain pointer (a coalesced final-output store)
template <typename T>
__device__ __forceinline__ void nt_store(T *p, T v)
{
    __builtin_nontemporal_store(v, p);
}
template <typename T>
__device__ __forceinline__ void buf_store(BufRsrc r, uint32_t off, T v)
{
    if constexpr (sizeof(T) == 8) {
        typedef uint32_t v2u __attribute__((ext_vector_type(2)));
        const v2u d = {(uint32_t)(uint64_t)v, (uint32_t)((uint64_t)v >> 32)};
        __builtin_amdgcn_raw_buffer_store_b64(d, r, (int)off, 0, CPOL_NT);
    } else if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)v, r, (int)off, 0, CPOL_NT);
    } else if constexpr (sizeof(T) == 2) {
        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)v, r, (int)off, 0, CPOL_NT);
    } else {
        __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v, r, (int)off, 0, CPOL_NT);
    }
}

// ---- LDS layout with one padding slot every ITEMS elements -------------------
// Thread t of a blocked arrangement touches elements [t*ITEMS, (t+1)*ITEMS);
// the padding makes the stride ITEMS+1 (odd), so lanes hit distinct banks.
template <int ITEMS>
__device__ __forceinline__ uint32_t pidx(uint32_t i) { return i + i / ITEMS; }

template <int ITEMS>
__host__ __device__ constexpr uint32_t padded_len(uint32_t n) { return n + n / ITEMS + 1; }

// ---- stable LDS block sort ---------------------------------------------------------
// Sorts (order-value, payload) pairs held in LDS: sk[] keys in order
// representation, si[] 32-bit payloads (ignored when !IDX), both in the pidx
// layout.  The n_eff elements are n_eff / run_limit independent arrays of
// run_limit elements each (run_limit == n_eff: one array).  n_eff and
// run_limit are powers of two, ITEMS <= run_limit <= n_eff <= NT*ITEMS, and
// the unused tail of every array has been filled with key_max (it stays
// behind every real element because every step is stable).  Threads >=
// n_eff/ITEMS only take part in the barriers.
// Step 1: each thread sorts its ITEMS consecutive elements in registers with
//         odd-even transposition (stable: swaps only strictly greater pairs);
// Step 2: log2(run_limit/ITEMS) rounds of stable merge-path merges; a merge
//         never crosses a multiple of run_limit.
// The result is left in LDS in the same layout.
template <typename U, bool IDX, int NT, int ITEMS>
__device__ __forceinline__ void block_sort(U *sk, uint32_t *si, uint32_t n_eff, uint32_t run_limit)
{
    const uint32_t t = threadIdx.x;
    const bool active = t * ITEMS < n_eff;
    // this thread's ITEMS consecutive elements: pidx(t*ITEMS + i) =
    // t*(ITEMS+1) + i, one base address plus constant offsets
    U *const mk = sk + t * (ITEMS + 1);
    uint32_t *const mi = si + t * (ITEMS + 1);
    U k[ITEMS];
    uint32_t v[ITEMS];
    if (active) {
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            k[i] = mk[i];
            if (IDX) v[i] = mi[i];
        }
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
#pragma unroll
            for (int i = (r & 1); i + 1 < ITEMS; i += 2) {
                bool sw = k[i + 1] < k[i];
                U a = k[i], b = k[i + 1];
                k[i] = sw ? b : a;
                k[i + 1] = sw ? a : b;
                if (IDX) {
                    uint32_t x = v[i], y = v[i + 1];
                    v[i] = sw ? y : x;
                    v[i + 1] = sw ? x : y;
                }
            }
        }
    }
    for (uint32_t run = ITEMS; run < run_limit; run *= 2) {
        block_sync();
        if (active) {
#pragma unroll
            for (int i = 0; i < ITEMS; i++) {
                mk[i] = k[i];
                if (IDX) mi[i] = v[i];
            }
        }
        block_sync();
        if (active) {
            // The thread's ITEMS outputs are CH independent merge-path
            // chains of ITEMS/CH outputs each.  More chains overlap the LDS
            // latencies of their searches and merges but each pays its own
            // binary search; on gfx950 one chain measured fastest (MultiSort
            // of 1024-key arrays 0.46 -> 0.33 ms per 64M keys against two
            // chains, four were slower still).
            constexpr int CH = 1;
            constexpr int PER = ITEMS / CH;
            uint32_t ai[CH], bi[CH], aend[CH], bend[CH];
            U ka[CH], kb[CH];
#pragma unroll
            for (int h = 0; h < CH; h++) {
                const uint32_t pos = t * ITEMS + h * PER;
                const uint32_t base = pos & ~(2 * run - 1);
                const uint32_t diag = pos - base;
                const uint32_t A = base, B = base + run;
                // merge path: number of A elements among the first diag outputs
                uint32_t lo = diag > run ? diag - run : 0;
                uint32_t hi = diag < run ? diag : run;
                while (lo < hi) {
                    uint32_t mid = (lo + hi) >> 1;
                    U bk = sk[pidx<ITEMS>(B + diag - 1 - mid)];
                    U ak = sk[pidx<ITEMS>(A + mid)];
                    if (!(bk < ak)) lo = mid + 1; else hi = mid;
                }
                ai[h] = A + lo;
                bi[h] = B + (diag - lo);
                aend[h] = B;
                bend[h] = B + run;
            }
            // serial stable merges; ties take A.  One LDS load per step (the
            // successor of the element taken) plus the payload of the
            // element taken.  Positions A+run and B+run stay inside the
            // padded array.
#pragma unroll
            for (int h = 0; h < CH; h++) {
                ka[h] = sk[pidx<ITEMS>(ai[h])];
                kb[h] = sk[pidx<ITEMS>(bi[h])];
            }
#pragma unroll
            for (int i = 0; i < PER; i++) {
#pragma unroll
                for (int h = 0; h < CH; h++) {
                    const bool takeA = (bi[h] >= bend[h]) || (ai[h] < aend[h] && !(kb[h] < ka[h]));
                    k[h * PER + i] = takeA ? ka[h] : kb[h];
                    if (IDX) v[h * PER + i] = si[pidx<ITEMS>(takeA ? ai[h] : bi[h])];
                    ai[h] += takeA ? 1u : 0u;
                    bi[h] += takeA ? 0u : 1u;
                    const U nk = sk[pidx<ITEMS>(takeA ? ai[h] : bi[h])];
                    ka[h] = takeA ? nk : ka[h];
                    kb[h] = takeA ? kb[h] : nk;
                }
            }
        }
    }
    block_sync();
    if (active) {
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            mk[i] = k[i];
            if (IDX) mi[i] = v[i];
        }
    }
    block_sync();
}

__host__ __device__ constexpr uint32_t next_pow2(uint32_t x)
{
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

__host__ __device__ constexpr int ilog2(uint32_t x) { return x <= 1 ? 0 : 1 + ilog2(x >> 1); }

__device__ __forceinline__ uint32_t next_pow2_dev(uint32_t x)
{
    return x <= 1 ? 1u : 1u << (32 - __clz(x - 1));
}

}  // namespace ts
