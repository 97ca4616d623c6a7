// ts_multisort.hip -- batched equal-size sort (MultiSort) for gfx950.
//
// The reference gives each CUDA thread one array and heap-sorts it in an
// interleaved layout (turtle-multi-sort.cuh:86-182); at 1024 keys per array
// that is ~20K dependent compare/swaps per thread.  Here one workgroup owns
// one array: a coalesced load into LDS, the stable block sort of
// ts_device.h (register odd-even transposition + LDS merge path), and a
// coalesced store.  Every array is sorted, including a remainder that is not
// a multiple of the reference's BlockSize (turtle-multi-sort.cuh:200-202
// skips it), and the input buffer is the one that is sorted
// (turtle-multi-sort.cuh:152 sorts an unfilled scratch buffer).
#include <type_traits>

#include <hip/hip_ext.h>

#include "ts_device.h"
#include "ts_internal.h"

namespace ts {

// IDX: a 32-bit payload per key moves with it (stably); iin == nullptr
// gives each key its position within its array.  A workgroup sorts
// `per_wg` consecutive arrays, each in its own power-of-two region of
// n_eff = next_pow2(arr_size) slots (per_wg * n_eff <= NT * ITEMS): short
// arrays are packed so that every lane has keys to sort.
template <typename U, bool FL, bool IDX, int NT, int ITEMS, bool PACK>
__global__ void __launch_bounds__(NT) ms_kernel(const U *__restrict__ in, U *__restrict__ out,
                                                 const uint32_t *__restrict__ iin, uint32_t *__restrict__ iout,
                                                 uint32_t arr_size, uint32_t per_wg, uint64_t num_arrays, U flip)
{
    constexpr uint32_t CAPA = (uint32_t)NT * ITEMS;
    __shared__ U sk[padded_len<ITEMS>(CAPA)];
    __shared__ uint32_t si[IDX ? padded_len<ITEMS>(CAPA) : 1];
    if constexpr (!PACK) {
        // one array per workgroup
        const size_t base = (size_t)blockIdx.x * arr_size;
        uint32_t n_eff = next_pow2_dev(arr_size);
        if (n_eff < (uint32_t)ITEMS) n_eff = ITEMS;
        U kv[ITEMS];
        uint32_t iv[ITEMS];
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            const uint32_t i = threadIdx.x + (uint32_t)r * NT;
            kv[r] = i < arr_size ? in[base + i] : (U)0;
            iv[r] = IDX && iin && i < arr_size ? iin[base + i] : i;
        }
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            const uint32_t i = threadIdx.x + (uint32_t)r * NT;
            if (i < n_eff) {
                sk[pidx<ITEMS>(i)] = i < arr_size ? to_ord<U, FL>(kv[r], flip) : key_max<U>();
                if (IDX) si[pidx<ITEMS>(i)] = iv[r];
            }
        }
        block_sync();
        block_sort<U, IDX, NT, ITEMS>(sk, si, n_eff, n_eff);
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            const uint32_t i = threadIdx.x + (uint32_t)r * NT;
            if (i < arr_size) {
                nt_store(&out[base + i], from_ord<U, FL>(sk[pidx<ITEMS>(i)], flip));
                if (IDX) nt_store(&iout[base + i], si[pidx<ITEMS>(i)]);
            }
        }
        return;
    }
    const uint64_t arr0 = (uint64_t)blockIdx.x * per_wg;
    const uint32_t nar = num_arrays - arr0 < per_wg ? (uint32_t)(num_arrays - arr0) : per_wg;
    uint32_t n_eff = next_pow2_dev(arr_size);
    if (n_eff < (uint32_t)ITEMS) n_eff = ITEMS;
    const uint32_t lg = 31 - __clz(n_eff);
    const uint32_t tot = per_wg * n_eff;   // slots in use: per_wg regions
    const size_t base = (size_t)arr0 * arr_size;
    // ITEMS strided slots per thread; slot s is key s & (n_eff-1) of the
    // workgroup's array s >> lg.  All loads are issued before any is used,
    // so that their latencies overlap
    U kv[ITEMS];
    uint32_t iv[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const uint32_t s = threadIdx.x + (uint32_t)r * NT;
        const uint32_t q = s >> lg, i = s & (n_eff - 1);
        const bool ok = q < nar && i < arr_size;
        const size_t g = base + (size_t)q * arr_size + i;
        kv[r] = ok ? in[g] : (U)0;
        iv[r] = IDX && iin && ok ? iin[g] : i;
    }
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const uint32_t s = threadIdx.x + (uint32_t)r * NT;
        const uint32_t q = s >> lg, i = s & (n_eff - 1);
        if (s < tot) {
            sk[pidx<ITEMS>(s)] = q < nar && i < arr_size ? to_ord<U, FL>(kv[r], flip) : key_max<U>();
            if (IDX) si[pidx<ITEMS>(s)] = iv[r];
        }
    }
    block_sync();
    block_sort<U, IDX, NT, ITEMS>(sk, si, tot, n_eff);
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const uint32_t s = threadIdx.x + (uint32_t)r * NT;
        const uint32_t q = s >> lg, i = s & (n_eff - 1);
        if (q < nar && i < arr_size) {
            const size_t g = base + (size_t)q * arr_size + i;
            nt_store(&out[g], from_ord<U, FL>(sk[pidx<ITEMS>(s)], flip));
            if (IDX) nt_store(&iout[g], si[pidx<ITEMS>(s)]);
        }
    }
}

// ---- arrays of at most 32 keys: one lane per array --------------------------------
// The reference's MultiSort gives each thread one array (its heap sort,
// turtle-multi-sort.cuh:86-182); for arrays this short that is the right
// grain.  Here a workgroup of NT lanes takes NT consecutive arrays: their
// keys are loaded with coalesced loads into LDS (array q at q*P, P = the
// array length rounded up to an odd number of words, so that lanes reading
// their own arrays hit distinct banks), each lane sorts its array in
// registers with Batcher's odd-even merge network of N = next_pow2(len)
// wires (keys past the length are the largest key), and the arrays go back
// through LDS with coalesced stores.  With an index the network compares
// (key, position) pairs, a total order, so the result is the stable order,
// and the payload is gathered by the sorted positions.
// Order values are held as W (32 bits for keys of up to 4 bytes).
template <typename W, int N, bool IDX>
__device__ __forceinline__ void cmpx(W (&k)[N], uint32_t (&p)[N], int i, int j)
{
    if constexpr (IDX) {
        const bool sw = k[j] < k[i] || (k[j] == k[i] && p[j] < p[i]);
        const W a = k[i], b = k[j];
        const uint32_t x = p[i], y = p[j];
        k[i] = sw ? b : a;
        k[j] = sw ? a : b;
        p[i] = sw ? y : x;
        p[j] = sw ? x : y;
    } else {
        const W a = k[i], b = k[j];
        k[i] = a < b ? a : b;
        k[j] = a < b ? b : a;
    }
}

// Batcher's odd-even merge sort on N (a power of two) registers
template <typename W, int N, bool IDX>
__device__ __forceinline__ void oddeven_merge_sort(W (&k)[N], uint32_t (&p)[N])
{
#pragma unroll
    for (int q = 1; q < N; q <<= 1)
#pragma unroll
        for (int d = q; d >= 1; d >>= 1)
#pragma unroll
            for (int j = d % q; j + d < N; j += 2 * d)
#pragma unroll
                for (int i = 0; i < d; i++)
                    if (i + j + d < N && (i + j) / (2 * q) == (i + j + d) / (2 * q)) cmpx<W, N, IDX>(k, p, i + j, i + j + d);
}

template <typename U, typename W, bool FL, bool IDX, int N, int NT>
__global__ void __launch_bounds__(NT) ms_small(const U *__restrict__ in, U *__restrict__ out,
                                                const uint32_t *__restrict__ iin, uint32_t *__restrict__ iout,
                                                uint32_t arr_size, uint64_t inv, uint64_t num_arrays, U flip)
{
    constexpr uint32_t PMAX = 33;
    __shared__ W sk[NT * PMAX];
    __shared__ uint32_t si[IDX ? NT * PMAX : 1];
    const uint32_t A = arr_size, P = A | 1u;
    const uint64_t arr0 = (uint64_t)blockIdx.x * NT;
    const uint32_t nar = num_arrays - arr0 < (uint64_t)NT ? (uint32_t)(num_arrays - arr0) : (uint32_t)NT;
    const uint32_t ne = nar * A;
    const size_t base = (size_t)arr0 * A;
    const uint32_t t = threadIdx.x;
    // coalesced load, element e of the workgroup's run = key e % A of its
    // array e / A (e / A = (e * inv) >> 32 with inv = ceil(2^32 / A), exact
    // for e < 2^32 / A); eight loads in flight per step
    for (uint32_t r0 = 0; r0 < A; r0 += 8) {
        U kv[8];
        uint32_t iv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint32_t e = t + (r0 + u) * NT;
            const bool ok = r0 + u < A && e < ne;
            kv[u] = ok ? in[base + e] : (U)0;
            iv[u] = IDX && iin && ok ? iin[base + e] : 0u;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint32_t e = t + (r0 + u) * NT;
            if (r0 + u < A && e < ne) {
                const uint32_t q = (uint32_t)(((uint64_t)e * inv) >> 32), j = e - q * A;
                sk[q * P + j] = (W)to_ord<U, FL>(kv[u], flip);
                if (IDX) si[q * P + j] = iin ? iv[u] : j;
            }
        }
    }
    block_sync();
    W k[N];
    uint32_t p[N];
    const bool act = t < nar;
#pragma unroll
    for (int j = 0; j < N; j++) {
        k[j] = act && (uint32_t)j < A ? sk[t * P + j] : (W)key_max<U>();
        p[j] = (uint32_t)j;
    }
    oddeven_merge_sort<W, N, IDX>(k, p);
    uint32_t v[N];
    if (IDX) {
#pragma unroll
        for (int j = 0; j < N; j++) v[j] = act && (uint32_t)j < A ? si[t * P + (p[j] < A ? p[j] : 0u)] : 0u;
    }
    block_sync();
    if (act) {
#pragma unroll
        for (int j = 0; j < N; j++) {
            if ((uint32_t)j < A) {
                sk[t * P + j] = k[j];
                if (IDX) si[t * P + j] = v[j];
            }
        }
    }
    block_sync();
    for (uint32_t r = 0; r < A; r++) {
        const uint32_t e = t + r * NT;
        if (e < ne) {
            const uint32_t q = (uint32_t)(((uint64_t)e * inv) >> 32), j = e - q * A;
            nt_store(&out[base + e], from_ord<U, FL>((U)sk[q * P + j], flip));
            if (IDX) nt_store(&iout[base + e], si[q * P + j]);
        }
    }
}

// ---- arrays of 513..1024 keys of up to 4 bytes: one wave per array, LDS radix ------
// Each wave sorts one array (four waves per workgroup, no block barrier)
// by stable LSD passes over (key - min), lane l holding positions i*64 + l
// (16 items): the first pass ranks the keys straight from the registers
// they were loaded into, later ones read them back from LDS in position
// order.  Per pass a 32-bit per-wave counter per 9-bit digit gives every
// key its rank with one LDS atomic (its return value: the wave's keys of a
// digit in (item, lane) order, i.e. in position order; see rank16_add), a
// 64-lane scan of the counters gives each digit's first slot, and the keys
// are placed in LDS.  Positions past the array's end hold the largest key
// and stay behind its keys.  The scans and the key range run on DPP moves.
//
// An array whose key span (max - min) is wider than two 9-bit digits is
// sorted by its top 18 span bits only ((key - min) >> lowb, two passes);
// keys that then agree in those bits (a group; for 1024 spread keys about
// two pairs per array) are neighbours, found by comparing each key with
// the one before it across lanes (DPP whole-wave shifts); the lane holding
// a group's first position orders the group in LDS by insertion sort
// (stable).  A group longer than GMAX makes the array take the exact passes
// instead (as does a span of at most 18 bits): 9-bit digits over the whole
// span.
template <typename U, bool FL, bool IDX>
__global__ void __launch_bounds__(256) ms_radix(const U *__restrict__ in, U *__restrict__ out,
                                                 const uint32_t *__restrict__ iin, uint32_t *__restrict__ iout,
                                                 uint32_t arr_size, uint64_t num_arrays, U flip)
{
    static_assert(sizeof(U) <= 4, "keys of up to 4 bytes");
    constexpr int WPB = 4, IT = 16;
    constexpr int IT_FULL = 8;               // arrays hold more than 512 keys: items 0..7 always valid
    constexpr uint32_t NS = WAVE * IT;       // 1024 slots per wave
    constexpr uint32_t RB = 9, NBIN = 1u << RB, TOPB = 2 * RB, GMAX = 64;
    constexpr uint32_t CPL = NBIN / WAVE;    // counters per lane in the scan
    __shared__ U sk[WPB][NS];
    __shared__ uint32_t si[IDX ? WPB : 1][IDX ? NS : 1];
    __shared__ uint32_t cnt[WPB][NBIN];      // digit counters per wave
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const uint64_t arr = (uint64_t)blockIdx.x * WPB + w;
    if (arr >= num_arrays) return;           // wave-uniform; no block barrier below
    const size_t base = (size_t)arr * arr_size;
    // range-checked views of this array: loads past its end read 0, stores
    // past it are dropped
    const BufRsrc rin = buf_rsrc(in + base, arr_size * (uint32_t)sizeof(U));
    const BufRsrc rout = buf_rsrc(out + base, arr_size * (uint32_t)sizeof(U));
    const BufRsrc riin = buf_rsrc(IDX && iin ? iin + base : nullptr, IDX && iin ? arr_size * 4u : 0u);
    const BufRsrc riout = buf_rsrc(IDX ? iout + base : nullptr, IDX ? arr_size * 4u : 0u);
    U *const ks = sk[w];
    uint32_t *const vs = si[IDX ? w : 0];
    uint32_t *const cw = cnt[w];
    // load: keys (order representation) and indices of the lane's positions
    // i*64 + l into registers; positions past the array's end hold the
    // largest key.  Also returns the lane's smallest and largest key.
    auto load = [&](U (&k)[IT], uint32_t (&v)[IT], U &mn, U &mx) {
        uint32_t l = lane();
        asm volatile("" : "+v"(l));   // keeps per-lane addresses out of the enclosing loops
#pragma unroll
        for (int i = 0; i < IT; i++) {
            const uint32_t j = (uint32_t)i * WAVE + l;
            k[i] = buf_load<U>(rin, j * (uint32_t)sizeof(U));
            if (IDX) v[i] = iin ? buf_load<uint32_t>(riin, j * 4u) : j;
        }
        mn = key_max<U>();
        mx = 0;
#pragma unroll
        for (int i = 0; i < IT; i++) {
            const uint32_t j = (uint32_t)i * WAVE + l;
            const bool ok = i < IT_FULL || j < arr_size;
            const U key = to_ord<U, FL>(k[i], flip);
            mn = ok && key < mn ? key : mn;
            mx = ok && key > mx ? key : mx;
            k[i] = ok ? key : key_max<U>();
        }
    };
    // one stable pass over the keys held in registers (position order):
    // digit(key) of every item, ranks through the counters, 64-lane scan,
    // placement in LDS
    auto pass_regs = [&](const U (&k)[IT], const uint32_t (&v)[IT], auto digit) {
        uint32_t l = lane();
        asm volatile("" : "+v"(l));
#pragma unroll
        for (uint32_t c = 0; c < CPL; c++) cw[l * CPL + c] = 0;
        wave_sync();
        uint32_t d[IT], r[IT];
#pragma unroll
        for (int i = 0; i < IT; i++) {
            d[i] = digit(k[i]);
            r[i] = atomicAdd(&cw[d[i]], 1u);   // stable rank (see rank16_add)
        }
        wave_sync();
        // exclusive scan: lane l owns the CPL consecutive digits from l*CPL
        {
            uint32_t c[CPL], tot = 0;
#pragma unroll
            for (uint32_t q = 0; q < CPL; q++) {
                c[q] = cw[l * CPL + q];
                tot += c[q];
            }
            uint32_t run = wave_incl_scan(tot) - tot;
#pragma unroll
            for (uint32_t q = 0; q < CPL; q++) {
                cw[l * CPL + q] = run;
                run += c[q];
            }
        }
        wave_sync();
        // the counts add up to the wave's 1024 slots, so pos < NS
#pragma unroll
        for (int i = 0; i < IT; i++) {
            const uint32_t pos = cw[d[i]] + r[i];
            ks[pos] = k[i];
            if (IDX) vs[pos] = v[i];
        }
        wave_sync();
    };
    // the same over the keys LDS holds in position order
    auto pass_lds = [&](auto digit) {
        uint32_t l = lane();
        asm volatile("" : "+v"(l));
        U k[IT];
        uint32_t v[IT];
#pragma unroll
        for (int i = 0; i < IT; i++) {
            k[i] = ks[(uint32_t)i * WAVE + l];
            if (IDX) v[i] = vs[(uint32_t)i * WAVE + l];
        }
        pass_regs(k, v, digit);
    };
    U mn = 0;
    uint32_t span = 0, nfull = 0, lowb = 0;
    bool top = false;
    // offset from the minimum, clamped to the span: padding positions (the
    // largest key) tie with the array's maximum and, being later, stay behind
    auto rel = [&](U key) {
        const uint32_t q = (uint32_t)(U)(key - mn);
        return q < span ? q : span;
    };
    auto top_of = [&](U key) { return rel(key) >> lowb; };
#pragma nounroll
    for (uint32_t attempt = 0; attempt < 2; attempt++) {   // wave-uniform
        {
            U k[IT], lmn, lmx;
            uint32_t v[IT];
            load(k, v, lmn, lmx);
            if (attempt == 0) {
                lmn = (U)wave_extreme<false, uint32_t>((uint32_t)lmn);   // DPP reductions, no LDS traffic
                lmx = (U)wave_extreme<true, uint32_t>((uint32_t)lmx);
                mn = (U)__builtin_amdgcn_readfirstlane((int)(uint32_t)lmn);
                span = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(U)(lmx - mn));
                const uint32_t bits = span ? 32u - (uint32_t)__clz(span) : 0u;
                nfull = (bits + RB - 1) / RB;
                top = nfull > 2;
                lowb = top ? bits - TOPB : 0u;
            }
            if (nfull == 0) {   // every key equal: stage them for the write-out
                uint32_t l = lane();
                asm volatile("" : "+v"(l));
#pragma unroll
                for (int i = 0; i < IT; i++) {
                    ks[(uint32_t)i * WAVE + l] = k[i];
                    if (IDX) vs[(uint32_t)i * WAVE + l] = v[i];
                }
                wave_sync();
            } else {
                pass_regs(k, v, [&](U key) { return (rel(key) >> lowb) & (NBIN - 1); });
            }
        }
        const uint32_t npass = top ? 2u : nfull;
#pragma nounroll
        for (uint32_t p = 1; p < npass; p++) {
            const uint32_t sh = lowb + p * RB;
            pass_lds([&](U key) { return (rel(key) >> sh) & (NBIN - 1); });
        }
        if (top) {
            // groups: runs of keys with equal top bits (position order
            // j = i*64 + l; the key before j is lane l-1 of item i, or lane 63
            // of item i-1; padding positions hold the largest key and never
            // join).  eq bit i: the key before item i's has the same top bits.
            // The neighbour comes from a DPP whole-wave shift by one lane,
            // lane 0 taking lane 63 of the item before.
            // The keys (and indices) read here are written out from these
            // registers, except for items where some lane holds a group
            // member (the fix-up below may have moved those).
            uint32_t l = lane();
            asm volatile("" : "+v"(l));
            U kr[IT];
            uint32_t vr[IDX ? IT : 1];
            uint32_t eq = 0, last = 0;
#pragma unroll
            for (int i = 0; i < IT; i++) {
                const uint32_t j = (uint32_t)i * WAVE + l;
                kr[i] = ks[j];
                if (IDX) vr[i] = vs[j];
                const uint32_t t = top_of(kr[i]);
                const uint32_t before = (uint32_t)__builtin_amdgcn_update_dpp((int)last, (int)t, 0x138, 0xf, 0xf, false);
                const bool ok = (i > 0 || l > 0) && (i < IT_FULL || j < arr_size);
                if (ok && before == t) eq |= 1u << i;
                last = (uint32_t)__builtin_amdgcn_readlane((int)t, WAVE - 1);
            }
            // a group's head has no eq bit and the key after it has one (lane
            // l+1 of item i, or lane 0 of item i+1: the shift the other way)
            const uint32_t eq_first = (uint32_t)__builtin_amdgcn_readfirstlane((int)eq);
            const uint32_t after = (uint32_t)__builtin_amdgcn_update_dpp((int)(eq_first >> 1), (int)eq, 0x130, 0xf, 0xf,
                                                                         false);
            const uint32_t heads = after & ~eq;
            bool big = false;
            if (ballot(heads != 0)) {   // wave-uniform: rare for spread keys
                // the head's lane orders its group in LDS by insertion sort
                // (stable); a group longer than GMAX makes the array take the
                // exact passes
#pragma nounroll
                for (uint32_t g = heads; g; g &= g - 1) {
                    const uint32_t gs = (uint32_t)__builtin_ctz(g) * WAVE + l;
                    const uint32_t t = top_of(ks[gs]);
                    uint32_t ge = gs + 1;
#pragma nounroll
                    while (ge < arr_size && ge - gs <= GMAX && top_of(ks[ge]) == t) ge++;
                    if (ge - gs > GMAX) {
                        big = true;
                        break;
                    }
#pragma nounroll
                    for (uint32_t q = gs + 1; q < ge; q++) {
                        const U kq = ks[q];
                        const uint32_t vq = IDX ? vs[q] : 0u;
                        uint32_t r = q;
#pragma nounroll
                        while (r > gs && ks[r - 1] > kq) {
                            ks[r] = ks[r - 1];
                            if (IDX) vs[r] = vs[r - 1];
                            r--;
                        }
                        ks[r] = kq;
                        if (IDX) vs[r] = vq;
                    }
                }
            }
            if (ballot(big)) {   // wave-uniform
                top = false;
                lowb = 0;
                continue;
            }
            wave_sync();
            const uint32_t grp = eq | after;   // this lane's items that are group members
#pragma unroll
            for (int i = 0; i < IT; i++) {
                const uint32_t j = (uint32_t)i * WAVE + l;
                U key = kr[i];
                uint32_t val = IDX ? vr[IDX ? i : 0] : 0u;
                if (ballot((grp >> i) & 1u)) {   // wave-uniform
                    key = ks[j];
                    if (IDX) val = vs[j];
                }
                buf_store<U>(rout, j * (uint32_t)sizeof(U), from_ord<U, FL>(key, flip));
                if (IDX) buf_store<uint32_t>(riout, j * 4u, val);
            }
            break;
        }
        // exact passes: write-out in position order from LDS (stores past the
        // end dropped)
        uint32_t l = lane();
        asm volatile("" : "+v"(l));
#pragma unroll
        for (int i = 0; i < IT; i++) {
            const uint32_t j = (uint32_t)i * WAVE + l;
            buf_store<U>(rout, j * (uint32_t)sizeof(U), from_ord<U, FL>(ks[j], flip));
            if (IDX) buf_store<uint32_t>(riout, j * 4u, vs[j]);
        }
        break;
    }
}

// Launches carry the caller's timing events (start, stop; either may be
// null) in the kernel's own dispatch: timing a launch with separate event
// records costs several microseconds of queue time per launch.
template <typename U, bool FL, bool IDX, int N>
static void small_launch(uint32_t arr_size, const U *in, U *out, const uint32_t *iin, uint32_t *iout, uint64_t na,
                         U f, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    typedef typename std::conditional<sizeof(U) == 8, uint64_t, uint32_t>::type W;
    constexpr int NT = sizeof(U) == 8 ? 128 : 256;
    const uint64_t inv = ((1ull << 32) + arr_size - 1) / arr_size;
    hipExtLaunchKernelGGL((ms_small<U, W, FL, IDX, N, NT>), dim3((uint32_t)((na + NT - 1) / NT)), dim3(NT), 0, st, e0,
                          e1, 0, in, out, iin, iout, arr_size, inv, na, f);
}

template <typename U, bool FL, bool IDX>
static hipError_t ms_launch(uint64_t flip, uint32_t arr_size, const void *in, void *out, const uint32_t *iin,
                            uint32_t *iout, uint64_t num_arrays, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1)
{
    const U f = (U)flip;
    const U *i = (const U *)in;
    U *o = (U *)out;
    uint32_t ne = next_pow2(arr_size);
    // grid.x is limited to 2^31-1 workgroups; larger batches go in chunks
    const uint64_t chunk = 1ull << 30;
    for (uint64_t a0 = 0; a0 < num_arrays; a0 += chunk) {
        const uint64_t na = (num_arrays - a0) < chunk ? (num_arrays - a0) : chunk;
        const uint64_t e0 = a0 * arr_size;
        const U *ii = i + e0;
        U *oo = o + e0;
        const uint32_t *xi = iin ? iin + e0 : nullptr;
        uint32_t *xo = IDX ? iout + e0 : nullptr;
        // the first chunk's launch starts the timing, the last one's stops it
        hipEvent_t t0 = a0 == 0 ? ev0 : nullptr, t1 = a0 + na == num_arrays ? ev1 : nullptr;
        if (ne <= 32) {
            // one lane per array, a register network of ne wires
            switch (ne < 2 ? 2 : ne) {
            case 2: small_launch<U, FL, IDX, 2>(arr_size, ii, oo, xi, xo, na, f, st, t0, t1); break;
            case 4: small_launch<U, FL, IDX, 4>(arr_size, ii, oo, xi, xo, na, f, st, t0, t1); break;
            case 8: small_launch<U, FL, IDX, 8>(arr_size, ii, oo, xi, xo, na, f, st, t0, t1); break;
            case 16: small_launch<U, FL, IDX, 16>(arr_size, ii, oo, xi, xo, na, f, st, t0, t1); break;
            default: small_launch<U, FL, IDX, 32>(arr_size, ii, oo, xi, xo, na, f, st, t0, t1); break;
            }
        } else if (ne == 1024 && sizeof(U) <= 4) {
            // 513..1024 keys of up to 4 bytes: LDS radix, one wave per array
            if constexpr (sizeof(U) <= 4)
                hipExtLaunchKernelGGL((ms_radix<U, FL, IDX>), dim3((uint32_t)((na + 3) / 4)), dim3(256), 0, st, t0, t1,
                                      0, ii, oo, xi, xo, arr_size, na, f);
        } else if (ne <= 1024) {
            // one wave of 16 keys per lane, 1024 / ne arrays per wave (no
            // cross-wave barriers; 16 keys per thread take one merge round
            // fewer than 8; 65,536 x 1024 int32 keys: 0.34 -> 0.29 ms)
            const uint32_t per = 1024u / ne;
            if (per > 1)
                hipExtLaunchKernelGGL((ms_kernel<U, FL, IDX, 64, 16, true>), dim3((uint32_t)((na + per - 1) / per)),
                                      dim3(64), 0, st, t0, t1, 0, ii, oo, xi, xo, arr_size, per, na, f);
            else
                hipExtLaunchKernelGGL((ms_kernel<U, FL, IDX, 64, 16, false>), dim3((uint32_t)na), dim3(64), 0, st, t0,
                                      t1, 0, ii, oo, xi, xo, arr_size, 1u, na, f);
        } else if (ne <= 2048)
            hipExtLaunchKernelGGL((ms_kernel<U, FL, IDX, 128, 16, false>), dim3((uint32_t)na), dim3(128), 0, st, t0,
                                  t1, 0, ii, oo, xi, xo, arr_size, 1u, na, f);
        else if (ne <= 4096)
            hipExtLaunchKernelGGL((ms_kernel<U, FL, IDX, 256, 16, false>), dim3((uint32_t)na), dim3(256), 0, st, t0,
                                  t1, 0, ii, oo, xi, xo, arr_size, 1u, na, f);
        else if (ne <= 8192)
            hipExtLaunchKernelGGL((ms_kernel<U, FL, IDX, 512, 16, false>), dim3((uint32_t)na), dim3(512), 0, st, t0,
                                  t1, 0, ii, oo, xi, xo, arr_size, 1u, na, f);
        else
            return hipErrorInvalidValue;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename U, bool FL>
static hipError_t ms_pick(uint64_t flip, uint32_t arr_size, const void *in, void *out, const uint32_t *iin,
                          uint32_t *iout, uint64_t num_arrays, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    return iout ? ms_launch<U, FL, true>(flip, arr_size, in, out, iin, iout, num_arrays, st, e0, e1)
                : ms_launch<U, FL, false>(flip, arr_size, in, out, nullptr, nullptr, num_arrays, st, e0, e1);
}

uint32_t multisort_max_size(int) { return 8192; }

hipError_t multisort_run(int kt, uint64_t flip, uint32_t arr_size, const void *in, void *out, const uint32_t *iin,
                         uint32_t *iout, uint64_t num_arrays, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    if (arr_size == 0 || num_arrays == 0) return hipSuccess;
    if (arr_size > multisort_max_size(kt)) return hipErrorInvalidValue;
    switch (kt) {
    case 1: return ms_pick<uint8_t, false>(flip, arr_size, in, out, iin, iout, num_arrays, st, e0, e1);
    case 2: return ms_pick<uint16_t, false>(flip, arr_size, in, out, iin, iout, num_arrays, st, e0, e1);
    case 4: return ms_pick<uint32_t, false>(flip, arr_size, in, out, iin, iout, num_arrays, st, e0, e1);
    case 8: return ms_pick<uint64_t, false>(flip, arr_size, in, out, iin, iout, num_arrays, st, e0, e1);
    case 4 | KT_FLOAT: return ms_pick<uint32_t, true>(flip, arr_size, in, out, iin, iout, num_arrays, st, e0, e1);
    case 8 | KT_FLOAT: return ms_pick<uint64_t, true>(flip, arr_size, in, out, iin, iout, num_arrays, st, e0, e1);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ts
