// ts_merge.hip -- stable merge of two sorted runs in HBM for gfx950.
//
// The reference merges the sorted chunks of Sync() with one thread per
// element: a run-2 element's output rank is its own offset plus the upper
// bound of its key in run 1, a run-1 element's its offset plus the lower
// bound in run 2 (mergeSortedChunks, turtle-sort.cu:190-281; schedule
// turtle-sort.cuh:192-243), i.e. a stable merge in which run 1 goes first on
// equal keys, at log2(n) dependent global loads per element.
//
// Here a workgroup produces TILE consecutive outputs.  Where the tile
// boundaries cross the merge path (the number of run-1 elements among the
// first d outputs) is found first, one wave per boundary, with a 64-way
// search: each round every lane tests one evenly spaced candidate and a
// ballot keeps the interval between the last passing and the first failing
// candidate, so a run of 2^26 keys takes 5 rounds of loads instead of 26.
// Every 64th boundary is searched first over the whole runs; the crossing
// only grows with the diagonal, and so does the count of run-2 elements, so
// the boundaries in between search the interval their two neighbours leave
// (at most 64 tiles wide: 3 rounds, about half the scattered line fetches).
// The tile's slices of both runs are loaded into LDS with coalesced loads,
// every thread merges ITEMS outputs from its own diagonal (a binary search
// in LDS, then a serial stable merge), and the tile is written back through
// LDS with consecutive lanes on consecutive addresses.  One read and one
// write of every key (and index): 2 (K + I) bytes per element.
#include "ts_device.h"
#include "ts_internal.h"

namespace ts {

constexpr int MERGE_NT = 256;
constexpr int MERGE_ITEMS = 16;
constexpr uint32_t MERGE_TILE = MERGE_NT * MERGE_ITEMS;

// boundaries between two searched over the whole runs
constexpr uint64_t CORANK_STRIDE = 64;

// Merge-path crossing of diagonal d (run-1 elements among the first d
// outputs), searched by one wave: the smallest i in [lo, hi] with
// !(a[i] <= b[d - 1 - i]), or hi if none.  [lo, hi] holds the crossing and
// lies inside [max(0, d - nb), min(d, na)], where both loads are in range.
template <typename U, bool FL>
__device__ __forceinline__ uint64_t corank_wave(const U *a, const U *b, uint64_t d, uint64_t lo, uint64_t hi, U flip)
{
    const uint32_t l = lane();
    while (hi > lo) {   // wave-uniform
        const uint64_t len = hi - lo;
        if (len <= (uint64_t)WAVE) {
            // every candidate of the interval, one per lane
            const uint64_t i = lo + l;
            const bool valid = i < hi;
            bool fail = false;
            if (valid) fail = !(to_ord<U, FL>(a[i], flip) <= to_ord<U, FL>(b[d - 1 - i], flip));
            const uint64_t m = ballot(valid && fail);
            return m ? lo + (uint64_t)__builtin_ctzll(m) : hi;
        }
        // 64 evenly spaced candidates strictly inside [lo, hi)
        const uint64_t step = len / (WAVE + 1);
        const uint64_t i = lo + (l + 1) * step;
        const bool fail = !(to_ord<U, FL>(a[i], flip) <= to_ord<U, FL>(b[d - 1 - i], flip));
        const uint64_t m = ballot(fail);
        if (m == 0) {
            lo = lo + (uint64_t)WAVE * step + 1;   // every candidate passes: beyond the last
        } else {
            const uint32_t k = (uint32_t)__builtin_ctzll(m);
            hi = lo + (k + 1) * step;              // the first failing candidate
            lo = k ? lo + k * step + 1 : lo;       // past the last passing one
        }
    }
    return lo;
}

// The crossings of the tile boundaries, found before the merge by one wave
// each, so that no merge workgroup starts with a chain of dependent loads.
// The crossing of boundary k (diagonal k*TILE) is kept in the output
// itself, where only the two workgroups that need it read it before they
// write anything: the first 8 bytes of tile k and the last 8 bytes of tile
// k-1 (a tile's keys take TILE * sizeof(U) >= 4096 bytes; only the last
// tile can be shorter).
// FINE = false: boundaries CORANK_STRIDE, 2*CORANK_STRIDE, ... over the
// whole runs.  FINE: the others; wave w takes offset w % (CORANK_STRIDE-1)
// + 1 after boundary K0 = (w / (CORANK_STRIDE-1)) * CORANK_STRIDE and
// searches between the crossings c0 of K0 and c1 of K1 = K0 +
// CORANK_STRIDE (the end of both runs past the last tile): with diagonals
// D0 < d < D1, c0 <= c(d) <= c1 and D0 - c0 <= d - c(d) <= D1 - c1.
template <typename U, bool FL, bool FINE>
__global__ void __launch_bounds__(256) k_merge_corank(const U *__restrict__ a, uint64_t na, const U *__restrict__ b,
                                                      uint64_t nb, unsigned char *__restrict__ out, uint64_t ntiles,
                                                      U flip)
{
    const uint64_t w = (uint64_t)blockIdx.x * (256 / WAVE) + threadIdx.x / WAVE;
    const uint64_t k0 = FINE ? (w / (CORANK_STRIDE - 1)) * CORANK_STRIDE : 0;
    const uint64_t k = FINE ? k0 + w % (CORANK_STRIDE - 1) + 1 : (w + 1) * CORANK_STRIDE;
    if (k >= ntiles) return;   // wave-uniform
    const uint64_t d = k * MERGE_TILE;
    uint64_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
    if constexpr (FINE) {
        const uint64_t k1 = k0 + CORANK_STRIDE;
        const uint64_t D0 = k0 * MERGE_TILE, D1 = k1 < ntiles ? k1 * MERGE_TILE : na + nb;
        // (read from the copies in the last 8 bytes of the tiles before
        // them, which are whole tiles)
        const uint64_t c0 = k0 == 0 ? 0ull : *(const uint64_t *)(out + D0 * sizeof(U) - 8);
        const uint64_t c1 = k1 < ntiles ? *(const uint64_t *)(out + D1 * sizeof(U) - 8) : na;
        lo = c0 > lo ? c0 : lo;
        if (d + c1 > D1 && d + c1 - D1 > lo) lo = d + c1 - D1;
        hi = c1 < hi ? c1 : hi;
        hi = d - D0 + c0 < hi ? d - D0 + c0 : hi;
    }
    const uint64_t c = corank_wave<U, FL>(a, b, d, lo, hi, flip);
    if (lane() == 0) {
        // (a last tile of fewer than 8 bytes has no room for its copy: its
        // workgroup searches the crossing itself)
        if ((na + nb - d) * sizeof(U) >= 8) *(uint64_t *)(out + d * sizeof(U)) = c;
        *(uint64_t *)(out + d * sizeof(U) - 8) = c;
    }
}

template <typename U, bool FL, bool IDX>
__global__ void __launch_bounds__(MERGE_NT) k_merge(const U *__restrict__ a, const uint32_t *__restrict__ ia, uint64_t na,
                                                    const U *__restrict__ b, const uint32_t *__restrict__ ib, uint64_t nb,
                                                    U *__restrict__ out, uint32_t *__restrict__ iout, U flip)
{
    constexpr int IT = MERGE_ITEMS;
    constexpr uint32_t T = MERGE_TILE;
    __shared__ U sk[padded_len<IT>(T)];
    __shared__ uint32_t si[IDX ? padded_len<IT>(T) : 1];
    __shared__ uint64_t cor[2];
    const uint64_t ntot = na + nb;
    const uint64_t d0 = (uint64_t)blockIdx.x * T;
    const uint64_t d1 = d0 + T < ntot ? d0 + T : ntot;
    const uint32_t tid = threadIdx.x;
    // the crossings k_merge_corank left in this tile (read before the tile
    // is written); a last tile of fewer than 8 bytes has no copy of its
    // first crossing, which wave 0 searches instead
    const unsigned char *ob = (const unsigned char *)out;
    if (d0 != 0 && (ntot - d0) * sizeof(U) < 8) {
        if (tid < (uint32_t)WAVE) {   // wave 0, every lane
            const uint64_t c = corank_wave<U, FL>(a, b, d0, d0 > nb ? d0 - nb : 0ull, d0 < na ? d0 : na, flip);
            if (tid == 0) cor[0] = c;
        }
    } else if (tid == 0) {
        cor[0] = d0 == 0 ? 0ull : *(const uint64_t *)(ob + d0 * sizeof(U));
    }
    if (tid == 0) cor[1] = d1 == ntot ? na : *(const uint64_t *)(ob + d1 * sizeof(U) - 8);
    block_sync();
    const uint64_t a0 = cor[0], a1 = cor[1];
    const uint64_t b0 = d0 - a0, b1 = d1 - a1;
    const uint32_t la = (uint32_t)(a1 - a0), lb = (uint32_t)(b1 - b0), nt = la + lb;
    // the tile's slices: run 1 at [0, la), run 2 at [la, nt); all loads are
    // issued before any is used
    U kv[IT];
    uint32_t iv[IT];
#pragma unroll
    for (int r = 0; r < IT; r++) {
        const uint32_t e = tid + (uint32_t)r * MERGE_NT;
        const bool fa = e < la;
        const uint64_t src = fa ? a0 + e : b0 + (e - la);
        const bool ok = e < nt;
        kv[r] = ok ? (fa ? a[src] : b[src]) : (U)0;
        if (IDX) iv[r] = ok ? (fa ? ia[src] : ib[src]) : 0u;
    }
#pragma unroll
    for (int r = 0; r < IT; r++) {
        const uint32_t e = tid + (uint32_t)r * MERGE_NT;
        if (e < nt) {
            sk[e] = to_ord<U, FL>(kv[r], flip);
            if (IDX) si[e] = iv[r];
        }
    }
    block_sync();
    // this thread's outputs [dd, dd + IT) of the tile
    const uint32_t dd = tid * IT;
    U k[IT];
    uint32_t v[IT];
    if (dd < nt) {
        uint32_t lo = dd > lb ? dd - lb : 0, hi = dd < la ? dd : la;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (!(sk[la + dd - 1 - mid] < sk[mid])) lo = mid + 1;
            else hi = mid;
        }
        uint32_t ai = lo, bi = la + (dd - lo);
        const uint32_t bend = nt;
        U ka = ai < la ? sk[ai] : (U)0, kb = bi < bend ? sk[bi] : (U)0;
#pragma unroll
        for (int i = 0; i < IT; i++) {
            // ties take run 1 (stable)
            const bool takeA = bi >= bend || (ai < la && !(kb < ka));
            k[i] = takeA ? ka : kb;
            if (IDX) v[i] = si[takeA ? ai : bi];
            ai += takeA ? 1u : 0u;
            bi += takeA ? 0u : 1u;
            const uint32_t nx = takeA ? ai : bi;
            const U nk = nx < nt ? sk[nx] : (U)0;
            ka = takeA ? nk : ka;
            kb = takeA ? kb : nk;
        }
    }
    block_sync();
    if (dd < nt) {
        U *mk = sk + tid * (IT + 1);            // pidx<IT>(dd + i) = tid * (IT + 1) + i
        uint32_t *mi = si + tid * (IT + 1);
#pragma unroll
        for (int i = 0; i < IT; i++) {
            if (dd + i < nt) {
                mk[i] = k[i];
                if (IDX) mi[i] = v[i];
            }
        }
    }
    block_sync();
#pragma unroll
    for (int r = 0; r < IT; r++) {
        const uint32_t e = tid + (uint32_t)r * MERGE_NT;
        if (e < nt) {
            nt_store(&out[d0 + e], from_ord<U, FL>(sk[pidx<IT>(e)], flip));
            if (IDX) nt_store(&iout[d0 + e], si[pidx<IT>(e)]);
        }
    }
}

template <typename U, bool FL>
static hipError_t merge_launch(uint64_t flip, const void *a, const uint32_t *ia, uint64_t na, const void *b,
                               const uint32_t *ib, uint64_t nb, void *out, uint32_t *iout, hipStream_t st)
{
    const uint64_t n = na + nb;
    const uint64_t grid = (n + MERGE_TILE - 1) / MERGE_TILE;
    if (grid == 0) return hipSuccess;
    if (grid > 0x7fffffffull) return hipErrorInvalidValue;
    static_assert(MERGE_TILE * sizeof(U) >= 16, "two crossings fit in a tile");
    if (grid > 1) {
        // the coarse boundaries, then the rest between them (four waves per
        // workgroup, one boundary each)
        const uint64_t ncoarse = (grid - 1) / CORANK_STRIDE;
        if (ncoarse)
            hipLaunchKernelGGL((k_merge_corank<U, FL, false>), dim3((uint32_t)((ncoarse + 3) / 4)), dim3(256), 0, st,
                               (const U *)a, na, (const U *)b, nb, (unsigned char *)out, grid, (U)flip);
        const uint64_t nfine = (grid + CORANK_STRIDE - 1) / CORANK_STRIDE * (CORANK_STRIDE - 1);
        hipLaunchKernelGGL((k_merge_corank<U, FL, true>), dim3((uint32_t)((nfine + 3) / 4)), dim3(256), 0, st,
                           (const U *)a, na, (const U *)b, nb, (unsigned char *)out, grid, (U)flip);
    }
    if (iout)
        hipLaunchKernelGGL((k_merge<U, FL, true>), dim3((uint32_t)grid), dim3(MERGE_NT), 0, st, (const U *)a, ia, na,
                           (const U *)b, ib, nb, (U *)out, iout, (U)flip);
    else
        hipLaunchKernelGGL((k_merge<U, FL, false>), dim3((uint32_t)grid), dim3(MERGE_NT), 0, st, (const U *)a, nullptr,
                           na, (const U *)b, nullptr, nb, (U *)out, nullptr, (U)flip);
    return hipGetLastError();
}

hipError_t merge_run(int kt, uint64_t flip, const void *a, const uint32_t *ia, uint64_t na, const void *b,
                     const uint32_t *ib, uint64_t nb, void *out, uint32_t *iout, hipStream_t st)
{
    switch (kt) {
    case 1: return merge_launch<uint8_t, false>(flip, a, ia, na, b, ib, nb, out, iout, st);
    case 2: return merge_launch<uint16_t, false>(flip, a, ia, na, b, ib, nb, out, iout, st);
    case 4: return merge_launch<uint32_t, false>(flip, a, ia, na, b, ib, nb, out, iout, st);
    case 8: return merge_launch<uint64_t, false>(flip, a, ia, na, b, ib, nb, out, iout, st);
    case 4 | KT_FLOAT: return merge_launch<uint32_t, true>(flip, a, ia, na, b, ib, nb, out, iout, st);
    case 8 | KT_FLOAT: return merge_launch<uint64_t, true>(flip, a, ia, na, b, ib, nb, out, iout, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ts
