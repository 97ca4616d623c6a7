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
#include "ts_device.h"
#include "ts_internal.h"

namespace ts {

// IDX: a 32-bit payload per key moves with it (stably); iin == nullptr
// gives each key its position within its array.
template <typename U, bool FL, bool IDX, int NT, int ITEMS>
__global__ void __launch_bounds__(NT) ms_kernel(const U *__restrict__ in, U *__restrict__ out,
                                                 const uint32_t *__restrict__ iin, uint32_t *__restrict__ iout,
                                                 uint32_t arr_size, U flip)
{
    constexpr uint32_t CAPA = (uint32_t)NT * ITEMS;
    __shared__ U sk[padded_len<ITEMS>(CAPA)];
    __shared__ uint32_t si[IDX ? padded_len<ITEMS>(CAPA) : 1];
    const size_t base = (size_t)blockIdx.x * arr_size;
    uint32_t n_eff = next_pow2_dev(arr_size);
    if (n_eff < (uint32_t)ITEMS) n_eff = ITEMS;
    // n_eff <= NT*ITEMS: ITEMS strided positions per thread; all loads are
    // issued before any is used, so that their latencies overlap
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
            out[base + i] = from_ord<U, FL>(sk[pidx<ITEMS>(i)], flip);
            if (IDX) iout[base + i] = si[pidx<ITEMS>(i)];
        }
    }
}

template <typename U, bool FL, bool IDX>
static hipError_t ms_launch(uint64_t flip, uint32_t arr_size, const void *in, void *out, const uint32_t *iin,
                            uint32_t *iout, uint64_t num_arrays, hipStream_t st)
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
        // one workgroup per array, sized so that every thread holds ITEMS
        // keys of it (one wave up to 1024 keys: no cross-wave barriers, and
        // 16 keys per thread take one merge round fewer than 8; 65,536 x
        // 1024 int32 keys: 0.34 -> 0.29 ms)
        if (ne <= 256)
            hipLaunchKernelGGL((ms_kernel<U, FL, IDX, 64, 4>), dim3((uint32_t)na), dim3(64), 0, st, ii, oo, xi, xo,
                               arr_size, f);
        else if (ne <= 512)
            hipLaunchKernelGGL((ms_kernel<U, FL, IDX, 64, 8>), dim3((uint32_t)na), dim3(64), 0, st, ii, oo, xi, xo,
                               arr_size, f);
        else if (ne <= 1024)
            hipLaunchKernelGGL((ms_kernel<U, FL, IDX, 64, 16>), dim3((uint32_t)na), dim3(64), 0, st, ii, oo, xi, xo,
                               arr_size, f);
        else if (ne <= 2048)
            hipLaunchKernelGGL((ms_kernel<U, FL, IDX, 128, 16>), dim3((uint32_t)na), dim3(128), 0, st, ii, oo, xi,
                               xo, arr_size, f);
        else if (ne <= 4096)
            hipLaunchKernelGGL((ms_kernel<U, FL, IDX, 256, 16>), dim3((uint32_t)na), dim3(256), 0, st, ii, oo, xi,
                               xo, arr_size, f);
        else if (ne <= 8192)
            hipLaunchKernelGGL((ms_kernel<U, FL, IDX, 512, 16>), dim3((uint32_t)na), dim3(512), 0, st, ii, oo, xi,
                               xo, arr_size, f);
        else
            return hipErrorInvalidValue;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename U, bool FL>
static hipError_t ms_pick(uint64_t flip, uint32_t arr_size, const void *in, void *out, const uint32_t *iin,
                          uint32_t *iout, uint64_t num_arrays, hipStream_t st)
{
    return iout ? ms_launch<U, FL, true>(flip, arr_size, in, out, iin, iout, num_arrays, st)
                : ms_launch<U, FL, false>(flip, arr_size, in, out, nullptr, nullptr, num_arrays, st);
}

uint32_t multisort_max_size(int) { return 8192; }

hipError_t multisort_run(int kt, uint64_t flip, uint32_t arr_size, const void *in, void *out, const uint32_t *iin,
                         uint32_t *iout, uint64_t num_arrays, hipStream_t st)
{
    if (arr_size == 0 || num_arrays == 0) return hipSuccess;
    if (arr_size > multisort_max_size(kt)) return hipErrorInvalidValue;
    switch (kt) {
    case 1: return ms_pick<uint8_t, false>(flip, arr_size, in, out, iin, iout, num_arrays, st);
    case 2: return ms_pick<uint16_t, false>(flip, arr_size, in, out, iin, iout, num_arrays, st);
    case 4: return ms_pick<uint32_t, false>(flip, arr_size, in, out, iin, iout, num_arrays, st);
    case 8: return ms_pick<uint64_t, false>(flip, arr_size, in, out, iin, iout, num_arrays, st);
    case 4 | KT_FLOAT: return ms_pick<uint32_t, true>(flip, arr_size, in, out, iin, iout, num_arrays, st);
    case 8 | KT_FLOAT: return ms_pick<uint64_t, true>(flip, arr_size, in, out, iin, iout, num_arrays, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ts
