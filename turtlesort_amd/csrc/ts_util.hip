// ts_util.hip -- synthetic input generation on the device, and the check
// of the LDS atomic ordering that the stable ranks rely on.
// splitmix64 is counter based: draw i (0-based) of a stream seeded with s is
// mix(s + (i+1) * 0x9e3779b97f4a7c15), so every element is generated
// independently and matches the sequential host generator draw for draw.
#include "ts_device.h"
#include "ts_internal.h"

namespace ts {

__device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1) * 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

template <typename U>
__global__ void fill_keys_kernel(U *keys, uint64_t n, uint64_t seed, int shift)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        keys[i] = (U)(splitmix_at(seed, i) >> shift);
}

// idx[i] = i, or i mod period (the position inside equal arrays of
// `period` elements) when period > 0
__global__ void fill_iota_kernel(uint32_t *idx, uint64_t n, uint64_t period)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        idx[i] = (uint32_t)(period ? i % period : i);
}

static uint32_t grid_for(uint64_t n)
{
    uint64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    return g == 0 ? 1u : (uint32_t)g;
}

hipError_t fill_keys(int kb, void *keys, uint64_t n, uint64_t seed, int shift, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const uint32_t g = grid_for(n);
    switch (kb) {
    case 1: hipLaunchKernelGGL(fill_keys_kernel<uint8_t>, dim3(g), dim3(256), 0, st, (uint8_t *)keys, n, seed, shift); break;
    case 2: hipLaunchKernelGGL(fill_keys_kernel<uint16_t>, dim3(g), dim3(256), 0, st, (uint16_t *)keys, n, seed, shift); break;
    case 4: hipLaunchKernelGGL(fill_keys_kernel<uint32_t>, dim3(g), dim3(256), 0, st, (uint32_t *)keys, n, seed, shift); break;
    case 8: hipLaunchKernelGGL(fill_keys_kernel<uint64_t>, dim3(g), dim3(256), 0, st, (uint64_t *)keys, n, seed, shift); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t fill_iota(uint32_t *idx, uint64_t n, hipStream_t st, uint64_t period)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_iota_kernel, dim3(grid_for(n)), dim3(256), 0, st, idx, n, period);
    return hipGetLastError();
}

// ---- LDS atomic lane order ------------------------------------------------------
// The scatter, the radix LDS sort and the pivot rank keys with one LDS
// atomic each (rank16_add), which is stable only if the lanes of one wave's
// atomic to the same word get their return values in ascending lane order.
// This kernel compares those ranks with ranks computed from wave64 match
// masks (the number of lower lanes with the same counter plus the
// counter's value before the instruction) over patterns of 1, 2, 4, 16 and
// 512 distinct counters, 16-bit halves included, and sets the flag on any
// difference.
__device__ uint32_t g_rank_order_bad;

__global__ void __launch_bounds__(512) k_rank_order_check()
{
    constexpr uint32_t NC = 512;
    __shared__ uint32_t words[8][NC / 2];   // atomic counters, two per word
    __shared__ uint32_t ref[8][NC];         // counts from match masks
    const uint32_t w = threadIdx.x / WAVE, l = lane();
    for (uint32_t i = l; i < NC; i += WAVE) {
        ref[w][i] = 0;
        if (i < NC / 2) words[w][i] = 0;
    }
    wave_sync();
    uint32_t bad = 0;
    for (uint32_t r = 0; r < 64; r++) {
        const uint32_t mask = (r % 5 == 0) ? 0u : (r % 5 == 1) ? 1u : (r % 5 == 2) ? 3u : (r % 5 == 3) ? 15u : NC - 1;
        const uint32_t c = (uint32_t)splitmix_at(0x51ab1e00ull + blockIdx.x, (uint64_t)r * 512 + threadIdx.x) & mask;
        const uint32_t got = rank16_add(words[w], c, 1u);
        const uint64_t peers = match_peers<9>(c, true);
        const uint32_t rk = mask_rank(peers);
        const uint32_t old = ref[w][c];
        wave_sync();
        if (rk == 0) ref[w][c] = old + (uint32_t)__popcll(peers);
        wave_sync();
        bad |= got != old + rk ? 1u : 0u;
    }
    if (bad) atomic_or(&g_rank_order_bad, 1u);
}

hipError_t check_lds_rank_order(bool *ok)
{
    *ok = false;
    const uint32_t zero = 0;
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(g_rank_order_bad), &zero, sizeof zero);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rank_order_check, dim3(256), dim3(512), 0, 0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    uint32_t bad = 1;
    if ((e = hipMemcpyFromSymbol(&bad, HIP_SYMBOL(g_rank_order_bad), sizeof bad)) != hipSuccess) return e;
    *ok = bad == 0;
    return hipSuccess;
}

}  // namespace ts
