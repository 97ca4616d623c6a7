// ts_util.hip -- synthetic input generation on the device.
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

}  // namespace ts
