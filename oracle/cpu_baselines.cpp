// cpu_baselines.cpp -- host-core baselines timed beside the GPU sort.
//
// TEST INFRASTRUCTURE ONLY (loaded by tests/ and the cpu_baseline leg of
// bench.py).  Each baseline sorts a {key, idx} struct by key, the way the
// reference's benchmark program does (README.md:123-127, 176-181):
//   std::sort (1 thread), std::sort(std::execution::par_unseq, ...) and the
//   libstdc++ parallel-mode __gnu_parallel::sort on all cores (OpenMP);
//   for batches of arrays/segments, a per-segment std::sort loop, plain and
//   under omp parallel for.
// libstdc++ selects a TBB backend for par_unseq only when <tbb/tbb.h> exists;
// without it the PSTL serial backend runs, and ts_cpu_pstl_backend() says so.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <execution>
#include <parallel/algorithm>
#include <thread>
#include <vector>
#include <omp.h>
#include <sched.h>

namespace {

template <typename K>
struct Rec { K key; int32_t idx; };

template <typename K>
double time_sort(int which, const K* keys, const int32_t* idx, int64_t n, int threads,
                 K* out_keys, int32_t* out_idx)
{
    std::vector<Rec<K>> v((size_t)n);
    for (int64_t i = 0; i < n; i++) { v[i].key = keys[i]; v[i].idx = idx ? idx[i] : (int32_t)i; }
    auto cmp = [](const Rec<K>& a, const Rec<K>& b) { return a.key < b.key; };
    auto t0 = std::chrono::steady_clock::now();
    if (which == 0) {
        std::sort(v.begin(), v.end(), cmp);
    } else if (which == 1) {
        std::sort(std::execution::par_unseq, v.begin(), v.end(), cmp);
    } else if (which == 2) {
        omp_set_num_threads(threads > 0 ? threads : omp_get_num_procs());
        __gnu_parallel::sort(v.begin(), v.end(), cmp);
    } else {
        std::stable_sort(v.begin(), v.end(), cmp);
    }
    auto t1 = std::chrono::steady_clock::now();
    if (out_keys) for (int64_t i = 0; i < n; i++) out_keys[i] = v[i].key;
    if (out_idx) for (int64_t i = 0; i < n; i++) out_idx[i] = v[i].idx;
    return std::chrono::duration<double>(t1 - t0).count();
}

// Per-segment sorts of a batch (the CPU form of MultiSort and of the
// batched uneven segments): segment s is [off[s], off[s+1]).  With idx, a
// {key, idx} struct per element, else the keys alone.  par: the segments
// are spread over `threads` OpenMP threads (dynamic schedule).
template <typename K>
double time_segmented(bool par, const K* keys, const int32_t* idx, const int64_t* off, int64_t nseg, int threads)
{
    const int64_t n = off[nseg];
    std::vector<Rec<K>> v(idx ? (size_t)n : 0);
    std::vector<K> k(idx ? 0 : (size_t)n);
    if (idx) for (int64_t i = 0; i < n; i++) { v[i].key = keys[i]; v[i].idx = idx[i]; }
    else std::memcpy(k.data(), keys, (size_t)n * sizeof(K));
    auto cmp = [](const Rec<K>& a, const Rec<K>& b) { return a.key < b.key; };
    omp_set_num_threads(par ? (threads > 0 ? threads : omp_get_num_procs()) : 1);
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 16) if (par)
    for (int64_t s = 0; s < nseg; s++) {
        if (idx) std::sort(v.begin() + off[s], v.begin() + off[s + 1], cmp);
        else std::sort(k.begin() + off[s], k.begin() + off[s + 1]);
    }
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // namespace

extern "C" {

// par: 0 a loop of std::sort over the segments on one thread, 1 the same
// loop under omp parallel for.  Returns seconds spent sorting, < 0 on error.
double ts_cpu_segmented_sort(int par, int dtype, const void* keys, const int32_t* idx, const int64_t* off,
                             int64_t nseg, int threads)
{
    const bool p = par != 0;
    switch (dtype) {
    case 1: return time_segmented<int8_t>(p, (const int8_t*)keys, idx, off, nseg, threads);
    case 2: return time_segmented<uint8_t>(p, (const uint8_t*)keys, idx, off, nseg, threads);
    case 3: return time_segmented<int16_t>(p, (const int16_t*)keys, idx, off, nseg, threads);
    case 4: return time_segmented<uint16_t>(p, (const uint16_t*)keys, idx, off, nseg, threads);
    case 5: return time_segmented<int32_t>(p, (const int32_t*)keys, idx, off, nseg, threads);
    case 6: return time_segmented<uint32_t>(p, (const uint32_t*)keys, idx, off, nseg, threads);
    case 7: return time_segmented<int64_t>(p, (const int64_t*)keys, idx, off, nseg, threads);
    case 8: return time_segmented<uint64_t>(p, (const uint64_t*)keys, idx, off, nseg, threads);
    default: return -1.0;
    }
}

// which: 0 std::sort, 1 std::sort(par_unseq), 2 __gnu_parallel::sort, 3 std::stable_sort
// dtype codes: as the oracle (1..8).  Returns seconds spent sorting, < 0 on error.
double ts_cpu_sort(int which, int dtype, const void* keys, const int32_t* idx, int64_t n,
                   int threads, void* out_keys, int32_t* out_idx)
{
    switch (dtype) {
    case 1: return time_sort<int8_t>(which, (const int8_t*)keys, idx, n, threads, (int8_t*)out_keys, out_idx);
    case 2: return time_sort<uint8_t>(which, (const uint8_t*)keys, idx, n, threads, (uint8_t*)out_keys, out_idx);
    case 3: return time_sort<int16_t>(which, (const int16_t*)keys, idx, n, threads, (int16_t*)out_keys, out_idx);
    case 4: return time_sort<uint16_t>(which, (const uint16_t*)keys, idx, n, threads, (uint16_t*)out_keys, out_idx);
    case 5: return time_sort<int32_t>(which, (const int32_t*)keys, idx, n, threads, (int32_t*)out_keys, out_idx);
    case 6: return time_sort<uint32_t>(which, (const uint32_t*)keys, idx, n, threads, (uint32_t*)out_keys, out_idx);
    case 7: return time_sort<int64_t>(which, (const int64_t*)keys, idx, n, threads, (int64_t*)out_keys, out_idx);
    case 8: return time_sort<uint64_t>(which, (const uint64_t*)keys, idx, n, threads, (uint64_t*)out_keys, out_idx);
    default: return -1.0;
    }
}

// 0 = serial, 1 = TBB, 2 = OpenMP, -1 = unknown
int ts_cpu_pstl_backend(void)
{
#if defined(_PSTL_PAR_BACKEND_TBB)
    return 1;
#elif defined(_PSTL_PAR_BACKEND_OPENMP)
    return 2;
#elif defined(_PSTL_PAR_BACKEND_SERIAL)
    return 0;
#else
    return -1;
#endif
}

int ts_cpu_hw_threads(void) { return (int)std::thread::hardware_concurrency(); }

int ts_cpu_affinity_count(void)
{
    cpu_set_t s;
    if (sched_getaffinity(0, sizeof(s), &s) != 0) return -1;
    return CPU_COUNT(&s);
}

}  // extern "C"
