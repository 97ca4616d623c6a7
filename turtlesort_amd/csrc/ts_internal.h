// ts_internal.h -- declarations shared between the kernel files and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ts_engine.h"

namespace ts {

// Optional launch timing: begin/end are called on the host right before and
// right after every kernel launch of an engine run (kind = K_*), so that an
// implementation can record HIP events on the same stream.
struct LaunchTimer {
    virtual void begin(int kind, hipStream_t st) = 0;
    virtual void end(int kind, hipStream_t st) = 0;
    virtual ~LaunchTimer() {}
};

// level-synchronous engine (ts_levels.hip)
bool engine_supports(int splitters);   // splitter counts compiled in
uint32_t engine_tile(int kb);          // elements per partition tile
uint32_t engine_cap(int kb);           // largest segment the LDS sort takes
// Enqueues on st: (first) reset + roots, then levels [L0, L0+nlev), then
// (last) the LDS sorts.  kt: key type code (KT_FLOAT | key bytes).  dev_off:
// nseg+1 segment offsets (uint32) in device memory, or nullptr for the single
// segment [0, single_len).
hipError_t engine_levels(const Params &p, int kt, bool idx, int splitters, int cus, int wg_per_cu,
                         const uint32_t *dev_off, uint32_t nseg, uint32_t single_len, uint32_t L0, uint32_t nlev,
                         bool first, bool last, hipStream_t st, LaunchTimer *timer = nullptr);

// Clears the LDS-sort size-class counters before the LDS phase runs again.
hipError_t engine_clear_classes(const Params &p, hipStream_t st);

// MultiSort kernels (ts_multisort.hip): arrays of arr_size keys, in -> out.
// kt: key type code.  iout != nullptr: a 32-bit payload moves with every
// key, read from iin, or each key's position in its array when iin is
// nullptr.  Returns hipErrorInvalidValue when arr_size exceeds
// multisort_max_size(kt).  e0, e1 (or nullptr): events recorded at the
// start and the end of the launches, carried by the launches themselves.
hipError_t multisort_run(int kt, uint64_t flip, uint32_t arr_size, const void *in, void *out, const uint32_t *iin,
                         uint32_t *iout, uint64_t num_arrays, hipStream_t st, hipEvent_t e0 = nullptr,
                         hipEvent_t e1 = nullptr);
uint32_t multisort_max_size(int kb);

// Stable merge of two sorted runs (ts_merge.hip): out[0, na+nb) = a merged
// with b, a's elements first among equal keys; ia/ib/iout (all or none) move
// a 32-bit payload with every key.
hipError_t merge_run(int kt, uint64_t flip, const void *a, const uint32_t *ia, uint64_t na, const void *b,
                     const uint32_t *ib, uint64_t nb, void *out, uint32_t *iout, hipStream_t st);

// Checks on the current device that the lanes of one wave's LDS atomic to
// the same word receive their return values in ascending lane order (the
// stable ranks of the scatter, the radix LDS sort and the pivot rely on
// it); synchronous (ts_util.hip).
hipError_t check_lds_rank_order(bool *ok);

// synthetic inputs (ts_util.hip)
hipError_t fill_keys(int kb, void *keys, uint64_t n, uint64_t seed, int shift, hipStream_t st);
hipError_t fill_iota(uint32_t *idx, uint64_t n, hipStream_t st, uint64_t period = 0);   // period > 0: i mod period

}  // namespace ts
