// ts_levels.hip -- the key-width independent kernels of the sort engine
// (reset, roots) and the dispatch of engine_levels to the per-key-type
// instantiations (ts_levels_kb*.hip, ts_levels_kf*.hip).
#include "ts_levels_impl.h"

namespace ts {

__global__ void k_reset(Ctrl *c)
{
    uint32_t *w = (uint32_t *)c;
    for (uint32_t i = threadIdx.x; i < sizeof(Ctrl) / 4; i += blockDim.x) w[i] = 0;
}

// Clears the LDS-sort size-class counters, so that the LDS phase can run
// again over the whole task list after further partition levels.
__global__ void k_clear_classes(Ctrl *c)
{
    for (uint32_t i = threadIdx.x; i < (uint32_t)LDS_CLASSES; i += blockDim.x) {
        c->cls_cnt[i] = 0;
        c->cls_cur[i] = 0;
    }
}

hipError_t engine_clear_classes(const Params &p, hipStream_t st)
{
    hipLaunchKernelGGL(k_clear_classes, dim3(1), dim3(64), 0, st, p.ctrl);
    return hipGetLastError();
}

// Roots of a batch of segments [off[s], off[s+1]) (off == nullptr: the
// single segment [0, single_len)): large ones become level-0 segments,
// small ones LDS tasks.  Appends are aggregated per wave.  reset: the launch
// is one workgroup, which first clears the control block itself (instead of
// a k_reset launch before it).
__global__ void k_roots(Params p, const uint32_t *off, uint32_t nseg, uint32_t single_len, uint32_t reset)
{
    if (reset) {   // grid-uniform
        uint32_t *w = (uint32_t *)p.ctrl;
        for (uint32_t i = threadIdx.x; i < sizeof(Ctrl) / 4; i += blockDim.x) w[i] = 0;
        // every thread's clears are complete at device scope before any wave
        // of the workgroup appends through device-scope atomics on the block
        __threadfence();
        __syncthreads();
    }
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
    const bool big = len > p.cutoff;
    const bool small = len > 0 && !big;
    const uint64_t mb = ballot(big), ms = ballot(small);
    uint32_t sbase = 0, lbase = 0;
    if (lane() == 0) {
        const uint32_t nb = (uint32_t)__popcll(mb), ns = (uint32_t)__popcll(ms);
        if (nb) sbase = atomic_add(&p.ctrl->seg_alloc, nb);
        if (ns) lbase = lds_reserve(p, ns);
    }
    sbase = __shfl(sbase, 0, WAVE);
    lbase = __shfl(lbase, 0, WAVE);
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
        sg->mode = SPLIT_SAMPLE;
        sg->lo = 0;
        sg->hi = ~0ull;   // the pivot clamps it to the key width
    } else if (small && lbase != ~0u) {
        LdsTask *t = &p.ldst[lbase + mask_rank(ms)];
        t->begin = beg;
        t->len = len;
        t->depth = 0;
    }
}

bool engine_supports(int splitters) { return splitters == 3 || splitters == 7 || splitters == 15 || splitters == 31; }
uint32_t engine_tile(int kb) { return kb >= 8 ? 512u * 8u : 512u * 16u; }
uint32_t engine_cap(int kb) { return kb >= 8 ? 512u * 8u : 512u * 16u; }

hipError_t engine_levels(const Params &p, int kt, bool idx, int splitters, int cus, int wg_per_cu,
                         const uint32_t *dev_off, uint32_t nseg, uint32_t single_len, uint32_t L0, uint32_t nlev,
                         bool first, bool last, hipStream_t st, LaunchTimer *timer)
{
    levels_fn f = nullptr;
    switch (kt) {
    case 1: f = pick_kb1(idx, splitters); break;
    case 2: f = pick_kb2(idx, splitters); break;
    case 4: f = pick_kb4(idx, splitters); break;
    case 8: f = pick_kb8(idx, splitters); break;
    case 4 | KT_FLOAT: f = pick_kf4(idx, splitters); break;
    case 8 | KT_FLOAT: f = pick_kf8(idx, splitters); break;
    default: break;
    }
    if (!f) return hipErrorInvalidValue;
    return f(p, cus, wg_per_cu, dev_off, nseg, single_len, L0, nlev, first, last, st, timer);
}

}  // namespace ts
