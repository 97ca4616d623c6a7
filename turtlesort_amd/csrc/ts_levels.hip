// ts_levels.hip -- the key-width independent kernels of the sort engine
// (reset, roots, level opening, LDS-task planning) and the dispatch of
// engine_levels to the per-width instantiations (ts_levels_kb*.hip).
#include "ts_levels_impl.h"

namespace ts {

__global__ void k_reset(Ctrl *c)
{
    uint32_t *w = (uint32_t *)c;
    for (uint32_t i = threadIdx.x; i < sizeof(Ctrl) / 4; i += blockDim.x) w[i] = 0;
}

// Roots of a batch of segments [off[s], off[s+1]) (off == nullptr: the
// single segment [0, single_len)): large ones become level-0 segments,
// small ones LDS tasks.  Appends are aggregated per wave.
__global__ void k_roots(Params p, const uint32_t *off, uint32_t nseg, uint32_t single_len)
{
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
    } else if (small && lbase != ~0u) {
        LdsTask *t = &p.ldst[lbase + mask_rank(ms)];
        t->begin = beg;
        t->len = len;
        t->depth = 0;
    }
}

// opens level L: segments [sb, se), no tiles yet
__global__ void k_level(Params p, uint32_t L)
{
    if (threadIdx.x == 0) {
        const uint32_t sb = L == 0 ? 0u : p.levels[L - 1].se;
        uint32_t se = p.ctrl->seg_alloc;
        if (se > p.segcap) se = p.segcap;   // overflow already flagged
        p.levels[L].sb = sb;
        p.levels[L].se = se;
        p.levels[L].tiles = 0;
        p.ctrl->tile_alloc = 0;
        if (se > sb && L + 1 > p.ctrl->max_depth) p.ctrl->max_depth = L + 1;
    }
}

// LDS-sort plan: first slot and first workgroup step of every size class
// (class c < crad: cap / 2^c tasks per workgroup step; c >= crad: tasks of
// c - crad + 1 whole waves, nw / (c - crad + 1) per step)
__global__ void k_ldsplan(Params p, uint32_t crad, uint32_t nw, uint32_t cap)
{
    if (threadIdx.x == 0) {
        Ctrl *c = p.ctrl;
        uint32_t off = 0, wg = 0;
        for (uint32_t k = 0; k < (uint32_t)LDS_CLASSES; k++) {
            c->cls_off[k] = off;
            c->cls_wg[k] = wg;
            c->cls_cur[k] = 0;
            const uint32_t n = c->cls_cnt[k];
            off += n;
            uint32_t P = k < crad ? cap >> k : (k - crad + 1 <= nw ? nw / (k - crad + 1) : 1u);
            if (P == 0) P = 1;
            if (n) wg += (n + P - 1) / P;
        }
        c->cls_wg[LDS_CLASSES] = wg;
    }
}

bool engine_supports(int splitters) { return splitters == 3 || splitters == 7 || splitters == 15 || splitters == 31; }
uint32_t engine_tile(int kb) { return kb >= 8 ? 512u * 8u : 512u * 16u; }
uint32_t engine_cap(int kb) { return kb >= 8 ? 512u * 8u : 512u * 16u; }

hipError_t engine_levels(const Params &p, int kb, bool idx, int splitters, int cus, int wg_per_cu,
                         const uint32_t *dev_off, uint32_t nseg, uint32_t single_len, uint32_t L0, uint32_t nlev,
                         bool first, bool last, hipStream_t st, LaunchTimer *timer)
{
    levels_fn f = nullptr;
    switch (kb) {
    case 1: f = pick_kb1(idx, splitters); break;
    case 2: f = pick_kb2(idx, splitters); break;
    case 4: f = pick_kb4(idx, splitters); break;
    case 8: f = pick_kb8(idx, splitters); break;
    default: break;
    }
    if (!f) return hipErrorInvalidValue;
    return f(p, cus, wg_per_cu, dev_off, nseg, single_len, L0, nlev, first, last, st, timer);
}

}  // namespace ts
