"""Inputs built against the sort engine's pivot sampling (CPU, numpy).

The engine's pivot samples 1024 keys of a segment of `len` keys as 128 runs
of 8 consecutive keys, run r starting at r*st + (st - 8)//2 with
st = len // 128 (k_pivot in turtlesort_amd/csrc/ts_levels_impl.h), and takes
its splitters at the sample's quantiles.  The partition scatter is stable,
so a child segment holds its parent's keys in their parent order.

planted_sample_input(n, levels) writes the value L at the sample positions
of the level-L segment and distinct larger values everywhere else: at every
level the whole sample is one value, every splitter equals it, only those
1024 keys leave (as pivot-equal keys) and a single child keeps all the rest.
A sort that only ever splits at sample quantiles needs about n/1024 levels
for it.
"""
import numpy as np

SAMPLE, RUN = 1024, 8


def sample_positions(length):
    """indices (into a segment of `length` keys) of the engine's pivot sample"""
    nrun = SAMPLE // RUN
    st = length // nrun
    starts = np.arange(nrun, dtype=np.int64) * st + (st - RUN) // 2
    return (starts[:, None] + np.arange(RUN, dtype=np.int64)[None, :]).ravel()


def planted_sample_input(n, levels, dtype=np.int32, seed=12345):
    """keys of length n that defeat sample-quantile pivots for `levels` levels"""
    keys = np.empty(n, dtype=np.int64)
    rest = np.arange(n, dtype=np.int64)        # positions of the level's single segment, in order
    for lev in range(levels):
        if rest.size < 2 * SAMPLE:
            break
        pos = sample_positions(rest.size)
        keys[rest[pos]] = lev
        keep = np.ones(rest.size, dtype=bool)
        keep[pos] = False
        rest = rest[keep]
    # the keys never sampled: distinct values above every planted one
    rng = np.random.default_rng(seed)
    hi = np.iinfo(dtype).max
    vals = levels + 1 + rng.choice(hi - levels - 1, size=rest.size, replace=False)
    keys[rest] = vals
    return keys.astype(dtype)


def sample_only_levels(keys, S, cutoff, max_levels):
    """Model of a sort that always splits at sample quantiles: follows the
    largest segment level by level (splitters = the sample's keys at ranks
    (j+1)*1024/(S+1), bins strictly between splitters) and returns the
    number of levels after which it is at or below `cutoff` keys, or None
    if that takes more than max_levels."""
    seg = np.asarray(keys)
    for lev in range(max_levels):
        if seg.size <= cutoff:
            return lev
        smp = np.sort(seg[sample_positions(seg.size)])
        split = smp[((np.arange(S) + 1) * SAMPLE) // (S + 1)]
        lt = np.searchsorted(split, seg, side="left")     # splitters < key
        le = np.searchsorted(split, seg, side="right")    # splitters <= key
        free = le == lt                                   # not equal to any splitter
        bins = lt[free]
        if bins.size == 0:
            return lev + 1
        big = np.bincount(bins).argmax()
        seg = seg[free][bins == big]                      # stable: keeps the parent order
    return None
