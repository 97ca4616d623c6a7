"""CPU checks of the planted-sample input (tests/adversary.py): it really
defeats sample-quantile splitting, which is what the GPU test of the
range-split fallback (test_gpu_robust.py) relies on."""
import numpy as np

from adversary import planted_sample_input, sample_only_levels, sample_positions
from turtlesort_amd import workloads as wl


def test_sample_positions_match_the_engine_rule():
    # 128 runs of 8 consecutive keys, run r at r*st + (st-8)//2, st = len//128
    p = sample_positions(128 * 100)
    assert p.size == 1024 and np.all(np.diff(p) > 0)
    assert p[0] == 46 and p[7] == 53 and p[8] == 146 and p[-1] == 127 * 100 + 46 + 7
    assert p[-1] < 128 * 100


def test_planted_input_defeats_sample_pivots():
    n = 1 << 20
    keys = planted_sample_input(n, 80)
    # every planted level's 1024 keys hold one value; the rest are distinct
    vals, cnt = np.unique(keys, return_counts=True)
    assert np.all(cnt[:80] == 1024) and np.all(cnt[80:] == 1)
    # a sort that only splits at sample quantiles is still above the cutoff
    # after 65 levels (the engine's former level limit was 64) ...
    assert sample_only_levels(keys, 31, 1024, 65) is None
    # ... while uniform keys of the same size need 3
    assert sample_only_levels(wl.gen_keys(np.int32, n, 5, wl.SHIFT_31BIT), 31, 1024, 65) == 3


def test_planted_input_model_planted_value_leaves_each_level():
    # the model itself: on the planted input the largest segment after L
    # levels holds exactly n - 1024 L keys (one pivot bin of 1024 per level)
    n = 1 << 16
    keys = planted_sample_input(n, 10)
    for lev in (1, 3, 6):
        seg = keys
        for _ in range(lev):
            smp = np.sort(seg[sample_positions(seg.size)])
            assert smp[0] == smp[-1]          # the whole sample is one value
            seg = seg[seg > smp[0]]
        assert seg.size == n - 1024 * lev
