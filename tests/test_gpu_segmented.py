"""Batched uneven segments in one engine run: each segment equal to
the oracle's sort of it (keys bit-exact, stable index)."""
import numpy as np
import pytest

from turtlesort_amd import TurtleSort, workloads as wl

from conftest import d2h, h2d

pytestmark = pytest.mark.gpu


def check_segments(orc, keys, offsets, ok, oi):
    ek, _ = orc.segmented_sort(keys, np.arange(keys.size, dtype=np.int32), offsets)
    np.testing.assert_array_equal(ok, ek)
    for s in range(len(offsets) - 1):
        a, b = offsets[s], offsets[s + 1]
        if b > a:
            np.testing.assert_array_equal(oi[a:b], a + np.argsort(keys[a:b], kind="stable"))


@pytest.mark.parametrize("dtype", [np.int64, np.int32, np.uint16, np.int8])
def test_segmented_small(orc, torch_dev, dtype):
    torch = torch_dev
    lens = wl.loguniform_lengths(17, 400, 1, 40000)
    lens[::7] = 0                               # empty segments too
    off = wl.segment_offsets(lens)
    n = int(off[-1])
    keys = wl.gen_keys(dtype, n, 23, 0)
    sorter = TurtleSort(dtype, n)
    tdt = {np.dtype(np.int64): torch.int64, np.dtype(np.int32): torch.int32,
           np.dtype(np.int8): torch.int8, np.dtype(np.uint16): torch.int16}[np.dtype(dtype)]
    kin = h2d(torch, keys.view(np.int16) if dtype == np.uint16 else keys)
    iin = torch.arange(n, dtype=torch.int32, device="cuda")
    kout = torch.empty_like(kin)
    iout = torch.empty_like(iin)
    sorter.segmented_sort_device(kin, kout, off, iin, iout)
    sorter.synchronize()
    ok = d2h(torch, kout)
    if dtype == np.uint16:
        ok = ok.view(np.uint16)
    check_segments(orc, keys, off, ok, d2h(torch, iout))
    sorter.close()


def test_segmented_config4_shape(torch_dev):
    """20,000 segments log-uniform in [16, 262144], int64 keys + int32 idx:
    per-segment sortedness, stability and permutation checked on the GPU."""
    torch = torch_dev
    lens = wl.loguniform_lengths(wl.SEED_C4_LENGTHS, wl.C4_SEGMENTS, wl.C4_LO, wl.C4_HI)
    off = wl.segment_offsets(lens)
    n = int(off[-1])
    sorter = TurtleSort(np.int64, n)
    from turtlesort_amd import fill_keys
    kin = torch.empty(n, dtype=torch.int64, device="cuda")
    fill_keys(kin, wl.SEED_C4_KEYS, wl.SHIFT_63BIT)
    iin = torch.arange(n, dtype=torch.int32, device="cuda")
    kout = torch.empty_like(kin)
    iout = torch.empty_like(iin)
    sorter.segmented_sort_device(kin, kout, off, iin, iout)
    st = sorter.synchronize()
    seg_id = torch.repeat_interleave(torch.arange(len(lens), device="cuda"), h2d(torch, lens))
    same = seg_id[1:] == seg_id[:-1]
    assert bool((~same | (kout[1:] >= kout[:-1])).all())
    assert bool((~same | (kout[1:] != kout[:-1]) | (iout[1:] > iout[:-1])).all())
    assert torch.equal(kin[iout.long()], kout)
    assert torch.equal(seg_id[iout.long()], seg_id)        # indices stay in their segment
    assert bool((torch.bincount(iout.long(), minlength=n) == 1).all())
    assert st["error"] == 0
    sorter.close()


def test_segmented_graph_replays(orc, torch_dev):
    """The hipGraph form replayed several times, with other device work and
    host synchronisation between replays; every replay equals the eager run
    and the oracle."""
    torch = torch_dev
    lens = wl.loguniform_lengths(31, 300, 16, 100_000)
    off = wl.segment_offsets(lens)
    n = int(off[-1])
    keys = wl.gen_keys(np.int64, n, 37, wl.SHIFT_63BIT)
    sorter = TurtleSort(np.int64, n)
    kin = h2d(torch, keys)
    iin = torch.arange(n, dtype=torch.int32, device="cuda")
    kout = torch.empty_like(kin)
    iout = torch.empty_like(iin)
    g = sorter.segmented_graph(kin, kout, off, iin, iout)   # captured on a fresh handle
    k_eager = torch.empty_like(kin)
    i_eager = torch.empty_like(iin)
    sorter.segmented_sort_device(kin, k_eager, off, iin, i_eager)
    sorter.synchronize()
    check_segments(orc, keys, off, d2h(torch, k_eager), d2h(torch, i_eager))
    for rep in range(4):
        print(f"case graph replay {rep}", flush=True)
        kout.fill_(-1)
        iout.fill_(-1)
        torch.cuda.synchronize()
        g.launch()
        st = sorter.synchronize()
        assert st["error"] == 0
        assert torch.equal(kout, k_eager) and torch.equal(iout, i_eager), f"replay {rep} differs"
        _ = (kin * 3).sum()                          # eager work between replays
        torch.cuda.synchronize()
    g.close()
    sorter.close()


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.uint8, np.int16])
def test_lds_size_classes(orc, torch_dev, dtype):
    """Segments at every LDS-sort size-class boundary (merge classes of
    64..512 keys, radix classes of 1..8 waves of 1024 or 512 slots), each
    sorted in LDS directly: keys equal the oracle's, the index is the stable
    order.  Two key ranges: full (several radix passes) and 8 values (one)."""
    torch = torch_dev
    lens = [1, 2, 63, 64, 65, 511, 512, 513, 1023, 1024, 1025, 1500, 2047, 2048, 2049, 3071, 3072, 3073,
            4095, 4096, 4097, 5000, 6143, 6144, 6145, 7000, 8191, 8192]
    if np.dtype(dtype).itemsize == 8:
        lens = [x for x in lens if x <= 4096]
    off = wl.segment_offsets(np.array(lens, dtype=np.int64))
    n = int(off[-1])
    sorter = TurtleSort(dtype, n)
    sorter.set_tuning(lds_cutoff=4096 if np.dtype(dtype).itemsize == 8 else 8192)
    for shift in (0, 61):
        print(f"case dtype={np.dtype(dtype)} shift={shift}", flush=True)
        keys = wl.gen_keys(dtype, n, 91 + shift, shift)
        kin = h2d(torch, keys.view(np.int16) if dtype == np.uint16 else keys)
        iin = torch.arange(n, dtype=torch.int32, device="cuda")
        kout = torch.empty_like(kin)
        iout = torch.empty_like(iin)
        sorter.segmented_sort_device(kin, kout, off, iin, iout)
        st = sorter.synchronize()
        assert st["error"] == 0 and st["segments"] == 0     # every segment went straight to the LDS sort
        check_segments(orc, keys, off, d2h(torch, kout), d2h(torch, iout))
    sorter.close()


def test_segmented_level_batches(monkeypatch, orc, torch_dev):
    """a segmented batch with the partition levels capped to one per
    read-back (TURTLE_LEVEL_BATCH=1): the LDS phase runs again after every
    extra level; the result equals the oracle's"""
    torch = torch_dev
    monkeypatch.setenv("TURTLE_LEVEL_BATCH", "1")
    lens = wl.loguniform_lengths(53, 60, 1, 400_000)
    off = wl.segment_offsets(lens)
    n = int(off[-1])
    keys = wl.gen_keys(np.int64, n, 61, wl.SHIFT_63BIT)
    sorter = TurtleSort(np.int64, n)
    sorter.set_tuning(splitters=7, lds_cutoff=1024)         # several levels
    kin = h2d(torch, keys)
    iin = torch.arange(n, dtype=torch.int32, device="cuda")
    kout = torch.full_like(kin, -1)
    iout = torch.full_like(iin, -1)
    sorter.segmented_sort_device(kin, kout, off, iin, iout)
    st = sorter.synchronize()
    assert st["error"] == 0 and st["max_depth"] >= 3
    check_segments(orc, keys, off, d2h(torch, kout), d2h(torch, iout))
    sorter.close()


@pytest.mark.parametrize("n2,n3", [(0, 5), (1, 5), (3, 3), (2, 4), (7, 1), (5, 0), (4, 9)])
def test_radix_mixed_steps(orc, torch_dev, n2, n3):
    """Radix-class tasks of 2 waves (1025..2048 keys) and 3 waves
    (2049..3072 keys) in one batch: the first min(ceil(n3/2), n2) steps
    of the 3-wave class also sort one 2-wave task each (an odd last step
    holds one 3-wave task and its 2-wave partner); the rest run in steps
    of their own class.  Plus a few 1-wave and 4-wave tasks, shuffled."""
    torch = torch_dev
    rng = np.random.default_rng(1000 + 10 * n2 + n3)
    lens = np.concatenate([rng.integers(1025, 2049, n2), rng.integers(2049, 3073, n3),
                           rng.integers(513, 1025, 3), rng.integers(3073, 4097, 2)]).astype(np.int64)
    rng.shuffle(lens)
    off = wl.segment_offsets(lens)
    n = int(off[-1])
    sorter = TurtleSort(np.int32, n)
    sorter.set_tuning(lds_cutoff=8192)
    for shift in (0, 40):
        print(f"case n2={n2} n3={n3} shift={shift}", flush=True)
        keys = wl.gen_keys(np.int32, n, 7 * n2 + n3 + shift, shift)
        kin = h2d(torch, keys)
        iin = torch.arange(n, dtype=torch.int32, device="cuda")
        kout = torch.empty_like(kin)
        iout = torch.empty_like(iin)
        sorter.segmented_sort_device(kin, kout, off, iin, iout)
        st = sorter.synchronize()
        assert st["error"] == 0 and st["segments"] == 0
        check_segments(orc, keys, off, d2h(torch, kout), d2h(torch, iout))
    sorter.close()


def test_autotune_segments(orc, torch_dev):
    """the segmented autotuner returns one of its candidates, and a batch
    sorted with it equals the oracle"""
    torch = torch_dev
    lens = wl.loguniform_lengths(29, 200, 16, 30000)
    off = wl.segment_offsets(lens)
    n = int(off[-1])
    sorter = TurtleSort(np.int64, n)
    t, ms = sorter.autotune_segments(off, with_index=True, reps=1)
    cap = 4096
    assert t["splitters"] in (7, 15, 31) and t["lds_cutoff"] in (cap // 4, cap // 2, cap)
    assert t["tiles_per_task"] in (1, 2) and ms > 0
    sorter.set_tuning(**t)
    keys = wl.gen_keys(np.int64, n, 3, 0)
    kin = h2d(torch, keys)
    iin = torch.arange(n, dtype=torch.int32, device="cuda")
    kout, iout = torch.empty_like(kin), torch.empty_like(iin)
    sorter.segmented_sort_device(kin, kout, off, iin, iout)
    sorter.synchronize()
    check_segments(orc, keys, off, d2h(torch, kout), d2h(torch, iout))
    sorter.close()
