"""Robustness of the engine on the GPU: termination on an input built
against its pivot sampling, growth of the work tables after an overflow
(forced with TURTLE_SEGCAP), graphs that own their tables, a graph captured
with too few levels, asynchronous entry points, and the 64M int32+idx
workload at full size."""
import time

import numpy as np
import pytest

from adversary import planted_sample_input
from turtlesort_amd import SortServer, TurtleError, TurtleSort, fill_iota, fill_keys, workloads as wl
from turtlesort_amd import _lib

from conftest import d2h, h2d

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"case {msg}", flush=True)


def stable_idx(keys):
    return np.argsort(keys, kind="stable").astype(np.int32)


@pytest.mark.parametrize("dtype,S", [(np.int32, 31), (np.int64, 31), (np.int32, 3)])
def test_planted_sample_input_terminates(orc, torch_dev, dtype, S):
    """Every level's sample is one planted value: sample quantiles alone
    would need ~n/1024 levels.  The child that keeps more than half of its
    parent is split by key range next, so the sort ends in a few levels."""
    n = 1 << 22 if S == 31 else 1 << 20
    case(f"planted dtype={np.dtype(dtype)} S={S} n={n}")
    keys = planted_sample_input(n, 200, dtype=dtype)
    sorter = TurtleSort(dtype, n)
    sorter.set_tuning(splitters=S, lds_cutoff=1024)
    k = keys.copy()
    idx = np.arange(n, dtype=np.int32)
    sorter.start_sorting(k, idx)
    sorter.sync()
    ok, _ = orc.turtle_sort(keys)
    np.testing.assert_array_equal(k, ok)
    np.testing.assert_array_equal(idx, stable_idx(keys))
    # the device form reports how deep it went
    torch = torch_dev
    kin = h2d(torch, keys)
    kout = torch.empty_like(kin)
    sorter.sort_device(kin, kout)
    st = sorter.synchronize()
    assert st["error"] == 0
    # the depth bound of the range split: log2(n / cutoff) halvings plus two
    # levels per range step of log2(S+1) key bits (32-bit keys: 26 at S = 31,
    # 42 at S = 3; 64-bit keys at S = 31: 38); the sample alone would need
    # ~n/1024 levels.  A CPU model of the levels gives 11 for the int32 case.
    bits = 8 * np.dtype(dtype).itemsize
    steps = -(-bits // int(np.log2(S + 1)))
    bound = int(np.log2(n // 1024)) + 2 * steps
    print(f"max_depth {st['max_depth']} (bound {bound})", flush=True)
    assert st["max_depth"] <= bound, st
    np.testing.assert_array_equal(d2h(torch, kout), ok)
    sorter.close()


def test_segcap_overflow_grows_and_reruns(orc, torch_dev, monkeypatch):
    """A 4-segment table overflows at the first partition level; the level
    check grows the tables and sorts again from the untouched input, for
    StartSorting/Sync and for an in-place device call (whose input survives
    in the handle's input buffer)."""
    torch = torch_dev
    n = 2_000_000
    monkeypatch.setenv("TURTLE_SEGCAP", "4")
    keys = wl.gen_keys(np.int32, n, 404, wl.SHIFT_31BIT)
    ok, _ = orc.turtle_sort(keys)
    sorter = TurtleSort(np.int32, n)
    case("segcap=4 StartSorting")
    k = keys.copy()
    idx = np.arange(n, dtype=np.int32)
    sorter.start_sorting(k, idx)
    sorter.sync()
    np.testing.assert_array_equal(k, ok)
    np.testing.assert_array_equal(idx, stable_idx(keys))
    sorter.close()
    case("segcap=4 in-place device form")
    sorter = TurtleSort(np.int32, n)
    kd = h2d(torch, keys)
    idd = torch.arange(n, dtype=torch.int32, device="cuda")
    sorter.sort_device(kd, kd, idd, idd)
    st = sorter.synchronize()
    assert st["error"] == 0
    np.testing.assert_array_equal(d2h(torch, kd), ok)
    np.testing.assert_array_equal(d2h(torch, idd), stable_idx(keys))
    sorter.close()


def test_segcap_overflow_in_server(orc, torch_dev, monkeypatch):
    """The request loop's slots overflow their first tables the same way."""
    monkeypatch.setenv("TURTLE_SEGCAP", "4")
    case("segcap=4 server")
    srv = SortServer(pool_bytes=4 << 30, slot_keys=1 << 21, slots_per_dtype=1, dtypes=[np.int32])
    reqs = []
    for i in range(4):
        keys = wl.gen_keys(np.int32, 1_500_000 + i, 900 + i, wl.SHIFT_31BIT)
        k = keys.copy()
        idx = np.arange(k.size, dtype=np.int32)
        reqs.append((keys, k, idx, srv.submit(k, idx)))
    for keys, k, idx, t in reqs:
        srv.wait(t)
        ok, _ = orc.turtle_sort(keys)
        np.testing.assert_array_equal(k, ok)
        np.testing.assert_array_equal(idx, stable_idx(keys))
    srv.close()


def _segments(torch, n_seg, lo, hi, seed):
    lens = wl.loguniform_lengths(seed, n_seg, lo, hi)
    off = wl.segment_offsets(lens)
    n = int(off[-1])
    keys = wl.gen_keys(np.int64, n, seed + 1, wl.SHIFT_63BIT)
    return keys, off, h2d(torch, keys)


def _check_segments(orc, keys, off, kout, iout):
    ek, _ = orc.segmented_sort(keys, np.arange(keys.size, dtype=np.int32), off)
    np.testing.assert_array_equal(kout, ek)
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        np.testing.assert_array_equal(iout[a:b], a + stable_idx(keys[a:b]))


def test_graph_tables_survive_handle_growth(orc, torch_dev, monkeypatch):
    """A graph owns its tables: an eager run that grows the handle's tables
    (forced by a 4-segment first sizing) leaves the graph's replays exact."""
    torch = torch_dev
    keys, off, kin = _segments(torch, 300, 16, 200_000, 71)
    n = keys.size
    monkeypatch.setenv("TURTLE_SEGCAP", "4")
    sorter = TurtleSort(np.int64, n)           # the handle's tables: 4 segments
    monkeypatch.delenv("TURTLE_SEGCAP")
    iin = torch.arange(n, dtype=torch.int32, device="cuda")
    kout, iout = torch.empty_like(kin), torch.empty_like(iin)
    case("graph before growth")
    g = sorter.segmented_graph(kin, kout, off, iin, iout)
    g.launch()
    assert sorter.synchronize()["error"] == 0
    _check_segments(orc, keys, off, d2h(torch, kout), d2h(torch, iout))
    case("eager run that grows the handle's tables")
    k2, i2 = torch.empty_like(kin), torch.empty_like(iin)
    sorter.segmented_sort_device(kin, k2, off, iin, i2)
    assert sorter.synchronize()["error"] == 0
    _check_segments(orc, keys, off, d2h(torch, k2), d2h(torch, i2))
    case("graph replays after the growth")
    for _ in range(3):
        kout.fill_(0)
        iout.fill_(-1)
        torch.cuda.synchronize()
        g.launch()
        assert sorter.synchronize()["error"] == 0
        assert torch.equal(kout, k2) and torch.equal(iout, i2)
    g.close()
    sorter.close()


def test_graph_depth_record_after_deeper_eager_run(orc, torch_dev):
    """turtle_synchronize after a replay reads the graph's own last level,
    not the level count of an eager run in between."""
    torch = torch_dev
    keys, off, kin = _segments(torch, 64, 9000, 30000, 81)     # one partition level
    n_big = 1 << 22
    sorter = TurtleSort(np.int64, max(n_big, keys.size))
    iin = torch.arange(keys.size, dtype=torch.int32, device="cuda")
    kout, iout = torch.empty_like(kin), torch.empty_like(iin)
    g = sorter.segmented_graph(kin, kout, off, iin, iout)
    case("eager deep run between replays")
    big = torch.empty(n_big, dtype=torch.int64, device="cuda")
    fill_keys(big, 5, wl.SHIFT_63BIT)
    bo = torch.empty_like(big)
    sorter.sort_device(big, bo)
    st = sorter.synchronize()
    assert st["error"] == 0 and st["max_depth"] >= 2
    g.launch()
    st = sorter.synchronize()
    assert st["error"] == 0
    _check_segments(orc, keys, off, d2h(torch, kout), d2h(torch, iout))
    g.close()
    sorter.close()


def test_graph_with_too_few_levels_reports_overflow(torch_dev, monkeypatch):
    """A graph captured with one partition level over segments that need
    two reports TURTLE_ERR_QUEUE_OVERFLOW instead of returning quietly."""
    torch = torch_dev
    keys, off, kin = _segments(torch, 8, 300_000, 400_000, 91)   # two levels at S=31, cutoff 4096
    sorter = TurtleSort(np.int64, keys.size)
    iin = torch.arange(keys.size, dtype=torch.int32, device="cuda")
    kout, iout = torch.empty_like(kin), torch.empty_like(iin)
    monkeypatch.setenv("TURTLE_GRAPH_LEVELS", "1")
    g = sorter.segmented_graph(kin, kout, off, iin, iout)
    monkeypatch.delenv("TURTLE_GRAPH_LEVELS")
    case("one-level graph over two-level segments")
    for _ in range(2):
        g.launch()
        with pytest.raises(TurtleError) as ei:
            sorter.synchronize()
        assert ei.value.status == _lib.ERR_QUEUE_OVERFLOW
    g.close()
    # the eager form finishes the same segments
    sorter.segmented_sort_device(kin, kout, off, iin, iout)
    assert sorter.synchronize()["error"] == 0
    sorter.close()


def test_destroy_refused_while_graph_alive(torch_dev):
    torch = torch_dev
    keys, off, kin = _segments(torch, 4, 100, 5000, 13)
    sorter = TurtleSort(np.int64, keys.size)
    kout = torch.empty_like(kin)
    g = sorter.segmented_graph(kin, kout, off)
    lib = _lib.load()
    assert lib.turtle_destroy(sorter._h) == _lib.ERR_STATE
    g.launch()
    sorter.synchronize()
    sorter.close()          # closes the graph first, then the handle
    assert g._g is None and sorter._h is None


def test_device_sort_returns_before_the_device_runs(torch_dev):
    """turtle_sort_device only queues: behind a 0.2 s spin on the same
    stream it returns at once (it used to wait for its partition levels)."""
    torch = torch_dev
    if not hasattr(torch.cuda, "_sleep"):
        pytest.skip("torch.cuda._sleep not available")
    n = 1 << 24
    sorter = TurtleSort(np.int32, n)
    kin = torch.empty(n, dtype=torch.int32, device="cuda")
    fill_keys(kin, 3, wl.SHIFT_31BIT)
    iin = torch.empty_like(kin)
    fill_iota(iin)
    kout, iout = torch.empty_like(kin), torch.empty_like(iin)
    sorter.sort_device(kin, kout, iin, iout)
    sorter.synchronize()
    case("sort_device behind a spin kernel")
    torch.cuda._sleep(400_000_000)          # ~0.2 s at ~2 GHz
    t0 = time.perf_counter()
    sorter.sort_device(kin, kout, iin, iout)
    t_call = time.perf_counter() - t0
    assert sorter.query()                    # still running
    st = sorter.synchronize()
    t_all = time.perf_counter() - t0
    assert st["error"] == 0
    assert t_call < 0.05 and t_all > 0.1, (t_call, t_all)
    ref_k, ref_i = torch.sort(kin.to(torch.int64), stable=True)
    assert torch.equal(kout.to(torch.int64), ref_k) and torch.equal(iout.to(torch.int64), ref_i)
    sorter.close()


def test_start_sorting_returns_before_the_sort_ends(torch_dev):
    """StartSorting copies the input and queues the sort; the sort's device
    time lands in Sync."""
    n = wl.C2_N
    keys = wl.gen_keys(np.int32, n, wl.SEED_C2_UNIFORM, wl.SHIFT_31BIT)
    sorter = TurtleSort(np.int32, n)
    k = keys.copy()
    idx = np.arange(n, dtype=np.int32)
    case("StartSorting 64M")
    sorter.start_sorting(k, idx)
    busy = sorter.query()
    sorter.sync()
    assert busy
    assert np.all(k[1:] >= k[:-1]) and np.array_equal(keys[idx], k)
    sorter.close()


def test_full_size_64M_int32_idx(torch_dev):
    """The benchmark workload at full size (3 levels at S=31 with the 2^26
    tuning entry): sorted, stable, a permutation, keys_out == keys_in[idx]."""
    torch = torch_dev
    n = wl.C2_N
    sorter = TurtleSort(np.int32, n)
    kin = torch.empty(n, dtype=torch.int32, device="cuda")
    fill_keys(kin, wl.SEED_C2_UNIFORM, wl.SHIFT_31BIT)
    iin = torch.empty_like(kin)
    fill_iota(iin)
    kout, iout = torch.empty_like(kin), torch.empty_like(iin)
    case("64M int32+idx uniform")
    sorter.sort_device(kin, kout, iin, iout)
    st = sorter.synchronize()
    assert st["error"] == 0 and st["max_depth"] == 3
    k = kout.to(torch.int64)
    assert bool((k[1:] >= k[:-1]).all())
    assert bool(((k[1:] != k[:-1]) | (iout[1:] > iout[:-1])).all())
    assert bool(torch.equal(kin[iout.long()], kout))
    assert bool((torch.bincount(iout.long(), minlength=n) == 1).all())
    sorter.close()


def test_lds_atomic_lane_order(torch_dev):
    """The device property the stable ranks rely on (ranks from one LDS
    atomic per key equal ranks from wave64 match masks)."""
    assert _lib.load().turtle_check_device() == _lib.OK
