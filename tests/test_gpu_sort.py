"""Parity of the plain sort (StartSorting/Sync and the device-resident form)
with the CPU oracle.  Keys must be bit-identical to the oracle's; the GPU
path is stable, so the index output must equal a stable sort's (numpy
argsort kind="stable"), which also proves the (key, idx) multiset is kept."""
import numpy as np
import pytest

from turtlesort_amd import TurtleSort, workloads as wl

pytestmark = pytest.mark.gpu

DTYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64]
EDGE_SIZES = [0, 1, 2, 31, 32, 33, 127, 128, 129, 4095, 4096, 4097, 8191, 8192, 8193, 65537]


def make_keys(dtype, n, dist, seed):
    dt = np.dtype(dtype)
    bits = 8 * dt.itemsize
    if dist == "full":
        k = wl.gen_keys(dt, n, seed, 0)
    elif dist == "dup15":
        k = wl.gen_keys(dt, n, seed, 64 - min(15, bits - (1 if dt.kind == "i" else 0)))
    elif dist == "few16":
        k = wl.gen_keys(dt, n, seed, 60)
    elif dist == "asc":
        k = np.sort(wl.gen_keys(dt, n, seed, 0), kind="stable")
    elif dist == "desc":
        k = np.sort(wl.gen_keys(dt, n, seed, 0), kind="stable")[::-1].copy()
    elif dist == "equal":
        k = np.full(n, wl.gen_keys(dt, 1, seed, 0)[0], dtype=dt)
    elif dist == "extremes":
        info = np.iinfo(dt)
        k = np.where(wl.gen_keys(np.uint8, n, seed, 0) & 1, info.min, info.max).astype(dt)
    elif dist == "organ":
        h = wl.gen_keys(dt, (n + 1) // 2, seed, 0)
        h.sort()
        k = np.concatenate([h, h[::-1]])[:n].copy()
    else:
        raise ValueError(dist)
    return np.ascontiguousarray(k)


def check_against_oracle(orc, keys, out_keys, out_idx):
    ok, oi = orc.turtle_sort(keys, np.arange(keys.size, dtype=np.int32))
    assert out_keys.dtype == keys.dtype
    np.testing.assert_array_equal(out_keys, ok)              # bit-exact keys
    # the oracle's index order among equal keys is one valid schedule of the
    # reference; compare the (key, idx) multisets
    assert np.array_equal(np.sort(oi), np.arange(keys.size))
    np.testing.assert_array_equal(keys[oi], ok)
    if out_idx is not None:
        stable = np.argsort(keys, kind="stable").astype(np.int32)
        np.testing.assert_array_equal(out_idx, stable)         # stable => exact


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("dist", ["full", "few16", "asc", "desc", "equal", "extremes", "dup15", "organ"])
def test_start_sorting_edges(orc, torch_dev, dtype, dist):
    sorter = TurtleSort(dtype, 70000)
    for i, n in enumerate(EDGE_SIZES):
        print(f"case dtype={np.dtype(dtype)} dist={dist} n={n}", flush=True)
        keys = make_keys(dtype, n, dist, 1000 + i)
        k = keys.copy()
        idx = np.arange(n, dtype=np.int32)
        sorter.start_sorting(k, idx)
        sec = sorter.sync()
        assert sec >= 0
        check_against_oracle(orc, keys, k, idx)
    sorter.close()


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.uint8])
def test_keys_only(orc, torch_dev, dtype):
    sorter = TurtleSort(dtype, 300000)
    for n in [0, 1, 100, 8193, 300000]:
        keys = make_keys(dtype, n, "full", 77 + n)
        k = keys.copy()
        sorter.start_sorting(k)
        sorter.sync()
        ok, _ = orc.turtle_sort(keys)
        np.testing.assert_array_equal(k, ok)
    sorter.close()


def test_config1_heavy_duplicates(orc, torch_dev):
    """n = 1,048,576 int32 keys in [0, 32767], idx = iota, via StartSorting."""
    n = wl.C1_N
    keys = wl.gen_keys(np.int32, n, wl.SEED_C1, wl.SHIFT_15BIT)
    assert keys.min() >= 0 and keys.max() <= 32767
    sorter = TurtleSort(np.int32, n)
    k = keys.copy()
    idx = np.arange(n, dtype=np.int32)
    sorter.start_sorting(k, idx)
    sorter.sync()
    check_against_oracle(orc, keys, k, idx)
    sorter.close()


@pytest.mark.parametrize("dist", ["full", "few16", "asc", "desc"])
def test_medium_sizes(orc, torch_dev, dist):
    sorter = TurtleSort(np.int32, 2_000_000)
    for n in [200_001, 2_000_000]:
        keys = make_keys(np.int32, n, dist, 5 + n)
        k = keys.copy()
        idx = np.arange(n, dtype=np.int32)
        sorter.start_sorting(k, idx)
        sorter.sync()
        check_against_oracle(orc, keys, k, idx)
    sorter.close()


def test_device_form_repeat_and_inplace(torch_dev):
    """Same input twice gives identical output (determinism); in-place works."""
    torch = torch_dev
    n = 3_000_000
    sorter = TurtleSort(np.int32, n)
    kin = torch.empty(n, dtype=torch.int32, device="cuda")
    from turtlesort_amd import fill_keys, fill_iota
    fill_keys(kin, 99, 33)
    iin = torch.empty(n, dtype=torch.int32, device="cuda")
    fill_iota(iin)
    outs = []
    for _ in range(2):
        ko = torch.empty_like(kin)
        io = torch.empty_like(iin)
        sorter.sort_device(kin, ko, iin, io)
        sorter.synchronize()
        outs.append((ko, io))
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    ref_k, ref_i = torch.sort(kin.to(torch.int64), stable=True)
    assert torch.equal(outs[0][0].to(torch.int64), ref_k)
    assert torch.equal(outs[0][1].to(torch.int64), ref_i)
    k2 = kin.clone()
    i2 = iin.clone()
    sorter.sort_device(k2, k2, i2, i2)
    sorter.synchronize()
    assert torch.equal(k2, outs[0][0]) and torch.equal(i2, outs[0][1])
    sorter.close()


@pytest.mark.parametrize("n", [3_000_000, 8_000_000])
def test_device_form_repeats_exact(torch_dev, n):
    """Regression test for two fixed races.  (1) The outputs are prefilled on
    the null stream right before each sort: the sort must run after that fill
    (NULL stream = the null stream), so no position may keep the fill value.
    (2) Every run must give torch's stable order exactly: a block barrier
    that did not wait for LDS loads let the merge's index gather read values
    of the next merge round, which duplicated a few indices per run."""
    torch = torch_dev
    from turtlesort_amd import fill_keys, fill_iota
    sorter = TurtleSort(np.int32, n)
    kin = torch.empty(n, dtype=torch.int32, device="cuda")
    iin = torch.empty(n, dtype=torch.int32, device="cuda")
    fill_keys(kin, 5 + n, 33)
    fill_iota(iin)
    ref_k, ref_i = torch.sort(kin.to(torch.int64), stable=True)
    for rep in range(4):
        print(f"case n={n} rep={rep}", flush=True)
        ko = torch.full_like(kin, -7)
        io = torch.full_like(iin, -7)
        sorter.sort_device(kin, ko, iin, io)
        st = sorter.synchronize()
        assert st["error"] == 0
        assert torch.equal(ko.to(torch.int64), ref_k), f"keys differ on run {rep}"
        bad = int((io.to(torch.int64) != ref_i).sum())
        assert bad == 0, f"{bad} indices differ from the stable order on run {rep}"
    sorter.close()


@pytest.mark.parametrize("batch", [1, 2])
def test_level_batches_rerun_lds_phase(monkeypatch, torch_dev, batch):
    """TURTLE_LEVEL_BATCH caps the partition levels per batch: the host reads
    back after each batch, runs more levels and then the LDS phase again over
    all tasks.  The result equals torch's stable sort and the unbatched run."""
    torch = torch_dev
    from turtlesort_amd import fill_keys, fill_iota
    n = 6_000_000
    sorter = TurtleSort(np.int32, n)
    sorter.set_tuning(splitters=7)                        # 4 levels at 6M keys
    kin = torch.empty(n, dtype=torch.int32, device="cuda")
    iin = torch.empty(n, dtype=torch.int32, device="cuda")
    fill_keys(kin, 321, 33)
    fill_iota(iin)
    k1, i1 = torch.empty_like(kin), torch.empty_like(iin)
    sorter.sort_device(kin, k1, iin, i1)
    sorter.synchronize()
    monkeypatch.setenv("TURTLE_LEVEL_BATCH", str(batch))
    print(f"case batch={batch}", flush=True)
    k2, i2 = torch.full_like(kin, -1), torch.full_like(iin, -1)
    sorter.sort_device(kin, k2, iin, i2)
    st = sorter.synchronize()
    assert st["error"] == 0 and st["max_depth"] >= 3
    ref_k, ref_i = torch.sort(kin.to(torch.int64), stable=True)
    assert torch.equal(k2.to(torch.int64), ref_k) and torch.equal(i2.to(torch.int64), ref_i)
    assert torch.equal(k1, k2) and torch.equal(i1, i2)
    sorter.close()


@pytest.mark.parametrize("splitters", [3, 7, 15, 31])
def test_tuning_variants(orc, torch_dev, splitters):
    """Every splitter count and a few LDS cutoffs / tile granularities give
    the oracle's keys and the stable index order."""
    sorter = TurtleSort(np.int32, 600_000)
    for cutoff, tiles in [(8192, 1), (2048, 2), (1024, 4)]:
        sorter.set_tuning(splitters=splitters, lds_cutoff=cutoff, tiles_per_task=tiles)
        for dist in ["full", "dup15", "asc", "few16"]:
            n = 600_000
            print(f"case S={splitters} cutoff={cutoff} tiles={tiles} dist={dist} n={n}", flush=True)
            keys = make_keys(np.int32, n, dist, 11 + cutoff)
            k = keys.copy()
            idx = np.arange(n, dtype=np.int32)
            sorter.start_sorting(k, idx)
            sorter.sync()
            check_against_oracle(orc, keys, k, idx)
    sorter.close()


def test_launch_timing(torch_dev):
    """per-kernel launch timing: every kind when on, only the selected kinds
    with a mask; the sums match the launches of the runs"""
    torch = torch_dev
    from turtlesort_amd import fill_keys, fill_iota
    n = 4_000_000
    sorter = TurtleSort(np.int32, n)
    kin = torch.empty(n, dtype=torch.int32, device="cuda")
    iin = torch.empty(n, dtype=torch.int32, device="cuda")
    fill_keys(kin, 17, 33)
    fill_iota(iin)
    kout, iout = torch.empty_like(kin), torch.empty_like(iin)
    sorter.set_profiling(True)
    for _ in range(2):
        sorter.sort_device(kin, kout, iin, iout)
    st = sorter.synchronize()
    prof = sorter.get_profile()
    per_run = prof["scatter"]["launches"] // 2              # levels launched per run
    assert per_run >= st["max_depth"] >= 1 and prof["scatter"]["launches"] == 2 * per_run
    assert prof["count"]["launches"] == 2 * per_run and prof["pivot"]["launches"] == 2 * per_run
    assert prof["lds"]["launches"] == 4 and all(prof[k]["ms"] > 0 for k in ("scatter", "count", "lds"))
    sorter.set_profiling(True, kinds=["scatter"])
    sorter.sort_device(kin, kout, iin, iout)
    sorter.synchronize()
    prof = sorter.get_profile()
    assert prof["scatter"]["launches"] == per_run and prof["scatter"]["ms"] > 0
    assert all(v["launches"] == 0 for k, v in prof.items() if k != "scatter")
    sorter.set_profiling(False)
    sorter.close()
