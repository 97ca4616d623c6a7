"""Seeded random sweep over key type, size, distribution, index and entry
point (host StartSorting/Sync, device sort, segmented batch, MultiSort with
and without an index), every case compared with the oracle: keys bit-exact,
the index the stable order of the oracle's key map."""
import numpy as np
import pytest

from turtlesort_amd import TurtleSort, workloads as wl

from conftest import d2h, h2d

pytestmark = pytest.mark.gpu

TYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64]


def draw_keys(rng, dt, n, seed):
    kind = int(rng.integers(0, 5))
    bits = 8 * np.dtype(dt).itemsize
    if kind == 0:
        return wl.gen_keys(dt, n, seed, 0), "full"
    if kind == 1:
        return wl.gen_keys(dt, n, seed, 64 - int(rng.integers(1, min(bits, 20)))), "narrow"
    if kind == 2:
        k = np.sort(wl.gen_keys(dt, n, seed, 0), kind="stable")
        return (k if rng.integers(0, 2) else k[::-1].copy()), "monotone"
    if kind == 3:
        v = wl.gen_keys(dt, 3, seed, 0)
        return v[rng.integers(0, 3, n)], "three values"
    k = wl.gen_keys(dt, n, seed, 0)
    k[rng.integers(0, max(n, 1), n // 3)] = k[0] if n else k
    return k, "one value a third of the time"


def uview(a):
    return a.view(np.dtype(f"u{a.dtype.itemsize}"))


def sview(a):
    """signed integer view of the same bits (every torch device op takes it)"""
    return a.view(np.dtype(f"i{a.dtype.itemsize}"))


@pytest.mark.parametrize("seed", range(6))
def test_random_sweep(orc, torch_dev, seed):
    torch = torch_dev
    rng = np.random.default_rng(1000 + seed)
    for case in range(12):
        dt = np.dtype(TYPES[int(rng.integers(0, len(TYPES)))])
        n = int(wl.loguniform_lengths(seed * 100 + case, 1, 1, 3_000_000)[0])
        keys, dist = draw_keys(rng, dt, n, seed * 1000 + case)
        form = int(rng.integers(0, 4))
        print(f"case seed={seed} case={case} dtype={dt} n={n} dist={dist} form={form}", flush=True)
        order = np.argsort(orc.order(keys), kind="stable").astype(np.int32)
        ek, _ = orc.turtle_sort(keys)
        sorter = TurtleSort(dt, max(n, 1))
        if form == 0:                                 # host StartSorting/Sync with an index
            k = keys.copy()
            x = np.arange(n, dtype=np.int32)
            sorter.start_sorting(k, x)
            sorter.sync()
            np.testing.assert_array_equal(uview(k), uview(ek))
            np.testing.assert_array_equal(x, order)
        elif form == 1:                               # device sort, keys only
            kin = h2d(torch, sview(keys))
            kout = torch.empty_like(kin)
            sorter.sort_device(kin, kout)
            assert sorter.synchronize()["error"] == 0
            np.testing.assert_array_equal(uview(d2h(torch, kout)), uview(ek))
        elif form == 2:                               # segmented batch with an index
            lens = wl.loguniform_lengths(seed * 7 + case, int(rng.integers(1, 40)), 1, max(n, 1))
            off = np.minimum(wl.segment_offsets(lens), n)
            off[-1] = n
            kin = h2d(torch, sview(keys))
            kout = torch.empty_like(kin)
            iin = torch.arange(n, dtype=torch.int32, device="cuda")
            iout = torch.empty_like(iin)
            sorter.segmented_sort_device(kin, kout, off, iin, iout)
            assert sorter.synchronize()["error"] == 0
            sk, _ = orc.segmented_sort(keys, np.arange(n, dtype=np.int32), off)
            np.testing.assert_array_equal(uview(d2h(torch, kout)), uview(sk))
            oi = d2h(torch, iout)
            for a, b in zip(off[:-1], off[1:]):
                if b > a:
                    np.testing.assert_array_equal(oi[a:b], a + np.argsort(orc.order(keys[a:b]), kind="stable"))
        else:                                         # MultiSort with positions
            arr = int(rng.integers(1, 10000))
            m = n - n % arr
            d = keys[:m].copy()
            pos = sorter.multi_sort_index(arr, d)
            if m:
                perm = np.argsort(orc.order(keys[:m]).reshape(-1, arr), axis=1, kind="stable")
                exp = np.take_along_axis(keys[:m].reshape(-1, arr), perm, axis=1).reshape(-1)
                np.testing.assert_array_equal(uview(d), uview(exp))
                np.testing.assert_array_equal(pos.reshape(-1, arr), perm)
        sorter.close()
