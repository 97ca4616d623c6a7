"""MultiSort parity: every array equal to the oracle's heap-sort result
(turtle-multi-sort.cuh:15-182); the full 65,536 x 1024 configuration through
per-row comparison with torch.sort."""
import numpy as np
import pytest

from turtlesort_amd import TurtleSort, fill_keys, workloads as wl

from conftest import d2h, h2d

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64])
def test_multisort_host_api(orc, torch_dev, dtype):
    sorter = TurtleSort(dtype, 1 << 20)
    for arr, num in [(1, 7), (2, 33), (10, 640), (32, 100), (300, 21), (512, 40), (1000, 37), (1024, 256),
                     (1500, 11), (2038, 50), (2048, 13), (4096, 9), (8192, 3), (8193, 2), (20000, 3)]:
        print(f"case dtype={np.dtype(dtype)} arr={arr} num={num}", flush=True)
        d = wl.gen_keys(dtype, arr * num, arr * 7 + num, 0)
        expect = orc.multisort(d, arr)
        got = d.copy()
        sec = sorter.multi_sort(arr, got, block_size=32)
        assert sec >= 0
        np.testing.assert_array_equal(got, expect)
    sorter.close()


def test_multisort_config3_full(torch_dev):
    """65,536 arrays x 1024 int32 uniform [0, 2^31), one launch"""
    torch = torch_dev
    num, arr = wl.C3_ARRAYS, wl.C3_SIZE
    sorter = TurtleSort(np.int32, num * arr, track_index=False)
    kin = torch.empty(num * arr, dtype=torch.int32, device="cuda")
    fill_keys(kin, wl.SEED_C3, wl.SHIFT_31BIT)
    kout = torch.empty_like(kin)
    sorter.multi_sort_device(arr, kin, kout)
    sorter.synchronize()
    ref, _ = torch.sort(kin.view(num, arr), dim=1)
    assert torch.equal(kout.view(num, arr), ref)
    # in place
    k2 = kin.clone()
    sorter.multi_sort_device(arr, k2, k2)
    sorter.synchronize()
    assert torch.equal(k2, kout)
    sorter.close()


def test_multisort_rejects_bad_sizes(torch_dev):
    from turtlesort_amd import TurtleError
    sorter = TurtleSort(np.int32, 1000)
    d = np.zeros(2000, dtype=np.int32)
    with pytest.raises(TurtleError):
        sorter.multi_sort(1000, d)          # 2 x 1000 > max 1000
    with pytest.raises(TypeError):
        sorter.multi_sort(10, np.zeros(100, dtype=np.int64))
    sorter.close()


# ---------------------------------------------------------------- with an index
def row_expect(orc, d, arr):
    """per array: keys as the oracle sorts them, and the stable permutation"""
    rows = d.reshape(-1, arr)
    perm = np.argsort(orc.order(d).reshape(-1, arr), axis=1, kind="stable")
    return np.take_along_axis(rows, perm, axis=1).reshape(-1), perm.astype(np.int32)


@pytest.mark.parametrize("dtype", [np.int8, np.uint16, np.int32, np.uint64, np.float32, np.float64])
def test_multisort_index_host(orc, torch_dev, dtype):
    """turtle_multisort_index: keys equal the oracle's per array; the index
    is each key's position (local) or the payload moved stably with it.
    The 2038 x 4000 case spans several 16 MiB staging chunks (3 for 1-byte
    keys, 6 for 8-byte keys with the index)."""
    sorter = TurtleSort(dtype, 2038 * 4000)
    ui = np.dtype(f"u{np.dtype(dtype).itemsize}")
    for arr, num in [(1, 9), (7, 300), (64, 33), (2038, 4000), (8192, 5), (8193, 3), (30000, 2)]:
        print(f"case dtype={np.dtype(dtype)} arr={arr} num={num}", flush=True)
        d = wl.gen_keys(dtype, arr * num, arr * 3 + num, 0 if np.dtype(dtype).kind == "f" else 60)
        ek, perm = row_expect(orc, d, arr)
        got = d.copy()
        li = sorter.multi_sort_index(arr, got)
        np.testing.assert_array_equal(got.view(ui), ek.view(ui))
        np.testing.assert_array_equal(li.reshape(-1, arr), perm)
        got = d.copy()
        pay = (np.arange(d.size, dtype=np.int64) * 7919 % (2**31 - 1)).astype(np.int32)
        out = sorter.multi_sort_index(arr, got, pay.copy())
        np.testing.assert_array_equal(got.view(ui), ek.view(ui))
        np.testing.assert_array_equal(out.reshape(-1, arr), np.take_along_axis(pay.reshape(-1, arr), perm, axis=1))
        # the keys-only host form through the same chunks
        got = d.copy()
        sorter.multi_sort(arr, got)
        np.testing.assert_array_equal(got.view(ui), ek.view(ui))
    sorter.close()


@pytest.mark.parametrize("arr", [100, 1024, 8192, 9000])
def test_multisort_index_device(orc, torch_dev, arr):
    """device form: local positions and a payload, int64 keys with heavy
    duplicates (stability decides the index)"""
    torch = torch_dev
    num = max(4, (1 << 20) // arr)
    n = num * arr
    print(f"case device arr={arr} num={num}", flush=True)
    d = wl.gen_keys(np.int64, n, arr, 58)
    ek, perm = row_expect(orc, d, arr)
    sorter = TurtleSort(np.int64, n)
    kin = h2d(torch, d)
    kout = torch.empty_like(kin)
    iout = torch.empty(n, dtype=torch.int32, device="cuda")
    sorter.multi_sort_index_device(arr, kin, kout, iout)
    sorter.synchronize()
    np.testing.assert_array_equal(d2h(torch, kout), ek)
    np.testing.assert_array_equal(d2h(torch, iout).reshape(-1, arr), perm)
    pay = torch.arange(n, dtype=torch.int32, device="cuda") * 3
    iout2 = torch.empty_like(iout)
    sorter.multi_sort_index_device(arr, kin, kout, iout2, idx_in=pay)
    sorter.synchronize()
    np.testing.assert_array_equal(d2h(torch, iout2).reshape(-1, arr),
                                  np.take_along_axis(d2h(torch, pay).reshape(-1, arr), perm, axis=1))
    sorter.close()


# ------------------------------------------- arrays of 513..1024 4-byte keys
def _radix_rows(kind, arr, num, dtype, seed):
    """rows that steer the one-wave LDS radix of 513..1024-key arrays: the
    top 18 bits of (key - min) first, keys sharing those bits ordered inside
    their group, and the exact 9-bit passes when a group exceeds 64 keys or
    the span fits 18 bits"""
    rng = np.random.default_rng(seed)
    ui = np.dtype(f"u{np.dtype(dtype).itemsize}")
    spread = rng.integers(0, 1 << 32, size=(num, arr), dtype=np.uint64)
    if kind == "spread":
        rows = spread
    elif kind == "pairs":            # many keys share their top bits with a neighbour
        rows = spread.copy()
        rows[:, 1::2] = (rows[:, 0::2][:, : arr // 2] & ~np.uint64(0x3FFF)) | rng.integers(
            0, 1 << 14, size=(num, arr // 2), dtype=np.uint64)
    elif kind in ("group64", "group65", "group200"):   # one group of that size
        g = int(kind[5:])
        rows = spread.copy()
        base = rng.integers(0, 1 << 31, size=(num, 1), dtype=np.uint64) & ~np.uint64(0x3FFF)
        perm = rng.permuted(np.tile(np.arange(arr), (num, 1)), axis=1)
        vals = base + rng.integers(0, 1 << 13, size=(num, g), dtype=np.uint64)
        np.put_along_axis(rows, perm[:, :g], vals, axis=1)
        # the span covers 32 bits (two keys outside the group)
        np.put_along_axis(rows, perm[:, g:g + 2], np.array([[0, 0xFFFFFFFF]], dtype=np.uint64), axis=1)
    elif kind in ("dups", "dups_big"):   # equal keys: ties decided by position
        rows = rng.integers(0, 40 if kind == "dups" else 8, size=(num, arr), dtype=np.uint64) << np.uint64(26)
    elif kind == "narrow":            # span of 17 bits: exact passes
        rows = rng.integers(0, 1 << 17, size=(num, arr), dtype=np.uint64) + np.uint64(123456789)
    else:
        raise ValueError(kind)
    return rows.astype(np.uint32).view(ui).view(dtype).reshape(-1)


@pytest.mark.parametrize("dtype", [np.int32, np.uint32, np.float32])
@pytest.mark.parametrize("kind", ["spread", "pairs", "group64", "group65", "group200", "dups", "dups_big", "narrow"])
def test_multisort_radix_groups(orc, torch_dev, dtype, kind):
    torch = torch_dev
    ui = np.dtype("u4")
    sorter = TurtleSort(dtype, 1024 * 300)
    for arr in (513, 700, 1000, 1024):
        num = 203 if arr < 1024 else 300
        print(f"case radix-groups dtype={np.dtype(dtype)} kind={kind} arr={arr} num={num}", flush=True)
        d = _radix_rows(kind, arr, num, dtype, arr * 31 + len(kind))
        ek, perm = row_expect(orc, d, arr)
        kin = h2d(torch, d.view(np.int32))
        kout = torch.empty_like(kin)
        sorter.multi_sort_device(arr, kin, kout)
        sorter.synchronize()
        np.testing.assert_array_equal(d2h(torch, kout).view(ui), ek.view(ui))
        iout = torch.empty(d.size, dtype=torch.int32, device="cuda")
        sorter.multi_sort_index_device(arr, kin, kout, iout)
        sorter.synchronize()
        np.testing.assert_array_equal(d2h(torch, kout).view(ui), ek.view(ui))
        np.testing.assert_array_equal(d2h(torch, iout).reshape(-1, arr), perm)
    sorter.close()
