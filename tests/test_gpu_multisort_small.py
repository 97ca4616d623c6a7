"""MultiSort of short arrays: up to 32 keys one lane per array (a register
odd-even merge network), 33..512 keys several arrays per wave.  Every array
equal to the oracle's per-array sort; with an index, the stable order.  The
reference's own published shapes (README.md:265-384): 6,400,000 arrays of
10 int, 3,200,000 arrays of 32 char."""
import numpy as np
import pytest

from turtlesort_amd import TurtleSort, fill_keys, workloads as wl

from conftest import d2h, h2d

pytestmark = pytest.mark.gpu

DTYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64]
SIZES = list(range(1, 34)) + [40, 63, 64, 65, 100, 255, 256, 257, 511, 512, 513]


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(f"i{a.itemsize}") if a.dtype.kind in "uf" else a


def row_expect(orc, d, arr):
    rows = d.reshape(-1, arr)
    perm = np.argsort(orc.order(d).reshape(-1, arr), axis=1, kind="stable")
    return np.take_along_axis(rows, perm, axis=1).reshape(-1), perm.astype(np.int32)


@pytest.mark.parametrize("dtype", DTYPES)
def test_short_arrays_device(orc, torch_dev, dtype):
    torch = torch_dev
    dt = np.dtype(dtype)
    ui = np.dtype(f"u{dt.itemsize}")
    sorter = TurtleSort(dtype, 1 << 20)
    for arr in SIZES:
        num = 1000 + arr        # a partial last workgroup
        n = arr * num
        print(f"case short dtype={dt} arr={arr} num={num}", flush=True)
        shift = 0 if dt.kind == "f" else 60       # float: every bit pattern; ints: 16 values (ties)
        d = wl.gen_keys(dtype, n, arr * 11 + 3, shift)
        ek, perm = row_expect(orc, d, arr)
        kin = h2d(torch, bits(d))
        kout = torch.empty_like(kin)
        sorter.multi_sort_device(arr, kin, kout)
        sorter.synchronize()
        np.testing.assert_array_equal(d2h(torch, kout).view(ui), ek.view(ui))
        # local positions: the stable order
        iout = torch.empty(n, dtype=torch.int32, device="cuda")
        k2 = torch.empty_like(kin)
        sorter.multi_sort_index_device(arr, kin, k2, iout)
        sorter.synchronize()
        np.testing.assert_array_equal(d2h(torch, k2).view(ui), ek.view(ui))
        np.testing.assert_array_equal(d2h(torch, iout).reshape(-1, arr), perm)
        # a payload moved with the keys, in place
        pay = torch.arange(n, dtype=torch.int32, device="cuda") * 5 + 1
        k3 = kin.clone()
        p3 = pay.clone()
        sorter.multi_sort_index_device(arr, k3, k3, p3, idx_in=p3)
        sorter.synchronize()
        np.testing.assert_array_equal(d2h(torch, k3).view(ui), ek.view(ui))
        np.testing.assert_array_equal(d2h(torch, p3).reshape(-1, arr),
                                      np.take_along_axis(d2h(torch, pay).reshape(-1, arr), perm, axis=1))
    sorter.close()


@pytest.mark.parametrize("dtype", [np.int32, np.uint8])
def test_short_arrays_host(orc, torch_dev, dtype):
    sorter = TurtleSort(dtype, 1 << 20)
    for arr, num in [(1, 5), (3, 1001), (10, 6400), (17, 999), (32, 3200), (33, 100), (200, 77)]:
        print(f"case short host dtype={np.dtype(dtype)} arr={arr} num={num}", flush=True)
        d = wl.gen_keys(dtype, arr * num, arr + num, 0)
        got = d.copy()
        sorter.multi_sort(arr, got, block_size=64)
        np.testing.assert_array_equal(got, orc.multisort(d, arr))
    sorter.close()


@pytest.mark.parametrize("dtype,arr,num", [(np.int32, 10, 6_400_000), (np.uint8, 32, 3_200_000)])
def test_reference_readme_shapes(torch_dev, dtype, arr, num):
    """MultiSort<int,10,64>(6400000) and 3,200,000 x 32 char: every row
    equal to torch's sort of it"""
    torch = torch_dev
    tdt = {np.dtype(np.int32): torch.int32, np.dtype(np.uint8): torch.uint8}[np.dtype(dtype)]
    n = arr * num
    sorter = TurtleSort(dtype, n, track_index=False)
    kin = torch.empty(n, dtype=tdt, device="cuda")
    fill_keys(kin, 1234, 0)
    kout = torch.empty_like(kin)
    print(f"case readme shape {np.dtype(dtype)} {num} x {arr}", flush=True)
    sorter.multi_sort_device(arr, kin, kout)
    sorter.synchronize()
    ref, _ = torch.sort(kin.view(num, arr), dim=1)
    assert torch.equal(kout.view(num, arr), ref)
    sorter.close()
