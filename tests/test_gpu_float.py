"""float and double keys (the reference instantiates both,
turtle-sort.cu:1136-1174, and its only test sorts floats, main.cpp:6-30).
Keys must be bit-identical to the oracle's; the index is the stable order of
the oracle's key order.  On NaN-free input with one sign of zero the result
is also checked against numpy's / torch's own sorts."""
import numpy as np
import pytest

from turtlesort_amd import TurtleSort, workloads as wl

from conftest import d2h, h2d

pytestmark = pytest.mark.gpu

FLOATS = [np.float32, np.float64]
SIZES = [0, 1, 2, 33, 129, 4097, 8193, 65537, 1_000_003]


def uview(a):
    return a.view(np.dtype(f"u{a.dtype.itemsize}"))


def make_float_keys(dtype, n, dist, seed):
    dt = np.dtype(dtype)
    rng = np.random.default_rng(seed)
    if dist == "bits":          # every bit pattern: NaNs, infinities, subnormals, both zeros
        return wl.gen_keys(dt, n, seed, 0)
    if dist == "normal":
        return rng.standard_normal(n).astype(dt)
    if dist == "specials":      # few values, heavy duplicates, NaNs and zeros of both signs
        f = np.finfo(dt)
        v = np.array([np.inf, -np.inf, np.nan, -np.nan, 0.0, -0.0, f.max, -f.max, f.tiny, -f.smallest_subnormal,
                      1.0, -2.5], dtype=dt)
        return v[rng.integers(0, v.size, n)]
    if dist == "rand314":       # main.cpp:25
        return (rng.integers(0, 2**31 - 1, n).astype(np.float32) * np.float32(3.14)).astype(dt)
    raise ValueError(dist)


def check_float(orc, keys, out_keys, out_idx):
    ok, _ = orc.turtle_sort(keys)
    np.testing.assert_array_equal(uview(out_keys), uview(ok))           # bit-exact
    if out_idx is not None:
        stable = np.argsort(orc.order(keys), kind="stable").astype(np.int32)
        np.testing.assert_array_equal(out_idx, stable)


@pytest.mark.parametrize("dtype", FLOATS)
@pytest.mark.parametrize("dist", ["bits", "normal", "specials", "rand314"])
def test_float_start_sorting(orc, torch_dev, dtype, dist):
    sorter = TurtleSort(dtype, max(SIZES))
    for i, n in enumerate(SIZES):
        print(f"case dtype={np.dtype(dtype)} dist={dist} n={n}", flush=True)
        keys = make_float_keys(dtype, n, dist, 500 + i)
        k = keys.copy()
        idx = np.arange(n, dtype=np.int32)
        sorter.start_sorting(k, idx)
        sorter.sync()
        check_float(orc, keys, k, idx)
        if dist in ("normal", "rand314"):                           # NaN-free, +0 only
            np.testing.assert_array_equal(uview(k), uview(np.sort(keys)))
    sorter.close()


@pytest.mark.parametrize("dtype", FLOATS)
def test_float_keys_only_and_segments(orc, torch_dev, dtype):
    torch = torch_dev
    lens = wl.loguniform_lengths(41, 300, 1, 60000)
    off = wl.segment_offsets(lens)
    n = int(off[-1])
    keys = make_float_keys(dtype, n, "bits", 7)
    print(f"case segmented dtype={np.dtype(dtype)} n={n}", flush=True)
    sorter = TurtleSort(dtype, n)
    kin = h2d(torch, keys)
    iin = torch.arange(n, dtype=torch.int32, device="cuda")
    kout = torch.empty_like(kin)
    iout = torch.empty_like(iin)
    sorter.segmented_sort_device(kin, kout, off, iin, iout)
    st = sorter.synchronize()
    assert st["error"] == 0
    ek, _ = orc.segmented_sort(keys, np.arange(n, dtype=np.int32), off)
    np.testing.assert_array_equal(uview(d2h(torch, kout)), uview(ek))
    oi = d2h(torch, iout)
    for s in range(len(off) - 1):
        a, b = int(off[s]), int(off[s + 1])
        np.testing.assert_array_equal(oi[a:b], a + np.argsort(orc.order(keys[a:b]), kind="stable"))
    print(f"case keys-only dtype={np.dtype(dtype)}", flush=True)
    k = keys.copy()
    sorter.start_sorting(k)
    sorter.sync()
    check_float(orc, keys, k, None)
    sorter.close()


@pytest.mark.parametrize("dtype", FLOATS)
def test_float_device_full_size(torch_dev, dtype):
    """64M float32 / 32M float64 normal keys with index, on the device:
    keys and index equal torch's stable sort bit for bit"""
    torch = torch_dev
    n = (64 << 20) if dtype == np.float32 else (32 << 20)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    print(f"case device dtype={np.dtype(dtype)} n={n}", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(1234)
    kin = torch.randn(n, dtype=tdt, device="cuda", generator=g)
    kin[kin == 0] = 1.0                                        # one sign of zero
    iin = torch.arange(n, dtype=torch.int32, device="cuda")
    kout = torch.empty_like(kin)
    iout = torch.empty_like(iin)
    sorter = TurtleSort(dtype, n)
    sorter.sort_device(kin, kout, iin, iout)
    st = sorter.synchronize()
    assert st["error"] == 0
    ev, ei = torch.sort(kin, stable=True)
    ib = torch.int32 if dtype == np.float32 else torch.int64
    assert torch.equal(kout.view(ib), ev.view(ib))
    assert torch.equal(iout.long(), ei)
    sorter.close()


@pytest.mark.parametrize("dtype", FLOATS)
def test_float_multisort(orc, torch_dev, dtype):
    """the reference's test (main.cpp:6-47) shape and fill: 32,000 arrays of
    2038 rand() * 3.14f keys, every array equal to numpy's sort of it; and a
    smaller batch of every bit pattern equal to the oracle's"""
    arr, num = 2038, 32000
    sorter = TurtleSort(dtype, arr * num)
    print(f"case multisort main.cpp dtype={np.dtype(dtype)}", flush=True)
    d = make_float_keys(dtype, arr * num, "rand314", 3).reshape(num, arr)
    got = d.copy()
    sorter.multi_sort(arr, got.reshape(-1), block_size=32)
    np.testing.assert_array_equal(uview(got), uview(np.sort(d, axis=1)))
    for a, m in [(1, 5), (100, 64), (2038, 40), (8192, 3), (8193, 2)]:
        print(f"case multisort bits dtype={np.dtype(dtype)} arr={a} num={m}", flush=True)
        d = make_float_keys(dtype, a * m, "bits", a + m)
        got = d.copy()
        sorter.multi_sort(a, got, block_size=32)
        np.testing.assert_array_equal(uview(got), uview(orc.multisort(d, a)))
    sorter.close()
