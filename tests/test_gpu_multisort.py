"""MultiSort parity: every array equal to the oracle's heap-sort result
(turtle-multi-sort.cuh:15-182); the full 65,536 x 1024 configuration through
per-row comparison with torch.sort."""
import numpy as np
import pytest

from turtlesort_amd import TurtleSort, fill_keys, workloads as wl

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64])
def test_multisort_host_api(orc, torch_dev, dtype):
    sorter = TurtleSort(dtype, 1 << 20)
    for arr, num in [(1, 7), (2, 33), (10, 640), (32, 100), (1000, 37), (1024, 256), (2038, 50), (4096, 9),
                     (8192, 3), (8193, 2), (20000, 3)]:
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
