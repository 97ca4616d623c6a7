"""The allocation-free entry point (turtle_sort_device_noalloc) called
through the C ABI: its output equals turtle_sort_device's on the same input,
bit for bit (keys and the stable index), on two calls in a row whose second
is smaller, for several key types, in place and not."""
import ctypes

import numpy as np
import pytest

from turtlesort_amd import TurtleSort, _lib, dtype_code, workloads as wl

from conftest import d2h, h2d

pytestmark = pytest.mark.gpu

MAX_N = 1 << 20


def noalloc(lib, torch, dtype, kin, kout, iin=None, iout=None):
    p = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return lib.turtle_sort_device_noalloc(dtype_code(dtype), p(kin), p(iin), p(kout), p(iout), kin.numel(), st)


@pytest.mark.parametrize("dtype,dist", [(np.int32, wl.SHIFT_15BIT), (np.int64, 0), (np.uint16, 0),
                                        (np.float32, 0), (np.int8, 0)])
def test_noalloc_matches_sort_device(orc, torch_dev, dtype, dist):
    torch = torch_dev
    lib = _lib.load()
    sorter = TurtleSort(dtype, MAX_N)
    for n, seed in ((MAX_N, 11), (300_001, 12)):          # the second call is smaller
        print(f"case noalloc dtype={np.dtype(dtype)} n={n}", flush=True)
        keys = wl.gen_keys(dtype, n, seed, dist)
        kin = h2d(torch, keys.view(f"i{keys.itemsize}") if keys.dtype.kind in "uf" else keys)
        iin = torch.arange(n, dtype=torch.int32, device="cuda")
        # reference: the handle-based device sort
        k_ref, i_ref = torch.empty_like(kin), torch.empty_like(iin)
        sorter.sort_device(kin, k_ref, iin, i_ref)
        assert sorter.synchronize()["error"] == 0
        k_na, i_na = torch.full_like(kin, 7), torch.full_like(iin, -1)
        assert noalloc(lib, torch, dtype, kin, k_na, iin, i_na) == 0, lib.turtle_last_error()
        torch.cuda.synchronize()
        assert torch.equal(k_na, k_ref) and torch.equal(i_na, i_ref)
        # keys only, in place
        k_ip = kin.clone()
        assert noalloc(lib, torch, dtype, k_ip, k_ip) == 0, lib.turtle_last_error()
        torch.cuda.synchronize()
        assert torch.equal(k_ip, k_ref)
        # and against the oracle
        ok, _ = orc.turtle_sort(keys)
        np.testing.assert_array_equal(d2h(torch, k_na).view(keys.dtype), ok)
    sorter.close()


def test_noalloc_rejects_oversize_and_bad_args(torch_dev):
    torch = torch_dev
    lib = _lib.load()
    big = torch.empty(MAX_N + 1, dtype=torch.int32, device="cuda")
    assert noalloc(lib, torch, np.int32, big, big) == _lib.ERR_CAPACITY
    k = torch.empty(16, dtype=torch.int32, device="cuda")
    i = torch.empty(16, dtype=torch.int32, device="cuda")
    assert noalloc(lib, torch, np.int32, k, k, i, None) == _lib.ERR_INVALID
    assert lib.turtle_sort_device_noalloc(99, None, None, None, None, 4, None) == _lib.ERR_DTYPE
    assert noalloc(lib, torch, np.int32, k[:0], k[:0]) == 0
