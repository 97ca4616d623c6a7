"""The device-wide stable merge (turtle_merge_device) against the oracle's
restatement of the reference's rank merge (turtle-sort.cu:190-281): keys
bit-exact and the carried index equal, run 1 first on equal keys."""
import numpy as np
import pytest

from turtlesort_amd import merge_device, workloads as wl

from conftest import d2h, h2d

pytestmark = pytest.mark.gpu

DTYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64]
SIZES = [(0, 0), (0, 1), (1, 0), (1, 1), (4095, 1), (4096, 4096), (4097, 3), (100_000, 7), (3, 100_001),
         (1_000_000, 2_500_000)]


def bits(a):
    """same-width signed view (torch has no unsigned 16/32/64-bit tensors here)"""
    a = np.ascontiguousarray(a)
    return a.view(f"i{a.itemsize}") if a.dtype.kind in "uf" else a


def make_runs(dt, na, nb, seed, shift):
    k = wl.gen_keys(dt, na + nb, seed, shift)
    if np.dtype(dt).kind == "f":
        k = (wl.gen_keys(np.int32, na + nb, seed, 56).astype(dt) - 128) / 4   # few values: ties across runs
    return np.sort(k[:na], kind="stable"), np.sort(k[na:], kind="stable")


@pytest.mark.parametrize("dtype", DTYPES)
def test_merge_against_oracle(orc, torch_dev, dtype):
    torch = torch_dev
    dt = np.dtype(dtype)
    for i, (na, nb) in enumerate(SIZES):
        for shift in (0, 60):                     # full range, 16 values (ties across the runs)
            print(f"case merge dtype={dt} na={na} nb={nb} shift={shift}", flush=True)
            a, b = make_runs(dt, na, nb, 500 + i, shift)
            ia = np.arange(na, dtype=np.int32)
            ib = np.arange(na, na + nb, dtype=np.int32) + 7
            ek, ei = orc.merge_ranks(a, ia, b, ib)
            ta, tb = h2d(torch, bits(a)), h2d(torch, bits(b))
            tia, tib = h2d(torch, ia), h2d(torch, ib)
            ko = torch.full((na + nb,), 3, dtype=ta.dtype, device="cuda")
            io = torch.full((na + nb,), -1, dtype=torch.int32, device="cuda")
            merge_device(ta, tb, ko, tia, tib, io, dtype=dt)
            np.testing.assert_array_equal(d2h(torch, ko).view(dt), ek)
            np.testing.assert_array_equal(d2h(torch, io), ei)
            # keys only
            k2 = torch.empty_like(ko)
            merge_device(ta, tb, k2, dtype=dt)
            np.testing.assert_array_equal(d2h(torch, k2).view(dt), ek)


def test_merge_full_size_int32_idx(torch_dev):
    """two sorted runs of 32M int32 keys + index: the merge equals a stable
    sort of their concatenation"""
    torch = torch_dev
    from turtlesort_amd import fill_keys
    n = 1 << 25
    x = torch.empty(2 * n, dtype=torch.int32, device="cuda")
    fill_keys(x, 77, wl.SHIFT_15BIT)             # [0, 32767]: ties inside and across the runs
    a, ia = torch.sort(x[:n], stable=True)
    b, ib = torch.sort(x[n:], stable=True)
    ia, ib = ia.to(torch.int32), (ib + n).to(torch.int32)
    ko, io = torch.empty(2 * n, dtype=torch.int32, device="cuda"), torch.empty(2 * n, dtype=torch.int32, device="cuda")
    print("case merge 2 x 32M int32+idx", flush=True)
    merge_device(a.contiguous(), b.contiguous(), ko, ia, ib, io)
    torch.cuda.synchronize()
    ref_k, ref_i = torch.sort(x, stable=True)
    # the stable sort of x puts the first run's equal keys first, as the merge does
    assert torch.equal(ko, ref_k)
    assert torch.equal(io.long(), ref_i)


SHORT_TAIL = [(4096 + 3, 0), (3, 4096), (64 * 4096 - 4, 5), (100, 2 * 4096 - 93), (129 * 4096 - 1000, 1003),
              (40_000, 64 * 4096 + 6 - 40_000)]


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.int32, np.float32, np.uint64])
def test_merge_short_last_tile(orc, torch_dev, dtype):
    """outputs whose last tile of 4096 holds 1..7 keys (fewer than 8 bytes
    for narrow keys): the merge writes nothing past the output, and the
    crossings of the 64-boundary blocks on either side of such a tile are
    right"""
    torch = torch_dev
    dt = np.dtype(dtype)
    for i, (na, nb) in enumerate(SHORT_TAIL):
        n = na + nb
        assert 1 <= n % 4096 <= 7
        print(f"case merge short tail dtype={dt} na={na} nb={nb}", flush=True)
        a, b = make_runs(dt, na, nb, 900 + i, 4)
        ia = np.arange(na, dtype=np.int32)
        ib = np.arange(na, na + nb, dtype=np.int32)
        ek, ei = orc.merge_ranks(a, ia, b, ib)
        ta, tb = h2d(torch, bits(a)), h2d(torch, bits(b))
        tia, tib = h2d(torch, ia), h2d(torch, ib)
        kfull = torch.full((n + 64,), 5, dtype=ta.dtype, device="cuda")
        ifull = torch.full((n + 64,), -9, dtype=torch.int32, device="cuda")
        merge_device(ta, tb, kfull[:n], tia, tib, ifull[:n], dtype=dt)
        np.testing.assert_array_equal(d2h(torch, kfull[:n]).view(dt), ek)
        np.testing.assert_array_equal(d2h(torch, ifull[:n]), ei)
        assert bool((kfull[n:] == 5).all()), "keys written past the output"
        assert bool((ifull[n:] == -9).all()), "index written past the output"
