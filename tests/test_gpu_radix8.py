"""8-byte keys in the LDS radix: two passes over the top 18 bits of each
task's key span, then keys that share those bits (groups) ordered by
comparison inside their group; a group longer than the limit (64) makes the
step redo the full passes.  Inputs built to hit each branch: spread keys,
pairs and triples that share their top bits, equal keys (stability), groups
around the limit, a cluster that forces the full passes, and a batch mixing
wide and narrow tasks.  Keys bit-exact against the oracle, index against a
stable sort."""
import numpy as np
import pytest

from turtlesort_amd import TurtleSort, workloads as wl

pytestmark = pytest.mark.gpu


def check(orc, keys, out_keys, out_idx):
    ok, _ = orc.turtle_sort(keys)
    u = np.dtype(f"u{keys.dtype.itemsize}")
    np.testing.assert_array_equal(out_keys.view(u), ok.view(u))
    np.testing.assert_array_equal(out_idx, np.argsort(orc.order(keys), kind="stable").astype(np.int32))


def sort_host(dtype, keys):
    s = TurtleSort(dtype, max(keys.size, 1))
    k = keys.copy()
    idx = np.arange(keys.size, dtype=np.int32)
    s.start_sorting(k, idx)
    s.sync()
    s.close()
    return k, idx


def shuffled(rng, a):
    return a[rng.permutation(a.size)]


def grouped(rng, n, size, shift):
    """n keys in groups of `size` that agree above bit `shift` and differ below it"""
    base = rng.integers(0, 1 << 61, size=(n + size - 1) // size, dtype=np.int64) << 1
    low = rng.integers(0, 1 << shift, size=base.size * size, dtype=np.int64)
    return (np.repeat(base, size) ^ low)[:n]


CASES = {
    "spread": lambda rng, n: rng.integers(-(1 << 62), 1 << 62, size=n, dtype=np.int64),
    "pairs": lambda rng, n: grouped(rng, n, 2, 8),
    "equal_triples": lambda rng, n: np.repeat(rng.integers(0, 1 << 62, size=n // 3 + 1, dtype=np.int64), 3)[:n],
    "groups_63": lambda rng, n: grouped(rng, n, 63, 10),
    "groups_64": lambda rng, n: grouped(rng, n, 64, 10),
    "groups_65": lambda rng, n: grouped(rng, n, 65, 10),
    "cluster_forces_full_passes": lambda rng, n: np.concatenate(
        [rng.integers(0, 1000, size=n - 4, dtype=np.int64), np.array([1 << 60, -(1 << 61), 7, 1 << 59])]),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("n", [300, 1500, 4096])
def test_radix8_single_task(orc, torch_dev, case, n):
    rng = np.random.default_rng(n * 31 + len(case))
    keys = shuffled(rng, CASES[case](rng, n).astype(np.int64))
    print(f"case {case} n={n}", flush=True)
    k, idx = sort_host(np.int64, keys)
    check(orc, keys, k, idx)


@pytest.mark.parametrize("dtype", [np.uint64, np.float64])
def test_radix8_other_types(orc, torch_dev, dtype):
    rng = np.random.default_rng(5)
    raw = np.concatenate([grouped(rng, 2000, 2, 6), grouped(rng, 1500, 70, 4), rng.integers(0, 1 << 62, 500)])
    keys = shuffled(rng, raw.astype(np.int64)).view(dtype)
    if dtype == np.float64:
        keys = keys.copy()
        keys[::97] = np.nan
        keys[::89] = -0.0
        keys[::83] = 0.0
    print(f"case other types {np.dtype(dtype)}", flush=True)
    k, idx = sort_host(dtype, keys)
    check(orc, keys, k, idx)


def test_radix8_mixed_batch(orc, torch_dev):
    """one sort whose LDS steps mix tasks of wide and narrow spans, groups
    and clusters"""
    rng = np.random.default_rng(11)
    parts = [rng.integers(-(1 << 62), 1 << 62, size=150_000, dtype=np.int64),
             grouped(rng, 60_000, 3, 20), grouped(rng, 40_000, 90, 6),
             rng.integers(0, 5000, size=30_000, dtype=np.int64)]
    keys = shuffled(rng, np.concatenate(parts))
    print(f"case mixed batch n={keys.size}", flush=True)
    k, idx = sort_host(np.int64, keys)
    check(orc, keys, k, idx)
