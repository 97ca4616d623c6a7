"""Pins of the CPU oracle (oracle/turtle_oracle.c) against facts that do not
come from the oracle itself: a published PRNG vector, the 0-1 principle,
brute force over small inputs, a textbook stable merge, and library sorts."""
import ctypes
import itertools

import numpy as np
import pytest

from turtlesort_amd import workloads as wl

INT_DTYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64]


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- generators
def test_splitmix64_published_vector(orc):
    # splitmix64 seeded with 1234567: first five outputs as published with
    # the reference implementation (Vigna)
    L = orc.load()
    s = ctypes.c_uint64(1234567)
    got = [L.ts_oracle_splitmix64(ctypes.byref(s)) for _ in range(5)]
    assert got == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                   4593380528125082431, 16408922859458223821]


@pytest.mark.parametrize("dtype,shift", [(np.int32, 49), (np.int32, 33), (np.int32, 60), (np.int64, 1),
                                         (np.uint8, 0), (np.int16, 0), (np.int64, 0)])
def test_generators_agree(orc, dtype, shift):
    a = orc.gen_keys(dtype, 10000, 0xABCDEF, shift)
    b = wl.gen_keys(dtype, 10000, 0xABCDEF, shift)
    np.testing.assert_array_equal(a, b)


def test_config_ranges():
    assert wl.gen_keys(np.int32, 100000, 1, wl.SHIFT_15BIT).max() <= 32767
    k = wl.gen_keys(np.int32, 100000, 1, wl.SHIFT_31BIT)
    assert k.min() >= 0 and k.max() > 2**30
    assert set(np.unique(wl.gen_keys(np.int32, 100000, 1, wl.SHIFT_16VAL))) == set(range(16))
    assert wl.gen_keys(np.int64, 1000, 1, wl.SHIFT_63BIT).min() >= 0


def test_loguniform_agrees(orc):
    a = orc.gen_loguniform(wl.SEED_C4_LENGTHS, wl.C4_LO, wl.C4_HI, wl.C4_SEGMENTS)
    b = wl.loguniform_lengths(wl.SEED_C4_LENGTHS, wl.C4_SEGMENTS, wl.C4_LO, wl.C4_HI)
    np.testing.assert_array_equal(a, b)
    assert a.min() >= wl.C4_LO and a.max() <= wl.C4_HI
    # E[len] = (hi - lo) / ln(hi / lo) ~ 27,012 for [16, 262144]
    assert 24000 < a.mean() < 30000


# ---------------------------------------------------------------- order map
@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.uint16])
def test_order_map_exhaustive(orc, dtype):
    info = np.iinfo(dtype)
    x = np.arange(info.min, info.max + 1, dtype=np.int64).astype(dtype)
    rng = np.random.default_rng(3)
    rng.shuffle(x)
    m = np.empty(x.size, dtype=np.uint64)
    L = orc.load()
    code = orc.DTYPES[np.dtype(dtype)]
    L.ts_oracle_to_ord(code, p(x), p(m), x.size)
    np.testing.assert_array_equal(np.argsort(m, kind="stable"), np.argsort(x, kind="stable"))
    back = np.empty_like(x)
    L.ts_oracle_from_ord(code, p(m), p(back), x.size)
    np.testing.assert_array_equal(back, x)


@pytest.mark.parametrize("dtype", [np.int32, np.uint32, np.int64, np.uint64])
def test_order_map_wide(orc, dtype):
    info = np.iinfo(dtype)
    x = np.concatenate([wl.gen_keys(dtype, 20000, 9, 0),
                        np.array([info.min, info.min + 1, -1 if info.min < 0 else 1, 0, info.max - 1, info.max],
                                 dtype=dtype)])
    m = np.empty(x.size, dtype=np.uint64)
    L = orc.load()
    code = orc.DTYPES[np.dtype(dtype)]
    L.ts_oracle_to_ord(code, p(x), p(m), x.size)
    np.testing.assert_array_equal(np.argsort(m, kind="stable"), np.argsort(x, kind="stable"))


# ---------------------------------------------------------------- networks
def network(orc, width):
    L = orc.load()
    t = ctypes.POINTER(ctypes.c_ubyte)()
    nc = L.ts_oracle_network_table(width, ctypes.byref(t))
    return [(t[2 * c], t[2 * c + 1]) for c in range(nc)]


@pytest.mark.parametrize("width,count", [(4, 5), (8, 19), (16, 60), (32, 185)])
def test_network_sizes(orc, width, count):
    net = network(orc, width)
    assert len(net) == count
    assert all(0 <= a < b < width for a, b in net)


@pytest.mark.parametrize("width", [4, 8, 16, 32])
def test_network_zero_one_principle(orc, width):
    """A comparator network sorts every input iff it sorts every 0-1 input
    (Knuth TAOCP 5.3.4).  All 2^width 0-1 inputs, bit-sliced 64 per word."""
    net = network(orc, width)
    total_words = max(1, (1 << width) // 64)
    low = min(width, 6)
    patterns = []
    for w in range(low):
        v = 0
        for i in range(64):
            if (i >> w) & 1:
                v |= 1 << i
        patterns.append(np.uint64(v))
    chunk = min(total_words, 1 << 14)
    full = np.uint64(0xFFFFFFFFFFFFFFFF)
    for start in range(0, total_words, chunk):
        j = np.arange(start, start + chunk, dtype=np.uint64)
        wires = []
        for w in range(width):
            if w < low:
                wires.append(np.full(chunk, patterns[w], dtype=np.uint64))
            else:
                bit = (j >> np.uint64(w - 6)) & np.uint64(1)
                wires.append(np.where(bit == 1, full, np.uint64(0)))
        if width < 6:
            mask = np.uint64((1 << (1 << width)) - 1)
            wires = [x & mask for x in wires]
        for a, b in net:
            lo_ = wires[a] & wires[b]
            hi_ = wires[a] | wires[b]
            wires[a], wires[b] = lo_, hi_
        for k in range(width - 1):
            assert not np.any(wires[k] & ~wires[k + 1]), f"width {width} fails at wire {k}"


def test_network_sort_partial_widths(orc):
    L = orc.load()
    rng = np.random.default_rng(5)
    for num in range(2, 33):
        for _ in range(20):
            k = rng.integers(0, 6, num).astype(np.uint64)
            i = np.arange(num, dtype=np.int32)
            k0 = k.copy()
            assert L.ts_oracle_network_sort(p(k), p(i), num) == 0
            np.testing.assert_array_equal(k, np.sort(k0))
            np.testing.assert_array_equal(k0[i], k)


def test_network_detects_planted_error(orc):
    """The 0-1 check is not vacuous: dropping one comparator breaks it."""
    net = network(orc, 8)[:-1]
    bad = False
    for bits in range(256):
        w = [(bits >> i) & 1 for i in range(8)]
        for a, b in net:
            w[a], w[b] = min(w[a], w[b]), max(w[a], w[b])
        if any(w[i] > w[i + 1] for i in range(7)):
            bad = True
            break
    assert bad


# ---------------------------------------------------------------- odd-even
def test_oddeven_all_permutations(orc):
    L = orc.load()
    for n in range(1, 9):
        base = [i // 2 for i in range(n)]          # duplicates included
        for perm in set(itertools.permutations(base)):
            k = np.array(perm, dtype=np.uint64)
            i = np.arange(n, dtype=np.int32)
            L.ts_oracle_oddeven_sort(p(k), p(i), n)
            np.testing.assert_array_equal(k, np.sort(np.array(perm, dtype=np.uint64)))
            # odd-even transposition only swaps strictly greater pairs: stable
            np.testing.assert_array_equal(i, np.argsort(np.array(perm), kind="stable"))


# ---------------------------------------------------------------- merge rule
def textbook_merge(k1, i1, k2, i2):
    """sequential stable merge; on equal keys run 1 comes first"""
    ko, io = [], []
    a = b = 0
    while a < len(k1) or b < len(k2):
        if b >= len(k2) or (a < len(k1) and k1[a] <= k2[b]):
            ko.append(k1[a]); io.append(i1[a]); a += 1
        else:
            ko.append(k2[b]); io.append(i2[b]); b += 1
    return np.array(ko, dtype=np.uint64), np.array(io, dtype=np.int32)


def test_merge_ranks_textbook(orc):
    L = orc.load()
    rng = np.random.default_rng(11)
    for trial in range(200):
        n1, n2 = rng.integers(0, 40, 2)
        k1 = np.sort(rng.integers(0, 8, n1).astype(np.uint64))
        k2 = np.sort(rng.integers(0, 8, n2).astype(np.uint64))
        i1 = np.arange(n1, dtype=np.int32)
        i2 = np.arange(n1, n1 + n2, dtype=np.int32)
        ko = np.zeros(n1 + n2, dtype=np.uint64)
        io = np.zeros(n1 + n2, dtype=np.int32)
        L.ts_oracle_merge_ranks(p(k1), p(i1), n1, p(k2), p(i2), n2, p(ko), p(io))
        ek, ei = textbook_merge(k1, i1, k2, i2)
        np.testing.assert_array_equal(ko, ek)
        np.testing.assert_array_equal(io, ei)


# ---------------------------------------------------------------- partition
def test_partition3_predicates(orc):
    L = orc.load()
    rng = np.random.default_rng(2)
    for trial in range(100):
        n = int(rng.integers(3, 3000))
        k = rng.integers(0, int(rng.choice([3, 10, 1000])), n).astype(np.uint64)
        k0 = k.copy()
        i = np.arange(n, dtype=np.int32)
        tk = np.zeros(n, dtype=np.uint64)
        ti = np.zeros(n, dtype=np.int32)
        counts = np.zeros(7, dtype=np.int64)
        piv = np.zeros(3, dtype=np.uint64)
        L.ts_oracle_partition3(p(k), p(i), n, p(tk), p(ti), p(counts), p(piv))
        # pivots: sort3(a[mid], a[start], a[stop])  (turtle-sort.cu:644-650)
        np.testing.assert_array_equal(piv, np.sort(np.array([k0[n // 2], k0[0], k0[n - 1]])))
        assert counts.sum() == n
        off = np.concatenate([[0], np.cumsum(counts)])
        pl, pv, pr = piv
        preds = [lambda x: (x < pl) & (x != pv) & (x != pr),
                 lambda x: x == pl,
                 lambda x: (x > pl) & (x < pv) & (x != pr),
                 lambda x: (x == pv) & (x != pl),
                 lambda x: (x > pv) & (x < pr) & (x != pl),
                 lambda x: (x == pr) & (x != pl) & (x != pv),
                 lambda x: (x > pr)]
        for b in range(7):
            seg = k[off[b]:off[b + 1]]
            assert np.all(preds[b](seg)), f"bin {b}"
            assert np.sum(preds[b](k0)) == counts[b]
        np.testing.assert_array_equal(k0[i], k)    # (key, idx) pairs kept


# ---------------------------------------------------------------- whole sort
@pytest.mark.parametrize("dtype", INT_DTYPES)
def test_turtle_sort_matches_library_sort(orc, dtype):
    rng = np.random.default_rng(int(np.dtype(dtype).itemsize))
    info = np.iinfo(dtype)
    for n in [0, 1, 2, 3, 31, 32, 33, 100, 128, 129, 1000, 8199, 8200, 8201, 20000, 100003]:
        for hi in [1, 15, 32767, None]:
            top = info.max if hi is None else min(hi, info.max)
            k = rng.integers(info.min if hi is None else 0, top, n, endpoint=True, dtype=dtype)
            ok, oi = orc.turtle_sort(k, np.arange(n, dtype=np.int32))
            np.testing.assert_array_equal(ok, np.sort(k))
            np.testing.assert_array_equal(np.sort(oi), np.arange(n))
            np.testing.assert_array_equal(k[oi], ok)


def test_turtle_sort_brute_force_small(orc):
    """every arrangement of n<=6 keys over 3 values"""
    for n in range(0, 7):
        for tup in itertools.product(range(3), repeat=n):
            k = np.array(tup, dtype=np.int32)
            ok, oi = orc.turtle_sort(k, np.arange(n, dtype=np.int32))
            assert list(ok) == sorted(tup)
            assert sorted(oi.tolist()) == list(range(n))


def test_turtle_sort_readme_sample(orc):
    # README.md:54-76 sample input {5,4,3,9,8,1}; the sorted result is
    # 1 3 4 5 8 9 (the README's printed "1 3 3 5 7" is not a sort of it)
    ok, _ = orc.turtle_sort(np.array([5, 4, 3, 9, 8, 1], dtype=np.int32))
    assert ok.tolist() == [1, 3, 4, 5, 8, 9]


def test_turtle_sort_adversarial_orders(orc):
    n = 50000
    base = wl.gen_keys(np.int32, n, 4, 0)
    for k in [np.sort(base), np.sort(base)[::-1].copy(), np.full(n, 7, np.int32),
              np.concatenate([np.sort(base[: n // 2]), np.sort(base[n // 2:])[::-1]])]:
        ok, oi = orc.turtle_sort(k, np.arange(n, dtype=np.int32))
        np.testing.assert_array_equal(ok, np.sort(k))
        np.testing.assert_array_equal(k[oi], ok)


def test_checker_detects_planted_errors(orc):
    """The parity comparison used by the GPU tests fails on a corrupted
    result (mistakes planted in a copy of the oracle's output)."""
    n = 5000
    k = wl.gen_keys(np.int32, n, 8, 49)
    ok, oi = orc.turtle_sort(k, np.arange(n, dtype=np.int32))
    bad = ok.copy()
    bad[10], bad[4000] = bad[4000], bad[10]
    assert not np.array_equal(bad, np.sort(k))
    bad_i = oi.copy()
    bad_i[0] = bad_i[1]
    assert not np.array_equal(np.sort(bad_i), np.arange(n))
    swapped = oi.copy()
    j = int(np.argmax(ok[1:] != ok[:-1]))       # a boundary between two key values
    swapped[j], swapped[j + 1] = swapped[j + 1], swapped[j]
    assert not np.array_equal(k[swapped], ok)


# ---------------------------------------------------------------- MultiSort
@pytest.mark.parametrize("dtype", [np.int32, np.uint8, np.int64])
def test_multisort_heapsort(orc, dtype):
    for arr in [1, 2, 10, 32, 1024, 2038]:
        num = 64
        d = wl.gen_keys(dtype, arr * num, 31 + arr, 0)
        out = orc.multisort(d, arr)
        np.testing.assert_array_equal(out.reshape(num, arr), np.sort(d.reshape(num, arr), axis=1))


def test_segmented_oracle(orc):
    lens = wl.loguniform_lengths(3, 300, 1, 5000)
    off = wl.segment_offsets(lens)
    k = wl.gen_keys(np.int64, int(off[-1]), 5, 1)
    ok, oi = orc.segmented_sort(k, np.arange(k.size, dtype=np.int32), off)
    for s in range(len(lens)):
        a, b = off[s], off[s + 1]
        np.testing.assert_array_equal(ok[a:b], np.sort(k[a:b]))
        np.testing.assert_array_equal(np.sort(oi[a:b]), np.arange(a, b))


# ---------------------------------------------------------------- CPU baselines
def test_cpu_baselines_sort(orc):
    k = wl.gen_keys(np.int32, 100000, 6, 49)
    i = np.arange(k.size, dtype=np.int32)
    for which in range(4):
        sec, ok, oi = orc.cpu_sort(which, k, i, want_output=True)
        np.testing.assert_array_equal(ok, np.sort(k))
        np.testing.assert_array_equal(k[oi], ok)
    sec, ok, oi = orc.cpu_sort(3, k, i, want_output=True)
    np.testing.assert_array_equal(oi, np.argsort(k, kind="stable"))


# ---------------------------------------------------------------- float keys
def float_specials(dt):
    f = np.finfo(dt)
    return np.array([np.inf, -np.inf, f.max, -f.max, f.tiny, -f.tiny, f.smallest_subnormal,
                     -f.smallest_subnormal, 1.0, -1.0, 0.5, -0.5], dtype=dt)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_float_order_map_vs_numpy(orc, dtype):
    """NaN-free keys with one sign of zero: the order map sorts exactly as
    numpy does, and maps back to the same bits"""
    dt = np.dtype(dtype)
    ui = np.dtype(f"u{dt.itemsize}")
    bits = wl.gen_keys(dtype, 50000, 77, 0)
    x = np.concatenate([bits[~np.isnan(bits)], float_specials(dt), np.zeros(5, dtype=dt)])
    x = x[~((x == 0) & np.signbit(x))]                      # +0 only
    np.random.default_rng(5).shuffle(x)
    m = orc.order(x)
    np.testing.assert_array_equal(np.argsort(m, kind="stable"), np.argsort(x, kind="stable"))
    back = np.empty_like(x)
    orc.load().ts_oracle_from_ord(orc.DTYPES[dt], p(m), p(back), x.size)
    np.testing.assert_array_equal(back.view(ui), x.view(ui))
    ok, _ = orc.turtle_sort(x)
    np.testing.assert_array_equal(ok.view(ui), np.sort(x).view(ui))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_float_sort_nan_and_zero_policy(orc, dtype):
    """NaNs of both signs and both zeros: the non-NaN keys equal numpy's sort
    by value, -0 precedes +0, sign-set NaNs come first and the rest last,
    and the key bits are a permutation of the input's"""
    dt = np.dtype(dtype)
    ui = np.dtype(f"u{dt.itemsize}")
    rng = np.random.default_rng(11)
    vals = np.concatenate([float_specials(dt), np.array([0.0, -0.0, np.nan, -np.nan], dtype=dt),
                           rng.standard_normal(20).astype(dt)])
    x = vals[rng.integers(0, vals.size, 20000)]
    # NaN payloads of both signs
    xb = x.view(ui).copy()
    nanb = np.isnan(x)
    xb[nanb] |= rng.integers(1, 1 << 20, nanb.sum()).astype(ui)
    x = xb.view(dt)
    ok, oi = orc.turtle_sort(x, np.arange(x.size, dtype=np.int32))
    nneg = int((np.isnan(x) & np.signbit(x)).sum())
    npos = int((np.isnan(x) & ~np.signbit(x)).sum())
    assert np.isnan(ok[:nneg]).all() and np.signbit(ok[:nneg]).all()
    assert np.isnan(ok[ok.size - npos:]).all() and not np.signbit(ok[ok.size - npos:]).any()
    mid = ok[nneg:ok.size - npos]
    np.testing.assert_array_equal(mid, np.sort(x[~np.isnan(x)]))          # by value
    z = mid[mid == 0]
    assert z.size > 0 and np.signbit(z).any() and not np.signbit(z).all()
    nz = int(np.signbit(z).sum())
    assert np.signbit(z[:nz]).all() and not np.signbit(z[nz:]).any()      # -0 before +0
    np.testing.assert_array_equal(np.sort(ok.view(ui)), np.sort(x.view(ui)))
    np.testing.assert_array_equal(x[oi].view(ui), ok.view(ui))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_float_order_map_brute_small(orc, dtype):
    """every ordered pair of a set of specials compares under the map as the
    IEEE comparison does wherever that is ordered (brute force)"""
    dt = np.dtype(dtype)
    x = np.concatenate([float_specials(dt), np.array([0.0, 3.0, -7.25, 1e-30, -1e-30], dtype=dt)])
    m = orc.order(x)
    for i in range(x.size):
        for j in range(x.size):
            assert (m[i] < m[j]) == (x[i] < x[j])


def test_merge_ranks_typed_wrapper(orc):
    """oracle.merge_ranks over every key type: the stable merge equals a
    stable sort of run 1 followed by run 2 (by the key type's order), and
    the payload follows (run 1 first on equal keys)."""
    rng = np.random.default_rng(12)
    for dt in [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64, np.float32,
               np.float64]:
        for n1, n2 in [(0, 0), (0, 5), (7, 0), (33, 50), (300, 17)]:
            if np.dtype(dt).kind == "f":
                raw = rng.integers(-4, 4, n1 + n2).astype(dt) / 2
            else:
                info = np.iinfo(dt)
                raw = rng.integers(max(info.min, -6), min(info.max, 6), n1 + n2).astype(dt)
            a, b = np.sort(raw[:n1]), np.sort(raw[n1:])
            ia = np.arange(n1, dtype=np.int32)
            ib = np.arange(n1, n1 + n2, dtype=np.int32)
            ko, io = orc.merge_ranks(a, ia, b, ib)
            cat = np.concatenate([a, b])
            perm = np.argsort(cat, kind="stable")
            np.testing.assert_array_equal(ko, cat[perm])
            np.testing.assert_array_equal(io, perm.astype(np.int32))
            k2, i2 = orc.merge_ranks(a, None, b, None)
            assert i2 is None
            np.testing.assert_array_equal(k2, cat[perm])
