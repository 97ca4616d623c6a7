"""Request loop with continuous batching: every result equals the oracle's
sort of that request (keys bit-exact, stable index), and is bit-identical
whether the request ran alone or mixed with others of every key type and
size, submitted concurrently from several threads."""
import threading

import numpy as np
import pytest

from turtlesort_amd import SortServer, TurtleError, workloads as wl

pytestmark = pytest.mark.gpu

DTYPES = [np.uint8, np.int16, np.int32, np.int64]


def make_requests(count, seed, lo=1, hi=300_000):
    sizes = wl.loguniform_lengths(seed, count, lo, hi)
    coins = wl.gen_keys(np.uint8, 2 * count, seed + 1, 0)
    reqs = []
    for i, n in enumerate(sizes):
        dt = DTYPES[coins[2 * i] % 4]
        keys = wl.gen_keys(dt, int(n), seed * 1000 + i, 0)
        idx = np.arange(n, dtype=np.int32)[::-1].copy() if coins[2 * i + 1] & 1 else None
        reqs.append((keys, idx))
    return reqs


def run_mixed(server, reqs, threads=4):
    outs = [None] * len(reqs)

    def client(part):
        tickets = []
        for i in part:
            k = reqs[i][0].copy()
            x = None if reqs[i][1] is None else reqs[i][1].copy()
            tickets.append((i, k, x, server.submit(k, x)))
        for i, k, x, t in tickets:
            server.wait(t)
            outs[i] = (k, x)

    parts = [list(range(j, len(reqs), threads)) for j in range(threads)]
    ths = [threading.Thread(target=client, args=(p,)) for p in parts]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    return outs


def test_server_mix_equals_alone_and_oracle(orc, torch_dev):
    reqs = make_requests(48, 4242)
    server = SortServer(pool_bytes=8 << 30, slot_keys=1 << 21, slots_per_dtype=2)
    mixed = run_mixed(server, reqs)
    st = server.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["batches"] < len(reqs)            # requests were batched together
    server.close()
    alone_server = SortServer(pool_bytes=8 << 30, slot_keys=1 << 21, slots_per_dtype=1)
    for i, (keys, idx) in enumerate(reqs):
        print(f"case request={i} dtype={keys.dtype} n={keys.size} idx={idx is not None}", flush=True)
        k = keys.copy()
        x = None if idx is None else idx.copy()
        alone_server.wait(alone_server.submit(k, x))
        mk, mx = mixed[i]
        np.testing.assert_array_equal(mk, k)                       # alone == in the mix
        if idx is not None:
            np.testing.assert_array_equal(mx, x)
        ok, _ = orc.turtle_sort(keys)
        np.testing.assert_array_equal(k, ok)                       # oracle keys
        if idx is not None:
            np.testing.assert_array_equal(x, idx[np.argsort(keys, kind="stable")])
    alone_server.close()


def test_server_admission(torch_dev):
    server = SortServer(pool_bytes=1 << 30, slot_keys=1 << 16, slots_per_dtype=1, dtypes=[np.int32])
    with pytest.raises(TurtleError) as e:
        server.submit(np.zeros(1 << 17, dtype=np.int32))             # larger than a slot
    assert e.value.status == 10
    with pytest.raises(TurtleError) as e:
        server.submit(np.zeros(10, dtype=np.int64))                 # type not served
    assert e.value.status == 4
    k = np.arange(1000, dtype=np.int32)[::-1].copy()
    assert server.wait(server.submit(k)) >= 0
    assert np.array_equal(k, np.arange(1000, dtype=np.int32))
    assert server.stats()["rejected"] == 1
    server.close()


def test_server_float_requests(orc, torch_dev):
    """float and double requests (every bit pattern: NaNs, infinities, both
    zeros) through the request loop, mixed with int32 ones: keys equal the
    oracle's bit for bit, the index is the stable order of the oracle's map"""
    sizes = wl.loguniform_lengths(77, 18, 1, 200_000)
    reqs = []
    for i, n in enumerate(sizes):
        dt = (np.float32, np.float64, np.int32)[i % 3]
        keys = wl.gen_keys(dt, int(n), 9000 + i, 0)
        idx = np.arange(n, dtype=np.int32) if i % 2 else None
        reqs.append((keys, idx))
    server = SortServer(pool_bytes=4 << 30, slot_keys=1 << 20, slots_per_dtype=2,
                        dtypes=[np.float32, np.float64, np.int32])
    outs = run_mixed(server, reqs, threads=3)
    assert server.stats()["failed"] == 0
    server.close()
    for i, (keys, idx) in enumerate(reqs):
        print(f"case request={i} dtype={keys.dtype} n={keys.size}", flush=True)
        k, x = outs[i]
        ok, _ = orc.turtle_sort(keys)
        ui = np.dtype(f"u{keys.dtype.itemsize}")
        np.testing.assert_array_equal(k.view(ui), ok.view(ui))
        if idx is not None:
            np.testing.assert_array_equal(x, np.argsort(orc.order(keys), kind="stable"))


def test_inflight_budget_admission(orc, torch_dev):
    """Requests are charged n * (4 K + 17) device bytes while their batch
    runs and join a batch only when those bytes fit the in-flight budget:
    with room for one 1M-key int32 request, four slots still sort one
    request at a time; a request larger than the budget is refused."""
    per = 1_000_000 * (4 * 4 + 17)
    srv = SortServer(pool_bytes=8 << 30, slot_keys=1 << 21, slots_per_dtype=4, dtypes=[np.int32],
                     inflight_bytes=per + per // 2)
    reqs = []
    for i in range(6):
        keys = wl.gen_keys(np.int32, 1_000_000, 77 + i, 0)
        k = keys.copy()
        reqs.append((keys, k, srv.submit(k)))
    infos = []
    for keys, k, t in reqs:
        infos.append(srv.wait_info(t))
        ok, _ = orc.turtle_sort(keys)
        np.testing.assert_array_equal(k, ok)
    st = srv.stats()
    assert st["max_bytes_in_flight"] <= per + per // 2
    assert st["max_in_flight_batches"] == 1
    assert st["bytes_in_flight"] == 0 and st["completed"] == 6
    assert all(i["batch_requests"] == 1 and i["status"] == 0 for i in infos)
    # the later requests waited for the budget
    assert max(i["queued_seconds"] for i in infos) > 0
    with pytest.raises(TurtleError) as e:
        srv.submit(np.zeros(2_000_000, dtype=np.int32))
    assert e.value.status == 10   # TURTLE_ERR_BUDGET
    srv.close()
