"""ctypes wrapper of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")

DTYPES = {np.dtype(np.int8): 1, np.dtype(np.uint8): 2, np.dtype(np.int16): 3, np.dtype(np.uint16): 4,
          np.dtype(np.int32): 5, np.dtype(np.uint32): 6, np.dtype(np.int64): 7, np.dtype(np.uint64): 8,
          np.dtype(np.float32): 9, np.dtype(np.float64): 10}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = ctypes.CDLL(LIB)
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    L.ts_oracle_turtle_sort_typed.argtypes = [i32, vp, vp, i64]
    L.ts_oracle_turtle_sort.argtypes = [vp, vp, i64]
    L.ts_oracle_multisort_typed.argtypes = [i32, vp, i64, i64]
    L.ts_oracle_segmented_sort_typed.argtypes = [i32, vp, vp, vp, i64]
    L.ts_oracle_network_sort.argtypes = [vp, vp, i32]
    L.ts_oracle_network_table.argtypes = [i32, ctypes.POINTER(ctypes.POINTER(ctypes.c_ubyte))]
    L.ts_oracle_oddeven_sort.argtypes = [vp, vp, i32]
    L.ts_oracle_partition3.argtypes = [vp, vp, i64, vp, vp, vp, vp]
    L.ts_oracle_merge_ranks.argtypes = [vp, vp, i64, vp, vp, i64, vp, vp]
    L.ts_oracle_heapsort.argtypes = [vp, i64]
    L.ts_oracle_splitmix64.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
    L.ts_oracle_splitmix64.restype = ctypes.c_uint64
    L.ts_oracle_gen_keys.argtypes = [i32, ctypes.c_uint64, i32, vp, i64]
    L.ts_oracle_gen_loguniform.argtypes = [ctypes.c_uint64, i64, i64, vp, i64]
    L.ts_oracle_to_ord.argtypes = [i32, vp, vp, i64]
    L.ts_oracle_from_ord.argtypes = [i32, vp, vp, i64]
    L.ts_oracle_dtype_size.argtypes = [i32]
    L.ts_cpu_sort.argtypes = [i32, i32, vp, vp, i64, i32, vp, vp]
    L.ts_cpu_sort.restype = ctypes.c_double
    L.ts_cpu_segmented_sort.argtypes = [i32, i32, vp, vp, vp, i64, i32]
    L.ts_cpu_segmented_sort.restype = ctypes.c_double
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def order(keys):
    """order values (uint64) of the oracle's key map: their unsigned order is
    the sort order of the keys"""
    k = np.ascontiguousarray(keys)
    m = np.empty(k.size, dtype=np.uint64)
    assert load().ts_oracle_to_ord(DTYPES[k.dtype], _p(k), _p(m), k.size) == 0
    return m


def turtle_sort(keys, idx=None):
    """The reference's StartSorting+Sync algorithm on the CPU; returns copies."""
    L = load()
    k = np.ascontiguousarray(keys).copy()
    i = None if idx is None else np.ascontiguousarray(idx, dtype=np.int32).copy()
    rc = L.ts_oracle_turtle_sort_typed(DTYPES[k.dtype], _p(k), _p(i), k.size)
    assert rc == 0
    return k, i


def multisort(data, arr_size):
    L = load()
    d = np.ascontiguousarray(data).copy()
    rc = L.ts_oracle_multisort_typed(DTYPES[d.dtype], _p(d), d.size // arr_size, arr_size)
    assert rc == 0
    return d


def segmented_sort(keys, idx, offsets):
    L = load()
    k = np.ascontiguousarray(keys).copy()
    i = None if idx is None else np.ascontiguousarray(idx, dtype=np.int32).copy()
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    rc = L.ts_oracle_segmented_sort_typed(DTYPES[k.dtype], _p(k), _p(i), _p(off), off.size - 1)
    assert rc == 0
    return k, i


def gen_keys(dtype, n, seed, shift):
    L = load()
    out = np.empty(n, dtype=dtype)
    assert L.ts_oracle_gen_keys(DTYPES[np.dtype(dtype)], ctypes.c_uint64(seed), shift, _p(out), n) == 0
    return out


def gen_loguniform(seed, lo, hi, n):
    L = load()
    out = np.empty(n, dtype=np.int64)
    assert L.ts_oracle_gen_loguniform(ctypes.c_uint64(seed), lo, hi, _p(out), n) == 0
    return out


def cpu_sort(which, keys, idx=None, threads=0, want_output=False):
    """which: 0 std::sort, 1 std::sort(par_unseq), 2 __gnu_parallel::sort, 3 std::stable_sort.
    Returns seconds (and the sorted keys/idx when want_output)."""
    L = load()
    k = np.ascontiguousarray(keys)
    ok = np.empty_like(k) if want_output else None
    oi = np.empty(k.size, dtype=np.int32) if want_output else None
    sec = L.ts_cpu_sort(which, DTYPES[k.dtype], _p(k), _p(idx), k.size, threads, _p(ok), _p(oi))
    assert sec >= 0
    return (sec, ok, oi) if want_output else sec


def merge_ranks(k1, i1, k2, i2):
    """The reference's stable rank merge of two sorted runs
    (turtle-sort.cu:190-281): run-1 keys go before equal run-2 keys.  Keys of
    any dtype (compared by their order values); i1/i2 may be None."""
    L = load()
    a, b = np.ascontiguousarray(k1), np.ascontiguousarray(k2)
    oa, ob = order(a), order(b)
    ko = np.empty(a.size + b.size, dtype=np.uint64)
    with_idx = i1 is not None
    ia = np.ascontiguousarray(i1, dtype=np.int32) if with_idx else None
    ib = np.ascontiguousarray(i2, dtype=np.int32) if with_idx else None
    io = np.empty(a.size + b.size, dtype=np.int32) if with_idx else None
    L.ts_oracle_merge_ranks(_p(oa), _p(ia), a.size, _p(ob), _p(ib), b.size, _p(ko), _p(io))
    out = np.empty(a.size + b.size, dtype=a.dtype)
    assert L.ts_oracle_from_ord(DTYPES[a.dtype], _p(ko), _p(out), out.size) == 0
    return out, io


def cpu_segmented_sort(par, keys, idx, offsets, threads=0):
    """seconds of a per-segment std::sort loop (par: under omp parallel for)"""
    L = load()
    k = np.ascontiguousarray(keys)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    ix = None if idx is None else np.ascontiguousarray(idx, dtype=np.int32)
    sec = L.ts_cpu_segmented_sort(int(par), DTYPES[k.dtype], _p(k), _p(ix), _p(off), off.size - 1, threads)
    assert sec >= 0
    return sec
