"""ctypes binding of libturtle_hip.so (the C ABI declared in include/turtle_c.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C turtlesort_amd``).  There is no fallback: if the library is missing
or cannot be loaded, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libturtle_hip.so")

# status codes (turtle_status)
OK = 0
ERR_INVALID = 1
ERR_CAPACITY = 2
ERR_OOM = 3
ERR_DTYPE = 4
ERR_UNSUPPORTED = 5
ERR_DEVICE = 6
ERR_QUEUE_OVERFLOW = 7
ERR_TIMEOUT = 8
ERR_STATE = 9
ERR_BUDGET = 10

# dtype codes (turtle_dtype)
INT8, UINT8, INT16, UINT16, INT32, UINT32, INT64, UINT64, FLOAT32, FLOAT64 = range(1, 11)


class TurtleError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"turtle status {status}: {msg}")
        self.status = status


class Stats(ctypes.Structure):
    _fields_ = [
        ("device_ms", ctypes.c_double),
        ("n", ctypes.c_uint64),
        ("bytes_count", ctypes.c_uint64),
        ("bytes_scatter", ctypes.c_uint64),
        ("bytes_lds", ctypes.c_uint64),
        ("tasks_pivot", ctypes.c_uint64),
        ("tasks_count", ctypes.c_uint64),
        ("tasks_scatter", ctypes.c_uint64),
        ("tasks_lds", ctypes.c_uint64),
        ("segments", ctypes.c_uint64),
        ("max_depth", ctypes.c_uint32),
        ("error", ctypes.c_uint32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [
        ("splitters", ctypes.c_int32),
        ("lds_cutoff", ctypes.c_int32),
        ("tiles_per_task", ctypes.c_int32),
        ("workgroups_per_cu", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# kernel kinds (turtle_kernel_kind)
KERNEL_KINDS = ("reset", "roots", "level", "pivot", "count", "scan", "scatter", "children", "lds", "ldsprep")


class Profile(ctypes.Structure):
    _fields_ = [
        ("ms", ctypes.c_double * len(KERNEL_KINDS)),
        ("launches", ctypes.c_uint64 * len(KERNEL_KINDS)),
    ]

    def as_dict(self):
        return {k: {"ms": self.ms[i], "launches": self.launches[i]} for i, k in enumerate(KERNEL_KINDS)}


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("pool_bytes", ctypes.c_int64),
        ("slot_keys", ctypes.c_int64),
        ("slots_per_dtype", ctypes.c_int32),
        ("dtype_mask", ctypes.c_int32),
        ("inflight_bytes", ctypes.c_int64),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
        ("failed", ctypes.c_uint64), ("batches", ctypes.c_uint64), ("keys", ctypes.c_uint64),
        ("pool_bytes", ctypes.c_uint64), ("slot_keys", ctypes.c_uint64),
        ("busy_seconds", ctypes.c_double),
        ("slots", ctypes.c_uint32), ("in_flight_batches", ctypes.c_uint32),
        ("max_in_flight_batches", ctypes.c_uint32), ("pad", ctypes.c_uint32),
        ("inflight_bytes", ctypes.c_uint64), ("bytes_in_flight", ctypes.c_uint64),
        ("max_bytes_in_flight", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("latency_seconds", ctypes.c_double), ("queued_seconds", ctypes.c_double),
        ("batch_seconds", ctypes.c_double), ("batch_keys", ctypes.c_int64),
        ("batch_requests", ctypes.c_int32), ("status", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_vp = ctypes.c_void_p
_i32 = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64

# name -> (restype, argtypes); every symbol of include/turtle_c.h
SIGNATURES = {
    "turtle_create": (_i32, [_i32, _i64, _i32, _i32, ctypes.POINTER(_vp)]),
    "turtle_destroy": (_i32, [_vp]),
    "turtle_compression_enabled": (_i32, [_vp]),
    "turtle_sort_start": (_i32, [_vp, _vp, _vp, _i64]),
    "turtle_sort_sync": (_i32, [_vp, ctypes.POINTER(ctypes.c_double)]),
    "turtle_multisort": (_i32, [_vp, _i32, _i64, _i32, _i32, _i64, _vp, ctypes.POINTER(ctypes.c_double)]),
    "turtle_sort_device": (_i32, [_vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "turtle_sort_device_noalloc": (_i32, [_i32, _vp, _vp, _vp, _vp, _i64, _vp]),
    "turtle_merge_device": (_i32, [_i32, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp]),
    "turtle_segmented_sort_device": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "turtle_multisort_device": (_i32, [_vp, _i32, _i64, _vp, _vp, _i64, _vp]),
    "turtle_multisort_index": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp, _i32, ctypes.POINTER(ctypes.c_double)]),
    "turtle_multisort_index_device": (_i32, [_vp, _i32, _i64, _vp, _vp, _vp, _vp, _i64, _vp]),
    "turtle_synchronize": (_i32, [_vp, ctypes.POINTER(Stats)]),
    "turtle_query": (_i32, [_vp]),
    "turtle_segmented_graph_create": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, ctypes.POINTER(_vp)]),
    "turtle_graph_launch": (_i32, [_vp, _vp]),
    "turtle_graph_destroy": (_i32, [_vp]),
    "turtle_set_profiling": (_i32, [_vp, _i32]),
    "turtle_set_profiling_kinds": (_i32, [_vp, ctypes.c_uint32]),
    "turtle_get_profile": (_i32, [_vp, ctypes.POINTER(Profile)]),
    "turtle_get_tuning": (_i32, [_vp, ctypes.POINTER(Tuning)]),
    "turtle_set_tuning": (_i32, [_vp, ctypes.POINTER(Tuning)]),
    "turtle_autotune": (_i32, [_vp, _i64, _i32, _i32, ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_double)]),
    "turtle_autotune_segmented": (_i32, [_vp, _vp, _i64, _i32, _i32, ctypes.POINTER(Tuning),
                                         ctypes.POINTER(ctypes.c_double)]),
    "turtle_server_create": (_i32, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]),
    "turtle_server_submit": (_i32, [_vp, _i32, _vp, _vp, _i64, ctypes.POINTER(_u64)]),
    "turtle_server_wait": (_i32, [_vp, _u64, ctypes.POINTER(ctypes.c_double)]),
    "turtle_server_wait_info": (_i32, [_vp, _u64, ctypes.POINTER(RequestInfo)]),
    "turtle_server_test": (_i32, [_vp, _u64, ctypes.POINTER(_i32)]),
    "turtle_server_get_stats": (_i32, [_vp, ctypes.POINTER(ServerStats)]),
    "turtle_server_destroy": (_i32, [_vp]),
    "turtle_fill_keys": (_i32, [_i32, _vp, _i64, _u64, _i32, _vp]),
    "turtle_fill_iota": (_i32, [_vp, _i64, _vp]),
    "turtle_status_string": (ctypes.c_char_p, [_i32]),
    "turtle_last_error": (ctypes.c_char_p, []),
    "turtle_dtype_size": (_i32, [_i32]),
    "turtle_device_cus": (_i32, []),
    "turtle_check_device": (_i32, []),
}

_lib = None


def load():
    """Load libturtle_hip.so once; raises if it is missing (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C turtlesort_amd")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != OK:
        lib = load()
        msg = lib.turtle_last_error().decode(errors="replace")
        raise TurtleError(status, msg or lib.turtle_status_string(status).decode())
    return status
