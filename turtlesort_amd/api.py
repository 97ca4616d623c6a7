"""Python mirror of the reference's Turtle::TurtleSort API over the C ABI.

Names and argument meaning follow turtle-sort.cuh:
  TurtleSort(dtype, max_elements, compress=False, track_index=True)  (63-113)
  .start_sorting(keys, idx=None)   StartSorting (124-184), host numpy arrays,
                                   sorted in place by .sync()
  .sync() -> seconds               Sync (188-258), wall time incl. transfers
  .multi_sort(arr_size, data, block_size=32, use_shared_memory=False)
                                   MultiSort (265-276), in place, -> seconds
  .memory_compression_successful() MemoryCompressionSuccessful (115-118)
plus device-resident forms that take torch tensors already on the GPU.
Errors raise TurtleError instead of exiting the process (the reference calls
exit() from gpuErrchk, turtle-globals.cuh:18-30).
"""
import ctypes
import weakref

import numpy as np

from . import _lib

_NP_TO_CODE = {
    np.dtype(np.int8): _lib.INT8, np.dtype(np.uint8): _lib.UINT8,
    np.dtype(np.int16): _lib.INT16, np.dtype(np.uint16): _lib.UINT16,
    np.dtype(np.int32): _lib.INT32, np.dtype(np.uint32): _lib.UINT32,
    np.dtype(np.int64): _lib.INT64, np.dtype(np.uint64): _lib.UINT64,
    np.dtype(np.float32): _lib.FLOAT32, np.dtype(np.float64): _lib.FLOAT64,
}
CODE_TO_NP = {v: k for k, v in _NP_TO_CODE.items()}
NAME_TO_CODE = {str(k): v for k, v in _NP_TO_CODE.items()}


def dtype_code(dtype):
    """numpy dtype / name / torch dtype / code -> turtle dtype code."""
    if isinstance(dtype, int):
        if dtype not in CODE_TO_NP:
            raise ValueError(f"unknown dtype code {dtype}")
        return dtype
    try:
        import torch
        if isinstance(dtype, torch.dtype):
            m = {torch.int8: _lib.INT8, torch.uint8: _lib.UINT8, torch.int16: _lib.INT16,
                 torch.int32: _lib.INT32, torch.int64: _lib.INT64,
                 torch.float32: _lib.FLOAT32, torch.float64: _lib.FLOAT64}
            for name in ("uint16", "uint32", "uint64"):
                td = getattr(torch, name, None)
                if td is not None:
                    m[td] = NAME_TO_CODE[name]
            return m[dtype]
    except ImportError:
        pass
    return _NP_TO_CODE[np.dtype(dtype)]


def _ptr(t):
    """device pointer of a torch tensor, or None"""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _check_tensors(specs):
    """Argument checks of the device forms, before any pointer reaches the C
    ABI.  specs: (tensor, name, key type code or None, min_numel, is_index).
    Every tensor's element type (int32 for an index; for keys the key
    width of the handle's type, since keys may be passed as a same-width
    bit view, e.g. uint32 keys in an int32 tensor), length and contiguity
    are checked first, then that all are on the GPU.  Raises TypeError or
    ValueError."""
    import torch
    for t, name, code, min_numel, index in specs:
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch tensor")
        if index:
            if t.dtype != torch.int32:
                raise TypeError(f"{name} must be int32, got {t.dtype}")
        elif code is not None and t.element_size() != CODE_TO_NP[code].itemsize:
            raise TypeError(f"{name} has dtype {t.dtype}, the sorter's key type is {CODE_TO_NP[code]}")
        if t.numel() < min_numel:
            raise ValueError(f"{name} holds {t.numel()} elements, needs at least {min_numel}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
    for t, name, _, _, _ in specs:
        if not t.is_cuda:
            raise ValueError(f"{name} must be a GPU tensor")


def validate_sort_args(key_code, track_index, n, keys_in, keys_out, idx_in=None, idx_out=None):
    """checks of a device sort over the first n elements of keys_in"""
    if (idx_in is None) != (idx_out is None):
        raise ValueError("idx_in and idx_out must both be given or both be None")
    if idx_in is not None and not track_index:
        raise ValueError("the sorter was created without index tracking")
    specs = [(keys_in, "keys_in", key_code, n, False), (keys_out, "keys_out", key_code, n, False)]
    if idx_in is not None:
        specs += [(idx_in, "idx_in", None, n, True), (idx_out, "idx_out", None, n, True)]
    _check_tensors(specs)


def validate_offsets(offsets):
    """segment offsets: 1-D int64, starting at >= 0, non-decreasing; -> array"""
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    if off.ndim != 1 or off.size < 1:
        raise ValueError("offsets must be a 1-D array of num_segments + 1 entries")
    if off[0] < 0 or (off.size > 1 and np.any(off[1:] < off[:-1])):
        raise ValueError("offsets must start at >= 0 and be non-decreasing")
    return off


def validate_multisort_args(key_code, arr_size, keys_in, keys_out, idx_in=None, idx_out=None):
    """checks of a device MultiSort; returns the number of arrays"""
    if int(arr_size) <= 0:
        raise ValueError("arr_size must be positive")
    n = keys_in.numel()
    if n % int(arr_size):
        raise ValueError(f"keys_in holds {n} keys, not a multiple of arr_size {arr_size}")
    specs = [(keys_in, "keys_in", key_code, n, False), (keys_out, "keys_out", key_code, n, False)]
    if idx_in is not None:
        specs.append((idx_in, "idx_in", None, n, True))
    if idx_out is not None:
        specs.append((idx_out, "idx_out", None, n, True))
    _check_tensors(specs)
    return n // int(arr_size)


def _stream(stream):
    """hipStream_t to run on: the given one, else torch's current stream, so
    that device work is ordered after the torch ops that produced the inputs"""
    if stream is not None:
        return ctypes.c_void_p(stream)
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class TurtleSort:
    def __init__(self, dtype, max_elements, compress=False, track_index=True):
        self._lib = _lib.load()
        self.dtype = dtype_code(dtype)
        self.np_dtype = CODE_TO_NP[self.dtype]
        self.max_n = int(max_elements)
        self.track_index = bool(track_index)
        h = ctypes.c_void_p()
        _lib.check(self._lib.turtle_create(self.dtype, self.max_n, int(self.track_index),
                                           int(bool(compress)), ctypes.byref(h)))
        self._h = h
        self._pending = None
        self._graphs = weakref.WeakSet()

    # ---- lifetime --------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            # graphs use the handle's buffers: they go first
            for g in list(getattr(self, "_graphs", ())):
                g.close()
            _lib.check(self._lib.turtle_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- reference API ------------------------------------------------------------
    def memory_compression_successful(self):
        return bool(self._lib.turtle_compression_enabled(self._h))

    def start_sorting(self, keys, idx=None):
        if not isinstance(keys, np.ndarray) or keys.dtype != self.np_dtype or not keys.flags.c_contiguous:
            raise TypeError(f"keys must be a C-contiguous numpy array of {self.np_dtype}")
        if idx is not None:
            if not isinstance(idx, np.ndarray) or idx.dtype != np.int32 or not idx.flags.c_contiguous:
                raise TypeError("idx must be a C-contiguous numpy int32 array")
            if idx.shape[0] != keys.shape[0]:
                raise ValueError("keys and idx must have the same length")
        _lib.check(self._lib.turtle_sort_start(
            self._h, keys.ctypes.data_as(ctypes.c_void_p),
            None if idx is None else idx.ctypes.data_as(ctypes.c_void_p), keys.shape[0]))
        self._pending = (keys, idx)   # keep the arrays alive until sync()

    def sync(self):
        sec = ctypes.c_double(0.0)
        try:
            _lib.check(self._lib.turtle_sort_sync(self._h, ctypes.byref(sec)))
        finally:
            self._pending = None
        return sec.value

    def multi_sort(self, arr_size, data, block_size=32, use_shared_memory=False):
        if not isinstance(data, np.ndarray) or data.dtype != self.np_dtype or not data.flags.c_contiguous:
            raise TypeError(f"data must be a C-contiguous numpy array of {self.np_dtype}")
        if arr_size <= 0 or data.size % arr_size:
            raise ValueError("data length must be a multiple of arr_size")
        sec = ctypes.c_double(0.0)
        _lib.check(self._lib.turtle_multisort(
            self._h, self.dtype, int(arr_size), int(block_size), int(bool(use_shared_memory)),
            data.size // arr_size, data.ctypes.data_as(ctypes.c_void_p), ctypes.byref(sec)))
        return sec.value

    def multi_sort_index(self, arr_size, data, idx=None):
        """MultiSort with index tracking (turtle_multisort_index): sorts the
        arrays in place and returns the index array moved with the keys --
        idx's values (also in place), or, when idx is None, each key's
        position within its array"""
        if not isinstance(data, np.ndarray) or data.dtype != self.np_dtype or not data.flags.c_contiguous:
            raise TypeError(f"data must be a C-contiguous numpy array of {self.np_dtype}")
        if arr_size <= 0 or data.size % arr_size:
            raise ValueError("data length must be a multiple of arr_size")
        local = idx is None
        if local:
            idx = np.empty(data.size, dtype=np.int32)
        elif idx.dtype != np.int32 or idx.size != data.size or not idx.flags.c_contiguous:
            raise TypeError("idx must be a C-contiguous int32 array of the data's length")
        sec = ctypes.c_double(0.0)
        _lib.check(self._lib.turtle_multisort_index(
            self._h, self.dtype, int(arr_size), data.size // arr_size, data.ctypes.data_as(ctypes.c_void_p),
            idx.ctypes.data_as(ctypes.c_void_p), int(local), ctypes.byref(sec)))
        self.last_seconds = sec.value
        return idx

    # ---- device-resident forms (torch tensors on the GPU) --------------------------
    def sort_device(self, keys_in, keys_out, idx_in=None, idx_out=None, stream=None):
        n = keys_in.numel()
        validate_sort_args(self.dtype, self.track_index, n, keys_in, keys_out, idx_in, idx_out)
        _lib.check(self._lib.turtle_sort_device(
            self._h, _ptr(keys_in), _ptr(idx_in), _ptr(keys_out), _ptr(idx_out), n, _stream(stream)))

    def segmented_sort_device(self, keys_in, keys_out, offsets, idx_in=None, idx_out=None, stream=None):
        off = validate_offsets(offsets)
        validate_sort_args(self.dtype, self.track_index, int(off[-1]), keys_in, keys_out, idx_in, idx_out)
        _lib.check(self._lib.turtle_segmented_sort_device(
            self._h, _ptr(keys_in), _ptr(idx_in), _ptr(keys_out), _ptr(idx_out),
            off.ctypes.data_as(ctypes.c_void_p), off.size - 1, _stream(stream)))

    def segmented_graph(self, keys_in, keys_out, offsets, idx_in=None, idx_out=None):
        """hipGraph of one segmented engine run over these buffers
        (turtle_segmented_graph_create); replay with .launch().  The graph
        keeps this sorter alive; close() of the sorter closes its graphs."""
        off = validate_offsets(offsets)
        validate_sort_args(self.dtype, self.track_index, int(off[-1]), keys_in, keys_out, idx_in, idx_out)
        g = ctypes.c_void_p()
        _lib.check(self._lib.turtle_segmented_graph_create(
            self._h, _ptr(keys_in), _ptr(idx_in), _ptr(keys_out), _ptr(idx_out),
            off.ctypes.data_as(ctypes.c_void_p), off.size - 1, ctypes.byref(g)))
        sg = SegmentGraph(self._lib, g, self)
        self._graphs.add(sg)
        return sg

    def multi_sort_device(self, arr_size, keys_in, keys_out, stream=None):
        num = validate_multisort_args(self.dtype, arr_size, keys_in, keys_out)
        _lib.check(self._lib.turtle_multisort_device(
            self._h, self.dtype, int(arr_size), _ptr(keys_in), _ptr(keys_out), num, _stream(stream)))

    def multi_sort_index_device(self, arr_size, keys_in, keys_out, idx_out, idx_in=None, stream=None):
        if idx_out is None:
            raise ValueError("idx_out is required")
        num = validate_multisort_args(self.dtype, arr_size, keys_in, keys_out, idx_in, idx_out)
        _lib.check(self._lib.turtle_multisort_index_device(
            self._h, self.dtype, int(arr_size), _ptr(keys_in), _ptr(idx_in), _ptr(keys_out), _ptr(idx_out), num,
            _stream(stream)))

    def query(self):
        """True while the sorter's last device work still runs (turtle_query)"""
        r = self._lib.turtle_query(self._h)
        if r < 0:
            _lib.check(-r)
        return bool(r)

    def synchronize(self):
        st = _lib.Stats()
        _lib.check(self._lib.turtle_synchronize(self._h, ctypes.byref(st)))
        return st.as_dict()

    # ---- per-kernel launch timing ---------------------------------------------------
    def set_profiling(self, on=True, kinds=None):
        """per-kernel launch timing on/off; kinds: names from
        _lib.KERNEL_KINDS to time only those launches"""
        if on and kinds is not None:
            mask = 0
            for k in kinds:
                mask |= 1 << _lib.KERNEL_KINDS.index(k)
            _lib.check(self._lib.turtle_set_profiling_kinds(self._h, mask))
        else:
            _lib.check(self._lib.turtle_set_profiling(self._h, int(bool(on))))

    def get_profile(self):
        """{kernel kind: {"ms": summed device ms, "launches": n}} since the last call"""
        pr = _lib.Profile()
        _lib.check(self._lib.turtle_get_profile(self._h, ctypes.byref(pr)))
        return pr.as_dict()

    # ---- tuning -------------------------------------------------------------------
    def get_tuning(self):
        t = _lib.Tuning()
        _lib.check(self._lib.turtle_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def autotune(self, n, with_index=True, reps=3):
        """fastest tuning on n synthetic keys in device memory -> (dict, ms)"""
        t = _lib.Tuning()
        ms = ctypes.c_double(0.0)
        _lib.check(self._lib.turtle_autotune(self._h, int(n), int(bool(with_index)), int(reps),
                                             ctypes.byref(t), ctypes.byref(ms)))
        return t.as_dict(), ms.value

    def autotune_segments(self, offsets, with_index=True, reps=3):
        """fastest tuning of a segmented batch laid out by `offsets` (host
        int64, offsets[0] == 0) on synthetic keys in device memory -> (dict, ms)"""
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        if off.ndim != 1 or off.size < 2:
            raise ValueError("offsets needs at least one segment")
        t = _lib.Tuning()
        ms = ctypes.c_double(0.0)
        _lib.check(self._lib.turtle_autotune_segmented(self._h, off.ctypes.data_as(ctypes.c_void_p), int(off.size - 1),
                                                       int(bool(with_index)), int(reps), ctypes.byref(t),
                                                       ctypes.byref(ms)))
        return t.as_dict(), ms.value

    def set_tuning(self, **kw):
        t = _lib.Tuning()
        _lib.check(self._lib.turtle_get_tuning(self._h, ctypes.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        _lib.check(self._lib.turtle_set_tuning(self._h, ctypes.byref(t)))


class SegmentGraph:
    """A captured segmented sort; launch() replays it on a stream."""

    def __init__(self, lib, g, sorter):
        self._lib = lib
        self._g = g
        self._sorter = sorter   # the graph replays into the sorter's buffers

    def launch(self, stream=None):
        _lib.check(self._lib.turtle_graph_launch(self._g, _stream(stream)))

    def close(self):
        if getattr(self, "_g", None) is not None and self._g.value:
            self._lib.turtle_graph_destroy(self._g)
            self._g = None
        self._sorter = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def merge_device(keys_a, keys_b, keys_out, idx_a=None, idx_b=None, idx_out=None, dtype=None, stream=None):
    """Stable merge of two sorted device runs into keys_out (turtle_merge_device):
    the reference's mergeSortedChunks rank rule, run a first on equal keys.
    idx_a, idx_b, idx_out: all None or all int32 tensors moved with the keys.
    dtype: the key type (default: the tensors' dtype; give it when the keys
    are a same-width bit view, e.g. uint32 keys in an int32 tensor)."""
    lib = _lib.load()
    code = dtype_code(keys_a.dtype if dtype is None else dtype)
    na, nb = keys_a.numel(), keys_b.numel()
    if (idx_a is None) != (idx_out is None) or (idx_b is None) != (idx_out is None):
        raise ValueError("idx_a, idx_b and idx_out must all be given or all be None")
    specs = [(keys_a, "keys_a", code, na, False), (keys_b, "keys_b", code, nb, False),
             (keys_out, "keys_out", code, na + nb, False)]
    if idx_out is not None:
        specs += [(idx_a, "idx_a", None, na, True), (idx_b, "idx_b", None, nb, True),
                  (idx_out, "idx_out", None, na + nb, True)]
    _check_tensors(specs)
    _lib.check(lib.turtle_merge_device(code, _ptr(keys_a), _ptr(idx_a), na, _ptr(keys_b), _ptr(idx_b), nb,
                                       _ptr(keys_out), _ptr(idx_out), _stream(stream)))


def fill_keys(keys, seed, shift, stream=None):
    """keys[i] = splitmix64 draw i of seed >> shift (device tensor, in place)."""
    lib = _lib.load()
    _lib.check(lib.turtle_fill_keys(dtype_code(keys.dtype), _ptr(keys), keys.numel(), int(seed) & (2**64 - 1),
                                    int(shift), _stream(stream)))


def fill_iota(idx, stream=None):
    lib = _lib.load()
    _lib.check(lib.turtle_fill_iota(_ptr(idx), idx.numel(), _stream(stream)))


class SortServer:
    """Request loop with continuous batching (turtle_server_* of the C ABI):
    submit() queues a host array (sorted in place, like start_sorting) and
    returns a ticket; wait(ticket) blocks until it is sorted and returns the
    latency in seconds.  The arrays must stay alive until waited for."""

    def __init__(self, pool_bytes=0, slot_keys=0, slots_per_dtype=0, dtypes=None, inflight_bytes=0):
        self._lib = _lib.load()
        cfg = _lib.ServerConfig()
        cfg.pool_bytes = int(pool_bytes)
        cfg.slot_keys = int(slot_keys)
        cfg.slots_per_dtype = int(slots_per_dtype)
        cfg.inflight_bytes = int(inflight_bytes)
        cfg.dtype_mask = 0 if dtypes is None else sum(1 << dtype_code(d) for d in dtypes)
        h = ctypes.c_void_p()
        _lib.check(self._lib.turtle_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._arrays = {}

    def submit(self, keys, idx=None):
        if not isinstance(keys, np.ndarray) or not keys.flags.c_contiguous:
            raise TypeError("keys must be a C-contiguous numpy array")
        if idx is not None and (idx.dtype != np.int32 or not idx.flags.c_contiguous or idx.size != keys.size):
            raise TypeError("idx must be a C-contiguous int32 array of the keys' length")
        t = ctypes.c_uint64(0)
        _lib.check(self._lib.turtle_server_submit(
            self._h, dtype_code(keys.dtype), keys.ctypes.data_as(ctypes.c_void_p),
            None if idx is None else idx.ctypes.data_as(ctypes.c_void_p), keys.size, ctypes.byref(t)))
        self._arrays[t.value] = (keys, idx)
        return t.value

    def wait(self, ticket):
        lat = ctypes.c_double(0.0)
        try:
            _lib.check(self._lib.turtle_server_wait(self._h, ctypes.c_uint64(ticket), ctypes.byref(lat)))
        finally:
            self._arrays.pop(ticket, None)
        return lat.value

    def wait_info(self, ticket):
        """as wait(), returning the request's queueing and batch figures"""
        info = _lib.RequestInfo()
        try:
            _lib.check(self._lib.turtle_server_wait_info(self._h, ctypes.c_uint64(ticket), ctypes.byref(info)))
        finally:
            self._arrays.pop(ticket, None)
        return info.as_dict()

    def stats(self):
        st = _lib.ServerStats()
        _lib.check(self._lib.turtle_server_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.turtle_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
