"""The C-ABI library loads and exports every symbol include/turtle_c.h
declares; host-side argument checks and error paths that need no GPU."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "turtle_c.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(turtle_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_all_declared_symbols():
    from turtlesort_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), f"{n} not exported"
    # the Python binding declares a signature for each of them
    assert set(names) == set(_lib.SIGNATURES)


def test_status_and_dtype_helpers():
    from turtlesort_amd import _lib
    lib = _lib.load()
    assert lib.turtle_status_string(0) == b"ok"
    assert lib.turtle_status_string(7) == b"work queue overflow"
    assert [lib.turtle_dtype_size(d) for d in range(1, 11)] == [1, 1, 2, 2, 4, 4, 8, 8, 4, 8]
    assert lib.turtle_dtype_size(0) == 0 and lib.turtle_dtype_size(11) == 0


def test_dtype_codes_match_oracle(orc):
    from turtlesort_amd import api
    for dt, code in orc.DTYPES.items():
        assert api.dtype_code(dt) == code


def test_argument_errors_without_device():
    """Bad arguments are rejected before any device call."""
    from turtlesort_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.turtle_create(99, 10, 1, 0, ctypes.byref(h)) == _lib.ERR_DTYPE
    assert lib.turtle_create(5, -1, 1, 0, ctypes.byref(h)) == _lib.ERR_INVALID
    assert lib.turtle_create(5, 1 << 31, 1, 0, ctypes.byref(h)) == _lib.ERR_INVALID
    assert lib.turtle_create(5, 10, 1, 0, None) == _lib.ERR_INVALID
    assert lib.turtle_sort_start(None, None, None, 0) == _lib.ERR_INVALID
    assert lib.turtle_sort_sync(None, None) == _lib.ERR_INVALID
    assert lib.turtle_fill_keys(42, None, 0, 0, 0, None) == _lib.ERR_DTYPE
    assert lib.turtle_fill_keys(5, None, -1, 0, 0, None) == _lib.ERR_INVALID
    assert lib.turtle_destroy(None) == _lib.OK


def test_create_reports_missing_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from turtlesort_amd import TurtleSort, TurtleError
    with pytest.raises(TurtleError) as e:
        TurtleSort(np.int32, 10)
    assert e.value.status == 6


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_cpp_dropin_header_compiles_with_gxx(tmp_path):
    """turtle-sort.hpp builds with plain g++ (no HIP headers) and links."""
    src = tmp_path / "t.cpp"
    src.write_text('#include "turtle-sort.hpp"\n'
                   'int main(){ Turtle::TurtleSort<long> s(10); std::vector<long> v{3,1,2};'
                   ' s.StartSorting(&v); s.Sync(); return v[0]==1?0:1; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                        "-o", str(tmp_path / "t"), "-L", os.path.join(ROOT, "turtlesort_amd"), "-lturtle_hip"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
