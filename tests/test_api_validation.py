"""The device wrappers of turtlesort_amd.api check their tensors before any
pointer reaches the C ABI (CPU only: the checks run before the GPU is used;
a CPU tensor that passes every other check is refused for not being on the
GPU)."""
import numpy as np
import pytest
import torch

from turtlesort_amd import _lib
from turtlesort_amd.api import validate_multisort_args, validate_offsets, validate_sort_args

I32 = _lib.INT32
I64 = _lib.INT64


def t(n, dtype=torch.int32):
    return torch.zeros(n, dtype=dtype)


def test_sort_args_dtype_mismatch():
    with pytest.raises(TypeError, match="dtype"):
        validate_sort_args(I64, True, 10, t(10, torch.int32), t(10, torch.int64))
    with pytest.raises(TypeError, match="dtype"):
        validate_sort_args(I32, True, 10, t(10), t(10, torch.int64))
    # a same-width bit view passes the type check (then fails placement)
    with pytest.raises(ValueError, match="GPU"):
        validate_sort_args(_lib.FLOAT32, True, 10, t(10), t(10))


def test_sort_args_short_output_or_index():
    with pytest.raises(ValueError, match="keys_out"):
        validate_sort_args(I32, True, 10, t(10), t(9))
    with pytest.raises(ValueError, match="idx_out"):
        validate_sort_args(I32, True, 10, t(10), t(10), t(10), t(5))


def test_sort_args_index_type_and_pairing():
    with pytest.raises(TypeError, match="int32"):
        validate_sort_args(I32, True, 10, t(10), t(10), t(10, torch.int64), t(10))
    with pytest.raises(ValueError, match="both"):
        validate_sort_args(I32, True, 10, t(10), t(10), t(10), None)
    with pytest.raises(ValueError, match="index tracking"):
        validate_sort_args(I32, False, 10, t(10), t(10), t(10), t(10))


def test_sort_args_contiguity_and_placement():
    strided = torch.zeros(20, dtype=torch.int32)[::2]
    with pytest.raises(ValueError, match="contiguous"):
        validate_sort_args(I32, True, 10, strided, t(10))
    with pytest.raises(ValueError, match="GPU"):
        validate_sort_args(I32, True, 10, t(10), t(10))
    with pytest.raises(TypeError, match="torch tensor"):
        validate_sort_args(I32, True, 10, np.zeros(10, np.int32), t(10))


def test_offsets():
    assert validate_offsets([0, 3, 3, 7]).dtype == np.int64
    with pytest.raises(ValueError):
        validate_offsets([0, 5, 4])
    with pytest.raises(ValueError):
        validate_offsets([-1, 4])
    with pytest.raises(ValueError):
        validate_offsets([])


def test_multisort_args():
    with pytest.raises(ValueError, match="multiple"):
        validate_multisort_args(I32, 7, t(20), t(20))
    with pytest.raises(ValueError, match="positive"):
        validate_multisort_args(I32, 0, t(20), t(20))
    with pytest.raises(ValueError, match="keys_out"):
        validate_multisort_args(I32, 10, t(20), t(10))
    with pytest.raises(TypeError, match="int32"):
        validate_multisort_args(I32, 10, t(20), t(20), None, t(20, torch.int64))
    with pytest.raises(ValueError, match="GPU"):
        validate_multisort_args(I32, 10, t(20), t(20))
