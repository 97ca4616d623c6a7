"""TurtleSort for AMD Instinct MI355X (gfx950).

The sort runs in hand-written HIP kernels in libturtle_hip.so (built in-tree
from turtlesort_amd/csrc); this package is the Python mirror of the
reference's Turtle::TurtleSort API over that library's C ABI.
"""
from ._lib import TurtleError, load  # noqa: F401
from .api import SortServer, TurtleSort, fill_iota, fill_keys, dtype_code, merge_device  # noqa: F401

__all__ = ["TurtleSort", "SortServer", "TurtleError", "fill_keys", "fill_iota", "dtype_code", "merge_device", "load"]
