"""The C++ callers of the reference API (README sample, README benchmark,
main.cpp MultiSort driver) built with g++ against include/turtle-sort.hpp and
run on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "examples", "build")


def run(name, *args, timeout=300):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (run __graft_entry__.build())")
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-2000:], r.stderr[-2000:])
    return r


def test_readme_sample(torch_dev):
    r = run("readme_sample")
    assert r.returncode == 0
    assert r.stdout.splitlines()[0].split() == ["1", "3", "4", "5", "8", "9"]
    assert "Memory compression successful=0" in r.stdout


def test_readme_benchmark(torch_dev):
    r = run("benchmark", 2_000_000, 2)
    assert r.returncode == 0
    assert r.stdout.count("completed successfully") == 2


def test_main_multisort(torch_dev):
    r = run("multisort_main", 2)
    assert r.returncode == 0
    assert r.stdout.count("sort success") == 2


def test_multisort_with_index(torch_dev):
    """MultiSort<double, 2038, 32>(num, keys, index): keys and index equal
    std::stable_sort of the (key, index) pairs of every array"""
    r = run("multisort_index", 4000)
    assert r.returncode == 0
    assert "sort success" in r.stdout
