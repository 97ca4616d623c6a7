"""The work-queue protocol of the persistent engine that preceded the
level-synchronous one (git 7b6f2b4, turtlesort_amd/csrc/ts_engine.hip),
emulated with one host thread per workgroup under ThreadSanitizer and
AddressSanitizer (tests/emulation/queue_protocol_emu.cpp).  CPU only.

What the runs establish:
- the queue protocol (slot generations, last-finisher counters, the
  outstanding-task count) has no data race and every run ends with no task
  outstanding and sorted output, for 1..8 workgroups, sizes 0..30,000, six
  key shapes, batched segments and injected scheduling jitter;
- the device entry points of that engine ran on the handle's non-blocking
  stream when given the null stream, so they were not ordered after the
  null-stream kernels that wrote their input: ThreadSanitizer reports that
  race between the producer and the engine's reads;
- the count and scatter passes classified the keys separately, so an input
  written between them makes the scatter's destinations leave the segment:
  the first version (no range checks) stores past the end of the next-level
  buffer (AddressSanitizer: heap-buffer-overflow in the scatter), which on
  the GPU is an illegal memory access; with the range checks the run ends
  with ERR_RANGE at the scatter's write-out check (site 9), not in a hang.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "emulation", "queue_protocol_emu.cpp")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


def _build(tmp_path_factory, sanitizer):
    out = str(tmp_path_factory.mktemp("emu") / f"emu_{sanitizer}")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", f"-fsanitize={sanitizer}", "-fno-omit-frame-pointer",
           "-o", out, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


@pytest.fixture(scope="module")
def emu_tsan(tmp_path_factory):
    return _build(tmp_path_factory, "thread")


@pytest.fixture(scope="module")
def emu_asan(tmp_path_factory):
    return _build(tmp_path_factory, "address")


def _run(exe, *args, timeout=240):
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=0 exitcode=66 history_size=4"
    env["ASAN_OPTIONS"] = "detect_leaks=0"
    return subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=timeout, env=env)


def _results(out):
    rows = []
    for line in out.splitlines():
        if line.startswith("result: "):
            rows.append(dict(kv.split("=", 1) for kv in line[len("result: "):].split()))
    return rows


def test_protocol_race_free_and_terminates(emu_tsan):
    r = _run(emu_tsan, "protocol", 30000, 8, 1)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = _results(r.stdout)
    assert len(rows) == 12 * 6 * 4 * 2
    for row in rows:
        assert row["error"] == "0" and row["outstanding"] == "0" and row["sorted"] == "1", row


def test_unordered_producer_is_a_data_race(emu_tsan):
    r = _run(emu_tsan, "race", 20000, 4, 1, 1)
    assert r.returncode == 66, r.stdout[-2000:] + r.stderr[-2000:]
    assert "WARNING: ThreadSanitizer: data race" in r.stderr
    assert "emu::produce" in r.stderr


def test_input_written_between_count_and_scatter_overflows_without_checks(emu_asan):
    r = _run(emu_asan, "between", 100000, 8, 1, 0)
    assert r.returncode != 0
    assert "AddressSanitizer: heap-buffer-overflow" in r.stderr, r.stderr[-4000:]
    report = r.stderr[r.stderr.index("heap-buffer-overflow"):]
    assert "WRITE of size 4" in report
    first_frame = [ln for ln in report.splitlines() if ln.strip().startswith("#0")][0]
    assert "emu::scatter" in first_frame


def test_input_written_between_count_and_scatter_with_checks_ends_in_range_error(emu_asan):
    r = _run(emu_asan, "between", 100000, 8, 1, 1)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "AddressSanitizer" not in r.stderr
    (row,) = _results(r.stdout)
    assert int(row["error"]) & 32 and row["bad_where"] == "9", row
    assert float(row["seconds"]) < 5.0
