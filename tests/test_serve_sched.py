"""The serving loop's scheduler on the CPU (tests/cpp/serve_sched_check.cpp): fh_serve.cpp built with
the host compiler against host stand-ins of the HIP runtime calls and library operations it makes,
run over every mode, bootstrap placement (in the tick, on the worker thread, on the yielding
worker thread) and ticks_in_flight.
Every request completes with the same result as in the first session, latencies are ordered,
bootstrap callbacks run where asked, and a failing bootstrap surfaces as an error."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "serve_sched_check.cpp")
HDR = os.path.join(ROOT, "tests", "cpp", "host_intrinsics.h")
SERVE = os.path.join(ROOT, "fideslib-fork_amd", "csrc", "fh_serve.cpp")
ROCM_INC = "/opt/rocm/include"


def test_serving_scheduler_on_cpu(tmp_path):
    if not os.path.isdir(os.path.join(ROCM_INC, "hip")):
        pytest.skip("HIP headers not installed")
    exe = str(tmp_path / "serve_sched_check")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + ROCM_INC, "-include", HDR, "-o", exe, SRC, SERVE,
           "-lpthread"]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = r.stdout.strip().splitlines()
    assert lines[-1] == "ok"
    assert len(lines) == 1 + 3 * 3 * 3  # one line per (mode, placement, ticks_in_flight) session
