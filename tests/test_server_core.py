"""Host-only tests of the proof server's scheduler (server_core.cpp).

The scheduler holds no HIP code; tests/cpu/server_core_test.cpp drives it
with mock provers whose completion it controls, so interleavings a GPU run
reaches only by chance are forced here: a proof finishing while a request of
a new class needs its prover's memory (the eviction must wait for the
copy-out), more classes cycled than the budget holds (shared tables released
with their last prover), tables shared by classes of one shape counted once,
device and launch failures, prepare() racing the dispatcher. The binary is
built with AddressSanitizer, so a read of a destroyed prover's result fails
the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "vortexstark_amd", "csrc")
SRC = os.path.join(ROOT, "tests", "cpu", "server_core_test.cpp")

SCENARIOS = ["evict_while_pending", "two_slots_alternating", "cycle_more_classes_than_fit",
             "shared_tables_counted_once", "rejects_what_cannot_fit", "device_and_launch_failures",
             "prepare_while_serving", "destructor_drains"]


@pytest.fixture(scope="module")
def binary(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("server_core") / "server_core_test")
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-pthread", "-I", CSRC, SRC, os.path.join(CSRC, "server_core.cpp"), "-o", out]
    subprocess.run(cmd, check=True)
    return out


@pytest.mark.parametrize("scenario", SCENARIOS)
def test_scheduler(binary, scenario):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([binary, scenario], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"PASS {scenario}" in r.stdout
