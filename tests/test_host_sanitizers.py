"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, on the CPU.

The shared-memory bootstrap (df-nvshmem-prototype_amd/csrc/bootstrap.cpp) is
compiled with g++ -fsanitize=address,undefined together with
tests/host/bootstrap_check.cpp, which forks ranks that run allgather, bcast
and barrier rounds through one segment and check every byte; the trace
ring's host index math (csrc/trace_ring.h) with tests/host/trace_ring_check.cpp.
Any heap or stack error, any undefined behaviour (recovery disabled) or any
wrong byte fails the run."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "df-nvshmem-prototype_amd" / "csrc"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("host_sanitizers") / "bootstrap_check"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fno-sanitize-recover=undefined", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{ROOT / 'include'}", f"-I{CSRC}", "-idirafter", "/opt/conda/include",
           str(ROOT / "tests" / "host" / "bootstrap_check.cpp"), str(CSRC / "bootstrap.cpp"), "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-3000:]
    return exe


@pytest.mark.parametrize("nranks,rounds", [(2, 300), (4, 200), (16, 50)])
def test_shm_bootstrap_under_sanitizers(checker, nranks, rounds):
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:exitcode=23", UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([str(checker), str(nranks), str(rounds)], capture_output=True, text=True, timeout=120, env=env)
    assert p.returncode == 0 and p.stdout.startswith("ok"), (p.stdout[-1000:], p.stderr[-3000:])
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr, p.stderr[-3000:]


def test_trace_ring_index_math_under_sanitizers(tmp_path):
    """published-record counts and wrap-joining copies of the trace ring, for
    capacities 1..1024 and claim counts that wrap the 32-bit index"""
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = tmp_path / "trace_ring_check"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fno-sanitize-recover=undefined", f"-I{CSRC}", str(ROOT / "tests" / "host" / "trace_ring_check.cpp"),
           "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-3000:]
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0 and p.stdout.startswith("ok"), (p.stdout[-1000:], p.stderr[-3000:])
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr, p.stderr[-3000:]
