"""The batched-put planner's host half (csrc/batch_plan.h) on the CPU, built
with AddressSanitizer and UndefinedBehaviorSanitizer into
tests/host/planner_check.cpp, which checks coverage, piece boundaries, offset
tables, sets, the automatic grid size, Graham's bound for the LPT lists and
the order of source-order runs over random request lists, and that a plan with
a planted mistake fails those checks."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "df-nvshmem-prototype_amd" / "csrc"


def test_planner_invariants_under_sanitizers(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = tmp_path / "planner_check"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fno-sanitize-recover=undefined", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{ROOT / 'include'}", f"-I{CSRC}", str(ROOT / "tests" / "host" / "planner_check.cpp"), "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="exitcode=23", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), "10"], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-3000:])
    assert "planted mistakes caught: yes" in r.stdout and r.stdout.strip().endswith("0 failed checks"), r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
