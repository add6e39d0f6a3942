"""IDEALIZE build (reference IDEALIZE=1, Makefile:14-16; nvshmem_device.cu:433-435,
450-456, 474-481, 502-508, 529-535, 544, 566): p, g, put, get, quiet and the
waits compile to nothing and g returns 0, so an application's compute can be
timed alone; host-launched extension copies still move data."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
HERE = Path(__file__).resolve().parent


def test_idealized_device_api_moves_nothing():
    p = subprocess.run([sys.executable, str(HERE / "idealize_worker.py")], capture_output=True, text=True,
                       timeout=120, env={"PATH": "/usr/bin:/bin", "NVSHMEM_HEAP_SIZE": "16M",
                                         "HSA_ENABLE_IPC_MODE_LEGACY": "0"})
    assert p.returncode == 0, p.stderr[-2000:]
    r = json.loads([l for l in p.stdout.splitlines() if l.startswith("{")][-1])
    assert r["p_moved"] == 0 and r["put_moved"] == 0
    assert r["g_values"] == [0]
    assert r["wait_errors"] == 0          # the wait returned at once instead of timing out
    assert r["grid_put_ok"]
