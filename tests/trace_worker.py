"""Runs the traced harness sequence with the tracing build of the library
(build/trace/libnvshmem_amd.so) in local mode (2 PEs, argv[2] workgroups
issuing it at once, default 1) and prints where the processed trace went.  Launched by test_gpu_trace.py in its own process so the
tracing and non-tracing builds never share a process."""
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "df-nvshmem-prototype_amd"))

import numpy as np  # noqa: E402

import nvshmem_amd as na  # noqa: E402


def main():
    reps = int(sys.argv[1])
    grid = int(sys.argv[2]) if len(sys.argv) > 2 else 1
    lib, hip = na.load(trace=True), na.Hip()
    assert lib.nvshmemx_start_local_pes(2, max(1, grid // 2)) == 0
    ints = lib.nvshmalloc(64 * 4)
    floats = lib.nvshmalloc(64 * 4)
    doubles = lib.nvshmalloc(64 * 8)
    flag = lib.nvshmalloc(64)
    for pe in range(2):
        for p, n in ((ints, 256), (floats, 256), (doubles, 512), (flag, 64)):
            hip.memset(lib.nvshmem_ptr(p, pe), 0, n)
        hip.htod(lib.nvshmem_ptr(flag, pe), np.array([7], np.int32))
    hip.sync()
    lib.nvshmemh_traced_ops(ints, floats, doubles, flag, 0, 1, reps, 64, grid, None)
    hip.sync()
    err = lib.nvshmemx_get_last_error()
    enabled = lib.nvshmemx_trace_enabled()
    lib.nvshmem_barrier_all()
    lib.nvstop_pes()
    print(json.dumps({"errors": err, "enabled": enabled, "khz": 0}), flush=True)


if __name__ == "__main__":
    main()
