"""Runs a traced harness kernel with the tracing build of the library
(build/trace/libnvshmem_amd.so) in local mode (2 PEs) and prints a JSON
status line.  argv: reps [grid [lanes block]] -- without `lanes` one thread
of each of `grid` workgroups (default 1) issues the traced operation
sequence; with it, `lanes` threads of each `block`-thread workgroup each
issue `reps` traced int puts.  Launched by test_gpu_trace.py in its own
process so the tracing and non-tracing builds never share a process."""
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
    lanes = int(sys.argv[3]) if len(sys.argv) > 3 else 0
    block = int(sys.argv[4]) if len(sys.argv) > 4 else 64
    lib, hip = na.load(trace=True), na.Hip()
    assert lib.nvshmemx_start_local_pes(2, max(1, grid // 2)) == 0
    ints = lib.nvshmalloc(1024 * 4)
    floats = lib.nvshmalloc(64 * 4)
    doubles = lib.nvshmalloc(64 * 8)
    flag = lib.nvshmalloc(64)
    for pe in range(2):
        for p, n in ((ints, 4096), (floats, 256), (doubles, 512), (flag, 64)):
            hip.memset(lib.nvshmem_ptr(p, pe), 0, n)
        hip.htod(lib.nvshmem_ptr(flag, pe), np.array([7], np.int32))
    hip.sync()
    if lanes:
        lib.nvshmemh_traced_lanes(ints, 1, lanes, reps, block, grid, None)
    else:
        lib.nvshmemh_traced_ops(ints, floats, doubles, flag, 0, 1, reps, 64, grid, None)
    hip.sync()
    err = lib.nvshmemx_get_last_error()
    enabled = lib.nvshmemx_trace_enabled()
    lib.nvshmem_barrier_all()
    lib.nvstop_pes()
    print(json.dumps({"errors": err, "enabled": enabled, "khz": 0}), flush=True)


if __name__ == "__main__":
    main()
