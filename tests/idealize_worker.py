"""Runs device API calls against the IDEALIZE build (build/idealize) in local
mode and reports what moved.  Launched by test_gpu_idealize.py in its own
process."""
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "df-nvshmem-prototype_amd"))

import numpy as np  # noqa: E402

import nvshmem_amd as na  # noqa: E402


def main():
    lib, hip = na.load(variant="idealize"), na.Hip()
    assert lib.nvshmemx_start_local_pes(2, 1) == 0
    lib.nvshmemx_set_wait_timeout_ms(200)
    n = 1000
    a, b, flag = lib.nvshmalloc(n * 4), lib.nvshmalloc(n * 4), lib.nvshmalloc(64)
    vals = hip.malloc(n * 4)
    out = hip.malloc(n * 4)
    hip.htod(vals, np.arange(1, n + 1, dtype=np.int32))
    for pe in range(2):
        for p in (a, b, flag):
            hip.memset(lib.nvshmem_ptr(p, pe), 0, 64 if p == flag else n * 4)
    hip.htod(lib.nvshmem_ptr(b, 1), np.full(n, 7, np.int32))
    hip.htod(lib.nvshmem_ptr(a, 0), np.arange(n, dtype=np.int32))
    res = {}
    lib.nvshmemh_pg(0, 0, a, vals, None, n, 1, None)            # p into PE 1
    hip.sync()
    res["p_moved"] = int(np.count_nonzero(hip.dtoh(lib.nvshmem_ptr(a, 1), n * 4, np.int32)))
    lib.nvshmemh_pg(1, 0, b, None, out, n, 1, None)             # g from PE 1 (holds 7s)
    hip.sync()
    res["g_values"] = sorted(set(hip.dtoh(out, n * 4, np.int32).tolist()))
    lib.nvshmemh_thread_putget(0, 0, a, lib.nvshmem_ptr(a, 0), n, 100, 1, None)   # per-thread put
    hip.sync()
    res["put_moved"] = int(np.count_nonzero(hip.dtoh(lib.nvshmem_ptr(a, 1), n * 4, np.int32)))
    w = hip.malloc(4)
    lib.nvshmemh_wait_once(lib.nvshmem_ptr(flag, 0), 99, 1, w, None)  # wait_until 99 never set
    hip.sync()
    res["wait_errors"] = lib.nvshmemx_get_last_error()
    lib.nvshmemx_putmem_on_stream(a, lib.nvshmem_ptr(a, 0), n * 4, 1, None)       # extension still moves data
    hip.sync()
    res["grid_put_ok"] = bool(np.array_equal(hip.dtoh(lib.nvshmem_ptr(a, 1), n * 4, np.int32),
                                             np.arange(n, dtype=np.int32)))
    lib.nvstop_pes()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
