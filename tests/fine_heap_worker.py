"""Local-mode checks of the fine-grained symmetric heap (NVSHMEM_HEAP_KIND=fine),
run by test_gpu_rma.py in a process of its own: grid and per-thread puts and
gets against the oracle, p + wait round trips, a 256 MiB put's rate.  Prints
"fine heap ok" at the end."""
import os
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "df-nvshmem-prototype_amd"))
sys.path.insert(0, str(ROOT / "tests"))

import numpy as np  # noqa: E402

import nvshmem_amd as na  # noqa: E402
import oracle_py  # noqa: E402
import test_gpu_rma as T  # noqa: E402


def main():
    os.environ.update({"NVSHMEM_HEAP_SIZE": "320M", "NVSHMEM_HEAP_KIND": "fine", "NVSHMEM_WAIT_TIMEOUT_MS": "5000"})
    lib, hip, oracle = na.load(), na.Hip(), oracle_py.Oracle()
    assert lib.nvshmemx_start_local_pes(2, 1) == 0
    assert lib.nvshmemx_heap_kind() == 1
    for path in ("grid", "thread"):
        for dtype, nelem, toff in (("int", 1, 0), ("float", 4099, 4), ("double", (1 << 20) + 3, 8)):
            chunk = 0 if path == "grid" else max(1, nelem // 3)
            print("case", path, dtype, nelem, toff, flush=True)
            lib.nvshmcleanup()
            T._run_put(lib, hip, oracle, path=path, dtype=dtype, nelem=nelem, toff=toff, soff=0, chunk=chunk,
                       seed=T._seed("fine", path, dtype))
            lib.nvshmcleanup()
            T._run_get(lib, hip, oracle, path=path, dtype=dtype, nelem=nelem, toff=toff, soff=0, chunk=chunk,
                       seed=T._seed(dtype, "fine", path))
    lib.nvshmcleanup()
    iters = 2000
    flag = lib.nvshmalloc(256)
    for pe in (0, 1):
        hip.memset(lib.nvshmem_ptr(flag, pe), 0, 256)
    samples = hip.malloc(iters * 8)
    print("case pingpong", flush=True)
    lib.nvshmemh_pingpong_local(flag, 0, 1, iters, None, 0, samples, None)
    hip.sync()
    assert lib.nvshmemx_get_last_error() == 0
    assert int(hip.dtoh(lib.nvshmem_ptr(flag, 0), 4, np.int32)[0]) == iters
    rtt = np.percentile(hip.dtoh(samples, iters * 8, np.int64) * 1e6 / lib.nvshmemx_wallclock_khz(), 50)
    # (barrier episodes on the fine-grained heap run between processes in
    # test_gpu_multiprocess.py: local mode launches them cooperatively, which
    # a child process cannot do while its parent holds the device's
    # cooperative-launch resources)
    # bulk rate into the fine-grained heap: 256 MiB float puts, PE 0 -> PE 1
    lib.nvshmcleanup()
    nbytes = 256 << 20
    dst = lib.nvshmalloc(nbytes)
    src = hip.malloc(nbytes)
    lib.nvshmemh_fill_pcg32(src, nbytes // 4, 3, 54, None)
    e0, e1 = hip.event(), hip.event()
    for _ in range(2):
        lib.nvshmemx_float_put_on_stream(dst, src, nbytes // 4, 1, None)
    hip.record(e0, None)
    for _ in range(10):
        lib.nvshmemx_float_put_on_stream(dst, src, nbytes // 4, 1, None)
    hip.record(e1, None)
    ms = hip.elapsed_ms(e0, e1) / 10
    assert np.array_equal(hip.dtoh(lib.nvshmem_ptr(dst, 1), nbytes, np.uint32), hip.dtoh(src, nbytes, np.uint32))
    print(f"fine heap: p+wait p50 {rtt:.0f} ns, 256 MiB put {2 * nbytes / (ms * 1e-3) / 1e9:.0f} GB/s", flush=True)
    lib.nvstop_pes()
    print("fine heap ok", flush=True)


if __name__ == "__main__":
    main()
