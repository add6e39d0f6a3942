#!/usr/bin/env python3
"""Benchmark of the device-initiated OpenSHMEM subset on one MI355X.

Workload (one "step"): PE 0 puts 1 GiB of float32 (PCG32-seeded words) into
PE 1's symmetric heap with the library's host-launched grid put
(nvshmemx_float_put_on_stream).  PE 0 is this process; PE 1 is a child process
on the same GPU whose heap is IPC-mapped into PE 0 (shared-memory bootstrap),
i.e. "2 PEs as 2 processes on one MI355X".  Inputs are resident in HBM before
the timed region.  `value` counts algorithmic HBM bytes, 2 * n * 4 per put
(one read + one write of the same HBM), over the average kernel time measured
with HIP events on the launch stream.

Also reported on the same line: get GB/s, p+wait round-trip p50/p99 between
the two processes, the hipMemcpy device-to-device ceiling, the roofline object
and the CPU baseline (the same put semantics as host MPI-3 RMA on 2 cores).

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--layout process|local]
"""
import argparse
import json
import os
import subprocess
import sys
import time
import uuid
from pathlib import Path

ROOT = Path(__file__).resolve().parent
PKG = ROOT / "df-nvshmem-prototype_amd"
sys.path.insert(0, str(PKG))

NBYTES = 1 << 30                   # 1 GiB put (largest config-2 size)
DTYPE = "float"
HEAP = "1088M"                     # 1 GiB destination + flags, below the 2 GiB IPC limit
SEED = 0x5EED
PINGPONG_ITERS = 10000
PEAK_HBM_GBPS = 8000.0             # MI355X HBM3E, vendor figure


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def worker(name):
    """PE 1: owns the destination heap; verifies what PE 0 put; answers the
    ping-pong.  Collective call sequence mirrors PE 0's."""
    import numpy as np
    import nvshmem_amd as na
    lib, hip = na.load(), na.Hip()
    lib.nvstart_pes()
    dst = lib.nvshmalloc(NBYTES)
    flag = lib.nvshmalloc(256)
    hip.memset(flag, 0, 256)
    hip.memset(dst, 0, NBYTES)
    hip.sync()
    lib.nvshmem_barrier_all()                       # 1: ready
    lib.nvshmem_barrier_all()                       # 2: PE 0 finished puts/gets
    ref = hip.malloc(NBYTES)
    lib.nvshmemh_fill_pcg32(ref, NBYTES // 4, SEED, 54, None)
    hip.sync()
    ok = bool(np.array_equal(hip.dtoh(dst, NBYTES, np.uint32), hip.dtoh(ref, NBYTES, np.uint32)))
    hip.free(ref)
    lib.nvshmem_barrier_all()                       # 3: start ping-pong
    lib.nvshmemh_pingpong_process(flag, 0, 0, PINGPONG_ITERS, None, 0, None, None)
    hip.sync()
    err = lib.nvshmemx_get_last_error()
    lib.nvshmem_barrier_all()                       # 4: done
    lib.nvstop_pes()
    print(json.dumps({"pe1_put_bitexact": ok, "pe1_errors": err}), flush=True)


def timed(hip, fn, steps, warmup):
    e0, e1 = hip.event(), hip.event()
    for _ in range(warmup):
        fn()
    hip.sync()
    hip.record(e0, None)
    for _ in range(steps):
        fn()
    hip.record(e1, None)
    ms = hip.elapsed_ms(e0, e1) / steps
    hip.event_destroy(e0)
    hip.event_destroy(e1)
    return ms


def cpu_baseline():
    """Same put semantics as host MPI-3 RMA (oracle/mpi_rma_baseline.c), 2 ranks
    on 2 host cores, bounded to ~10 s: 1 GiB float MPI_Put + flush, repeated."""
    exe = ROOT / "oracle" / "build" / "rma_baseline"
    mpirun = Path("/opt/conda/bin/mpirun")
    if not exe.exists() or not mpirun.exists():
        return {"value": None, "unit": "GB/s", "cores": 2, "kind": "port",
                "sample": f"unavailable ({exe.name} or mpirun missing)"}
    cmd = [str(mpirun), "-n", "2", "-bind-to", "core", str(exe), "putget", "put", "float", str(NBYTES), str(NBYTES),
           "1000", "10"]
    try:
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd=ROOT)
        rec = [json.loads(l) for l in out.stdout.splitlines() if l.startswith("{")][-1]
        return {"value": round(rec["gbps"], 2), "unit": "GB/s", "cores": 2, "kind": "port",
                "sample": f"{rec['reps']} x 1 GiB float MPI_Put+MPI_Win_flush over a shared window, 2 ranks "
                          f"(bit-exact: {rec['bitexact']})"}
    except Exception as exc:  # report, never fail the GPU measurement
        return {"value": None, "unit": "GB/s", "cores": 2, "kind": "port", "sample": f"failed: {exc}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--layout", choices=["process", "local"], default="process")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pe-worker", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pe_worker:
        worker(args.pe_worker)
        return
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: all PEs share one GPU by design; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import nvshmem_amd as na

    child = None
    name = f"nvshmem_bench_{os.getpid()}_{uuid.uuid4().hex[:8]}"
    if args.layout == "process":
        # start PE 1 before this process touches the GPU
        child = subprocess.Popen([sys.executable, __file__, "--pe-worker", name],
                                 env=na.shm_job_env(1, 2, name, {"NVSHMEM_HEAP_SIZE": HEAP}),
                                 stdout=subprocess.PIPE, text=True)
        os.environ.update(na.shm_job_env(0, 2, name, {"NVSHMEM_HEAP_SIZE": HEAP}))
    else:
        os.environ["NVSHMEM_HEAP_SIZE"] = HEAP
    result = {}
    try:
        lib, hip = na.load(), na.Hip()
        if args.layout == "process":
            lib.nvstart_pes()
        else:
            lib.nvshmemx_start_local_pes(2, 1)
        dst = lib.nvshmalloc(NBYTES)
        flag = lib.nvshmalloc(256)
        if args.layout == "local":
            hip.memset(lib.nvshmem_ptr(flag, 0), 0, 256)
            hip.memset(lib.nvshmem_ptr(flag, 1), 0, 256)
        else:
            hip.memset(flag, 0, 256)
        src = hip.malloc(NBYTES)
        back = hip.malloc(NBYTES)
        lib.nvshmemh_fill_pcg32(src, NBYTES // 4, SEED, 54, None)
        hip.sync()
        lib.nvshmem_barrier_all()                                        # 1
        n = NBYTES // 4
        _log(f"layout={args.layout} bootstrap={lib.nvshmemx_bootstrap_name().decode()} "
             f"npes={lib.nvshmem_n_pes()} timing {args.steps} puts of {NBYTES} B")
        put_ms = timed(hip, lambda: lib.nvshmemx_float_put_on_stream(dst, src, n, 1, None), args.steps, args.warmup)
        get_ms = timed(hip, lambda: lib.nvshmemx_float_get_on_stream(back, dst, n, 1, None), max(args.steps // 4, 5),
                       2)
        memcpy_ms = timed(hip, lambda: hip.check(hip.lib.hipMemcpyAsync(back, src, NBYTES, hip.D2D, None)),
                          max(args.steps // 4, 5), 2)
        hip.sync()
        get_ok = bool(np.array_equal(hip.dtoh(back, NBYTES, np.uint32), hip.dtoh(src, NBYTES, np.uint32)))
        if args.layout == "local":
            put_ok = bool(np.array_equal(hip.dtoh(lib.nvshmem_ptr(dst, 1), NBYTES, np.uint32),
                                         hip.dtoh(src, NBYTES, np.uint32)))
        lib.nvshmem_barrier_all()                                        # 2
        _log("puts/gets done; ping-pong")
        lib.nvshmem_barrier_all()                                        # 3
        samples = hip.malloc(PINGPONG_ITERS * 8)
        if args.layout == "process":
            lib.nvshmemh_pingpong_process(flag, 1, 1, PINGPONG_ITERS, None, 0, samples, None)
        else:
            lib.nvshmemh_pingpong_local(flag, 0, 1, PINGPONG_ITERS, None, 0, samples, None)
        hip.sync()
        s = hip.dtoh(samples, PINGPONG_ITERS * 8, np.int64)
        khz = lib.nvshmemx_wallclock_khz()
        err = lib.nvshmemx_get_last_error()
        lib.nvshmem_barrier_all()                                        # 4
        lib.nvstop_pes()
        if child is not None:
            out, _ = child.communicate(timeout=300)
            peer = json.loads([l for l in out.splitlines() if l.startswith("{")][-1])
            put_ok = peer["pe1_put_bitexact"]
            err |= peer["pe1_errors"]
        rtt_ns = s.astype(np.float64) * 1e6 / khz
        put_gbps = 2 * NBYTES / (put_ms * 1e-3) / 1e9
        result = dict(
            put_ms=put_ms, put_gbps=put_gbps, get_gbps=2 * NBYTES / (get_ms * 1e-3) / 1e9,
            memcpy_gbps=2 * NBYTES / (memcpy_ms * 1e-3) / 1e9, put_ok=put_ok, get_ok=get_ok, errors=err,
            rtt_p50=float(np.percentile(rtt_ns, 50)), rtt_p99=float(np.percentile(rtt_ns, 99)))
    finally:
        if child is not None and child.poll() is None:
            child.kill()
        shm = Path("/dev/shm") / name
        if shm.exists():
            shm.unlink()

    cpu = None if args.no_cpu_baseline else cpu_baseline()
    r = result
    line = {
        "metric": "put/get GB/s & HBM-roofline frac; p50 p+wait ns; batched vs per-call puts",
        "value": round(r["put_gbps"], 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["put_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (PCG32-seeded 32-bit words, resident in HBM)",
        "config": {"workload": "bulk put of 1 GiB float32, PE 0 -> PE 1, 2 PEs "
                               + ("as 2 processes on one MI355X (IPC-mapped heaps)" if args.layout == "process"
                                  else "in one process on one MI355X"),
                   "bytes_per_put": NBYTES, "layout": args.layout, "parallelism": "1 GPU, 2 PEs"},
        "roofline": {"bound": "hbm", "achieved": round(r["put_gbps"], 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(r["put_gbps"] / PEAK_HBM_GBPS, 4), "traffic": None},
        "cpu_baseline": cpu,
        "get_gbps": round(r["get_gbps"], 1),
        "hip_memcpy_d2d_gbps": round(r["memcpy_gbps"], 1),
        "p_wait_rtt_ns": {"p50": round(r["rtt_p50"], 1), "p99": round(r["rtt_p99"], 1),
                          "iters": PINGPONG_ITERS, "between": "2 processes" if args.layout == "process" else
                          "2 workgroup PEs"},
        "bitexact": {"put": r["put_ok"], "get": r["get_ok"]},
        "device_errors": r["errors"],
    }
    print(json.dumps(line), flush=True)
    if not (r["put_ok"] and r["get_ok"]) or r["errors"]:
        sys.exit(1)


if __name__ == "__main__":
    main()
