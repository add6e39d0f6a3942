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
and the CPU baseline (the same put semantics as host MPI-3 RMA on 2 cores);
the 4096-put batch (fused, tiles, persistent, posted, per-call), the learned
plan choice, the 8 B .. 1 GiB sweep, and the 2x2x2 Jacobi application at
256^3 per PE (one fused step per launch, two steps per launch, unfused
steps; one process and 8 processes).

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--layout process|local]
"""
import argparse
import json
import math
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
    payload = lib.nvshmalloc(PUTFLAG_SIZES[-1])
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
    lib.nvshmem_barrier_all()                       # 3b: barrier episodes
    lib.nvshmemh_barrier_latency(BARRIER_ITERS, None, None)
    hip.sync()
    for size in PUTFLAG_SIZES:                      # 3c: put + flag ping-pong per payload size
        hip.memset(flag, 0, 256)
        hip.sync()
        lib.nvshmem_barrier_all()
        lib.nvshmemh_pingpong_process(flag, 0, 0, PUTFLAG_ITERS, payload, size // 4, None, None)
        hip.sync()
    err = lib.nvshmemx_get_last_error()
    lib.nvshmem_barrier_all()                       # 4: done
    lib.nvstop_pes()
    print(json.dumps({"pe1_put_bitexact": ok, "pe1_errors": err}), flush=True)


BARRIER_ITERS, BARRIER_SKIP = 10000, 10     # barrier episodes timed on the device; the first few absorb launch skew
# the reference's put-latency microbenchmark: per-thread put of 4 B..64 KiB,
# quiet, flag p, quiet, wait; 20 warm-up + 200 timed round trips per size
PUTFLAG_SIZES = [4, 64, 1024, 4096, 16384, 65536]
STREAM_SIZES_LOG2 = [26, 27, 28, 29]  # 64..512 MiB bulk streams beside the 1 GiB one
PUTFLAG_ITERS, PUTFLAG_SKIP = 220, 20
BATCH_NPES, BATCH_WGS, BATCH_SRC, BATCH_DST, BATCH_REQS, BATCH_SEED = 64, 4, 32 << 20, 64 << 20, 4096, 7
POLICY_ONE_PIECE, POLICY_TILES = 11, 12  # batch policies: one workgroup piece per request; every request as tiles


def batch_worker(counter_pass=False):
    """Uneven-put workload on 64 workgroup-partitioned PEs in one process:
    4096 puts per step, log-uniform 64 B..4 MiB, seed 7.  Times one fused
    launch, the same under a captured hipGraph, the owner discipline (each PE's
    workgroups run the puts it initiated), every request cut into tiles (one
    workgroup a tile, launched and replayed from a hipGraph), the persistent
    kernel planned on the device by the issuing PEs (library-allocated and
    static scratch), the same puts posted on the device by the PEs' own
    workgroups and flushed, and the per-call path (one launch per put).
    Prints one JSON line.  counter_pass: only five fused launches, five
    persistent launches, five post-and-flush rounds, five tile launches and
    one 1 GiB device-to-device copy (the scale of a counter pass)."""
    import ctypes as C
    import numpy as np
    import nvshmem_amd as na
    os.environ["NVSHMEM_HEAP_SIZE"] = "96M"
    lib, hip = na.load(), na.Hip()
    lib.nvshmemx_start_local_pes(BATCH_NPES, BATCH_WGS)
    base = lib.nvshmalloc(BATCH_SRC + BATCH_DST)
    for pe in range(BATCH_NPES):
        p = lib.nvshmem_ptr(base, pe)
        lib.nvshmemh_fill_pcg32(p, BATCH_SRC // 4, 100 + pe, 54, None)
        hip.memset(p + BATCH_SRC, 0, BATCH_DST)
    hip.sync()
    reqs = na.gen_requests(lib, BATCH_SEED, BATCH_REQS, BATCH_NPES, BATCH_SRC, BATCH_DST)
    arr = na.make_requests(lib, reqs, base, 0, BATCH_SRC)
    n = len(arr)
    payload = int(reqs["nbytes"].sum())
    out = {"requests": n, "payload_bytes": payload, "pes": BATCH_NPES}

    def plan_of(launch=None, choices=None):
        plan = C.c_void_p()
        lib.nvshmemx_batch_plan_create(arr, n, choices, C.byref(launch) if launch is not None else None,
                                       C.byref(plan))
        return plan

    plan = plan_of()
    if counter_pass:
        for _ in range(5):
            lib.nvshmemx_batch_plan_launch(plan, None)
        raw = np.frombuffer(bytes(arr), np.uint8)
        d_reqs = hip.malloc(raw.size)
        hip.htod(d_reqs, raw)
        for _ in range(5):
            lib.nvshmemx_batch_put_persistent_on_stream(d_reqs, n, None)
        iss = np.ascontiguousarray(reqs["src_pe"], np.int32)
        d_iss = hip.malloc(iss.nbytes)
        hip.htod(d_iss, iss.view(np.uint8))
        for _ in range(5):
            lib.nvshmemh_post_nbi(d_reqs, d_iss, n, None)
            lib.nvshmemx_nbi_flush_on_stream(None)
        tile_choices = (na.BatchChoice * n)()
        for i in range(n):
            tile_choices[i].pieces, tile_choices[i].vec_bytes, tile_choices[i].mode = 1, 16, 2
        tiles = plan_of(None, tile_choices)
        for _ in range(5):
            lib.nvshmemx_batch_plan_launch(tiles, None)
        ref = hip.malloc(2 << 30)
        hip.check(hip.lib.hipMemcpyAsync(ref + (1 << 30), ref, 1 << 30, hip.D2D, None))
        hip.sync()
        out["bitexact"] = int(lib.nvshmemh_verify_requests(arr, n)) == 0
        lib.nvshmemx_batch_plan_destroy(plan)
        lib.nvstop_pes()
        print(json.dumps(out), flush=True)
        return
    _log(f"batch: {n} requests, {payload} payload bytes planned")
    out["fused_ms"] = timed(hip, lambda: lib.nvshmemx_batch_plan_launch(plan, None), 20, 3)
    out["pieces"] = int(lib.nvshmemx_batch_plan_pieces(plan))
    out["bitexact"] = int(lib.nvshmemh_verify_requests(arr, n)) == 0
    lib.nvshmemx_batch_plan_capture(plan, 10)
    out["graph_ms"] = timed(hip, lambda: lib.nvshmemx_batch_plan_graph_launch(plan, None), 3, 1) / 10
    lib.nvshmemx_batch_plan_destroy(plan)
    # every request cut into tiles, one workgroup a tile, in source order
    tile_choices = (na.BatchChoice * n)()
    for i in range(n):
        tile_choices[i].pieces, tile_choices[i].vec_bytes, tile_choices[i].mode = 1, 16, 2
    tiles = plan_of(None, tile_choices)
    for pe in range(BATCH_NPES):  # the check below then sees only what the tiles wrote
        hip.memset(lib.nvshmem_ptr(base, pe) + BATCH_SRC, 0, BATCH_DST)
    hip.sync()
    out["tiles_ms"] = timed(hip, lambda: lib.nvshmemx_batch_plan_launch(tiles, None), 20, 3)
    out["tiles"] = int(lib.nvshmemx_batch_plan_pieces(tiles))
    out["bitexact"] = out["bitexact"] and int(lib.nvshmemh_verify_requests(arr, n)) == 0
    lib.nvshmemx_batch_plan_capture(tiles, 10)
    out["tiles_graph_ms"] = timed(hip, lambda: lib.nvshmemx_batch_plan_graph_launch(tiles, None), 3, 1) / 10
    lib.nvshmemx_batch_plan_destroy(tiles)
    # each PE's puts on its own 128 workgroups (the default grid split by PE)
    owner = plan_of(na.BatchLaunch(8192, 0, 2, na.BATCH_OWNER))
    out["owner_ms"] = timed(hip, lambda: lib.nvshmemx_batch_plan_launch(owner, None), 10, 2)
    lib.nvshmemx_batch_plan_destroy(owner)
    # the persistent kernel: each workgroup PE plans its own puts on the
    # device from the request list in device memory (and the same kernel with
    # its scratch in the library's static array)
    raw = np.frombuffer(bytes(arr), np.uint8)
    d_reqs = hip.malloc(raw.size)
    hip.htod(d_reqs, raw)
    out["persistent_ms"] = timed(hip, lambda: lib.nvshmemx_batch_put_persistent_on_stream(d_reqs, n, None), 20, 3)
    out["persistent_noalloc_ms"] = timed(hip, lambda: lib.nvshmemx_batch_put_persistent_noalloc_on_stream(
        d_reqs, n, None), 20, 3)
    out["persistent_grid"] = int(lib.nvshmemx_batch_persistent_grid())
    out["bitexact"] = out["bitexact"] and int(lib.nvshmemh_verify_requests(arr, n)) == 0
    # the same puts posted on the device by the workgroup PEs that issue them
    # (nvshmemx_putmem_nbi from each PE's 4 workgroups), then one flush; the
    # posting kernel is timed alone as well
    iss = np.ascontiguousarray(reqs["src_pe"], np.int32)
    d_iss = hip.malloc(iss.nbytes)
    hip.htod(d_iss, iss.view(np.uint8))
    for pe in range(BATCH_NPES):  # the check below then sees only what the posts wrote
        hip.memset(lib.nvshmem_ptr(base, pe) + BATCH_SRC, 0, BATCH_DST)
    hip.sync()

    def post_and_flush():
        lib.nvshmemh_post_nbi(d_reqs, d_iss, n, None)
        lib.nvshmemx_nbi_flush_on_stream(None)
    out["posted_ms"] = timed(hip, post_and_flush, 20, 3)
    out["bitexact"] = out["bitexact"] and int(lib.nvshmemh_verify_requests(arr, n)) == 0
    e0, e1 = hip.event(), hip.event()
    post_ms = []
    for _ in range(20):
        hip.record(e0, None)
        lib.nvshmemh_post_nbi(d_reqs, d_iss, n, None)
        hip.record(e1, None)
        lib.nvshmemx_nbi_flush_on_stream(None)
        hip.sync()
        post_ms.append(hip.elapsed_ms(e0, e1))
    hip.event_destroy(e0)
    hip.event_destroy(e1)
    out["posted_post_kernel_ms"] = float(np.median(post_ms))
    hip.free(d_iss)
    hip.free(d_reqs)
    _log("batch: fused, graph, owner, persistent and posted timed; per-call path next")
    out["per_call_ms"] = timed(hip, lambda: lib.nvshmemx_put_requests_on_stream(arr, n, None), 3, 1)
    out["bitexact"] = out["bitexact"] and int(lib.nvshmemh_verify_requests(arr, n)) == 0
    out["errors"] = lib.nvshmemx_get_last_error()
    lib.nvstop_pes()
    print(json.dumps(out), flush=True)


def learned_worker(n_batches, train_steps):
    """Learned config choice on the uneven-put workload (64 workgroup PEs):
    1. per-request autotuner: every one of the 54 candidates measured at every
       point of a 1782-point feature grid (96,228 samples), and the
       per-request MLP trained on them on the GPU;
    2. batch-level autotuner: the 13 batch policies (the rule, the
       per-request network's choices, uniform wavefront / workgroup pieces of
       4 KiB .. 1 MiB, one piece per request, every request as tiles) of
       self-generated batches of 8
       to 8192 puts timed on the device, and the batch-level MLP trained on
       them (inputs: log2 batch bytes, log2 request count, mean and largest
       log2 request size);
    3. held-out batches of 4096 (`n_batches`), 128 and 16 puts: every policy
       timed on the device (the exhaustive batch-level judge) and the batch
       network's pick read from the same timing; the per-call path timed on
       the first 10 4096-put batches.
    Prints one JSON line."""
    import ctypes as C
    import numpy as np
    import nvshmem_amd as na
    os.environ["NVSHMEM_HEAP_SIZE"] = "96M"
    lib, hip = na.load(), na.Hip()
    lib.nvshmemx_start_local_pes(BATCH_NPES, BATCH_WGS)
    base = lib.nvshmalloc(BATCH_SRC + BATCH_DST)
    for pe in range(BATCH_NPES):
        lib.nvshmemh_fill_pcg32(lib.nvshmem_ptr(base, pe), BATCH_SRC // 4, 100 + pe, 54, None)
    hip.sync()
    out = {"batches": n_batches, "requests_per_batch": BATCH_REQS}
    t0 = time.time()
    grid = na.tune_grid()
    times = na.tune_measure(lib, grid, reps=3)
    feats = na.tune_features(lib, grid)
    out["tune_points"], out["tune_samples"] = int(len(times)), int(times.size)
    out["tune_seconds"] = round(time.time() - t0, 1)
    _log(f"learned: {len(times)} feature points x 54 candidates measured in {out['tune_seconds']} s")
    t0 = time.time()
    model = na.CostModel(lib, seed=1)
    T = np.log(np.maximum(times, 1e-3)).astype(np.float32)
    model.train(feats, T, train_steps, 3e-3)
    model.train(feats, T, train_steps // 2, 1e-3)
    pred = model.predict(feats)
    out["train_seconds"] = round(time.time() - t0, 1)
    out["train_mse_log"] = round(float(np.mean((pred - T) ** 2)), 5)
    best_t = times.min(axis=1)
    regret = times[np.arange(len(times)), pred.argmin(axis=1)] / best_t
    out["grid_top1_agreement"] = round(float(np.mean(pred.argmin(axis=1) == times.argmin(axis=1))), 3)
    out["grid_mean_regret"] = round(float(np.mean(regret)), 4)

    def batch_of(seed, nreq, hi):
        try:
            reqs = na.gen_requests(lib, seed, nreq, BATCH_NPES, BATCH_SRC, BATCH_DST, 6, hi)
        except ValueError:
            return None
        return na.make_requests(lib, reqs, base, 0, BATCH_SRC)

    # batch-level training set, timed by the library's exhaustive judge
    t0 = time.time()
    X, Y = [], []
    for nreq in (8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192):
        for hi in (16, 18, 20, 22):
            for sd in range(16):
                arr = batch_of(20000 + 1000 * nreq + 10 * hi + sd, nreq, hi)
                if arr is None:
                    continue
                ms, _ = na.measure_batch(lib, arr, len(arr), model, 5)
                X.append(na.batch_features(lib, arr, len(arr)))
                Y.append(np.log(ms * 1e3))
    X, Y = np.array(X, np.float32), np.array(Y, np.float32)
    npol = lib.nvshmemx_batch_policy_count()
    bmodel = na.CostModel(lib, seed=3, ncand=npol)
    untrained = na.CostModel(lib, seed=3, ncand=npol)
    bmodel.train(X, Y, 6000, 3e-3)
    bmodel.train(X, Y, 3000, 1e-3)
    bpred = bmodel.predict(X)
    out["batch_training"] = {
        "batches": int(len(X)), "policies": int(npol), "seconds": round(time.time() - t0, 1),
        "mse_log": round(float(np.mean((bpred - Y) ** 2)), 5),
        "mse_log_untrained": round(float(np.mean((untrained.predict(X) - Y) ** 2)), 5),
        "fit_regret": round(float(np.mean(np.exp(Y[np.arange(len(Y)), bpred.argmin(1)] - Y.min(1)))), 4)}
    untrained.close()
    _log(f"learned: batch policy model trained on {len(X)} batches")

    def timed_launches(fn):
        first = timed(hip, fn, 1, 1)
        return timed(hip, fn, int(min(64, max(2, math.ceil(0.05 / max(first, 1e-4))))), 0)

    # held-out batches: the judge (every policy timed on the device), the
    # batch network's pick, the per-request network, the rule
    held = {}
    per_call, bad = [], 0
    t0 = time.time()
    for nreq, nb, seed0 in ((BATCH_REQS, n_batches, 1000), (128, 100, 5128000), (16, 100, 5016000)):
        r = {"judge": [], "batch_net": [], "request_net": [], "rule": [], "one_piece": [], "tiles": []}
        picks = np.zeros(npol, np.int64)
        for b in range(nb):
            arr = batch_of(seed0 + b, nreq, 22)
            n = len(arr)
            ms, _ = na.measure_batch(lib, arr, n, model, 5)
            _, pol = bmodel.choose_batch(model, arr, n)
            picks[pol] += 1
            r["judge"].append(float(ms.min()))
            r["batch_net"].append(float(ms[pol]))
            r["request_net"].append(float(ms[1]))
            r["rule"].append(float(ms[0]))
            r["one_piece"].append(float(ms[POLICY_ONE_PIECE]))
            r["tiles"].append(float(ms[POLICY_TILES]))
            if nreq == BATCH_REQS and b < 10:
                per_call.append(timed(hip, lambda: lib.nvshmemx_put_requests_on_stream(arr, n, None), 1, 0))
            if b < 2:
                bad += int(lib.nvshmemh_verify_requests(arr, n))
            if (b + 1) % 100 == 0:
                _log(f"learned: {b + 1} of {nb} held-out {nreq}-put batches judged")
        j = np.array(r["judge"])
        rec = {"mean_ms": {k: round(float(np.mean(v)), 4) for k, v in r.items()},
               "batch_net_over_judge": round(float(np.mean(np.array(r["batch_net"]) / j)), 4),
               "batch_net_within_5pct_frac": round(float(np.mean(np.array(r["batch_net"]) <= 1.05 * j)), 3),
               "request_net_over_judge": round(float(np.mean(np.array(r["request_net"]) / j)), 4),
               "rule_over_judge": round(float(np.mean(np.array(r["rule"]) / j)), 4),
               "batch_net_gain_over_rule": round(float(np.mean(r["rule"]) / np.mean(r["batch_net"])), 4),
               "request_net_gain_over_rule": round(float(np.mean(r["rule"]) / np.mean(r["request_net"])), 4),
               "batches": nb, "policy_picks": picks.tolist()}
        if nreq == BATCH_REQS:
            rec["batch_net_speedup_vs_per_call"] = round(float(np.mean(per_call) / np.mean(r["batch_net"][:len(per_call)])), 2)
        held[str(nreq)] = rec
    out["held_out"] = held
    out["eval_seconds"] = round(time.time() - t0, 1)
    out["bitexact"] = bad == 0
    out["errors"] = lib.nvshmemx_get_last_error()
    bmodel.close()
    model.close()
    lib.nvstop_pes()
    print(json.dumps(out), flush=True)


JACOBI_N, JACOBI_STEPS = 256, 100


def jacobi_worker(layout, name=None, counter_pass=False):
    """3-D 7-point Jacobi, 8 PEs (2x2x2, periodic), 256^3 float64 cells per
    PE, 100 steps; each step is one fused stencil launch (halo read from this
    PE's receive slots, boundary values written into the neighbours' slots of
    the next parity) and the barrier (kernel boundary in local mode,
    nvshmem_barrier_all_offload between processes).  The same 100 steps are
    then timed two per launch (cells past a block read from the neighbours'
    fields), and the unfused steps for 20 steps each.  Prints one JSON line
    (PE 0)."""
    import numpy as np
    import nvshmem_amd as na
    lib, hip = na.load(), na.Hip()
    if layout == "local":
        os.environ["NVSHMEM_HEAP_SIZE"] = "320M"
        lib.nvshmemx_start_local_pes(8, 1)
        pes_here = list(range(8))
    else:
        lib.nvstart_pes()
        pes_here = [lib.nvshmem_my_pe()]
    n = JACOBI_N
    fb = lib.nvshmemh_jacobi_field_bytes(n)
    u, v = lib.nvshmalloc(fb), lib.nvshmalloc(fb)
    recv = lib.nvshmalloc(lib.nvshmemh_jacobi_recv_bytes(n))
    for pe in pes_here:
        for p in (u, v):
            hip.memset(lib.nvshmem_ptr(p, pe), 0, fb)
    pe0 = pes_here[0]
    lib.nvshmemh_jacobi_init(u, n, pe0, len(pes_here), 42, None)
    hip.sync()

    def total(field):
        return sum(float(na.jacobi_interior(hip, lib.nvshmem_ptr(field, pe), n, lib).sum(dtype=np.float64))
                   for pe in pes_here)
    s0 = total(u)
    lib.nvshmem_barrier_all()
    # planes per workgroup (0: the library's default, 6)
    chunk = int(os.environ.get("JACOBI_CHUNK", "0"))
    chunk2 = int(os.environ.get("JACOBI_CHUNK2", "0"))
    if counter_pass:  # five steps, two two-step launches and one 1 GiB device copy (the counters' scale)
        lib.nvshmemh_jacobi_run(u, v, recv, n, 5, chunk, 0, None)
        lib.nvshmemh_jacobi_run2(u, v, recv, n, 4, chunk2, None)
        ref = hip.malloc(2 << 30)
        hip.check(hip.lib.hipMemcpyAsync(ref + (1 << 30), ref, 1 << 30, hip.D2D, None))
        hip.sync()
        lib.nvstop_pes()
        return
    def warm_then_time(run):
        """one untimed run (the device leaves the clocks of an idle spell:
        the first few milliseconds of a run from idle go up to 30% slower),
        the initial field again, then the timed run"""
        run()
        lib.nvshmemh_jacobi_init(u, n, pe0, len(pes_here), 42, None)
        hip.sync()
        lib.nvshmem_barrier_all()
        return timed(hip, run, 1, 0) / JACOBI_STEPS

    ms = warm_then_time(lambda: lib.nvshmemh_jacobi_run(u, v, recv, n, JACOBI_STEPS, chunk, 0, None))
    res = u if JACOBI_STEPS % 2 == 0 else v
    lib.nvshmem_barrier_all()
    s1 = total(res)
    # the same steps two per launch (no halo exchange: the cells past a
    # block are read straight from the neighbours' fields), from the same
    # initial field; the result must carry the same bits, so the same sum
    ms2 = warm_then_time(lambda: lib.nvshmemh_jacobi_run2(u, v, recv, n, JACOBI_STEPS, chunk2, None))
    lib.nvshmem_barrier_all()
    s2 = total(u if JACOBI_STEPS % 2 == 0 else v)
    # unfused steps (pack+put, barrier, unpack, stencil), the faces stored
    # directly or sent through the library's contiguous and strided device
    # puts; an even number of steps leaves the result back in u
    sep = {}
    for how, api in (("direct", 0), ("device_put_api", 1)):
        sep[how] = round(timed(hip, lambda: lib.nvshmemh_jacobi_run_separate(u, v, recv, n, 20, chunk, api, None),
                               1, 0) / 20, 4)
    lib.nvshmem_barrier_all()
    err = lib.nvshmemx_get_last_error()
    out = {"layout": layout, "pes": 8, "cells_per_pe": n ** 3, "steps": JACOBI_STEPS, "chunk": chunk or "library default", "ms_per_step": ms,
           "sum_before": s0, "sum_after": s1, "unfused_ms_per_step": sep,
           "two_step": {"ms_per_step": ms2, "chunk": chunk2 or "library default", "sum_matches_fused": s2 == s1},
           "errors": err}
    lib.nvstop_pes()
    if pe0 == 0:
        print(json.dumps(out), flush=True)


def with_copy_bound(jacobi, copy_bound):
    """each Jacobi layout's HBM rate also as a fraction of 8 TB/s and of the
    copy bound measured in the same run"""
    if isinstance(jacobi, dict):
        for r in jacobi.values():
            if isinstance(r, dict) and "hbm_gbps" in r:
                r["frac_of_peak"] = round(r["hbm_gbps"] / PEAK_HBM_GBPS, 4)
                r["frac_of_copy_bound"] = round(r["hbm_gbps"] / copy_bound, 4)
            t2 = r.get("two_step") if isinstance(r, dict) else None
            if isinstance(t2, dict) and "launch_hbm_gbps" in t2:
                t2["launch_frac_of_peak"] = round(t2["launch_hbm_gbps"] / PEAK_HBM_GBPS, 4)
    return jacobi


def run_jacobi_leg():
    import nvshmem_amd as na
    env = {k: v for k, v in os.environ.items() if not k.startswith("NVSHMEM_")}
    out = {}
    try:
        p = subprocess.run([sys.executable, __file__, "--jacobi-worker", "local"], stdout=subprocess.PIPE, text=True,
                           timeout=600, env=env)
        lines = [l for l in p.stdout.splitlines() if l.startswith("{")]
        out["local"] = json.loads(lines[-1]) if lines else {"error": f"local jacobi worker exited {p.returncode}"}
        name = f"nvshmem_jacobi_{os.getpid()}_{uuid.uuid4().hex[:8]}"
        procs = [subprocess.Popen([sys.executable, __file__, "--jacobi-worker", "process"],
                                  env=na.shm_job_env(r, 8, name, {"NVSHMEM_HEAP_SIZE": "320M"}), stdout=subprocess.PIPE,
                                  stderr=subprocess.PIPE, text=True) for r in range(8)]
        outs = [pr.communicate(timeout=600) for pr in procs]
        shm = Path("/dev/shm") / name
        if shm.exists():
            shm.unlink()
        lines = [l for o, _ in outs for l in o.splitlines() if l.startswith("{")]
        out["process"] = json.loads(lines[-1]) if lines else {"error": outs[0][1][-800:]}
    except Exception as exc:
        out["error"] = str(exc)[-500:]
    for k in ("local", "process"):
        r = out.get(k)
        if isinstance(r, dict) and "ms_per_step" in r:
            # one read of u and one write of v per cell; halo: 6 faces of n^2
            # doubles written into the neighbours' slots and read back next step
            cells = 8 * JACOBI_N ** 3
            halo = 8 * 6 * JACOBI_N ** 2 * 8 * 2
            r["hbm_gbps"] = round((cells * 16 + halo) / (r["ms_per_step"] * 1e-3) / 1e9, 1)
            r["ms_per_step"] = round(r["ms_per_step"], 4)
            t2 = r.get("two_step")
            if isinstance(t2, dict) and t2.get("ms_per_step"):
                # one read of u and one write of v per cell per launch of two steps
                t2["ms_per_launch"] = round(2 * t2["ms_per_step"], 4)
                t2["launch_hbm_gbps"] = round(cells * 16 / (2 * t2["ms_per_step"] * 1e-3) / 1e9, 1)
                t2["speedup_vs_fused"] = round(r["ms_per_step"] / t2["ms_per_step"], 3)
                t2["ms_per_step"] = round(t2["ms_per_step"], 4)
            if k == "local":  # all 8 PEs' cells: the stencil conserves the total up to rounding
                r["rel_sum_drift"] = abs(r["sum_after"] - r["sum_before"]) / abs(r["sum_before"])
            r.pop("sum_before", None)
            r.pop("sum_after", None)
    return out


SWEEP_REPS = 100


def sweep_worker():
    """{int, float, double} x {put, get} x 2^3..2^30 bytes, 2 PEs as 2
    processes; PE 0 times every launch (HIP events, 100 per point) and checks
    every point bit-exact on the device; PE 1 owns the target heap.  Sources
    are PCG32 words seeded per (PE, dtype, size).  Prints one JSON line (PE 0)."""
    import ctypes as C
    import numpy as np
    import nvshmem_amd as na
    lib, hip = na.load(), na.Hip()
    lib.nvstart_pes()
    me = lib.nvshmem_my_pe()
    top = 1 << 30
    dst = lib.nvshmalloc(top)
    lib.nvshmem_barrier_all()
    if me == 0:
        src, back = hip.malloc(top), hip.malloc(top)
        e0, e1 = hip.event(), hip.event()
        out = {"reps": SWEEP_REPS, "bitexact": True, "points": {}}
        for di, (dtype, w) in enumerate((("int", 4), ("float", 4), ("double", 8))):
            put = getattr(lib, f"nvshmemx_{dtype}_put_on_stream")
            get = getattr(lib, f"nvshmemx_{dtype}_get_on_stream")
            for op in ("put", "get"):
                out["points"][f"{op}_{dtype}"] = []
            _log(f"sweep: {dtype}")
            for lg in range(3, 31):
                size = 1 << lg
                n = size // w
                lib.nvshmemh_fill_pcg32(src, (size + 3) // 4, 1000003 * me + 1009 * di + lg, 54, None)
                for op in ("put", "get"):
                    fn = (lambda: put(dst, src, n, 1, None)) if op == "put" else (lambda: get(back, dst, n, 1, None))
                    fn()
                    t = []
                    for _ in range(SWEEP_REPS):
                        hip.record(e0, None)
                        fn()
                        hip.record(e1, None)
                        hip.sync()
                        t.append(hip.elapsed_ms(e0, e1))
                    req = (na.PutRequest * 1)()
                    req[0].target, req[0].source, req[0].nbytes, req[0].pe, req[0].elem_bytes = (
                        dst, src if op == "put" else back, size, 1, w)
                    ok = int(lib.nvshmemh_verify_requests(req, 1)) == 0
                    out["bitexact"] = out["bitexact"] and ok
                    t = np.array(t)
                    out["points"][f"{op}_{dtype}"].append(
                        [size, round(2 * size / (t.mean() * 1e-3) / 1e9, 2), round(float(np.percentile(t, 50)) * 1e3, 2),
                         round(float(np.percentile(t, 99)) * 1e3, 2)])
        out["errors"] = lib.nvshmemx_get_last_error()
    lib.nvshmem_barrier_all()
    lib.nvstop_pes()
    if me == 0:
        print(json.dumps(out), flush=True)


def run_sweep_leg():
    import nvshmem_amd as na
    name = f"nvshmem_sweep_{os.getpid()}_{uuid.uuid4().hex[:8]}"
    try:
        procs = [subprocess.Popen([sys.executable, __file__, "--sweep-worker"],
                                  env=na.shm_job_env(r, 2, name, {"NVSHMEM_HEAP_SIZE": HEAP}), stdout=subprocess.PIPE,
                                  text=True) for r in range(2)]
        outs = [pr.communicate(timeout=600) for pr in procs]
        lines = [l for o, _ in outs for l in o.splitlines() if l.startswith("{")]
        return json.loads(lines[-1]) if lines else {"error": f"sweep workers exited {[pr.returncode for pr in procs]}"}
    except Exception as exc:
        return {"error": str(exc)[-500:]}
    finally:
        shm = Path("/dev/shm") / name
        if shm.exists():
            shm.unlink()


def run_learned_leg(n_batches, train_steps):
    env = {k: v for k, v in os.environ.items() if not k.startswith("NVSHMEM_")}
    try:
        # the worker's progress lines (stderr) pass through: the leg runs for minutes
        p = subprocess.run([sys.executable, __file__, "--learned-worker", "--learned-batches", str(n_batches),
                            "--train-steps", str(train_steps)], stdout=subprocess.PIPE, text=True, timeout=900,
                           env=env)
        lines = [l for l in p.stdout.splitlines() if l.startswith("{")]
        if p.returncode != 0 or not lines:
            return {"error": f"learned worker exited {p.returncode} (its messages are on stderr)"}
        return json.loads(lines[-1])
    except Exception as exc:
        return {"error": str(exc)[-500:]}


def run_batch_leg():
    try:
        env = {k: v for k, v in os.environ.items() if not k.startswith("NVSHMEM_")}
        p = subprocess.run([sys.executable, __file__, "--batch-worker"], stdout=subprocess.PIPE, text=True,
                           timeout=600, env=env)
        lines = [l for l in p.stdout.splitlines() if l.startswith("{")]
        if p.returncode != 0 or not lines:
            return {"error": f"batch worker exited {p.returncode} (its messages are on stderr)"}
        rec = json.loads(lines[-1])
    except Exception as exc:
        return {"error": str(exc)[-500:]}
    best = min(rec["fused_ms"], rec["graph_ms"])
    rec["fused_gbps"] = round(2 * rec["payload_bytes"] / (best * 1e-3) / 1e9, 1)
    rec["speedup_vs_per_call"] = round(rec["per_call_ms"] / best, 2)
    for k in ("fused_ms", "graph_ms", "owner_ms", "per_call_ms", "persistent_ms", "persistent_noalloc_ms", "tiles_ms",
              "tiles_graph_ms", "posted_ms", "posted_post_kernel_ms"):
        if k in rec:
            rec[k] = round(rec[k], 4)
    if "tiles_ms" in rec:
        rec["tiles_gbps"] = round(2 * rec["payload_bytes"] / (rec["tiles_ms"] * 1e-3) / 1e9, 1)
        rec["tiles_speedup_vs_per_call"] = round(rec["per_call_ms"] / rec["tiles_ms"], 2)
    if "persistent_ms" in rec:
        rec["persistent_gbps"] = round(2 * rec["payload_bytes"] / (rec["persistent_ms"] * 1e-3) / 1e9, 1)
        rec["persistent_speedup_vs_per_call"] = round(rec["per_call_ms"] / rec["persistent_ms"], 2)
    if "posted_ms" in rec:
        rec["posted_gbps"] = round(2 * rec["payload_bytes"] / (rec["posted_ms"] * 1e-3) / 1e9, 1)
        rec["posted_speedup_vs_per_call"] = round(rec["per_call_ms"] / rec["posted_ms"], 2)
    return rec


ANATOMY_BYTES, ANATOMY_LAUNCHES = 64 << 20, 32


def launch_anatomy(lib, hip, dst, src):
    """Where the time of back-to-back 64 MiB copies goes: the library's
    interleaved copy with every workgroup's start and end stamped from the
    constant clock, 32 launches rotated over 1 GiB.  Per launch: the span from
    the first workgroup's start to the last one's end, the last tenth of the
    workgroup ends (the drain), and the gap from the last end to the next
    launch's first start (kernel end, fences, next dispatch)."""
    import numpy as np
    words = ANATOMY_LAUNCHES * 16384 * 2
    stamps = hip.malloc(words * 8)
    try:
        grids = []

        def burst():  # the timed burst's stamps overwrite the warm-up's
            grids.append(lib.nvshmemh_copy_stamped(dst, src, ANATOMY_BYTES, NBYTES, ANATOMY_LAUNCHES, stamps, words,
                                                   None))
        ms = timed(hip, burst, 1, 1) / ANATOMY_LAUNCHES
        grid = grids[-1]
        t = hip.dtoh(stamps, words * 8, np.uint64).reshape(ANATOMY_LAUNCHES, grid, 2).astype(np.float64)
    finally:
        hip.free(stamps)
    ns = 1e6 / lib.nvshmemx_wallclock_khz()
    first, last = t[:, :, 0].min(axis=1), t[:, :, 1].max(axis=1)
    ends = np.sort(t[:, :, 1], axis=1)
    span = (last - first) * ns
    drain = (last - ends[:, int(grid * 0.9)]) * ns
    gap = (first[1:] - last[:-1]) * ns
    per_us = ms * 1e3
    return {"us_per_launch": round(per_us, 2), "gbps": round(2 * ANATOMY_BYTES / (ms * 1e-3) / 1e9, 1),
            "workgroup_span_us": round(float(np.median(span)) / 1e3, 2),
            "drain_last_tenth_us": round(float(np.median(drain)) / 1e3, 2),
            "gap_between_launches_us": round(float(np.median(gap)) / 1e3, 2),
            "launches": ANATOMY_LAUNCHES, "workgroups": grid}


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
    on 2 host cores: 1 GiB float MPI_Put + flush repeated for ~10 s (the
    reported value), then the p+wait round trip and put/get of 8 B..1 GiB
    (100 operations or 0.15 s per size), ~10 s more; then the Jacobi
    application and the 4096-put batch as 8 ranks, a few seconds."""
    exe = ROOT / "oracle" / "build" / "rma_baseline"
    mpirun = Path("/opt/conda/bin/mpirun")
    if not exe.exists() or not mpirun.exists():
        return {"value": None, "unit": "GB/s", "cores": 2, "kind": "port",
                "sample": f"unavailable ({exe.name} or mpirun missing)"}
    cmd = [str(mpirun), "-n", "2", "-bind-to", "core", str(exe), "putget", "put", "float", str(NBYTES), str(NBYTES),
           "1000", "10"]
    def run(args, timeout=240):
        _log(f"cpu baseline: 2 ranks {' '.join(args[:3])}")
        out = subprocess.run([str(mpirun), "-n", "2", "-bind-to", "core", str(exe)] + args, capture_output=True,
                             text=True, timeout=timeout, cwd=ROOT)
        return [json.loads(l) for l in out.stdout.splitlines() if l.startswith("{")]

    try:
        rec = run(cmd[6:])[-1]
        res = {"value": round(rec["gbps"], 2), "unit": "GB/s", "cores": 2, "kind": "port",
               "sample": f"{rec['reps']} x 1 GiB float MPI_Put+MPI_Win_flush over a shared window, 2 ranks "
                         f"(bit-exact: {rec['bitexact']})"}
    except Exception as exc:  # report, never fail the GPU measurement
        return {"value": None, "unit": "GB/s", "cores": 2, "kind": "port", "sample": f"failed: {exc}"}
    try:  # the same semantics at the other sizes and the p+wait round trip, ~10 s more
        pp = run(["pingpong", str(PINGPONG_ITERS)])[-1]
        res["pingpong_rtt_ns"] = {"p50": pp["rtt_p50_ns"], "p99": pp["rtt_p99_ns"], "iters": pp["iters"]}
        res["sweep"] = {}
        for op in ("put", "get"):
            recs = run(["putget", op, "float", "8", str(NBYTES), "100", "0.15"])
            res["sweep"][f"{op}_float"] = [[r["bytes"], round(r["gbps"], 3), round(r["p50_ns"] / 1e3, 2),
                                            round(r["p99_ns"] / 1e3, 2)] for r in recs]
            res["sweep"][f"{op}_float_bitexact"] = all(r["bitexact"] for r in recs)
    except Exception as exc:
        res["sweep_error"] = str(exc)[-300:]

    def run8(args, timeout=300):
        _log(f"cpu baseline: 8 ranks {' '.join(args[:3])}")
        out = subprocess.run([str(mpirun), "-n", "8", "-bind-to", "core", str(exe)] + args, capture_output=True,
                             text=True, timeout=timeout, cwd=ROOT)
        return [json.loads(l) for l in out.stdout.splitlines() if l.startswith("{")]

    try:  # the Jacobi application and the uneven-put batch as 8 host ranks
        check = run8(["jacobi", "32", "5", "1"])[-1]
        jac = run8(["jacobi", str(JACOBI_N), "10", "0"])[-1]
        res["jacobi"] = {"ranks": 8, "cores": 8, "n": JACOBI_N, "steps": jac["steps"],
                         "ms_per_step": jac["ms_per_step"], "bitexact_small": check["bitexact"],
                         "sample": "8 ranks x 256^3 f64, 10 steps: face MPI_Put + MPI_Win_fence + update; "
                                   "bit-exact check against the host replay at 32^3 x 5 steps"}
        bat = run8(["batch", str(BATCH_REQS), str(BATCH_SEED), "3"])[-1]
        res["batch"] = {"ranks": 8, "cores": 8, "pes": bat["pes"], "requests": bat["requests"],
                        "payload_bytes": bat["payload_bytes"], "best_ms": bat["best_ms"], "gbps": bat["gbps"],
                        "bitexact": bat["bitexact"],
                        "sample": "the 4096-put step (seed 7) as 8 ranks x 8 logical PEs, MPI_Put from each source "
                                  "PE's rank + MPI_Win_flush_all, best of 3"}
    except Exception as exc:
        res["apps_error"] = str(exc)[-300:]
    return res


def profiled_traffic(kernel):
    """HBM bytes per put launch from the committed FETCH_SIZE / WRITE_SIZE
    profile of this kernel (scaled by a plain copy of known size), or None when
    the profile was taken of a different kernel."""
    f = Path(__file__).resolve().parent / "profiles" / "put_1GiB_f32_traffic.json"
    try:
        prof = json.loads(f.read_text())
    except (OSError, ValueError):
        return None
    if prof.get("kernel") != kernel or prof.get("algorithmic_bytes") != 2 * NBYTES:
        return None
    return int(prof["traffic_bytes"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--layout", choices=["process", "local"], default="process")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--pe-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--batch-worker", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--counter-pass", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--learned-worker", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--learned-batches", type=int, default=1000)
    ap.add_argument("--train-steps", type=int, default=12000)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-jacobi", action="store_true")
    ap.add_argument("--jacobi-worker", choices=["local", "process"], default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--sweep-worker", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pe_worker:
        worker(args.pe_worker)
        return
    if args.batch_worker:
        batch_worker(args.counter_pass)
        return
    if args.learned_worker:
        learned_worker(args.learned_batches, args.train_steps)
        return
    if args.sweep_worker:
        sweep_worker()
        return
    if args.jacobi_worker:
        jacobi_worker(args.jacobi_worker, counter_pass=args.counter_pass)
        return
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: all PEs share one GPU by design; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    import ctypes as C
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
        payload = lib.nvshmalloc(PUTFLAG_SIZES[-1])
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
        # smaller bulk sizes as streams of back-to-back launches whose buffers
        # rotate over the whole 1 GiB source and target, so the 256 MB MALL
        # cannot serve them (same offsets on both sides: the copies stay exact)
        streams_by_size = {}
        for lg in STREAM_SIZES_LOG2:
            size = 1 << lg
            k = NBYTES // size
            reps = max(16, 2 * k)
            rot = [0]

            def put_rot():
                i = rot[0] % k
                rot[0] += 1
                lib.nvshmemx_float_put_on_stream(dst + i * size, src + i * size, size // 4, 1, None)

            def get_rot():
                i = rot[0] % k
                rot[0] += 1
                lib.nvshmemx_float_get_on_stream(back + i * size, dst + i * size, size // 4, 1, None)
            pm = timed(hip, put_rot, reps, k)
            gm = timed(hip, get_rot, reps, k)
            streams_by_size[str(size)] = {"put_gbps": round(2 * size / (pm * 1e-3) / 1e9, 1),
                                          "get_gbps": round(2 * size / (gm * 1e-3) / 1e9, 1), "launches": reps}
        hip.sync()
        get_ok = bool(np.array_equal(hip.dtoh(back, NBYTES, np.uint32), hip.dtoh(src, NBYTES, np.uint32)))
        if args.layout == "local":
            put_ok = bool(np.array_equal(hip.dtoh(lib.nvshmem_ptr(dst, 1), NBYTES, np.uint32),
                                         hip.dtoh(src, NBYTES, np.uint32)))
        # the copy's own HBM ceiling on this part: read-only and write-only
        # streams over the same 1 GiB (after the bit-exact checks: the write
        # stream overwrites `back`)
        sink = hip.malloc(64)
        read_ms = timed(hip, lambda: lib.nvshmemh_stream_read(src, NBYTES, sink, None), 20, 3)
        # write stream with each store policy (0 non-temporal, 1 sc1); the
        # faster one is the part's write ceiling
        write_ms = [timed(hip, lambda pol=pol: lib.nvshmemh_stream_write(back, NBYTES, pol, None), 20, 3)
                    for pol in (0, 1)]
        anatomy = launch_anatomy(lib, hip, back, src)
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
        lib.nvshmem_barrier_all()                                        # 3b
        bsamples = hip.malloc(BARRIER_ITERS * 8)
        lib.nvshmemh_barrier_latency(BARRIER_ITERS, bsamples, None)
        hip.sync()
        bs = hip.dtoh(bsamples, BARRIER_ITERS * 8, np.int64)[BARRIER_SKIP:]
        khz = lib.nvshmemx_wallclock_khz()
        putflag = {}
        psamples = hip.malloc(PUTFLAG_ITERS * 8)
        for size in PUTFLAG_SIZES:                                       # 3c
            hip.memset(lib.nvshmem_ptr(flag, 0), 0, 256)
            if args.layout == "local":
                hip.memset(lib.nvshmem_ptr(flag, 1), 0, 256)
            hip.sync()
            lib.nvshmem_barrier_all()
            if args.layout == "process":
                lib.nvshmemh_pingpong_process(flag, 1, 1, PUTFLAG_ITERS, payload, size // 4, psamples, None)
            else:
                lib.nvshmemh_pingpong_local(flag, 0, 1, PUTFLAG_ITERS, payload, size // 4, psamples, None)
            hip.sync()
            ps = hip.dtoh(psamples, PUTFLAG_ITERS * 8, np.int64)[PUTFLAG_SKIP:].astype(np.float64) * 1e6 / khz
            putflag[str(size)] = {"p50": round(float(np.percentile(ps, 50)), 1),
                                  "p99": round(float(np.percentile(ps, 99)), 1)}
        err = lib.nvshmemx_get_last_error()
        lib.nvshmem_barrier_all()                                        # 4
        lib.nvstop_pes()
        if child is not None:
            out, _ = child.communicate(timeout=300)
            peer = json.loads([l for l in out.splitlines() if l.startswith("{")][-1])
            put_ok = peer["pe1_put_bitexact"]
            err |= peer["pe1_errors"]
        rtt_ns = s.astype(np.float64) * 1e6 / khz
        bar_ns = bs.astype(np.float64) * 1e6 / khz
        dcfg = na.CopyConfig()
        lib.nvshmemx_default_copy_config(NBYTES, C.byref(dcfg))
        put_kernel = f"k_copy<16, {dcfg.unroll}, {dcfg.nontemporal}, {dcfg.mode}>"
        put_gbps = 2 * NBYTES / (put_ms * 1e-3) / 1e9
        result = dict(
            put_ms=put_ms, put_gbps=put_gbps, get_gbps=2 * NBYTES / (get_ms * 1e-3) / 1e9,
            memcpy_gbps=2 * NBYTES / (memcpy_ms * 1e-3) / 1e9, put_ok=put_ok, get_ok=get_ok, errors=err,
            put_kernel=put_kernel, read_gbps=NBYTES / (read_ms * 1e-3) / 1e9,
            write_gbps=NBYTES / (min(write_ms) * 1e-3) / 1e9,
            write_by_policy={"nt": round(NBYTES / (write_ms[0] * 1e-3) / 1e9, 1),
                             "sc1": round(NBYTES / (write_ms[1] * 1e-3) / 1e9, 1)},
            rtt_p50=float(np.percentile(rtt_ns, 50)), rtt_p99=float(np.percentile(rtt_ns, 99)),
            bar_p50=float(np.percentile(bar_ns, 50)), bar_p99=float(np.percentile(bar_ns, 99)), putflag=putflag,
            streams_by_size=streams_by_size, anatomy=anatomy)
    finally:
        if child is not None and child.poll() is None:
            child.kill()
        shm = Path("/dev/shm") / name
        if shm.exists():
            shm.unlink()

    def leg(name, skip, fn, *a):
        if skip:
            return None
        _log(f"{name} leg")
        t = time.time()
        rec = fn(*a)
        _log(f"{name} leg done in {time.time() - t:.0f} s")
        return rec
    batched = leg("batch", args.no_batch, run_batch_leg)
    learned = leg("learned", args.no_learned, run_learned_leg, args.learned_batches, args.train_steps)
    jacobi = leg("jacobi", args.no_jacobi, run_jacobi_leg)
    sweep = leg("sweep", args.no_sweep, run_sweep_leg)
    cpu = leg("cpu baseline", args.no_cpu_baseline, cpu_baseline)
    r = result
    copy_bound = 2 / (1 / r["read_gbps"] + 1 / r["write_gbps"])  # read and write streams back to back
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
                     "frac": round(r["put_gbps"] / PEAK_HBM_GBPS, 4), "traffic": profiled_traffic(r["put_kernel"]),
                     "kernel": r["put_kernel"]},
        "cpu_baseline": cpu,
        "get_gbps": round(r["get_gbps"], 1),
        "hip_memcpy_d2d_gbps": round(r["memcpy_gbps"], 1),
        "hbm_ceilings_gbps": {
            "read": round(r["read_gbps"], 1), "write": round(r["write_gbps"], 1),
            "write_by_store_policy": r["write_by_policy"],
            "copy_bound": round(copy_bound, 1),
            "put_frac_of_copy_bound": round(r["put_gbps"] * (1 / r["read_gbps"] + 1 / r["write_gbps"]) / 2, 4),
            "pattern": "1 GiB read-only stream of 16-byte non-temporal loads and write-only streams of 16-byte "
                       "stores (non-temporal and sc1; the faster counts); a copy moves every byte once each way, "
                       "so its bound is the two streams back to back"},
        "bulk_streams_gbps": {"sizes": r["streams_by_size"], "frac_of_peak_min": round(
            min(min(v["put_gbps"], v["get_gbps"]) for v in r["streams_by_size"].values()) / PEAK_HBM_GBPS, 4),
            "frac_of_copy_bound_min": round(
                min(min(v["put_gbps"], v["get_gbps"]) for v in r["streams_by_size"].values()) / copy_bound, 4),
            "pattern": "back-to-back float put (get) launches, PE 0 -> PE 1, buffers rotated over 1 GiB",
            "launch_anatomy_64MiB": r["anatomy"]},
        "p_wait_rtt_ns": {"p50": round(r["rtt_p50"], 1), "p99": round(r["rtt_p99"], 1),
                          "iters": PINGPONG_ITERS, "between": "2 processes" if args.layout == "process" else
                          "2 workgroup PEs"},
        "barrier_ns": {"p50": round(r["bar_p50"], 1), "p99": round(r["bar_p99"], 1),
                       "episodes": BARRIER_ITERS - BARRIER_SKIP,
                       "between": "2 processes" if args.layout == "process" else "2 workgroup PEs"},
        "put_flag_rtt_ns": {"sizes": r["putflag"], "iters": PUTFLAG_ITERS - PUTFLAG_SKIP,
                            "pattern": "per-thread put, quiet, p(flag), quiet, wait; both directions"},
        "bitexact": {"put": r["put_ok"], "get": r["get_ok"]},
        "device_errors": r["errors"],
        "batched": batched,
        "learned": learned,
        "jacobi": with_copy_bound(jacobi, copy_bound),
        "sweep": sweep,
    }
    print(json.dumps(line), flush=True)
    if not (r["put_ok"] and r["get_ok"]) or r["errors"]:
        sys.exit(1)


if __name__ == "__main__":
    main()
