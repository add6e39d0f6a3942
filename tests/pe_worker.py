"""One PE process of a multi-process test job (shared-memory bootstrap, heaps
IPC-mapped between processes on the same GPU).  Launched by
test_gpu_multiprocess.py; prints one JSON line with what it observed.

Scenarios (every PE runs the same collective sequence):
  putget   ring exchange: PE p puts seeded words into PE p+1, gets from PE p+1;
           checked bit-exact against the oracle's generator
  pingpong p + wait round trips between PE 0 and PE 1 (process kernels)
  barrier  rounds of: publish round r to every PE with p, offload barrier,
           check every slot >= r (sequence-counter ordering)
  batch    every PE plans the puts of a shared uneven-put batch whose source
           is its own heap (targets in the other processes' IPC-mapped heaps),
           launches it and replays it from a hipGraph; each PE's targets are
           checked bit-exact against the oracle's batch
  jacobi   the 2x2x2 Jacobi application, one PE per process (fused steps,
           unfused ones whose faces go through the device put API, or two
           steps per launch reading the neighbours' fields)
  config2  2 PEs: int/float/double x put/get x 8 B .. 2^max_log2 B through the
           host-launched grid copy and the device per-thread calls, each
           landing checked against the oracle with guard bytes
  persistent  like batch, through the persistent kernel (requests in device
           memory, each process's kernel the only PE of its grid)
  nbi      like persistent, the puts posted with nvshmemx_putmem_nbi from a
           kernel of the process and moved by nvshmemx_nbi_flush_on_stream
  barrier_timeout  PE 0 alone enters the offload barrier with a 50 ms budget;
           its next host barrier must end the process
"""
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "df-nvshmem-prototype_amd"))
sys.path.insert(0, str(ROOT / "tests"))

import numpy as np  # noqa: E402

import nvshmem_amd as na  # noqa: E402


def main():
    scenario = sys.argv[1]
    lib, hip = na.load(), na.Hip()
    lib.nvstart_pes()
    me, n = lib.nvshmem_my_pe(), lib.nvshmem_n_pes()
    out = {"pe": me, "npes": n, "mode": lib.nvshmemx_mode(), "bootstrap": lib.nvshmemx_bootstrap_name().decode(),
           "heap_kind": lib.nvshmemx_heap_kind()}
    if scenario == "putget":
        import oracle_py
        orc = oracle_py.Oracle()
        sizes = [4, 12, 4096, 65540, 1 << 20, (8 << 20) + 12]
        bufs = [lib.nvshmalloc(s) for s in sizes]
        gots = [lib.nvshmalloc(s) for s in sizes]
        srcs = []
        for i, s in enumerate(sizes):
            data = orc.pcg32(s // 4, 1000 * me + i).view(np.uint8)
            src = hip.malloc(s)
            hip.htod(src, data)
            srcs.append(src)
            hip.memset(bufs[i], 0, s)
        hip.sync()
        lib.nvshmem_barrier_all()
        right = (me + 1) % n
        for i, s in enumerate(sizes):
            lib.nvshmemx_putmem_on_stream(bufs[i], srcs[i], s, right, None)
        lib.nvshmem_barrier_all()
        left = (me - 1) % n
        ok_put = all(np.array_equal(hip.dtoh(bufs[i], s), orc.pcg32(s // 4, 1000 * left + i).view(np.uint8))
                     for i, s in enumerate(sizes))
        # get back what PE `right` received (it came from me)
        for i, s in enumerate(sizes):
            lib.nvshmemx_getmem_on_stream(gots[i], bufs[i], s, right, None)
        hip.sync()
        ok_get = all(np.array_equal(hip.dtoh(gots[i], s), orc.pcg32(s // 4, 1000 * me + i).view(np.uint8))
                     for i, s in enumerate(sizes))
        # device-side per-thread put into the right neighbour's second-to-last buffer
        lib.nvshmem_barrier_all()
        out.update(put_ok=ok_put, get_ok=ok_get)
    elif scenario == "pingpong":
        iters = int(sys.argv[2])
        flag = lib.nvshmalloc(256)
        hip.memset(flag, 0, 256)
        hip.sync()
        lib.nvshmem_barrier_all()
        samples = hip.malloc(iters * 8)
        if me in (0, 1):
            lib.nvshmemh_pingpong_process(flag, 1 - me, 1 if me == 0 else 0, iters, None, 0, samples, None)
        hip.sync()
        if me == 0:
            s = hip.dtoh(samples, iters * 8, np.int64).astype(np.float64) * 1e6 / lib.nvshmemx_wallclock_khz()
            out.update(p50_ns=float(np.percentile(s, 50)), p99_ns=float(np.percentile(s, 99)))
        out["final_flag"] = int(hip.dtoh(flag, 4, np.int32)[0])
        lib.nvshmem_barrier_all()
    elif scenario == "barrier":
        rounds = int(sys.argv[2])
        slots = lib.nvshmalloc(n * 4)
        hip.memset(slots, 0, n * 4)
        viol = hip.malloc(8)
        hip.memset(viol, 0, 8)
        hip.sync()
        lib.nvshmem_barrier_all()
        for r in range(1, rounds + 1):
            lib.nvshmemh_publish_round(slots, r, None)
            lib.nvshmem_barrier_all_offload()
            lib.nvshmemh_check_round(slots, r, viol, None)
            if r % 10 == 0 and lib.nvshmemx_get_last_error():
                break  # a timed-out barrier: stop after one budget, not one per round
        hip.sync()
        out["violations"] = int(hip.dtoh(viol, 8, np.uint64)[0])
        out["slots"] = hip.dtoh(slots, n * 4, np.int32).tolist()
        lib.nvshmem_barrier_all()
    elif scenario == "batch":
        import ctypes as C
        import oracle_py
        orc = oracle_py.Oracle()
        SRC, DST = 4 << 20, 24 << 20
        nreq, disc = int(sys.argv[2]), int(sys.argv[3])
        base = lib.nvshmalloc(SRC + DST)
        lib.nvshmemh_fill_pcg32(base, SRC // 4, 100 + me, 54, None)
        hip.memset(base + SRC, 0, DST)
        hip.sync()
        reqs = na.gen_requests(lib, 11, nreq, n, SRC, DST, 6, 20)
        mine = {k: v[reqs["src_pe"] == me] for k, v in reqs.items()}
        arr = na.make_requests(lib, mine, base, 0, SRC)
        k = len(arr)
        lib.nvshmem_barrier_all()
        plan = C.c_void_p()
        launch = na.BatchLaunch(0, 0, 2, disc)
        lib.nvshmemx_batch_plan_create(arr, k, None, C.byref(launch), C.byref(plan))
        lib.nvshmemx_batch_plan_launch(plan, None)
        lib.nvshmemx_batch_plan_capture(plan, 3)
        lib.nvshmemx_batch_plan_graph_launch(plan, None)
        hip.sync()
        lib.nvshmemx_batch_plan_destroy(plan)
        lib.nvshmem_barrier_all()
        imgs = [np.zeros(SRC + DST, np.uint8) for _ in range(n)]
        for pe in range(n):
            imgs[pe][:SRC] = orc.pcg32(SRC // 4, 100 + pe).view(np.uint8)
        orc.batch_apply(imgs, reqs, 0, SRC)
        out["batch_ok"] = bool(np.array_equal(hip.dtoh(base + SRC, DST), imgs[me][SRC:]))
        out["requests_issued"] = k
        lib.nvshmem_barrier_all()
    elif scenario in ("persistent", "nbi"):
        # every PE issues its share of a shared uneven-put batch (sources in
        # its own heap, targets in the others' IPC-mapped heaps) through the
        # persistent kernel, or posts it from a kernel of its own and flushes
        # (nbi); each PE checks its destinations against the oracle
        import oracle_py
        orc = oracle_py.Oracle()
        SRC, DST = 4 << 20, 24 << 20
        nreq = int(sys.argv[2])
        base = lib.nvshmalloc(SRC + DST)
        lib.nvshmemh_fill_pcg32(base, SRC // 4, 100 + me, 54, None)
        hip.memset(base + SRC, 0, DST)
        hip.sync()
        reqs = na.gen_requests(lib, 13, nreq, n, SRC, DST, 6, 20)
        mine = {k: v[reqs["src_pe"] == me] for k, v in reqs.items()}
        arr = na.make_requests(lib, mine, base, 0, SRC)
        raw = np.frombuffer(bytes(arr), np.uint8)
        d_reqs = hip.malloc(max(1, raw.size))
        if raw.size:
            hip.htod(d_reqs, raw)
        lib.nvshmem_barrier_all()
        if scenario == "nbi":
            out["nbi_capacity"] = lib.nvshmemx_nbi_capacity()
            lib.nvshmemh_post_nbi(d_reqs, None, len(arr), None)
            assert lib.nvshmemx_nbi_flush_on_stream(None) == 0
        else:
            assert lib.nvshmemx_batch_put_persistent_on_stream(d_reqs, len(arr), None) == 0
        hip.sync()
        lib.nvshmem_barrier_all()
        imgs = [np.zeros(SRC + DST, np.uint8) for _ in range(n)]
        for pe in range(n):
            imgs[pe][:SRC] = orc.pcg32(SRC // 4, 100 + pe).view(np.uint8)
        orc.batch_apply(imgs, reqs, 0, SRC)
        out["batch_ok"] = bool(np.array_equal(hip.dtoh(base + SRC, DST), imgs[me][SRC:]))
        out["requests_issued"] = len(arr)
        hip.free(d_reqs)
        lib.nvshmem_barrier_all()
    elif scenario == "config2":
        # every dtype x {put, get} x 2^3 .. 2^max_log2 bytes between the two
        # processes, through the host-launched typed calls (the grid copy) and
        # the device per-thread calls; each PE checks what landed in its own
        # memory against the oracle, guard bytes on both sides included
        import time
        import oracle_py
        orc = oracle_py.Oracle()
        max_log2, paths = int(sys.argv[2]), sys.argv[3].split(",")
        G, CANARY = 256, 0xA5
        top = 1 << max_log2
        region = top + 2 * G
        sym = lib.nvshmalloc(region)
        src = hip.malloc(region)            # private put source
        dst = hip.malloc(region)            # private get target
        peer = 1 - me
        checked, failures, t0 = 0, [], time.time()
        for dt_name, dt in na.DTYPES.items():
            w = na.DTYPE_BYTES[dt_name]
            for lg in range(3, max_log2 + 1):
                nbytes = 1 << lg
                n = nbytes // w
                mine = orc.pcg32(nbytes // 4, 7_000_000 + 10_000 * me + 100 * dt + lg).view(np.uint8)
                theirs = orc.pcg32(nbytes // 4, 7_000_000 + 10_000 * peer + 100 * dt + lg).view(np.uint8)
                hip.htod(src + G, mine)
                for path in paths:
                    hip.memset(sym, CANARY, nbytes + 2 * G)
                    hip.memset(dst, CANARY, nbytes + 2 * G)
                    hip.sync()
                    lib.nvshmem_barrier_all()
                    if path == "host":
                        getattr(lib, f"nvshmemx_{dt_name}_put_on_stream")(sym + G, src + G, n, peer, None)
                    else:  # one device call per thread, 4096 elements each
                        lib.nvshmemh_thread_putget(0, dt, sym + G, src + G, n, 4096, peer, None)
                    hip.sync()
                    lib.nvshmem_barrier_all()
                    img = np.full(nbytes + 2 * G, CANARY, np.uint8)
                    orc.put([img, img], G, theirs, me)     # what the peer's put leaves on this PE
                    if not np.array_equal(hip.dtoh(sym, nbytes + 2 * G), img):
                        failures.append(f"put {path} {dt_name} {nbytes}")
                    # the peer's symmetric copy now holds this PE's bytes: get them back
                    if path == "host":
                        getattr(lib, f"nvshmemx_{dt_name}_get_on_stream")(dst + G, sym + G, n, peer, None)
                    else:
                        lib.nvshmemh_thread_putget(1, dt, dst + G, sym + G, n, 4096, peer, None)
                    hip.sync()
                    peer_img = np.full(nbytes + 2 * G, CANARY, np.uint8)
                    orc.put([peer_img, peer_img], G, mine, peer)
                    got = np.full(nbytes + 2 * G, CANARY, np.uint8)
                    orc.get([peer_img, peer_img], got[G:G + nbytes], G, peer)
                    if not np.array_equal(hip.dtoh(dst, nbytes + 2 * G), got):
                        failures.append(f"get {path} {dt_name} {nbytes}")
                    checked += 2
                    lib.nvshmem_barrier_all()
                print(f"pe {me}: {dt_name} 2^{lg} B done ({time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
        out.update(checked=checked, failures=failures)
        hip.free(src)
        hip.free(dst)
    elif scenario == "barrier_timeout":
        # PE 0 enters the offload barrier with a 50 ms budget; no other PE
        # does, so the barrier gives up; PE 0's next nvshmem_barrier_all must
        # end the process (the others then time out in the bootstrap barrier)
        lib.nvshmem_barrier_all()
        if me == 0:
            lib.nvshmemx_set_wait_timeout_ms(50)
            lib.nvshmem_barrier_all_offload()
            hip.sync()
            print("offload barrier returned", flush=True)
        lib.nvshmem_barrier_all()
        print("continued", flush=True)
    elif scenario == "jacobi":
        n, steps, outdir = int(sys.argv[2]), int(sys.argv[3]), Path(sys.argv[4])
        how = sys.argv[5] if len(sys.argv) > 5 else "fused"
        fb = lib.nvshmemh_jacobi_field_bytes(n)
        u, v = lib.nvshmalloc(fb), lib.nvshmalloc(fb)
        recv = lib.nvshmalloc(lib.nvshmemh_jacobi_recv_bytes(n))
        hip.memset(u, 0, fb)
        hip.memset(v, 0, fb)
        lib.nvshmemh_jacobi_init(u, n, me, 1, 42, None)
        hip.sync()
        lib.nvshmem_barrier_all()
        if how == "fused":
            res = lib.nvshmemh_jacobi_run(u, v, recv, n, steps, 0, 0, None)
        elif how == "two_step":  # two steps per launch, no halo exchange
            res = lib.nvshmemh_jacobi_run2(u, v, recv, n, steps, 0, None)
        else:  # unfused steps, faces through the library's device puts
            res = lib.nvshmemh_jacobi_run_separate(u, v, recv, n, steps, 0, 1, None)
        hip.sync()
        lib.nvshmem_barrier_all()
        np.save(outdir / f"jacobi_pe{me}.npy", na.jacobi_interior(hip, res, n, lib))
    out["errors"] = lib.nvshmemx_get_last_error()
    lib.nvstop_pes()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
