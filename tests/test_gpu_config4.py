"""The uneven-put workload at its configured size: 64 workgroup-partitioned
PEs in one process, 4096 puts per step with sizes log-uniform on 64 B .. 4 MiB,
int32/f32 payloads, random (source, destination) PE pairs, 256-byte aligned
offsets, seed 7 -- about 1.55 GB of payload per step.

Every path's result is checked on the host: each PE's whole 64 MiB
destination region is read back and compared with the oracle's sequential
application of the same requests (oracle_batch_apply) on host images of all
64 heaps, so every put's bytes and every byte no put should touch (still zero)
are checked.  Nothing on the device takes part in the comparison.

Paths: the fused plan (one launch), its captured hipGraph replayed, the
owner discipline (each PE's workgroups run the puts it issued), every request
cut into tiles (one workgroup a tile; launched and replayed), the per-call
path (one grid put launch per request), and the persistent kernel in which
each PE cuts its own puts into pieces on the device (library-allocated
scratch)."""
import ctypes as C

import numpy as np
import pytest

import nvshmem_amd as na

pytestmark = pytest.mark.gpu
NPES, WGS, SRC, DST, NREQ, SEED = 64, 4, 32 << 20, 64 << 20, 4096, 7


@pytest.fixture(scope="module")
def full(local_pes, hip, oracle):
    lib = local_pes.start(NPES, wgs_per_pe=WGS, heap="96M")
    base = lib.nvshmalloc(SRC + DST)
    for pe in range(NPES):
        p = lib.nvshmem_ptr(base, pe)
        lib.nvshmemh_fill_pcg32(p, SRC // 4, 100 + pe, 54, None)
    hip.sync()
    reqs = na.gen_requests(lib, SEED, NREQ, NPES, SRC, DST)
    ref = oracle.batch_generate(SEED, NREQ, NPES, SRC, DST)
    for k in ref:
        assert np.array_equal(reqs[k], ref[k]), f"library workload generator differs from the oracle's in {k}"
    arr = na.make_requests(lib, reqs, base, 0, SRC)
    # expected destination regions: the oracle's puts on host images whose
    # sources are the oracle's PCG32 words for each PE
    imgs = []
    for pe in range(NPES):
        img = np.zeros(SRC + DST, np.uint8)
        img[:SRC] = oracle.pcg32(SRC // 4, 100 + pe).view(np.uint8)
        imgs.append(img)
    oracle.batch_apply(imgs, reqs, 0, SRC)
    want = [im[SRC:].copy() for im in imgs]
    del imgs
    print(f"config 4: {len(arr)} requests, {int(reqs['nbytes'].sum())} payload bytes", flush=True)
    return lib, base, arr, want


def _reset(lib, hip, base):
    for pe in range(NPES):
        hip.memset(lib.nvshmem_ptr(base, pe) + SRC, 0, DST)
    hip.sync()


def _check(lib, hip, base, want):
    for pe in range(NPES):
        got = hip.dtoh(lib.nvshmem_ptr(base, pe) + SRC, DST)
        if not np.array_equal(got, want[pe]):
            bad = np.flatnonzero(got != want[pe])
            raise AssertionError(f"PE {pe}: {bad.size} bytes differ, first at destination offset {bad[0]}")


def device_requests(hip, arr):
    """the request array copied to device memory (the persistent kernel reads
    its requests there)"""
    raw = np.frombuffer(bytes(arr), np.uint8)
    d = hip.malloc(max(1, raw.size))
    if raw.size:
        hip.htod(d, raw)
    return d


@pytest.mark.parametrize("path", ["fused", "graph", "owner", "per_call", "persistent", "persistent_graph", "tiles",
                                  "tiles_graph"])
def test_config4_full_size(full, hip, path):
    lib, base, arr, want = full
    n = len(arr)
    _reset(lib, hip, base)
    if path.startswith("persistent"):
        d_reqs = device_requests(hip, arr)
        try:
            lib.nvshmemx_batch_put_persistent_on_stream(d_reqs, n, None)
            hip.sync()
            if path == "persistent_graph":
                # capture two calls on a stream, replay twice with the
                # destinations cleared in between; eager calls in between too
                s = hip.stream()
                graph, exe = C.c_void_p(), C.c_void_p()
                hip.check(hip.lib.hipStreamBeginCapture(s, 1), "hipStreamBeginCapture")  # thread-local mode
                for _ in range(2):
                    lib.nvshmemx_batch_put_persistent_on_stream(d_reqs, n, s)
                hip.check(hip.lib.hipStreamEndCapture(s, C.byref(graph)), "hipStreamEndCapture")
                hip.check(hip.lib.hipGraphInstantiate(C.byref(exe), graph, None, None, 0), "hipGraphInstantiate")
                try:
                    for _ in range(2):
                        _reset(lib, hip, base)
                        hip.check(hip.lib.hipGraphLaunch(exe, s), "hipGraphLaunch")
                        hip.stream_sync(s)
                        _check(lib, hip, base, want)
                        _reset(lib, hip, base)
                        lib.nvshmemx_batch_put_persistent_on_stream(d_reqs, n, None)
                        hip.sync()
                        _check(lib, hip, base, want)
                finally:
                    hip.lib.hipGraphExecDestroy(exe)
                    hip.lib.hipGraphDestroy(graph)
                    hip.stream_destroy(s)
        finally:
            hip.free(d_reqs)
    elif path == "per_call":
        lib.nvshmemx_put_requests_on_stream(arr, n, None)
        hip.sync()
    else:
        launch = na.BatchLaunch(8192, 0, 2, na.BATCH_OWNER) if path == "owner" else None
        choices = None
        if path.startswith("tiles"):  # every request cut into tiles
            choices = (na.BatchChoice * n)()
            for i in range(n):
                choices[i].pieces, choices[i].vec_bytes, choices[i].mode = 1, 16, 2
        plan = C.c_void_p()
        assert lib.nvshmemx_batch_plan_create(arr, n, choices, C.byref(launch) if launch else None,
                                              C.byref(plan)) == 0
        try:
            if path in ("graph", "tiles_graph"):
                lib.nvshmemx_batch_plan_capture(plan, 2)
                for _ in range(2):  # replays, the destinations cleared in between
                    _reset(lib, hip, base)
                    lib.nvshmemx_batch_plan_graph_launch(plan, None)
                    hip.sync()
                    _check(lib, hip, base, want)
            else:
                lib.nvshmemx_batch_plan_launch(plan, None)
                hip.sync()
        finally:
            lib.nvshmemx_batch_plan_destroy(plan)
    assert lib.nvshmemx_get_last_error() == 0
    _check(lib, hip, base, want)
