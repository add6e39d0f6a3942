"""GPU parity of the data-movement path against the CPU oracle.

All PEs live in this process (local mode): PE p's heap is one stride of one
allocation and kernels translate symmetric addresses exactly as they would for
IPC-mapped peers.  Sources are seeded PCG32 words from the oracle; the expected
destination is the oracle's put/get applied to a host image of the same heap
region (canary bytes around the target must survive untouched).  Bit-exact.
"""
import ctypes as C
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import nvshmem_amd as na

pytestmark = pytest.mark.gpu

CANARY = 0xA5
GUARD = 256


def _seed(*parts):
    h = 0x9E3779B97F4A7C15
    for p in parts:
        h = (h * 0x100000001B3 ^ (hash(p) & 0xFFFFFFFF)) & 0xFFFFFFFFFFFFFFFF
    return h


def _src_bytes(oracle, nbytes, seed):
    words = oracle.pcg32((nbytes + 3) // 4, seed)
    return words.view(np.uint8)[:nbytes].copy()


def test_device_pcg_matches_oracle(local_pes, hip, oracle):
    lib = local_pes.start(2)
    for n in [1, 63, 64, 4095, 4096, 4097, (1 << 20) + 5]:
        d = hip.malloc(n * 4)
        try:
            lib.nvshmemh_fill_pcg32(d, n, 42, 54, None)
            hip.sync()
            got = hip.dtoh(d, n * 4, np.uint32)
        finally:
            hip.free(d)
        np.testing.assert_array_equal(got, oracle.pcg32(n, 42, 54))


def test_nvshmalloc_offsets_match_oracle(local_pes, oracle):
    """the bump allocator (reference nvshmem.cu:775-819, here with 256-byte
    alignment) hands out the oracle's offsets for the same call sequence,
    symmetric on every PE; an exact fit of the remaining heap succeeds, and
    nvshmcleanup starts again at offset 0"""
    lib = local_pes.start(4, heap="16M")
    heap = int(lib.nvshmemx_heap_size())
    rng = np.random.default_rng(11)
    sizes = [0, 1, 255, 256, 257, 4096, 12, 0, (1 << 20) + 3] + [int(s) for s in rng.integers(0, 200000, 40)]
    for rnd in range(2):
        lib.nvshmcleanup()
        h = oracle.heap(heap, 256)
        base = None
        for s in sizes:
            want = oracle.nvshmalloc(h, s)
            assert want >= 0, "sequence sized to fit the heap"
            p = lib.nvshmalloc(s)
            base = p if base is None else base
            assert p - base == want, f"round {rnd}: size {s} at {p - base}, oracle {want}"
            for pe in range(1, 4):
                assert lib.nvshmem_ptr(p, pe) - lib.nvshmem_ptr(base, pe) == want
        # the rest of the heap fits exactly
        rest = heap - (oracle.nvshmalloc(h, 0))
        want = oracle.nvshmalloc(h, rest)
        assert want >= 0 and want + rest == heap
        assert lib.nvshmalloc(rest) - base == want
        assert oracle.nvshmalloc(h, 1) == -1  # the library exits here (misuse test heap_exhausted)


def _check_region(hip, oracle, lib, dst_dev_region, image, nbytes_region):
    got = hip.dtoh(dst_dev_region, nbytes_region)
    if not np.array_equal(got, image):
        bad = np.flatnonzero(got != image)
        raise AssertionError(f"{bad.size} bytes differ, first at {bad[0]} (got {got[bad[0]]}, want {image[bad[0]]})")


def _run_put(lib, hip, oracle, *, path, dtype, nelem, toff, soff, pe_src=0, pe_dst=1, seed=1, chunk=0, cfg=None):
    """put `nelem` elements of `dtype` from a local buffer (PE pe_src's heap) into
    PE pe_dst's symmetric buffer at byte offset toff; compare with the oracle."""
    w = na.DTYPE_BYTES[dtype]
    nbytes = nelem * w
    region = nbytes + toff + 2 * GUARD
    sym = lib.nvshmalloc(region)
    src_sym = lib.nvshmalloc(nbytes + soff + 64)
    dst_on_pe = lib.nvshmem_ptr(sym, pe_dst)
    src_local = lib.nvshmem_ptr(src_sym, pe_src) + soff
    data = _src_bytes(oracle, nbytes, seed)
    hip.memset(dst_on_pe, CANARY, region)
    if nbytes:
        hip.htod(src_local, data)
    target = sym + GUARD + toff  # symmetric address
    if path == "grid":
        getattr(lib, f"nvshmemx_{dtype}_put_on_stream")(target, src_local, nelem, pe_dst, None)
    elif path == "config":
        lib.nvshmemx_putmem_config_on_stream(target, src_local, nbytes, pe_dst, C.byref(cfg), None)
    elif path == "thread":
        lib.nvshmemh_thread_putget(0, na.DTYPES[dtype], target, src_local, nelem, chunk, pe_dst, None)
    elif path in ("wave", "block"):
        lib.nvshmemh_coop_putget(0, 0 if path == "wave" else 1, target, src_local, nbytes, chunk * w, pe_dst, 256,
                                 None)
    else:
        raise ValueError(path)
    hip.sync()
    assert lib.nvshmemx_get_last_error() == 0
    # oracle: a put into PE pe_dst's image of the region
    images = [np.full(region, CANARY, np.uint8) for _ in range(2)]
    oracle.put(images, GUARD + toff, data, 1)
    _check_region(hip, oracle, lib, dst_on_pe, images[1], region)


def _run_get(lib, hip, oracle, *, path, dtype, nelem, toff, soff, pe_me=0, pe_src=1, seed=2, chunk=0):
    """get `nelem` elements from PE pe_src's symmetric buffer (byte offset soff)
    into a local buffer at byte offset toff; compare with the oracle."""
    w = na.DTYPE_BYTES[dtype]
    nbytes = nelem * w
    region = nbytes + toff + 2 * GUARD
    src_sym = lib.nvshmalloc(nbytes + soff + 64)
    dst_sym = lib.nvshmalloc(region)
    data = _src_bytes(oracle, nbytes, seed)
    if nbytes:
        hip.htod(lib.nvshmem_ptr(src_sym, pe_src) + soff, data)
    dst_local = lib.nvshmem_ptr(dst_sym, pe_me)
    hip.memset(dst_local, CANARY, region)
    target = dst_local + GUARD + toff
    source = src_sym + soff
    if path == "grid":
        getattr(lib, f"nvshmemx_{dtype}_get_on_stream")(target, source, nelem, pe_src, None)
    elif path == "thread":
        lib.nvshmemh_thread_putget(1, na.DTYPES[dtype], target, source, nelem, chunk, pe_src, None)
    elif path in ("wave", "block"):
        lib.nvshmemh_coop_putget(1, 0 if path == "wave" else 1, target, source, nbytes, chunk * w, pe_src, 256, None)
    hip.sync()
    assert lib.nvshmemx_get_last_error() == 0
    heap_img = [np.zeros(nbytes + soff + 64, np.uint8) for _ in range(2)]
    heap_img[1][soff:soff + nbytes] = data
    expect = np.full(region, CANARY, np.uint8)
    view = expect[GUARD + toff:GUARD + toff + nbytes]
    oracle.get(heap_img, view, soff, 1)
    _check_region(hip, oracle, lib, dst_local, expect, region)


@pytest.mark.parametrize("dtype", ["int", "float", "double"])
@pytest.mark.parametrize("op", ["put", "get"])
def test_grid_putget_sizes(local_pes, hip, oracle, dtype, op):
    """host-launched grid copies, 8 B .. 64 MiB by powers of two"""
    lib = local_pes.start(2, heap="512M")
    w = na.DTYPE_BYTES[dtype]
    nb = 8
    while nb <= (64 << 20):
        lib.nvshmcleanup()
        fn = _run_put if op == "put" else _run_get
        fn(lib, hip, oracle, path="grid", dtype=dtype, nelem=nb // w, toff=0, soff=0, seed=_seed(dtype, op, nb))
        nb *= 2


@pytest.mark.parametrize("op", ["put", "get"])
def test_grid_putget_1GiB(local_pes, hip, oracle, op):
    """the largest size of the put/get workload, 1 GiB of float32, on the
    library's default launch shape for it, against the oracle byte for byte"""
    lib = local_pes.start(2, heap="2100M")
    fn = _run_put if op == "put" else _run_get
    fn(lib, hip, oracle, path="grid", dtype="float", nelem=(1 << 30) // 4, toff=0, soff=0, seed=_seed("1GiB", op))


@pytest.mark.parametrize("mode,block,unroll,streams,workgroups", [
    (0, 256, 4, 0, 7), (0, 1024, 8, 0, 0), (1, 512, 8, 0, 13), (1, 64, 1, 0, 0),
    (2, 512, 1, 2, 0), (2, 256, 2, 1, 0), (2, 1024, 1, 4, 0), (2, 64, 8, 3, 0), (2, 256, 4, 5, 3)])
def test_copy_configs(local_pes, hip, oracle, mode, block, unroll, streams, workgroups):
    """every copy mode and launch shape of nvshmemx_putmem_config_on_stream on
    empty, ragged and relatively misaligned puts (mode 2 = tiles alternating
    between address runs; workgroups 0 = one tile per workgroup)"""
    lib = local_pes.start(2, heap="256M")
    cfg = na.CopyConfig(workgroups, block, 16, unroll, 3, mode, streams)
    for dtype in ["int", "double"]:
        w = na.DTYPE_BYTES[dtype]
        for nelem in [0, 1, 17, 4099, 65536 + 3, (3 << 20) // w + 5]:
            for toff, soff in [(0, 0), (w, 0), (8, 12) if w == 4 else (8, 0)]:
                lib.nvshmcleanup()
                _run_put(lib, hip, oracle, path="config", dtype=dtype, nelem=nelem, toff=toff, soff=soff, cfg=cfg,
                         seed=_seed("config", mode, block, dtype, nelem, toff, soff))


@pytest.mark.parametrize("fields", [(-1, -1, -1, -1, -1, -1, -1), (0, 512, 0, 0, -1, -1, 0), (-1, -1, 8, 2, 0, -1, -1)])
def test_copy_config_keeps_library_choices(local_pes, hip, oracle, fields):
    """negative fields (and zero sizes) of nvshmemx_copy_config_t keep the
    library's choice: all of them, only the block size overridden, only the
    access width, accesses in flight and store policy overridden; sizes on
    both sides of the 32 MiB shape boundary"""
    lib = local_pes.start(2, heap="256M")
    cfg = na.CopyConfig(*fields)
    for nelem in [0, 17, 4099, (3 << 20) // 4 + 5, (40 << 20) // 4 + 3]:
        lib.nvshmcleanup()
        _run_put(lib, hip, oracle, path="config", dtype="int", nelem=nelem, toff=4, soff=0, cfg=cfg,
                 seed=_seed("keep", fields, nelem))


def test_large_put_default_shape(local_pes, hip, oracle):
    """a 512 MiB + 3-element float put, misaligned by 4 bytes, on the library's
    default shape for large copies"""
    lib = local_pes.start(2, heap="1100M")
    lib.nvshmcleanup()
    _run_put(lib, hip, oracle, path="grid", dtype="float", nelem=(512 << 20) // 4 + 3, toff=4, soff=0,
             seed=_seed("large"))


@pytest.mark.parametrize("path", ["grid", "thread", "wave", "block"])
@pytest.mark.parametrize("op", ["put", "get"])
def test_ragged_and_misaligned(local_pes, hip, oracle, path, op):
    """empty, odd-length and relatively misaligned copies on every path"""
    lib = local_pes.start(2, heap="256M")
    fn = _run_put if op == "put" else _run_get
    for dtype in ["int", "double"]:
        w = na.DTYPE_BYTES[dtype]
        for nelem in [0, 1, 3, 17, 1023, 4099]:
            for toff, soff in [(0, 0), (w, 0), (0, w), (3 * w, w), (8, 12) if w == 4 else (8, 0)]:
                lib.nvshmcleanup()
                chunk = 0 if path == "grid" else max(1, nelem // 3)
                fn(lib, hip, oracle, path=path, dtype=dtype, nelem=nelem, toff=toff, soff=soff, chunk=chunk,
                   seed=_seed(path, op, dtype, nelem, toff, soff))


@pytest.mark.parametrize("path,chunk", [("thread", 0), ("thread", 4096), ("wave", 65536), ("block", 1 << 20)])
def test_device_api_bulk(local_pes, hip, oracle, path, chunk):
    """per-thread (reference semantics), wave- and block-cooperative device
    calls at 4 MiB, every dtype, both directions"""
    lib = local_pes.start(2, heap="256M")
    for dtype in ["int", "float", "double"]:
        w = na.DTYPE_BYTES[dtype]
        n = (4 << 20) // w if chunk else (256 << 10) // w
        c = chunk // w if chunk else 0
        lib.nvshmcleanup()
        _run_put(lib, hip, oracle, path=path, dtype=dtype, nelem=n, toff=0, soff=0, chunk=c, seed=_seed(path, dtype))
        lib.nvshmcleanup()
        _run_get(lib, hip, oracle, path=path, dtype=dtype, nelem=n, toff=0, soff=0, chunk=c, seed=_seed(dtype, path))


@pytest.mark.parametrize("dtype", ["int", "float", "double"])
def test_p_and_g(local_pes, hip, oracle, dtype):
    """single-element p and g with every bit pattern class (NaN payloads too)"""
    lib = local_pes.start(4)
    w = na.DTYPE_BYTES[dtype]
    n = 3000
    bits = oracle.pcg32(n * w // 4, _seed("pg", dtype)).view(np.uint8)
    vals_dev = hip.malloc(n * w)
    out_dev = hip.malloc(n * w)
    try:
        hip.htod(vals_dev, bits)
        sym = lib.nvshmalloc(n * w)
        for pe in range(4):
            hip.memset(lib.nvshmem_ptr(sym, pe), 0, n * w)
        lib.nvshmemh_pg(0, na.DTYPES[dtype], sym, vals_dev, None, n, 2, None)  # p into PE 2
        hip.sync()
        images = [np.zeros(n * w, np.uint8) for _ in range(4)]
        for i in range(n):
            if w == 4:
                oracle.lib.oracle_p32((C.c_void_p * 4)(*[im.ctypes.data for im in images]), i * 4,
                                      int(bits[i * 4:i * 4 + 4].view(np.uint32)[0]), 2)
            else:
                oracle.lib.oracle_p64((C.c_void_p * 4)(*[im.ctypes.data for im in images]), i * 8,
                                      int(bits[i * 8:i * 8 + 8].view(np.uint64)[0]), 2)
        for pe in range(4):
            np.testing.assert_array_equal(hip.dtoh(lib.nvshmem_ptr(sym, pe), n * w), images[pe])
        lib.nvshmemh_pg(1, na.DTYPES[dtype], sym, None, out_dev, n, 2, None)  # g from PE 2
        hip.sync()
        np.testing.assert_array_equal(hip.dtoh(out_dev, n * w), images[2])
        assert lib.nvshmemx_get_last_error() == 0
    finally:
        hip.free(vals_dev)
        hip.free(out_dev)


def test_pingpong_local(local_pes, hip):
    """p + wait round trips between two workgroup PEs: every round trip completes"""
    lib = local_pes.start(2)
    iters = 10000
    flag = lib.nvshmalloc(64)
    for pe in range(2):
        hip.memset(lib.nvshmem_ptr(flag, pe), 0, 64)
    samples = hip.malloc(iters * 8)
    try:
        lib.nvshmemh_pingpong_local(flag, 0, 1, iters, None, 0, samples, None)
        hip.sync()
        s = hip.dtoh(samples, iters * 8, np.int64)
    finally:
        hip.free(samples)
    assert lib.nvshmemx_get_last_error() == 0
    assert (s > 0).all()
    # both flags end at the last round number
    for pe in range(2):
        assert hip.dtoh(lib.nvshmem_ptr(flag, pe), 4, np.int32)[0] == iters
    khz = lib.nvshmemx_wallclock_khz()
    p50 = float(np.percentile(s, 50)) * 1e6 / khz
    print(f"local p+wait round trip p50 {p50:.0f} ns p99 {float(np.percentile(s, 99)) * 1e6 / khz:.0f} ns")


def test_quiet_flag_ordering(local_pes, hip):
    """payload published by block put + quiet + p is complete when the flag
    is seen (sequence counter k never observed with stale payload)"""
    lib = local_pes.start(2)
    n, rounds = 16384, 300
    data = lib.nvshmalloc(n * 4)
    stage = lib.nvshmalloc(n * 4)
    flag = lib.nvshmalloc(64)
    ack = lib.nvshmalloc(64)
    for pe in range(2):
        for p in (data, stage, flag, ack):
            hip.memset(lib.nvshmem_ptr(p, pe), 0, 64 if p in (flag, ack) else n * 4)
    viol = hip.malloc(16)
    try:
        hip.memset(viol, 0, 16)
        lib.nvshmemh_order_local(data, stage, flag, ack, n, rounds, 0, 1, viol, None)
        hip.sync()
        v = hip.dtoh(viol, 16, np.uint64)
    finally:
        hip.free(viol)
    assert lib.nvshmemx_get_last_error() == 0
    assert v[0] == 0 and v[1] == 0, v
    assert hip.dtoh(lib.nvshmem_ptr(data, 1), n * 4, np.int32).tolist() == [rounds] * n


@pytest.mark.parametrize("npes,wgs", [(8, 1), (8, 4), (64, 1)])
def test_workgroup_pe_barrier(local_pes, hip, npes, wgs):
    """device barrier across workgroup PEs of one persistent kernel: no PE
    observes a slot older than the round it is in"""
    lib = local_pes.start(npes, wgs_per_pe=wgs, heap="16M")
    slots = lib.nvshmalloc(npes * 4)
    for pe in range(npes):
        hip.memset(lib.nvshmem_ptr(slots, pe), 0, npes * 4)
    viol = hip.malloc(8)
    try:
        hip.memset(viol, 0, 8)
        lib.nvshmemh_wg_barrier_check(slots, 50, viol, 128, None)
        hip.sync()
        v = int(hip.dtoh(viol, 8, np.uint64)[0])
    finally:
        hip.free(viol)
    assert lib.nvshmemx_get_last_error() == 0
    assert v == 0


@pytest.mark.parametrize("until", [0, 1])
def test_wait_timeout_is_reported(local_pes, hip, until):
    """a wait nobody satisfies ends at the configured budget with the sticky
    error bit set (the reference spins forever)"""
    lib = local_pes.start(2, timeout_ms=5000)
    lib.nvshmemx_set_wait_timeout_ms(50)
    try:
        flag = lib.nvshmalloc(64)
        p = lib.nvshmem_ptr(flag, 0)
        hip.memset(p, 0, 64)
        out = hip.malloc(4)
        # while-equal wait on 0 (stays 0), or until-equal wait on 7 (never 7)
        lib.nvshmemh_wait_once(p, 0 if until == 0 else 7, until, out, None)
        hip.sync()
        hip.free(out)
        assert lib.nvshmemx_get_last_error() & na.ERR_WAIT_TIMEOUT
        lib.nvshmemx_clear_errors()
        assert lib.nvshmemx_get_last_error() == 0
    finally:
        lib.nvshmemx_set_wait_timeout_ms(5000)


@pytest.mark.parametrize("case", ["wait_timeout_barrier", "wait_timeout_stop"])
def test_recorded_timeout_ends_the_process(case):
    """a device wait that gave up (its condition still false) makes the next
    nvshmem_barrier_all or nvstop_pes exit non-zero with the reference's error
    convention, instead of the run carrying on with stale data and exiting 0
    (the reference's barrier prints and continues, nvshmem_kernels.cu:65-68)"""
    env = {k: v for k, v in os.environ.items() if not k.startswith("NVSHMEM_")}
    env["NVSHMEM_HEAP_SIZE"] = "64M"
    p = subprocess.run([sys.executable, str(Path(__file__).with_name("misuse_worker.py")), case], env=env,
                       capture_output=True, text=True, timeout=120)
    assert "wait returned" in p.stdout, p.stderr[-2000:]
    assert p.returncode != 0 and "launched" not in p.stdout
    fn = "nvshmem_barrier_all" if case == "wait_timeout_barrier" else "nvstop_pes"
    assert f"{fn}: a device wait outlived its 20 ms budget" in p.stderr, p.stderr[-2000:]


@pytest.mark.parametrize("case", ["get_private_source", "put_private_target", "put_past_heap_end",
                                  "put_into_slot_padding", "heap_exhausted", "too_many_pes", "symmetric_ok"])
def test_non_symmetric_ranges_are_refused(case):
    """host launchers check that the symmetric side of a copy lies in a PE's
    heap before launching: a private or out-of-heap range ends the process with
    a message instead of a device access outside every mapping"""
    env = {k: v for k, v in os.environ.items() if not k.startswith("NVSHMEM_")}
    env["NVSHMEM_HEAP_SIZE"] = "64M"
    p = subprocess.run([sys.executable, str(Path(__file__).with_name("misuse_worker.py")), case], env=env,
                       capture_output=True, text=True, timeout=120)
    if case == "symmetric_ok":
        assert p.returncode == 0 and "launched" in p.stdout, p.stderr[-2000:]
    else:
        assert p.returncode != 0 and "launched" not in p.stdout
        assert any(w in p.stderr for w in ("symmetric", "heap", "PEs requested")), p.stderr[-2000:]


@pytest.mark.parametrize("npes", [2, 8])
def test_barrier_latency_episodes(local_pes, hip, npes):
    """the offload barrier's protocol run back to back by one lane per PE:
    every episode completes (no time-out), each takes a positive time"""
    lib = local_pes.start(npes)
    iters = 2000
    samples = hip.malloc(iters * 8)
    try:
        lib.nvshmemh_barrier_latency(iters, samples, None)
        hip.sync()
        ticks = hip.dtoh(samples, iters * 8, np.int64)
    finally:
        hip.free(samples)
    assert lib.nvshmemx_get_last_error() == 0
    assert (ticks > 0).all()
    ns = ticks * 1e6 / lib.nvshmemx_wallclock_khz()
    print(f"{npes} PEs: barrier p50 {np.percentile(ns, 50):.0f} ns, p99 {np.percentile(ns, 99):.0f} ns")
    assert np.percentile(ns, 50) < 100_000



_STRIDED_DT = {"int": (0, 4), "float": (1, 4), "double": (2, 8)}


@pytest.mark.parametrize("path", ["device_block", "device_wave", "host"])
@pytest.mark.parametrize("dtype", ["int", "double"])
def test_iput_iget(local_pes, hip, oracle, path, dtype):
    """strided put into PE 1 and strided get back (OpenSHMEM iput/iget
    semantics), device-cooperative and host-launched, against the oracle;
    canary bytes between and around the addressed elements survive"""
    lib = local_pes.start(2, heap="64M")
    code, w = _STRIDED_DT[dtype]
    for tst, sst, n in [(1, 1, 1000), (3, 1, 777), (1, 5, 513), (7, 3, 64), (2, 2, 0), (1, 1, 1), (4, 9, 5000)]:
        lib.nvshmcleanup()
        span_t = (tst * max(n - 1, 0) + 1) * w
        span_s = (sst * max(n - 1, 0) + 1) * w
        region = span_t + 2 * GUARD
        sym = lib.nvshmalloc(region)
        src_sym = lib.nvshmalloc(span_s + 64)
        dst_on_pe1 = lib.nvshmem_ptr(sym, 1)
        src_local = lib.nvshmem_ptr(src_sym, 0)
        data = _src_bytes(oracle, span_s, _seed("iput", path, dtype, tst, sst, n))
        hip.memset(dst_on_pe1, CANARY, region)
        hip.htod(src_local, data)
        target = sym + GUARD
        if path == "host":
            getattr(lib, f"nvshmemx_{dtype}_iput_on_stream")(target, src_local, tst, sst, n, 1, None)
        else:
            lib.nvshmemh_strided_device(0 if path == "device_block" else 2, code, target, src_local, tst, sst, n,
                                        0, 0, 0, 0, 1, 256 if path == "device_block" else 64, None)
        hip.sync()
        assert lib.nvshmemx_get_last_error() == 0
        images = [np.full(region, CANARY, np.uint8) for _ in range(2)]
        if n:
            oracle.iput(images, GUARD, tst, data, sst, n, w, 1)
        _check_region(hip, oracle, lib, dst_on_pe1, images[1], region)
        # strided get of what PE 1 now holds into a local buffer laid out like the source
        back = hip.malloc(span_s + 2 * GUARD)
        try:
            hip.memset(back, CANARY, span_s + 2 * GUARD)
            if path == "host":
                getattr(lib, f"nvshmemx_{dtype}_iget_on_stream")(back + GUARD, target, sst, tst, n, 1, None)
            else:
                wave = path == "device_wave"
                lib.nvshmemh_strided_device(5 if wave else 1, code, back + GUARD, target, sst, tst, n, 0, 0, 0, 0, 1,
                                            64 if wave else 256, None)
            hip.sync()
            assert lib.nvshmemx_get_last_error() == 0
            want = np.full(span_s + 2 * GUARD, CANARY, np.uint8)
            if n:
                oracle.iget(images, want[GUARD:], sst, GUARD, tst, n, w, 1)
            _check_region(hip, oracle, lib, back, want, span_s + 2 * GUARD)
        finally:
            hip.free(back)


@pytest.mark.parametrize("path", ["device", "device96", "device160", "host"])
def test_put_get_2d(local_pes, hip, oracle, path):
    """pitched 2-D put into PE 1 and 2-D get back, narrow and wide rows,
    relatively misaligned pitches, against the oracle; the device path also
    with workgroups of 96 and 160 threads, whose last wavefront is partial"""
    block = {"device96": 96, "device160": 160}.get(path, 256)
    lib = local_pes.start(2, heap="64M")
    for width, height, tpitch, spitch in [(1, 1, 1, 1), (5, 7, 13, 9), (4096, 5, 5000, 4100),
                                          (100000, 3, 100008, 100001), (64, 300, 64, 128), (0, 4, 8, 8)]:
        lib.nvshmcleanup()
        span_t = (height - 1) * tpitch + width
        span_s = (height - 1) * spitch + width
        region = span_t + 2 * GUARD
        sym = lib.nvshmalloc(region)
        src_sym = lib.nvshmalloc(span_s + 64)
        dst_on_pe1 = lib.nvshmem_ptr(sym, 1)
        src_local = lib.nvshmem_ptr(src_sym, 0)
        data = _src_bytes(oracle, span_s, _seed("2d", path, width, height, tpitch, spitch))
        hip.memset(dst_on_pe1, CANARY, region)
        if span_s:
            hip.htod(src_local, data)
        target = sym + GUARD
        if path == "host":
            lib.nvshmemx_putmem_2d_on_stream(target, tpitch, src_local, spitch, width, height, 1, None)
        else:
            lib.nvshmemh_strided_device(3, 0, target, src_local, 0, 0, 0, tpitch, spitch, width, height, 1, block, None)
        hip.sync()
        assert lib.nvshmemx_get_last_error() == 0
        images = [np.full(region, CANARY, np.uint8) for _ in range(2)]
        oracle.put_2d(images, GUARD, tpitch, data, spitch, width, height, 1)
        _check_region(hip, oracle, lib, dst_on_pe1, images[1], region)
        back = hip.malloc(span_s + 2 * GUARD)
        try:
            hip.memset(back, CANARY, span_s + 2 * GUARD)
            if path == "host":
                lib.nvshmemx_getmem_2d_on_stream(back + GUARD, spitch, target, tpitch, width, height, 1, None)
            else:
                lib.nvshmemh_strided_device(4, 0, back + GUARD, target, 0, 0, 0, spitch, tpitch, width, height, 1,
                                            block, None)
            hip.sync()
            assert lib.nvshmemx_get_last_error() == 0
            want = np.full(span_s + 2 * GUARD, CANARY, np.uint8)
            for r in range(height):
                want[GUARD + r * spitch:GUARD + r * spitch + width] = images[1][GUARD + r * tpitch:
                                                                              GUARD + r * tpitch + width]
            _check_region(hip, oracle, lib, back, want, span_s + 2 * GUARD)
        finally:
            hip.free(back)


def test_fine_grained_heap_local():
    """NVSHMEM_HEAP_KIND=fine: the heap comes from fine-grained device memory;
    grid and per-thread puts and gets stay bit-exact (aligned and misaligned),
    p + wait round trips complete without a time-out (barrier episodes on
    this heap: test_gpu_multiprocess.py).
    Runs in a process of its own (tests/fine_heap_worker.py), so the
    fine-grained heap is allocated and freed away from the other tests."""
    env = {k: v for k, v in os.environ.items() if not k.startswith("NVSHMEM_")}
    p = subprocess.run([sys.executable, str(Path(__file__).with_name("fine_heap_worker.py"))], env=env,
                       capture_output=True, text=True, timeout=180)
    assert p.returncode == 0 and "fine heap ok" in p.stdout, (p.stdout[-1000:], p.stderr[-2000:])
    print([line for line in p.stdout.splitlines() if line.startswith("fine heap:")][-1])


@pytest.mark.parametrize("pe", [2, -1, 127])
def test_out_of_range_pe_is_refused_on_device(local_pes, hip, pe):
    """device calls that name a PE outside [0, npes) (2 PEs here; 127 is a
    slot of the base table that no PE fills) move nothing and record
    NVSHMEMX_ERR_BAD_PE: per-thread put and get, p and g (g returns 0), the
    wavefront- and workgroup-cooperative puts, and the strided and 2-D puts and
    gets (whose local side stays untouched too)"""
    lib = local_pes.start(2)
    n = 64
    sym = lib.nvshmalloc(n * 4)
    src = lib.nvshmalloc(n * 4)
    vals = hip.malloc(n * 4)
    out = hip.malloc(n * 4)
    loc = hip.malloc(n * 4)
    try:
        for q in (0, 1):
            hip.memset(lib.nvshmem_ptr(sym, q), 0x5A, n * 4)
            hip.memset(lib.nvshmem_ptr(src, q), 0x33, n * 4)
        hip.memset(out, 0x77, n * 4)
        hip.memset(loc, 0x44, n * 4)
        hip.sync()
        lsrc = lib.nvshmem_ptr(src, 0)
        calls = [
            lambda: lib.nvshmemh_thread_putget(0, 1, sym, lib.nvshmem_ptr(src, 0), n, 16, pe, None),
            lambda: lib.nvshmemh_thread_putget(1, 1, lib.nvshmem_ptr(src, 0), sym, n, 16, pe, None),
            lambda: lib.nvshmemh_pg(0, 0, sym, vals, out, n, pe, None),
            lambda: lib.nvshmemh_pg(1, 0, sym, vals, out, n, pe, None),
            lambda: lib.nvshmemh_coop_putget(0, 0, sym, lib.nvshmem_ptr(src, 0), n * 4, 64, pe, 256, None),
            lambda: lib.nvshmemh_coop_putget(0, 1, sym, lib.nvshmem_ptr(src, 0), n * 4, 128, pe, 256, None),
            # strided: op 0 iput_block, 1 iget_block, 2 iput_wave, 5 iget_wave; 2-D: 3 put, 4 get
            lambda: lib.nvshmemh_strided_device(0, 0, sym, lsrc, 1, 1, n, 0, 0, 0, 0, pe, 256, None),
            lambda: lib.nvshmemh_strided_device(1, 0, loc, sym, 1, 1, n, 0, 0, 0, 0, pe, 256, None),
            lambda: lib.nvshmemh_strided_device(2, 0, sym, lsrc, 1, 1, n, 0, 0, 0, 0, pe, 64, None),
            lambda: lib.nvshmemh_strided_device(5, 0, loc, sym, 1, 1, n, 0, 0, 0, 0, pe, 64, None),
            lambda: lib.nvshmemh_strided_device(3, 0, sym, lsrc, 0, 0, 0, 64, 64, 64, 4, pe, 256, None),
            lambda: lib.nvshmemh_strided_device(4, 0, loc, sym, 0, 0, 0, 64, 64, 64, 4, pe, 256, None),
        ]
        for call in calls:
            lib.nvshmemx_clear_errors()
            call()
            hip.sync()
            assert lib.nvshmemx_get_last_error() == na.ERR_BAD_PE
        for q in (0, 1):
            assert (hip.dtoh(lib.nvshmem_ptr(sym, q), n * 4) == 0x5A).all()
            assert (hip.dtoh(lib.nvshmem_ptr(src, q), n * 4) == 0x33).all()
        assert (hip.dtoh(out, n * 4, np.int32) == 0).all()      # g of a bad PE returns 0
        assert (hip.dtoh(loc, n * 4) == 0x44).all()
    finally:
        lib.nvshmemx_clear_errors()
        hip.free(vals)
        hip.free(out)
        hip.free(loc)


def test_private_address_is_refused_on_device(local_pes, hip):
    """in local mode every PE's heap is one slot of a single allocation and a
    symmetric address is translated by its offset within its slot; a private
    allocation's address whose offset within a slot lands on a live symmetric
    object is still not symmetric: device puts, p, g and wavefront puts
    naming it move nothing and record NVSHMEMX_ERR_BAD_ADDRESS"""
    lib = local_pes.start(2, heap="64M")
    sym = lib.nvshmalloc(4096)
    src = lib.nvshmalloc(4096)
    stride = lib.nvshmem_ptr(sym, 1) - lib.nvshmem_ptr(sym, 0)
    assert stride > 0 and stride & (stride - 1) == 0
    priv = hip.malloc(stride + 8192)
    out = hip.malloc(64)
    try:
        # the address in the private buffer that lies at sym's offset within a slot
        alias = priv + ((sym - priv) % stride)
        assert priv <= alias < priv + stride and (alias - sym) % stride == 0
        assert not (lib.nvshmem_ptr(sym, 0) <= alias < lib.nvshmem_ptr(sym, 0) + 2 * stride)
        for q in (0, 1):
            hip.memset(lib.nvshmem_ptr(sym, q), 0x5A, 4096)
        hip.memset(alias, 0x3C, 4096)
        hip.memset(out, 0x77, 64)
        hip.sync()
        local_src = lib.nvshmem_ptr(src, 0)
        calls = [
            lambda: lib.nvshmemh_thread_putget(0, 1, alias, local_src, 16, 16, 1, None),
            lambda: lib.nvshmemh_pg(0, 0, alias, out, out, 16, 1, None),
            lambda: lib.nvshmemh_pg(1, 0, alias, out, out, 16, 1, None),
            lambda: lib.nvshmemh_coop_putget(0, 0, alias, local_src, 4096, 4096, 1, 64, None),
            lambda: lib.nvshmemh_coop_putget(1, 1, out, alias, 64, 64, 0, 256, None),
        ]
        for call in calls:
            lib.nvshmemx_clear_errors()
            call()
            hip.sync()
            assert lib.nvshmemx_get_last_error() == na.ERR_BAD_ADDRESS
        for q in (0, 1):
            assert (hip.dtoh(lib.nvshmem_ptr(sym, q), 4096) == 0x5A).all()
        assert (hip.dtoh(alias, 4096) == 0x3C).all()
        assert (hip.dtoh(out, 64, np.int32) == 0).all()      # g of a bad address returns 0
    finally:
        lib.nvshmemx_clear_errors()
        hip.free(priv)
        hip.free(out)


def test_out_of_heap_range_is_refused_on_device(local_pes, hip):
    """device calls whose symmetric range leaves the heap move nothing and
    record NVSHMEMX_ERR_BAD_ADDRESS: per-thread put, p, g (returns 0), a
    wavefront put, and the strided and 2-D puts and gets at an address past the
    heap's end, and a wavefront put and a strided put that start inside the
    heap but run past its end.  (Local PEs' slots are padded to a power of
    two, so every address used here stays inside the allocation even if a
    check were missing.)"""
    lib = local_pes.start(2, heap="64M")
    heap = int(lib.nvshmemx_heap_size())
    sym = lib.nvshmalloc(4096)
    src = lib.nvshmalloc(4096)
    past = sym + heap + (8 << 20)
    out = hip.malloc(64)
    loc = hip.malloc(4096)
    try:
        for q in (0, 1):
            hip.memset(lib.nvshmem_ptr(sym, q), 0x5A, 4096)
            hip.memset(lib.nvshmem_ptr(past, q), 0x6B, 4096)
        hip.memset(out, 0x77, 64)
        hip.memset(loc, 0x44, 4096)
        hip.sync()
        local_src = lib.nvshmem_ptr(src, 0)
        big = heap // 32  # an element stride that takes 16 ints from sym about 7/8 of a heap past its end
        calls = [
            lambda: lib.nvshmemh_thread_putget(0, 1, past, local_src, 16, 16, 1, None),
            lambda: lib.nvshmemh_pg(0, 0, past, out, out, 16, 1, None),
            lambda: lib.nvshmemh_pg(1, 0, past, out, out, 16, 1, None),
            lambda: lib.nvshmemh_coop_putget(0, 0, past, local_src, 4096, 4096, 1, 64, None),
            lambda: lib.nvshmemh_coop_putget(0, 0, sym, local_src, heap + (4 << 20), heap + (4 << 20), 1, 64, None),
            lambda: lib.nvshmemh_strided_device(0, 0, past, local_src, 1, 1, 16, 0, 0, 0, 0, 1, 256, None),
            lambda: lib.nvshmemh_strided_device(1, 0, loc, past, 1, 1, 16, 0, 0, 0, 0, 1, 256, None),
            lambda: lib.nvshmemh_strided_device(2, 0, past, local_src, 1, 1, 16, 0, 0, 0, 0, 1, 64, None),
            lambda: lib.nvshmemh_strided_device(5, 0, loc, past, 1, 1, 16, 0, 0, 0, 0, 1, 64, None),
            lambda: lib.nvshmemh_strided_device(2, 0, sym, local_src, big, 1, 16, 0, 0, 0, 0, 1, 64, None),
            lambda: lib.nvshmemh_strided_device(3, 0, past, local_src, 0, 0, 0, 64, 64, 64, 4, 1, 256, None),
            lambda: lib.nvshmemh_strided_device(4, 0, loc, past, 0, 0, 0, 64, 64, 64, 4, 1, 256, None),
        ]
        for call in calls:
            lib.nvshmemx_clear_errors()
            call()
            hip.sync()
            assert lib.nvshmemx_get_last_error() == na.ERR_BAD_ADDRESS
        for q in (0, 1):
            assert (hip.dtoh(lib.nvshmem_ptr(sym, q), 4096) == 0x5A).all()
            assert (hip.dtoh(lib.nvshmem_ptr(past, q), 4096) == 0x6B).all()
        assert (hip.dtoh(out, 64, np.int32) == 0).all()      # g of a bad range returns 0
        assert (hip.dtoh(loc, 4096) == 0x44).all()
    finally:
        lib.nvshmemx_clear_errors()
        hip.free(out)
        hip.free(loc)
