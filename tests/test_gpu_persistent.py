"""GPU parity of the persistent batch kernel (nvshmemx_batch_put_persistent_on_stream)
against the oracle: the PEs that issued the puts cut them into chunks on the
device, and every workgroup copies chunks of all PEs' queues in PE order.

8 workgroup-partitioned PEs in one process; each PE's heap holds a 32 MiB
source region (PCG32 words, seed 100 + pe) and a 64 MiB destination region.
Every destination byte of every PE is compared with the oracle's sequential
application of the same requests.  Covered: a batch of the uneven-put
workload, byte-granular (every alignment class) requests mixed with
zero-byte ones, sources outside the heap, thousands of tiny requests (many
per chunk, the wavefront path), one request larger than many chunks (the
workgroup path, every workgroup copying chunks of one queue), repeated
calls (the kernel clears its own counters), an empty list, requests naming a
bad PE or leaving the heap (skipped, error bit set), and two processes each
issuing puts into the other's IPC-mapped heap."""
import ctypes as C

import numpy as np
import pytest

import nvshmem_amd as na
from test_gpu_config4 import device_requests
from test_gpu_multiprocess import run_job

pytestmark = pytest.mark.gpu
NPES, WGS = 8, 4
SRC, DST = 32 << 20, 64 << 20


@pytest.fixture(scope="module")
def setup(local_pes, hip, oracle):
    lib = local_pes.start(NPES, wgs_per_pe=WGS, heap="96M")
    base = lib.nvshmalloc(SRC + DST)
    images = []
    for pe in range(NPES):
        p = lib.nvshmem_ptr(base, pe)
        lib.nvshmemh_fill_pcg32(p, SRC // 4, 100 + pe, 54, None)
        img = np.zeros(SRC + DST, np.uint8)
        img[:SRC] = oracle.pcg32(SRC // 4, 100 + pe).view(np.uint8)
        images.append(img)
    hip.sync()
    return lib, base, images


def _reset(lib, hip, base):
    for pe in range(NPES):
        hip.memset(lib.nvshmem_ptr(base, pe) + SRC, 0, DST)
    hip.sync()


def _check(lib, hip, base, want):
    for pe in range(NPES):
        got = hip.dtoh(lib.nvshmem_ptr(base, pe) + SRC, DST)
        if not np.array_equal(got, want[pe][SRC:]):
            bad = np.flatnonzero(got != want[pe][SRC:])
            raise AssertionError(f"PE {pe}: {bad.size} bytes differ, first at dst offset {bad[0]}")


def _expected(oracle, images, reqs):
    imgs = [im.copy() for im in images]
    oracle.batch_apply(imgs, reqs, 0, SRC)
    return imgs


def _run(lib, hip, arr, n, times=1):
    d = device_requests(hip, arr)
    try:
        for _ in range(times):
            assert lib.nvshmemx_batch_put_persistent_on_stream(d, n, None) == 0
        hip.sync()
    finally:
        hip.free(d)


def _ragged(rng, n, max_nb, sizes=None):
    reqs = {"src_pe": rng.integers(0, NPES, n).astype(np.int32), "dst_pe": np.zeros(n, np.int32),
            "src_off": np.zeros(n, np.uint64), "dst_off": np.zeros(n, np.uint64), "nbytes": np.zeros(n, np.uint64),
            "dtype": np.zeros(n, np.int32)}
    cursor = np.zeros(NPES, np.int64)
    for i in range(n):
        d = (int(reqs["src_pe"][i]) + 1 + int(rng.integers(0, NPES - 1))) % NPES
        nb = int(rng.choice(sizes)) if sizes is not None else int(rng.integers(0, max_nb + 1))
        reqs["dst_pe"][i] = d
        reqs["nbytes"][i] = nb
        reqs["src_off"][i] = int(rng.integers(0, SRC - nb - 16))
        reqs["dst_off"][i] = int(cursor[d]) + int(rng.integers(0, 16))
        cursor[d] = int(reqs["dst_off"][i]) + nb + 1
    assert cursor.max() <= DST
    return reqs


def _per_pe(rng, sizes):
    """sizes[pe] = the byte counts PE pe issues (sources in its own heap),
    each to a random other PE, destinations packed with gaps"""
    rows = [(pe, nb) for pe in range(NPES) for nb in sizes[pe]]
    rows = [rows[i] for i in rng.permutation(len(rows))]
    n = len(rows)
    reqs = {"src_pe": np.zeros(n, np.int32), "dst_pe": np.zeros(n, np.int32), "src_off": np.zeros(n, np.uint64),
            "dst_off": np.zeros(n, np.uint64), "nbytes": np.zeros(n, np.uint64), "dtype": np.zeros(n, np.int32)}
    cursor = np.zeros(NPES, np.int64)
    for i, (pe, nb) in enumerate(rows):
        d = (pe + 1 + int(rng.integers(0, NPES - 1))) % NPES
        reqs["src_pe"][i], reqs["dst_pe"][i], reqs["nbytes"][i] = pe, d, nb
        reqs["src_off"][i] = int(rng.integers(0, SRC - nb - 16))
        reqs["dst_off"][i] = int(cursor[d]) + int(rng.integers(0, 16))
        cursor[d] = int(reqs["dst_off"][i]) + nb + 1
    assert cursor.max() <= DST
    return reqs


@pytest.mark.parametrize("case", ["workload", "ragged", "tiny", "one_large", "repeated", "leader_served",
                                  "served_and_queued"])
def test_persistent_matches_oracle(setup, hip, oracle, case):
    """leader_served: every PE issues at most 64 requests and 128 KiB, so its
    leader copies them itself (no queue): 1 to 64 requests, byte-granular
    sizes below 1600 B (wavefront or 16-lane groups), requests of 16 KiB and
    more among them (the whole workgroup), and a PE with none;
    served_and_queued: half the PEs served by their leaders, half with
    queues (more requests or more bytes than that)"""
    lib, base, images = setup
    rng = np.random.default_rng(31)
    if case == "leader_served":
        counts = [1, 5, 16, 17, 40, 64, 0, 3]
        sizes = [[int(rng.integers(0, 1600)) for _ in range(c)] for c in counts]
        sizes[1][0], sizes[3][5], sizes[5][63] = 16384, 70001, 20000
        sizes[7] = [65536, 16385, 100]
        assert all(len(z) <= 64 and sum(z) <= 128 << 10 for z in sizes)
        reqs = _per_pe(rng, sizes)
    elif case == "served_and_queued":
        sizes = [[int(rng.integers(1, 4096)) for _ in range(c)] for c in [10, 65, 30, 200, 64, 1, 2, 50]]
        sizes[6] = [131072, 1]  # one byte over the leader's limit: queued
        sizes[7][0] = 1 << 20
        reqs = _per_pe(rng, sizes)
    elif case == "workload":
        reqs = na.gen_requests(lib, 41, 1024, NPES, SRC, DST)
    elif case == "ragged":
        reqs = _ragged(rng, 700, 0, sizes=[0, 1, 2, 3, 5, 15, 17, 255, 1023, 4095, 4097, 16383, 16385, 70001,
                                           131071, 131073, 300007, 1 << 20])
    elif case == "tiny":
        reqs = _ragged(rng, 20000, 2048)
    elif case == "one_large":
        reqs = {"src_pe": np.array([3], np.int32), "dst_pe": np.array([5], np.int32),
                "src_off": np.array([4096 + 16], np.uint64), "dst_off": np.array([256 + 16], np.uint64),
                "nbytes": np.array([(30 << 20) + 12], np.uint64), "dtype": np.zeros(1, np.int32)}
    else:
        reqs = na.gen_requests(lib, 43, 512, NPES, SRC, DST)
    arr = na.make_requests(lib, reqs, base, 0, SRC)
    want = _expected(oracle, images, reqs)
    _reset(lib, hip, base)
    _run(lib, hip, arr, len(arr), times=3 if case == "repeated" else 1)
    assert lib.nvshmemx_get_last_error() == 0
    _check(lib, hip, base, want)


def test_persistent_private_sources(setup, hip, oracle):
    """sources outside every heap (a private allocation) count for PE i mod
    npes; the bytes land as the oracle's puts of those sources say"""
    lib, base, images = setup
    rng = np.random.default_rng(5)
    reqs = na.gen_requests(lib, 45, 256, NPES, SRC, DST)
    arr = na.make_requests(lib, reqs, base, 0, SRC)
    priv_bytes = oracle.pcg32((8 << 20) // 4, 999).view(np.uint8)
    priv = hip.malloc(8 << 20)
    try:
        hip.htod(priv, priv_bytes)
        imgs = [im.copy() for im in images]
        for i in range(len(arr)):
            if i % 3 == 0:  # every third request reads the private buffer instead
                off = int(rng.integers(0, (8 << 20) - int(arr[i].nbytes) + 1))
                arr[i].source = priv + off
                src = priv_bytes[off:off + int(arr[i].nbytes)]
            else:
                s_pe, s_off = int(reqs["src_pe"][i]), int(reqs["src_off"][i])
                src = images[s_pe][s_off:s_off + int(reqs["nbytes"][i])]
            d_pe, d_off = int(reqs["dst_pe"][i]), SRC + int(reqs["dst_off"][i])
            imgs[d_pe][d_off:d_off + src.size] = src
        _reset(lib, hip, base)
        _run(lib, hip, arr, len(arr))
        assert lib.nvshmemx_get_last_error() == 0
        _check(lib, hip, base, imgs)
    finally:
        hip.free(priv)


def test_persistent_empty_and_invalid(setup, hip, oracle):
    lib, base, images = setup
    _reset(lib, hip, base)
    assert lib.nvshmemx_batch_put_persistent_on_stream(None, 0, None) == 0
    hip.sync()
    _check(lib, hip, base, images)
    reqs = na.gen_requests(lib, 47, 64, NPES, SRC, DST)
    arr = na.make_requests(lib, reqs, base, 0, SRC)
    want = _expected(oracle, images, {k: v[3:] for k, v in reqs.items()})
    arr[0].pe = NPES + 3                      # no such PE
    arr[1].target = arr[1].target + (200 << 20)  # past the end of the heap
    # a private address at the same offset within a PE slot as a real target
    stride = lib.nvshmem_ptr(base, 1) - lib.nvshmem_ptr(base, 0)
    priv = hip.malloc(stride + (1 << 20))
    try:
        arr[2].target = priv + ((arr[2].target - priv) % stride)
        hip.memset(priv, 0x3C, stride + (1 << 20))
        _reset(lib, hip, base)
        _run(lib, hip, arr, len(arr))
        err = lib.nvshmemx_get_last_error()
        lib.nvshmemx_clear_errors()
        assert err & na.ERR_BAD_PE and err & na.ERR_BAD_ADDRESS, hex(err)
        _check(lib, hip, base, want)
        assert (hip.dtoh(priv, stride + (1 << 20)) == 0x3C).all()
    finally:
        hip.free(priv)


def test_persistent_between_processes():
    """two processes, each the only PE of its own kernel: every process's puts
    (sources in its own heap) land in the other's IPC-mapped heap as the
    oracle's batch says"""
    outs = run_job(2, "persistent", 300)
    for o in outs:
        assert o["errors"] == 0 and o["batch_ok"], o


def test_persistent_128_pes(local_pes, hip, oracle):
    """the most local PEs the library holds (128): leaders 64..127 take the
    second half of each lane's PEs when the queues' offsets are summed.  Runs
    after the tests of the module fixture: it restarts the library that
    fixture set up (as the tests after it do)."""
    npes, src, dst = 128, 1 << 20, 1 << 20
    lib = local_pes.start(npes, wgs_per_pe=1, heap="2M")
    base = lib.nvshmalloc(src + dst)
    images = []
    for pe in range(npes):
        p = lib.nvshmem_ptr(base, pe)
        lib.nvshmemh_fill_pcg32(p, src // 4, 500 + pe, 54, None)
        hip.memset(p + src, 0, dst)
        img = np.zeros(src + dst, np.uint8)
        img[:src] = oracle.pcg32(src // 4, 500 + pe).view(np.uint8)
        images.append(img)
    hip.sync()
    reqs = na.gen_requests(lib, 53, 1024, npes, src, dst, 6, 14)
    arr = na.make_requests(lib, reqs, base, 0, src)
    oracle.batch_apply(images, reqs, 0, src)
    _run(lib, hip, arr, len(arr), times=2)
    assert lib.nvshmemx_get_last_error() == 0
    for pe in range(npes):
        got = hip.dtoh(lib.nvshmem_ptr(base, pe) + src, dst)
        if not np.array_equal(got, images[pe][src:]):
            bad = np.flatnonzero(got != images[pe][src:])
            raise AssertionError(f"PE {pe}: {bad.size} bytes differ, first at dst offset {bad[0]}")


@pytest.mark.parametrize("npes", [6, 8])
def test_persistent_after_restart_with_another_grid(local_pes, hip, oracle, npes):
    """The kernel numbers its launches by tickets that count from the start of
    the library with a fixed grid (the local PEs times the most workgroups
    per PE that fit at once).  On 256 CUs, 6 PEs make a grid of 252
    workgroups where 8 or 128 make 256: after each restart, launches (and
    repeated ones) still find every queue, bytes exact, no wait timing
    out"""
    src, dst = 4 << 20, 4 << 20
    lib = local_pes.start(npes, wgs_per_pe=1, heap="16M")
    grid = lib.nvshmemx_batch_persistent_grid()
    assert grid % npes == 0
    base = lib.nvshmalloc(src + dst)
    images = []
    for pe in range(npes):
        p = lib.nvshmem_ptr(base, pe)
        lib.nvshmemh_fill_pcg32(p, src // 4, 700 + pe, 54, None)
        hip.memset(p + src, 0, dst)
        img = np.zeros(src + dst, np.uint8)
        img[:src] = oracle.pcg32(src // 4, 700 + pe).view(np.uint8)
        images.append(img)
    hip.sync()
    reqs = na.gen_requests(lib, 59 + npes, 300, npes, src, dst, 6, 18)
    arr = na.make_requests(lib, reqs, base, 0, src)
    oracle.batch_apply(images, reqs, 0, src)
    _run(lib, hip, arr, len(arr), times=3)
    assert lib.nvshmemx_get_last_error() == 0
    for pe in range(npes):
        got = hip.dtoh(lib.nvshmem_ptr(base, pe) + src, dst)
        if not np.array_equal(got, images[pe][src:]):
            bad = np.flatnonzero(got != images[pe][src:])
            raise AssertionError(f"PE {pe}: {bad.size} bytes differ, first at dst offset {bad[0]}")
