"""GPU parity of puts posted on the device by workgroup-partitioned PEs
(nvshmemx_putmem_nbi) and moved by nvshmemx_nbi_flush_on_stream, against the
oracle's sequential application of the same requests.

8 PEs of 4 workgroups each in one process; each PE's heap holds a 32 MiB
source region (PCG32 words, seed 100 + pe) and a 64 MiB destination region.
A posting kernel runs every PE's workgroups; each PE posts the requests it
issued (those whose source lies in its heap), then one flush moves them.
Covered: the uneven-put workload, byte-granular and zero-byte requests,
queues small enough for their leaders to copy, a queue that overflows its
share (the overflow is copied as it is posted), posts refused for a bad PE
or a target outside the heap, repeated post/flush rounds, a flush with
nothing posted, the typed int/float/double posts, gets posted beside puts
(nvshmemx_getmem_nbi, the target's PE fetching the same bytes), a captured
post+flush replayed, and two processes posting into each other's
IPC-mapped heaps."""
import ctypes as C

import numpy as np
import pytest

import nvshmem_amd as na
from test_gpu_config4 import device_requests
from test_gpu_multiprocess import run_job
from test_gpu_persistent import _per_pe, _ragged

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


class _Posted:
    """the request array, its issuing PEs and (optionally) which requests go
    as gets, in device memory"""

    def __init__(self, lib, hip, base, reqs, gets=None):
        self.arr = na.make_requests(lib, reqs, base, 0, SRC)
        self.n = len(self.arr)
        self.hip = hip
        self.d = device_requests(hip, self.arr)
        self.iss = self._dev(np.ascontiguousarray(reqs["src_pe"], np.int32))
        self.kind = self._dev(np.ascontiguousarray(gets, np.int32)) if gets is not None else None

    def _dev(self, a):
        d = self.hip.malloc(max(4, a.nbytes))
        if a.nbytes:
            self.hip.htod(d, a.view(np.uint8))
        return d

    def post(self, lib, stream=None):
        if self.kind is None:
            lib.nvshmemh_post_nbi(self.d, self.iss, self.n, stream)
        else:
            lib.nvshmemh_post_nbi_mixed(self.d, self.iss, self.kind, self.n, stream)

    def close(self):
        self.hip.free(self.d)
        self.hip.free(self.iss)
        if self.kind is not None:
            self.hip.free(self.kind)


def _post_and_flush(lib, hip, base, reqs, rounds=1, gets=None):
    p = _Posted(lib, hip, base, reqs, gets)
    try:
        for _ in range(rounds):
            p.post(lib)
            assert lib.nvshmemx_nbi_flush_on_stream(None) == 0
        hip.sync()
    finally:
        p.close()


@pytest.mark.parametrize("case", ["workload", "ragged", "leader_served", "served_and_queued", "repeated"])
def test_posted_puts_match_oracle(setup, hip, oracle, case):
    lib, base, images = setup
    rng = np.random.default_rng(71)
    if case == "workload":
        reqs = na.gen_requests(lib, 73, 1024, NPES, SRC, DST)
    elif case == "ragged":
        reqs = _ragged(rng, 700, 0, sizes=[0, 1, 2, 3, 5, 15, 17, 255, 1023, 4095, 4097, 16383, 16385, 70001,
                                           131071, 131073, 300007, 1 << 20])
    elif case == "leader_served":
        sizes = [[int(rng.integers(0, 1600)) for _ in range(c)] for c in [1, 5, 16, 17, 40, 64, 0, 3]]
        sizes[1][0], sizes[3][5], sizes[5][63] = 16384, 70001, 20000
        reqs = _per_pe(rng, sizes)
    elif case == "served_and_queued":
        sizes = [[int(rng.integers(1, 4096)) for _ in range(c)] for c in [10, 65, 30, 200, 64, 1, 2, 50]]
        sizes[6] = [131072, 1]
        sizes[7][0] = 1 << 20
        reqs = _per_pe(rng, sizes)
    else:
        reqs = na.gen_requests(lib, 79, 512, NPES, SRC, DST)
    want = _expected(oracle, images, reqs)
    _reset(lib, hip, base)
    _post_and_flush(lib, hip, base, reqs, rounds=3 if case == "repeated" else 1)
    assert lib.nvshmemx_get_last_error() == 0
    _check(lib, hip, base, want)


@pytest.mark.parametrize("case", ["workload", "ragged", "leader_served"])
def test_posted_gets_and_puts(setup, hip, oracle, case):
    """about half the requests posted as the get that moves the same bytes
    (by the target's PE, from the source's PE) and the rest as puts, all in
    one flush: the destinations equal the oracle's puts of every request"""
    lib, base, images = setup
    rng = np.random.default_rng(101)
    if case == "workload":
        reqs = na.gen_requests(lib, 107, 1024, NPES, SRC, DST)
    elif case == "ragged":
        reqs = _ragged(rng, 700, 0, sizes=[0, 1, 3, 17, 255, 4095, 4097, 16383, 16385, 70001, 131073, 1 << 20])
    else:
        sizes = [[int(rng.integers(0, 1600)) for _ in range(c)] for c in [1, 5, 16, 17, 40, 64, 0, 3]]
        sizes[1][0], sizes[3][5] = 16384, 70001
        reqs = _per_pe(rng, sizes)
    gets = rng.integers(0, 2, len(reqs["nbytes"])).astype(np.int32)
    want = _expected(oracle, images, reqs)
    _reset(lib, hip, base)
    _post_and_flush(lib, hip, base, reqs, gets=gets)
    assert lib.nvshmemx_get_last_error() == 0
    _check(lib, hip, base, want)


def test_overflowing_queue_is_copied_as_posted(setup, hip, oracle):
    """one PE posts more puts than its share of the queue between two
    flushes: the first share waits for the flush, the rest are copied by the
    posting lanes; every byte lands either way"""
    lib, base, images = setup
    share = lib.nvshmemx_nbi_capacity()
    assert share == 65536 // NPES
    n = share + 500
    rng = np.random.default_rng(83)
    reqs = {"src_pe": np.full(n, 3, np.int32), "dst_pe": np.zeros(n, np.int32), "src_off": np.zeros(n, np.uint64),
            "dst_off": np.zeros(n, np.uint64), "nbytes": np.full(n, 24, np.uint64), "dtype": np.zeros(n, np.int32)}
    for i in range(n):
        reqs["dst_pe"][i] = (3 + 1 + i % (NPES - 1)) % NPES
        reqs["src_off"][i] = int(rng.integers(0, SRC - 64))
        reqs["dst_off"][i] = 32 * (i // (NPES - 1)) + 3
    want = _expected(oracle, images, reqs)
    _reset(lib, hip, base)
    _post_and_flush(lib, hip, base, reqs)
    assert lib.nvshmemx_get_last_error() == 0
    _check(lib, hip, base, want)


def test_refused_posts_and_empty_flush(setup, hip, oracle):
    """posts naming a bad PE or a target past the heap are refused (error
    bits set) and the rest land, and so is a get whose source leaves the
    heap; a flush with nothing posted moves nothing"""
    lib, base, images = setup
    _reset(lib, hip, base)
    assert lib.nvshmemx_nbi_flush_on_stream(None) == 0
    hip.sync()
    _check(lib, hip, base, images)
    reqs = na.gen_requests(lib, 89, 64, NPES, SRC, DST)
    gets = np.zeros(len(reqs["nbytes"]), np.int32)
    gets[2] = 1  # posted as a get whose source is pushed past the heap
    p = _Posted(lib, hip, base, reqs, gets)
    try:
        p.arr[0].pe = NPES + 3
        p.arr[1].target = p.arr[1].target + (200 << 20)
        p.arr[2].source = p.arr[2].source + (200 << 20)
        raw = np.frombuffer(bytes(p.arr), np.uint8)
        hip.htod(p.d, raw)
        p.post(lib)
        assert lib.nvshmemx_nbi_flush_on_stream(None) == 0
        hip.sync()
    finally:
        p.close()
    err = lib.nvshmemx_get_last_error()
    lib.nvshmemx_clear_errors()
    assert err & na.ERR_BAD_PE and err & na.ERR_BAD_ADDRESS, hex(err)
    _check(lib, hip, base, _expected(oracle, images, {k: v[3:] for k, v in reqs.items()}))


@pytest.mark.parametrize("get", [0, 1])
@pytest.mark.parametrize("dtype,width", [(0, 4), (1, 4), (2, 8)])
def test_typed_posts(setup, hip, oracle, dtype, width, get):
    """nvshmemx_{int,float,double}_put_nbi / _get_nbi: len elements of the
    type (0, 1 and 777, the last at an odd element offset), posted by PE 0:
    puts from PE 0's source into PE 5's destination region, gets from PE 5's
    source into PE 0's; they land after the flush, len 0 moves nothing"""
    lib, base, images = setup
    _reset(lib, hip, base)
    want = [im.copy() for im in images]
    src_off, dst_off = 4096 + 3 * width, 100 * width
    for k, length in enumerate((0, 1, 777)):
        d = SRC + dst_off + k * 8192
        if get:
            lib.nvshmemh_post_nbi_typed(dtype, 1, lib.nvshmem_ptr(base, 0) + d, base + src_off, length, 5, None)
            want[0][d:d + length * width] = images[5][src_off:src_off + length * width]
        else:
            lib.nvshmemh_post_nbi_typed(dtype, 0, base + d, lib.nvshmem_ptr(base, 0) + src_off, length, 5, None)
            want[5][d:d + length * width] = images[0][src_off:src_off + length * width]
    assert lib.nvshmemx_nbi_flush_on_stream(None) == 0
    hip.sync()
    assert lib.nvshmemx_get_last_error() == 0
    _check(lib, hip, base, want)


def test_captured_post_and_flush_replays(setup, hip, oracle):
    """the posting kernel and the flush captured in one hipGraph and replayed
    three times, with eager work between the replays: the flush allocates
    nothing and clears nothing with a memset"""
    lib, base, images = setup
    reqs = na.gen_requests(lib, 97, 300, NPES, SRC, DST)
    want = _expected(oracle, images, reqs)
    p = _Posted(lib, hip, base, reqs)
    s = hip.stream()
    graph, exe = C.c_void_p(), C.c_void_p()
    try:
        hip.check(hip.lib.hipStreamBeginCapture(s, 1), "hipStreamBeginCapture")
        p.post(lib, s)
        rc = lib.nvshmemx_nbi_flush_on_stream(s)
        hip.check(hip.lib.hipStreamEndCapture(s, C.byref(graph)), "hipStreamEndCapture")
        assert rc == 0
        hip.check(hip.lib.hipGraphInstantiate(C.byref(exe), graph, None, None, 0), "hipGraphInstantiate")
        for _ in range(3):
            _reset(lib, hip, base)
            hip.check(hip.lib.hipGraphLaunch(exe, s), "hipGraphLaunch")
            hip.stream_sync(s)
            _check(lib, hip, base, want)
    finally:
        if exe.value:
            hip.lib.hipGraphExecDestroy(exe)
        if graph.value:
            hip.lib.hipGraphDestroy(graph)
        hip.stream_destroy(s)
        p.close()
    assert lib.nvshmemx_get_last_error() == 0


def test_posted_puts_between_processes():
    """two processes, each the only PE of its own flush: every process posts
    puts (sources in its own heap) into the other's IPC-mapped heap from a
    kernel of its own, flushes, and the bytes land as the oracle's batch says"""
    outs = run_job(2, "nbi", 300)
    for o in outs:
        assert o["errors"] == 0 and o["batch_ok"], o
        assert o["nbi_capacity"] == 65536, o  # a process holds one PE: the whole queue is its
