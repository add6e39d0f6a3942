"""The no-allocation entry point (nvshmemx_batch_put_persistent_noalloc_on_stream)
against the entry point it mirrors (nvshmemx_batch_put_persistent_on_stream),
through the C ABI: two calls in a row, the second with a smaller request
list, each compared bit for bit with the allocating call's destinations and
with the oracle; a call captured into a hipGraph as the very first use of the
static scratch in this process and replayed; and a list longer than the
static scratch holds, refused with nothing launched."""
import ctypes as C

import numpy as np
import pytest

import nvshmem_amd as na
from test_gpu_config4 import device_requests

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


def _dst(lib, hip, base):
    return [hip.dtoh(lib.nvshmem_ptr(base, pe) + SRC, DST) for pe in range(NPES)]


def _expected(oracle, images, reqs):
    imgs = [im.copy() for im in images]
    oracle.batch_apply(imgs, reqs, 0, SRC)
    return [im[SRC:] for im in imgs]


def test_noalloc_graph_capture_first_use(setup, hip, oracle):
    """the first use of the static scratch is a captured call: nothing to
    allocate, so it captures; three replays with eager work between them"""
    lib, base, images = setup
    reqs = na.gen_requests(lib, 61, 300, NPES, SRC, DST)
    arr = na.make_requests(lib, reqs, base, 0, SRC)
    want = _expected(oracle, images, reqs)
    d = device_requests(hip, arr)
    s = hip.stream()
    graph, exe = C.c_void_p(), C.c_void_p()
    try:
        hip.check(hip.lib.hipStreamBeginCapture(s, 1), "hipStreamBeginCapture")
        rc = lib.nvshmemx_batch_put_persistent_noalloc_on_stream(d, len(arr), s)
        hip.check(hip.lib.hipStreamEndCapture(s, C.byref(graph)), "hipStreamEndCapture")
        assert rc == 0
        hip.check(hip.lib.hipGraphInstantiate(C.byref(exe), graph, None, None, 0), "hipGraphInstantiate")
        for _ in range(3):
            _reset(lib, hip, base)
            hip.check(hip.lib.hipGraphLaunch(exe, s), "hipGraphLaunch")
            hip.stream_sync(s)
            got = _dst(lib, hip, base)
            for pe in range(NPES):
                assert np.array_equal(got[pe], want[pe]), f"replay: PE {pe} differs from the oracle"
    finally:
        if exe.value:
            hip.lib.hipGraphExecDestroy(exe)
        if graph.value:
            hip.lib.hipGraphDestroy(graph)
        hip.stream_destroy(s)
        hip.free(d)
    assert lib.nvshmemx_get_last_error() == 0


@pytest.mark.parametrize("nreq", [2048, 333])
def test_noalloc_matches_allocating_entry_point(setup, hip, oracle, nreq):
    """the second call (333 requests) follows a larger one in the same static
    scratch: nothing of the first may leak into it"""
    lib, base, images = setup
    reqs = na.gen_requests(lib, 60 + nreq, nreq, NPES, SRC, DST, 6, 20)
    arr = na.make_requests(lib, reqs, base, 0, SRC)
    want = _expected(oracle, images, reqs)
    d = device_requests(hip, arr)
    try:
        _reset(lib, hip, base)
        assert lib.nvshmemx_batch_put_persistent_on_stream(d, len(arr), None) == 0
        hip.sync()
        ref = _dst(lib, hip, base)
        _reset(lib, hip, base)
        assert lib.nvshmemx_batch_put_persistent_noalloc_on_stream(d, len(arr), None) == 0
        hip.sync()
        got = _dst(lib, hip, base)
    finally:
        hip.free(d)
    assert lib.nvshmemx_get_last_error() == 0
    for pe in range(NPES):
        assert np.array_equal(got[pe], ref[pe]), f"PE {pe}: no-allocation result differs from the allocating call's"
        assert np.array_equal(got[pe], want[pe]), f"PE {pe}: differs from the oracle"


def test_noalloc_refuses_more_than_its_scratch_holds(setup, hip):
    lib, base, _ = setup
    limit = lib.nvshmemx_batch_persistent_max_requests_noalloc()
    assert limit == 65536
    _reset(lib, hip, base)
    # the request memory is never read: the call is refused before any launch
    assert lib.nvshmemx_batch_put_persistent_noalloc_on_stream(C.c_void_p(16), limit + 1, None) == -1
    hip.sync()
    assert all(not g.any() for g in _dst(lib, hip, base))
    assert lib.nvshmemx_get_last_error() == 0
