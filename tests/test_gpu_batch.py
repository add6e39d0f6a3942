"""GPU parity of the fused batched-put engine against the oracle.

8 workgroup-partitioned PEs in one process; each PE's heap holds a 32 MiB
source region (PCG32 words, seed 100 + pe) and a 64 MiB destination region.
A batch drawn from the uneven-put workload is applied by the engine and by
the oracle (memcpy per request on host heap images); every destination byte of
every PE must match, bit-exact.  Covered: the library's rule, random per-request
choices (every piece count / access width / mode), the owner discipline
(PE-initiated), replays of a captured hipGraph, the per-call path, and
unaligned / ragged requests, an empty request list and zero-byte requests."""
import ctypes as C

import numpy as np
import pytest

import nvshmem_amd as na

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
        hip.memset(p + SRC, 0, DST)
        img = np.zeros(SRC + DST, np.uint8)
        img[:SRC] = oracle.pcg32(SRC // 4, 100 + pe).view(np.uint8)
        images.append(img)
    hip.sync()
    return lib, base, images


def _reset_dst(lib, hip, base):
    for pe in range(NPES):
        hip.memset(lib.nvshmem_ptr(base, pe) + SRC, 0, DST)
    hip.sync()


def _check(lib, hip, base, images):
    for pe in range(NPES):
        got = hip.dtoh(lib.nvshmem_ptr(base, pe) + SRC, DST)
        want = images[pe][SRC:]
        if not np.array_equal(got, want):
            bad = np.flatnonzero(got != want)
            raise AssertionError(f"PE {pe}: {bad.size} bytes differ, first at dst offset {bad[0]}")


def _expected(oracle, images, reqs):
    imgs = [im.copy() for im in images]
    oracle.batch_apply(imgs, reqs, 0, SRC)
    return imgs


def _run_plan(lib, hip, arr, n, choices=None, launch=None):
    plan = C.c_void_p()
    lib.nvshmemx_batch_plan_create(arr, n, choices, C.byref(launch) if launch is not None else None, C.byref(plan))
    lib.nvshmemx_batch_plan_launch(plan, None)
    hip.sync()
    assert lib.nvshmemx_get_last_error() == 0
    return plan


@pytest.mark.parametrize("variant", ["default", "random_choices", "owner", "per_call", "sc1_stores",
                                     "sc1_stores_nt_loads", "tiles", "mixed_with_tiles"])
def test_batch_matches_oracle(setup, hip, oracle, variant):
    lib, base, images = setup
    reqs = na.gen_requests(lib, 8, 512, NPES, SRC, DST)
    arr = na.make_requests(lib, reqs, base, 0, SRC)
    want = _expected(oracle, images, reqs)
    _reset_dst(lib, hip, base)
    n = len(arr)
    if variant == "per_call":
        lib.nvshmemx_put_requests_on_stream(arr, n, None)
        hip.sync()
    else:
        choices, launch = None, None
        if variant == "random_choices":
            rng = np.random.default_rng(5)
            choices = (na.BatchChoice * n)()
            for i in range(n):
                choices[i].pieces = int(rng.integers(1, 257))
                choices[i].vec_bytes = int(rng.choice([4, 8, 16]))
                choices[i].mode = int(rng.integers(0, 2))
            launch = na.BatchLaunch(300, 512, 0, na.BATCH_SHARED)
        if variant == "owner":
            launch = na.BatchLaunch(0, 256, 3, na.BATCH_OWNER)
        if variant == "tiles":  # every request cut into tiles, one workgroup a tile, in source order
            choices = (na.BatchChoice * n)()
            for i in range(n):
                choices[i].pieces, choices[i].vec_bytes, choices[i].mode = 1, 16, 2
        if variant == "mixed_with_tiles":  # tiles beside wavefront and workgroup pieces, random widths
            rng = np.random.default_rng(13)
            choices = (na.BatchChoice * n)()
            for i in range(n):
                choices[i].pieces = int(rng.integers(1, 257))
                choices[i].vec_bytes = int(rng.choice([4, 8, 16]))
                choices[i].mode = int(rng.integers(0, 3))
            launch = na.BatchLaunch(0, 0, 2, na.BATCH_OWNER)
        if variant.startswith("sc1_stores"):  # stores written through the L2 (bit 2), random access widths
            rng = np.random.default_rng(9)
            choices = (na.BatchChoice * n)()
            for i in range(n):
                choices[i].pieces = int(rng.integers(1, 65))
                choices[i].vec_bytes = int(rng.choice([4, 8, 16]))
                choices[i].mode = int(rng.integers(0, 2))
            launch = na.BatchLaunch(0, 0, 5 if variant.endswith("nt_loads") else 4, na.BATCH_SOURCE_ORDER)
        plan = _run_plan(lib, hip, arr, n, choices, launch)
        assert lib.nvshmemx_batch_plan_bytes(plan) == int(reqs["nbytes"].sum())
        lib.nvshmemx_batch_plan_destroy(plan)
    _check(lib, hip, base, want)


@pytest.mark.parametrize("kind", ["pieces", "tiles"])
def test_batch_graph_replays(setup, hip, oracle, kind):
    """a captured graph of 3 fused steps, replayed 3 times with the
    destinations cleared in between: every replay redoes all the work (the
    queue counters reset themselves; the graph holds no memset node)"""
    lib, base, images = setup
    reqs = na.gen_requests(lib, 9, 384, NPES, SRC, DST)
    arr = na.make_requests(lib, reqs, base, 0, SRC)
    want = _expected(oracle, images, reqs)
    choices = None
    if kind == "tiles":
        choices = (na.BatchChoice * len(arr))()
        for i in range(len(arr)):
            choices[i].pieces, choices[i].vec_bytes, choices[i].mode = 1, 16, 2
    plan = C.c_void_p()
    lib.nvshmemx_batch_plan_create(arr, len(arr), choices, None, C.byref(plan))
    lib.nvshmemx_batch_plan_capture(plan, 3)
    try:
        for _ in range(3):
            _reset_dst(lib, hip, base)
            lib.nvshmemx_batch_plan_graph_launch(plan, None)
            hip.sync()
            _check(lib, hip, base, want)
            # eager launch of the same plan between replays
            _reset_dst(lib, hip, base)
            lib.nvshmemx_batch_plan_launch(plan, None)
            hip.sync()
            _check(lib, hip, base, want)
    finally:
        lib.nvshmemx_batch_plan_destroy(plan)
    assert lib.nvshmemx_get_last_error() == 0


def test_batch_unaligned_ragged(setup, hip, oracle):
    """byte-granular offsets and lengths: every relative alignment class, as
    pieces and as tiles"""
    lib, base, images = setup
    rng = np.random.default_rng(17)
    n = 300
    reqs = {"src_pe": rng.integers(0, NPES, n).astype(np.int32), "dst_pe": np.zeros(n, np.int32),
            "src_off": np.zeros(n, np.uint64), "dst_off": np.zeros(n, np.uint64), "nbytes": np.zeros(n, np.uint64),
            "dtype": np.zeros(n, np.int32)}
    cursor = np.zeros(NPES, np.int64)
    for i in range(n):
        d = (int(reqs["src_pe"][i]) + 1 + int(rng.integers(0, NPES - 1))) % NPES
        nb = int(rng.choice([0, 1, 2, 3, 5, 15, 17, 255, 1023, 4097, 70001, 300007]))
        reqs["dst_pe"][i] = d
        reqs["nbytes"][i] = nb
        reqs["src_off"][i] = int(rng.integers(0, SRC - nb - 16))
        reqs["dst_off"][i] = int(cursor[d]) + int(rng.integers(0, 16))
        cursor[d] = int(reqs["dst_off"][i]) + nb + 1
    arr = na.make_requests(lib, reqs, base, 0, SRC)
    want = _expected(oracle, images, reqs)
    _reset_dst(lib, hip, base)
    choices = (na.BatchChoice * n)()
    for i in range(n):  # wavefront pieces, workgroup pieces and tiles (head, body and tail tiles of every width)
        choices[i].pieces, choices[i].vec_bytes, choices[i].mode = int(rng.integers(1, 9)), 16, int(i % 3)
    plan = _run_plan(lib, hip, arr, n, choices, None)
    lib.nvshmemx_batch_plan_destroy(plan)
    _check(lib, hip, base, want)


def test_batch_plan_select_picks_the_faster_plan(setup, hip, oracle):
    """the selection times every candidate plan of one exchange on the device and
    returns the fastest: a plan confined to one workgroup must lose to the
    default plan, in either position; its medians are positive and ordered
    like its choice; the destinations hold the exchange afterwards"""
    lib, base, images = setup
    reqs = na.gen_requests(lib, 21, 512, NPES, SRC, DST)
    arr = na.make_requests(lib, reqs, base, 0, SRC)
    want = _expected(oracle, images, reqs)
    _reset_dst(lib, hip, base)
    n = len(arr)
    fast, slow = C.c_void_p(), C.c_void_p()
    lib.nvshmemx_batch_plan_create(arr, n, None, None, C.byref(fast))
    one_wg = na.BatchLaunch(1, 256, 2, na.BATCH_SHARED)
    lib.nvshmemx_batch_plan_create(arr, n, None, C.byref(one_wg), C.byref(slow))
    try:
        for order, fast_at in (((slow, fast), 1), ((fast, slow), 0)):
            plans = (C.c_void_p * 2)(*order)
            ms = (C.c_float * 2)()
            best = lib.nvshmemx_batch_plan_select(plans, 2, 3, None, ms)
            assert best == fast_at, list(ms)
            assert 0 < ms[fast_at] < ms[1 - fast_at]
        hip.sync()
        assert lib.nvshmemx_get_last_error() == 0
        _check(lib, hip, base, want)
    finally:
        lib.nvshmemx_batch_plan_destroy(fast)
        lib.nvshmemx_batch_plan_destroy(slow)


@pytest.mark.parametrize("variant", ["plan", "graph", "per_call"])
def test_batch_empty_and_zero_byte_requests(setup, hip, oracle, variant):
    """an empty request list builds a plan of 0 bytes that launches (and is
    captured and replayed) without error and moves nothing; zero-byte
    requests mixed into a batch move nothing, and the other requests land as
    the oracle says"""
    lib, base, images = setup
    _reset_dst(lib, hip, base)
    if variant != "per_call":
        empty = C.c_void_p()
        lib.nvshmemx_batch_plan_create(None, 0, None, None, C.byref(empty))
        try:
            assert lib.nvshmemx_batch_plan_bytes(empty) == 0
            if variant == "graph":
                lib.nvshmemx_batch_plan_capture(empty, 2)
                lib.nvshmemx_batch_plan_graph_launch(empty, None)
            else:
                lib.nvshmemx_batch_plan_launch(empty, None)
            hip.sync()
        finally:
            lib.nvshmemx_batch_plan_destroy(empty)
        assert lib.nvshmemx_get_last_error() == 0
        _check(lib, hip, base, images)  # destinations still all zero
    reqs = na.gen_requests(lib, 23, 256, NPES, SRC, DST)
    reqs["nbytes"][::3] = 0
    arr = na.make_requests(lib, reqs, base, 0, SRC)
    want = _expected(oracle, images, reqs)
    n = len(arr)
    if variant == "per_call":
        lib.nvshmemx_put_requests_on_stream(arr, n, None)
        hip.sync()
    else:
        plan = C.c_void_p()
        lib.nvshmemx_batch_plan_create(arr, n, None, None, C.byref(plan))
        try:
            assert lib.nvshmemx_batch_plan_bytes(plan) == int(reqs["nbytes"].sum())
            if variant == "graph":
                lib.nvshmemx_batch_plan_capture(plan, 2)
                lib.nvshmemx_batch_plan_graph_launch(plan, None)
            else:
                lib.nvshmemx_batch_plan_launch(plan, None)
            hip.sync()
        finally:
            lib.nvshmemx_batch_plan_destroy(plan)
    assert lib.nvshmemx_get_last_error() == 0
    _check(lib, hip, base, want)
