"""Host logic of the per-request choice, on the CPU: the candidate set, the
alignment class, the concurrency feature, and the autotuner table's choice
(nearest measured point; among candidates within 5% of the cheapest, the
longest contiguous pieces).  Pure host functions of the C-ABI library: no
GPU is touched."""
import ctypes as C
import itertools

import numpy as np

import nvshmem_amd as na


def _cand(lib, i):
    c = na.BatchChoice()
    lib.nvshmemx_tune_candidate(i, C.byref(c))
    return c.pieces, c.vec_bytes, c.mode


def test_candidate_set(lib):
    n = lib.nvshmemx_tune_candidate_count()
    assert n == 54
    got = [_cand(lib, i) for i in range(n)]
    assert sorted(got) == sorted(itertools.product([1 << k for k in range(9)], [4, 8, 16], [0, 1]))


def test_align_class(lib):
    for src, dst, want in [(0x1000, 0x2000, 0), (0x1010, 0x2000, 0), (0x1008, 0x2000, 1), (0x1004, 0x2000, 2),
                           (0x100C, 0x2000, 2), (0x1018, 0x2010, 1), (0x1001, 0x2000, 2)]:
        assert lib.nvshmemx_align_class(C.c_void_p(src), C.c_void_p(dst)) == want, (hex(src), hex(dst))


def test_effective_concurrency(lib):
    f = lib.nvshmemx_effective_concurrency
    assert f(4096, 1 << 30, 7) == 7                     # an explicit count wins
    assert f(4096, 40960, 0) == 10                      # requests of this size carrying the batch
    assert f(3000, 10000, 0) == 3                       # rounded to nearest
    assert f(1 << 20, 1000, 0) == 1                     # at least one
    assert f(64, 1 << 40, 0) == 32768                   # clamped
    assert f(0, 4096, 0) == 4096                        # an empty request counts as one byte


def _features(lib, log2_bytes, width, conc, align):
    out = (C.c_float * 4)()
    lib.nvshmemx_tune_features(float(log2_bytes), width, conc, align, out)
    return list(out)


def _choose(lib, feats, times, sizes):
    n = len(sizes)
    reqs = (na.PutRequest * n)()
    for i, nb in enumerate(sizes):
        reqs[i].target, reqs[i].source, reqs[i].nbytes, reqs[i].pe, reqs[i].elem_bytes = 0x200000, 0x100000, nb, 0, 4
    out = (na.BatchChoice * n)()
    f = np.ascontiguousarray(feats, np.float32)
    t = np.ascontiguousarray(times, np.float32)
    assert lib.nvshmemx_tune_table_choose(f.ctypes.data, t.ctypes.data, len(t), reqs, n, 1, out) == 0
    return [(o.pieces, o.vec_bytes, o.mode) for o in out]


def test_table_choice_nearest_point(lib):
    """each request takes the best candidate of the measured point nearest to
    its features"""
    feats = [_features(lib, 10, 4, 1, 0), _features(lib, 20, 4, 1, 0)]
    times = np.full((2, 54), 100.0)
    times[0, 5] = 1.0
    times[1, 40] = 1.0
    assert _choose(lib, feats, times, [1 << 10, 1 << 20]) == [_cand(lib, 5), _cand(lib, 40)]
    # features are monotone in size: a 2^11-byte request is still nearest the 2^10 point
    assert _choose(lib, feats, times, [1 << 11]) == [_cand(lib, 5)]


def test_table_choice_prefers_longer_pieces_within_tolerance(lib):
    """candidate 8 (256 wavefront pieces) is cheapest; candidate 45 (one
    workgroup piece of 16-byte accesses) is 3% slower and wins by its longer
    pieces; at 6% slower it does not"""
    assert _cand(lib, 8) == (256, 4, 0) and _cand(lib, 45) == (1, 16, 1)
    feats = [_features(lib, 20, 4, 1, 0)]
    times = np.full((1, 54), 10.0)
    times[0, 8] = 1.0
    times[0, 45] = 1.03
    assert _choose(lib, feats, times, [1 << 20]) == [_cand(lib, 45)]
    times[0, 45] = 1.06
    assert _choose(lib, feats, times, [1 << 20]) == [_cand(lib, 8)]


def _reqs(sizes):
    arr = (na.PutRequest * len(sizes))()
    for i, s in enumerate(sizes):
        arr[i].target, arr[i].source, arr[i].nbytes, arr[i].pe, arr[i].elem_bytes = 0x10000, 0x20000, s, 0, 4
    return arr


def test_batch_features(lib):
    """closed form: log2 of the batch's bytes and of its request count, mean
    and largest log2 request size, each shifted and scaled as documented"""
    sizes = [64, 4096, 1 << 20, 3]
    f = na.batch_features(lib, _reqs(sizes), len(sizes))
    logs = np.log2(np.maximum(sizes, 1))
    want = [(np.log2(sum(sizes)) - 24) / 8, (np.log2(len(sizes)) - 7) / 4, (logs.mean() - 14) / 8, (logs.max() - 14) / 8]
    np.testing.assert_allclose(f, want, rtol=1e-6, atol=1e-6)


def test_batch_policy_choices(lib):
    """the rule's choices equal nvshmemx_batch_default_choice; a uniform
    policy cuts a request into ceil(bytes / granularity) pieces (a wavefront
    policy counts four wavefront pieces per piece), 1..256, 16-byte accesses;
    policy 11 is one workgroup piece per request, policy 12 every request as
    tiles (mode 2)"""
    assert lib.nvshmemx_batch_policy_count() == 13
    sizes = [1, 4095, 4096, 65536, 65537, 1 << 20, (1 << 22) - 4, 1 << 30]
    arr, n = _reqs(sizes), len(sizes)
    out = (na.BatchChoice * n)()
    lib.nvshmemx_batch_policy_choices(arr, n, 0, None, out)
    for i, s in enumerate(sizes):
        d = na.BatchChoice()
        lib.nvshmemx_batch_default_choice(s, 4, C.byref(d))
        assert (out[i].pieces, out[i].vec_bytes, out[i].mode) == (d.pieces, d.vec_bytes, d.mode)
    grans = {2: (4 << 10, 0), 3: (8 << 10, 0), 4: (16 << 10, 0), 5: (32 << 10, 0), 6: (64 << 10, 1), 7: (128 << 10, 1),
             8: (256 << 10, 1), 9: (512 << 10, 1), 10: (1 << 20, 1)}
    for p, (g, mode) in grans.items():
        lib.nvshmemx_batch_policy_choices(arr, n, p, None, out)
        per = g * (4 if mode == 0 else 1)
        for i, s in enumerate(sizes):
            assert (out[i].pieces, out[i].vec_bytes, out[i].mode) == (max(1, min(256, -(-s // per))), 16, mode), (p, s)
    lib.nvshmemx_batch_policy_choices(arr, n, 11, None, out)
    assert all((out[i].pieces, out[i].mode) == (1, 1) for i in range(n))
    lib.nvshmemx_batch_policy_choices(arr, n, 12, None, out)
    assert all((out[i].vec_bytes, out[i].mode) == (16, 2) for i in range(n))
