"""Pins the CPU oracle before it is trusted as the checker.

The reference ships no tests and no golden vectors, so each oracle function is
pinned by a published known-answer vector, a closed form, an invariant, or a
brute-force enumeration, chosen so that a plausible slip (off-by-one bound,
wrong sign, transposed argument) fails.  No GPU needed.
"""
import ctypes as C
import itertools
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


# ---------------------------------------------------------------- PCG32

def test_pcg32_published_known_answer(oracle):
    """pcg32-demo from the PCG minimal C reference: srandom(42, 54), first six
    outputs of the first block the demo prints (published with the library)."""
    kat = json.loads((GOLDEN / "pcg32_kat.json").read_text())
    got = oracle.pcg32(len(kat["outputs"]), kat["seed"], kat["seq"])
    assert [int(x) for x in got] == [int(x, 16) for x in kat["outputs"]]


def test_pcg32_fixture_hashes(oracle):
    """first words and a digest of longer streams, committed as fixtures"""
    import hashlib
    fx = json.loads((GOLDEN / "pcg32_streams.json").read_text())
    for case in fx["cases"]:
        w = oracle.pcg32(case["n"], case["seed"], case["seq"])
        assert [int(x) for x in w[:8]] == case["first8"]
        assert hashlib.sha256(w.tobytes()).hexdigest() == case["sha256"]


def test_uniform_pm1_is_exact_and_in_range(oracle):
    n = 1 << 16
    f = oracle.uniform_pm1(n, 42)
    w = oracle.pcg32(n, 42)
    assert f.dtype == np.float32 and (f >= -1).all() and (f < 1).all()
    # closed form: (w >> 8) * 2^-23 - 1, every value exactly representable
    np.testing.assert_array_equal(f.astype(np.float64), (w >> 8).astype(np.float64) * 2.0 ** -23 - 1.0)


def test_uniform01_f64_is_exact_and_in_range(oracle):
    n = 1 << 12
    d = oracle.uniform01_f64(n, 9)
    w = oracle.pcg32(2 * n, 9).astype(np.uint64)
    u = ((w[0::2] << np.uint64(32)) | w[1::2]) >> np.uint64(11)
    np.testing.assert_array_equal(d, u.astype(np.float64) * 2.0 ** -53)
    assert (d >= 0).all() and (d < 1).all()


# ---------------------------------------------------------------- heap

def test_heap_reference_arithmetic(oracle):
    """align 1 restates nvshmalloc exactly: offsets are prefix sums; a request
    fails iff tail + size > heap (an exact fit succeeds)"""
    h = oracle.heap(100, 1)
    assert [oracle.nvshmalloc(h, s) for s in (10, 0, 30, 60)] == [0, 10, 10, 40]
    assert oracle.nvshmalloc(h, 1) == -1          # 100 + 1 > 100
    oracle.lib.oracle_nvshmcleanup(h)
    assert oracle.nvshmalloc(h, 100) == 0          # exact fit after cleanup
    assert oracle.nvshmalloc(h, 0) == 100          # zero-size at the end still fits


def test_heap_aligned_and_symmetric(oracle):
    rng = np.random.default_rng(5)
    sizes = [int(x) for x in rng.integers(0, 5000, 200)]
    a, b = oracle.heap(1 << 20, 256), oracle.heap(1 << 20, 256)
    prev_end = 0
    for s in sizes:
        oa, ob = oracle.nvshmalloc(a, s), oracle.nvshmalloc(b, s)
        assert oa == ob                       # same call sequence -> same offsets on every PE
        if oa < 0:
            break
        assert oa % 256 == 0 and oa >= prev_end and oa - prev_end < 256
        prev_end = oa + s


# ---------------------------------------------------------------- translation

def test_translate_process_mode_is_reference_formula(oracle):
    bases = [0x7F0000000000, 0x7E1234560000, 0x7D0000001000]
    mine = bases[1]
    for off in [0, 1, 255, 4096, (1 << 30) - 8]:
        for pe in range(3):
            assert oracle.translate(bases, mine, (1 << 64) - 1, mine + off, pe) == bases[pe] + off


def test_translate_local_mode_any_pe_coordinates(oracle):
    stride = 1 << 28
    base0 = 0x7F0000000000 + 12345 * 256     # allocation base need not be stride-aligned
    bases = [base0 + i * stride for i in range(8)]
    for off in [0, 8, stride - 4]:
        for src_pe in range(8):
            for pe in range(8):
                got = oracle.translate(bases, base0, stride - 1, bases[src_pe] + off, pe)
                assert got == bases[pe] + off
                # composition: translating twice lands where translating once does
                assert oracle.translate(bases, base0, stride - 1, got, (pe + 3) % 8) == bases[(pe + 3) % 8] + off


# ---------------------------------------------------------------- data movement

def test_put_get_closed_form(oracle):
    rng = np.random.default_rng(11)
    heaps = [np.zeros(4096, np.uint8) for _ in range(3)]
    before = [h.copy() for h in heaps]
    src = rng.integers(0, 256, 777, dtype=np.uint8)
    oracle.put(heaps, 123, src, 2)
    assert np.array_equal(heaps[2][123:900], src)
    assert np.array_equal(heaps[2][:123], before[2][:123]) and np.array_equal(heaps[2][900:], before[2][900:])
    assert np.array_equal(heaps[0], before[0]) and np.array_equal(heaps[1], before[1])
    back = np.zeros(777, np.uint8)
    oracle.get(heaps, back, 123, 2)
    assert np.array_equal(back, src)


def test_p_g_closed_form(oracle):
    """p stores exactly the 4 or 8 bytes of the value, little-endian (the
    GPU's byte order), at the target offset of PE pe's heap and nothing else;
    g returns those bytes of PE pe's heap (reference nvshmem_device.cu:445-448,
    469-472).  A NaN's payload bits survive both."""
    L = oracle.lib
    rng = np.random.default_rng(12)
    heaps = [rng.integers(0, 256, 256, dtype=np.uint8) for _ in range(3)]
    before = [h.copy() for h in heaps]

    def ptrs():
        return (C.c_void_p * 3)(*[h.ctypes.data for h in heaps])

    def only_changed(pe, lo, hi):
        for q in range(3):
            changed = set(np.flatnonzero(heaps[q] != before[q]).tolist())
            assert changed <= (set(range(lo, hi)) if q == pe else set())

    L.oracle_p32(ptrs(), 17, 0xA1B2C3D4, 1)
    assert heaps[1][17:21].tolist() == [0xD4, 0xC3, 0xB2, 0xA1]
    only_changed(1, 17, 21)
    assert L.oracle_g32(ptrs(), 17, 1) == 0xA1B2C3D4
    assert L.oracle_g32(ptrs(), 17, 0) == int.from_bytes(before[0][17:21].tobytes(), "little")
    before = [h.copy() for h in heaps]
    nan_bits = 0x7FF4000000C0FFEE  # a signalling NaN with a payload
    L.oracle_p64(ptrs(), 101, nan_bits, 2)
    assert heaps[2][101:109].tobytes() == nan_bits.to_bytes(8, "little")
    only_changed(2, 101, 109)
    assert L.oracle_g64(ptrs(), 101, 2) == nan_bits
    assert L.oracle_g64(ptrs(), 101, 0) == int.from_bytes(before[0][101:109].tobytes(), "little")
    # a float's bits travel unchanged through p and g
    f = np.array([np.float32(-0.0), np.float32(np.inf)], np.float32).view(np.uint32)
    for bits in f.tolist():
        L.oracle_p32(ptrs(), 200, bits, 0)
        assert L.oracle_g32(ptrs(), 200, 0) == bits


# ---------------------------------------------------------------- waits

def test_wait_predicates_follow_readme(oracle):
    L = oracle.lib
    for obs, cmp_value in itertools.product([-2, -1, 0, 1, 7], repeat=2):
        # README:157-164: nvshmem_int_wait blocks while *ivar == cmp_value
        assert L.oracle_int_wait_done(obs, cmp_value) == (obs != cmp_value)
        # README:166-174: wait_until blocks until *ivar == cmp_value; cmp ignored
        for cmp in range(-1, 7):
            assert L.oracle_int_wait_until_done(obs, cmp, cmp_value) == (obs == cmp_value)
        ops = [obs == cmp_value, obs != cmp_value, obs > cmp_value, obs >= cmp_value, obs < cmp_value,
               obs <= cmp_value]
        assert [L.oracle_cmp_holds(obs, op, cmp_value) for op in range(6)] == [int(x) for x in ops]


# ---------------------------------------------------------------- barrier

def test_reference_barrier_value_sequence(oracle):
    L = oracle.lib
    vals = [L.oracle_ref_barrier_value(k) for k in range(65540)]
    assert vals[:3] == [1, 2, 3] and vals[65534] == 65535 and vals[65535] == 0 and vals[65536] == 1
    assert all(L.oracle_ref_barrier_half(k) == vals[k] % 2 for k in range(0, 65540, 997))


def _explore_barrier(npes, episodes, arrive, released, init):
    """Enumerate every interleaving of the per-PE steps [arrive (one fetch-add,
    which also returns the generation it observed), wait-until-released] for
    `episodes` barrier episodes and check that no PE leaves episode e before
    all PEs arrived in e.  Returns #states visited."""
    seen = set()
    stack = [((0,) * npes, (None,) * npes, (0,) * npes, init())]
    visited = 0
    while stack:
        phase, gen_seen, ep, state = stack.pop()
        key = (phase, gen_seen, ep, tuple(state))
        if key in seen:
            continue
        seen.add(key)
        visited += 1
        for p in range(npes):
            if ep[p] >= episodes:
                continue
            st = list(state)
            ph, gs, e = list(phase), list(gen_seen), list(ep)
            if ph[p] == 0:
                gs[p] = arrive(st)
                ph[p] = 1
            else:
                if not released(st, gs[p]):
                    continue
                # leaving episode e: every PE must have arrived in e
                arrived = [(ep[q] > e[p]) or (ep[q] == e[p] and phase[q] == 1) for q in range(npes)]
                assert all(arrived), f"PE {p} released early in episode {e[p]}"
                ph[p], e[p] = 0, e[p] + 1
            stack.append((tuple(ph), tuple(gs), tuple(e), tuple(st)))
    return visited


def _word_barrier_model(npes, release_at):
    """the count-and-generation word in Python, releasing after `release_at`
    + 1 arrivals (npes - 1 is the correct protocol)"""
    def init():
        return [0]

    def arrive(st):
        old = st[0]
        st[0] = old + 1
        if (old & 0xFFFF) == release_at:
            st[0] = (((old >> 16) + 1) & 0xFFFF) << 16
        return old >> 16

    def released(st, g):
        return (st[0] >> 16) != g
    return init, arrive, released


@pytest.mark.parametrize("npes,generation", [(1, 0), (2, 0), (3, 0), (4, 0), (3, 65535)])
def test_generation_barrier_exhaustive(oracle, npes, generation):
    """the count-and-generation barrier (oracle state machine) never releases
    a PE early, over every interleaving of 3 episodes; generation 65535 wraps
    to 0 inside the run"""
    b = oracle.gen_barrier(npes, generation)
    L = oracle.lib
    w0 = b[0]
    assert w0 == (generation & 0xFFFF) << 16

    def init():
        return [w0]

    def arrive(st):
        b[0] = st[0]
        g = L.oracle_gen_barrier_arrive(b)
        st[0] = b[0]
        return g

    def released(st, g):
        b[0] = st[0]
        return bool(L.oracle_gen_barrier_released(b, g))
    assert _explore_barrier(npes, 3, arrive, released, init) > 0


def test_generation_barrier_word_known_answer(oracle):
    """three arrivals of a 3-PE episode at generation 7: counts 1, 2, then the
    next generation with a zero count; every arrival saw generation 7"""
    b = oracle.gen_barrier(3, 7)
    seen = [oracle.lib.oracle_gen_barrier_arrive(b) for _ in range(2)]
    assert seen == [7, 7] and b[0] == (7 << 16) + 2
    assert not oracle.lib.oracle_gen_barrier_released(b, 7)
    assert oracle.lib.oracle_gen_barrier_arrive(b) == 7 and b[0] == 8 << 16
    assert oracle.lib.oracle_gen_barrier_released(b, 7)
    b = oracle.gen_barrier(2, 65535)
    oracle.lib.oracle_gen_barrier_arrive(b)
    oracle.lib.oracle_gen_barrier_arrive(b)
    assert b[0] == 0  # generation wraps to 0


def test_barrier_checker_catches_planted_mistake():
    """the exhaustive checker fails a barrier that releases one arrival early
    (mistake planted in a CPU model, never on the GPU)"""
    init, arrive, released = _word_barrier_model(3, release_at=1)
    with pytest.raises(AssertionError, match="released early"):
        _explore_barrier(3, 2, arrive, released, init)
    init, arrive, released = _word_barrier_model(3, release_at=2)
    assert _explore_barrier(3, 2, arrive, released, init) > 0


# ---------------------------------------------------------------- trace CSV

def test_trace_csv_known_answer(oracle):
    """header text of nvshmem.cu:390 and a row of the nvshmem.cu:411 format"""
    assert oracle.lib.oracle_trace_header().decode() == (
        "# timestamp, gridid, blockid, tid, type, local_address, remote_address, size, remote_pe, timespent\n")
    row = oracle.trace_row(123, 456, 7, 8, 0, 0x1000, 0x2000, 4, 1, 9)
    assert row == "123, 456, 7, 8, put_int, 0x1000, 0x2000, 4, 1, 9\n"
    row = oracle.trace_row(5, 6, 0, 0, 7, 0x10, 0, 1, 0, 250)
    assert row == "5, 6, 0, 0, wait_until_int, 0x10, (nil), 1, 0, 250\n"
    names = [oracle.lib.oracle_trace_op_name(i).decode() for i in range(8)]
    assert names == ["put_int", "put_float", "put_double", "get_int", "get_float", "get_double", "wait_int",
                     "wait_until_int"]


# ---------------------------------------------------------------- cost model (MLP)

def test_mlp_gradient_matches_finite_differences(oracle):
    """the oracle's analytic gradient against central differences (double),
    every parameter of a small model"""
    rng = np.random.default_rng(0)
    c = 5
    npar = oracle.mlp_num_params(c)
    P = rng.normal(0, 0.5, npar)
    X = rng.normal(0, 1, (7, 4))
    T = rng.normal(0, 1, (7, c))
    loss, g = oracle.mlp_loss_grad(P, X, T, c)
    h = 1e-6
    fd = np.zeros(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        fd[i] = (oracle.mlp_loss_grad(P + e, X, T, c)[0] - oracle.mlp_loss_grad(P - e, X, T, c)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)
    Y = oracle.mlp_forward(P, X, c)
    assert abs(loss - np.mean((Y - T) ** 2)) < 1e-12


def test_mlp_forward_closed_forms(oracle):
    """zero weights give the output biases; a single active path is linear"""
    c = 3
    npar = oracle.mlp_num_params(c)
    assert npar == (32 * 4 + 32) + (32 * 32 + 32) + (c * 32 + c)  # W1 b1 W2 b2 W3 b3
    P = np.zeros(npar)
    b3 = npar - c
    P[b3:] = [1.5, -2.0, 0.25]
    X = np.random.default_rng(1).normal(size=(4, 4))
    np.testing.assert_array_equal(oracle.mlp_forward(P, X, c), np.tile(P[b3:], (4, 1)))
    # W1[0][0] = 1, W2[0][0] = 1, W3[j][0] = j + 1: y_j = (j + 1) * relu(x0) + b3_j
    P[0] = 1.0
    w2 = 4 * 32 + 32
    P[w2] = 1.0
    w3 = w2 + 32 * 32 + 32
    for j in range(c):
        P[w3 + j * 32] = j + 1
    Y = oracle.mlp_forward(P, X, c)
    want = np.maximum(X[:, :1], 0) * np.arange(1, c + 1) + P[b3:]
    np.testing.assert_allclose(Y, want, rtol=0, atol=1e-15)


def test_jacobi_plane_waves_are_eigenvectors(oracle):
    """cos(2 pi k.x / G) is an eigenvector of the periodic 7-point update with
    eigenvalue 1/4 + 1/4 (cos 2 pi kx/G + cos 2 pi ky/G + cos 2 pi kz/G): pins the
    coefficients, the three axes (distinct k per axis) and the periodic wrap"""
    G = 8
    x, y, z = np.meshgrid(np.arange(G), np.arange(G), np.arange(G), indexing="ij")
    for k in [(1, 2, 3), (3, 0, 1), (0, 0, 4), (2, 5, 7)]:
        u = np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) / G)
        lam = 0.25 + 0.25 * sum(np.cos(2 * np.pi * ki / G) for ki in k)
        np.testing.assert_allclose(oracle.jacobi_step(u), lam * u, rtol=0, atol=1e-14)


def test_jacobi_exact_special_cases(oracle):
    """a constant (dyadic) field is a fixed point, bit for bit; a unit spike at a
    corner spreads to exactly 1/4 on itself and 1/8 on its six periodic
    neighbours and nothing else"""
    G = 6
    c = np.full((G, G, G), 0.375)
    np.testing.assert_array_equal(oracle.jacobi_step(c), c)
    u = np.zeros((G, G, G))
    u[0, 0, G - 1] = 1.0
    want = np.zeros((G, G, G))
    want[0, 0, G - 1] = 0.25
    for d in [(1, 0, G - 1), (G - 1, 0, G - 1), (0, 1, G - 1), (0, G - 1, G - 1), (0, 0, 0), (0, 0, G - 2)]:
        want[d] = 0.125
    np.testing.assert_array_equal(oracle.jacobi_step(u), want)


def test_jacobi_run_conserves_sum_and_stays_in_range(oracle):
    """the update's weights sum to one, so the periodic sum is conserved (up to
    rounding) and values stay in the initial [0, 1) range"""
    G = 12
    u0 = oracle.jacobi(G, 0, 42)
    assert u0.min() >= 0.0 and u0.max() < 1.0
    assert np.all(u0 * 2.0 ** 53 == np.floor(u0 * 2.0 ** 53))
    # the initial field is the closed form: cell (x, y, z) holds uniform value
    # number (x * G + y) * G + z of the seeded stream (word pair 2g, 2g + 1)
    np.testing.assert_array_equal(u0.ravel(), oracle.uniform01_f64(G ** 3, 42))
    u5 = oracle.jacobi(G, 5, 42)
    assert abs(u5.sum() - u0.sum()) < 1e-12 * u0.sum()
    assert u5.min() >= u0.min() and u5.max() <= u0.max()
    # steps compose: 5 steps = 2 steps then 3 more
    u2 = oracle.jacobi(G, 2, 42)
    np.testing.assert_array_equal(oracle.jacobi_step(oracle.jacobi_step(oracle.jacobi_step(u2))), u5)


def test_pcg32_jump_ahead_matches_the_stream(oracle):
    """words offset..offset+n-1 by jump-ahead equal the same slice of the
    sequentially generated stream (offsets 0, 1, odd, powers of two, large)"""
    import ctypes as C
    fill_at = oracle.lib.oracle_pcg32_fill_at
    fill_at.restype, fill_at.argtypes = None, [C.c_void_p, C.c_size_t, C.c_uint64, C.c_uint64, C.c_uint64]
    full = oracle.pcg32(70000, 42, 54)
    for off in [0, 1, 2, 3, 1023, 1024, 4097, 65536, 69000]:
        n = min(500, 70000 - off)
        got = np.zeros(n, np.uint32)
        fill_at(got.ctypes.data, n, 42, 54, off)
        np.testing.assert_array_equal(got, full[off:off + n])


def test_strided_and_2d_closed_forms(oracle):
    """iput/iget/put_2d against numpy strided slicing (an independent
    restatement): only the addressed elements change, every one of them to
    its source element; a planted transposition of strides is caught"""
    rng = np.random.default_rng(5)
    for w, dt in [(4, np.uint32), (8, np.uint64)]:
        for tst, sst, n in [(1, 1, 7), (3, 1, 10), (1, 4, 9), (5, 2, 13), (2, 7, 1)]:
            src = rng.integers(0, 2 ** 31, size=sst * n + 3).astype(dt)
            heaps = [np.zeros(4096, np.uint8) for _ in range(2)]
            oracle.iput(heaps, 16, tst, src, sst, n, w, 1)
            want = np.zeros(4096, np.uint8)
            view = want[16:16 + (tst * (n - 1) + 1) * w].view(dt)
            view[::tst] = src[::sst][:n]
            np.testing.assert_array_equal(heaps[1], want)
            assert not heaps[0].any()
            # iget reads back what iput wrote
            back = np.zeros(sst * n + 3, dt)
            oracle.iget(heaps, back, sst, 16, tst, n, w, 1)
            np.testing.assert_array_equal(back[::sst][:n], src[::sst][:n])
            if tst != sst and n > 1:
                wrong = np.zeros(4096, np.uint8)
                wrong[16:16 + (sst * (n - 1) + 1) * w].view(dt)[::sst] = src[::tst][:n] if len(src[::tst]) >= n else 0
                assert not np.array_equal(heaps[1], wrong)
    for width, height, tpitch, spitch in [(5, 3, 8, 5), (64, 4, 100, 70), (1, 9, 3, 2)]:
        src = rng.integers(0, 256, size=spitch * height + 1).astype(np.uint8)
        heaps = [np.zeros(2048, np.uint8) for _ in range(2)]
        oracle.put_2d(heaps, 32, tpitch, src, spitch, width, height, 0)
        want = np.zeros(2048, np.uint8)
        for r in range(height):
            want[32 + r * tpitch:32 + r * tpitch + width] = src[r * spitch:r * spitch + width]
        np.testing.assert_array_equal(heaps[0], want)
        assert not heaps[1].any()


def test_jacobi_threaded_run_equals_the_serial_run(oracle):
    """the threaded replay gives the serial run's field bit for bit: slabs of
    unequal size, more threads than planes, one thread, zero and odd step
    counts (which buffer holds the result)"""
    for G, steps, threads in [(24, 5, 3), (24, 4, 7), (9, 3, 16), (5, 2, 8), (12, 1, 1), (12, 0, 4), (16, 6, 5)]:
        want = oracle.jacobi(G, steps, 42)
        got = oracle.jacobi_mt(G, steps, 42, threads)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (G, steps, threads)
    # the serial run it is checked against moves: a threaded run that skipped a
    # step would equal neither neighbour
    a, b = oracle.jacobi_mt(16, 5, 42, 4), oracle.jacobi(16, 4, 42)
    assert not np.array_equal(a, b)


# ---------------------------------------------------------------- coverage

# Each function oracle.h declares, and the CPU test that pins it.
PINNED_BY = {
    "oracle_pcg32_fill": ("test_oracle", "test_pcg32_published_known_answer"),
    "oracle_pcg32_fill_at": ("test_oracle", "test_pcg32_jump_ahead_matches_the_stream"),
    "oracle_pcg32_uniform_pm1": ("test_oracle", "test_uniform_pm1_is_exact_and_in_range"),
    "oracle_pcg32_uniform01_f64": ("test_oracle", "test_uniform01_f64_is_exact_and_in_range"),
    "oracle_heap_init": ("test_oracle", "test_heap_reference_arithmetic"),
    "oracle_nvshmalloc": ("test_oracle", "test_heap_reference_arithmetic"),
    "oracle_nvshmcleanup": ("test_oracle", "test_heap_reference_arithmetic"),
    "oracle_translate": ("test_oracle", "test_translate_process_mode_is_reference_formula"),
    "oracle_put": ("test_oracle", "test_put_get_closed_form"),
    "oracle_get": ("test_oracle", "test_put_get_closed_form"),
    "oracle_p32": ("test_oracle", "test_p_g_closed_form"),
    "oracle_p64": ("test_oracle", "test_p_g_closed_form"),
    "oracle_g32": ("test_oracle", "test_p_g_closed_form"),
    "oracle_g64": ("test_oracle", "test_p_g_closed_form"),
    "oracle_iput": ("test_oracle", "test_strided_and_2d_closed_forms"),
    "oracle_iget": ("test_oracle", "test_strided_and_2d_closed_forms"),
    "oracle_put_2d": ("test_oracle", "test_strided_and_2d_closed_forms"),
    "oracle_int_wait_done": ("test_oracle", "test_wait_predicates_follow_readme"),
    "oracle_int_wait_until_done": ("test_oracle", "test_wait_predicates_follow_readme"),
    "oracle_cmp_holds": ("test_oracle", "test_wait_predicates_follow_readme"),
    "oracle_ref_barrier_value": ("test_oracle", "test_reference_barrier_value_sequence"),
    "oracle_ref_barrier_half": ("test_oracle", "test_reference_barrier_value_sequence"),
    "oracle_gen_barrier_init": ("test_oracle", "test_generation_barrier_exhaustive"),
    "oracle_gen_barrier_arrive": ("test_oracle", "test_generation_barrier_exhaustive"),
    "oracle_gen_barrier_released": ("test_oracle", "test_generation_barrier_exhaustive"),
    "oracle_trace_header": ("test_oracle", "test_trace_csv_known_answer"),
    "oracle_trace_op_name": ("test_oracle", "test_trace_csv_known_answer"),
    "oracle_trace_row": ("test_oracle", "test_trace_csv_known_answer"),
    "oracle_jacobi_init": ("test_oracle", "test_jacobi_run_conserves_sum_and_stays_in_range"),
    "oracle_jacobi_step": ("test_oracle", "test_jacobi_plane_waves_are_eigenvectors"),
    "oracle_jacobi_run": ("test_oracle", "test_jacobi_run_conserves_sum_and_stays_in_range"),
    "oracle_jacobi_run_mt": ("test_oracle", "test_jacobi_threaded_run_equals_the_serial_run"),
    "oracle_batch_generate": ("test_batch_cpu", "test_generator_matches_oracle"),
    "oracle_batch_apply": ("test_batch_cpu", "test_batch_apply_matches_sequential_puts"),
    "oracle_mlp_num_params": ("test_oracle", "test_mlp_forward_closed_forms"),
    "oracle_mlp_forward": ("test_oracle", "test_mlp_forward_closed_forms"),
    "oracle_mlp_loss_grad": ("test_oracle", "test_mlp_gradient_matches_finite_differences"),
}


def test_every_oracle_function_is_pinned():
    """every function oracle.h declares has a CPU pinning test, named here, that
    exists and calls it (directly or through its wrapper in oracle_py.py)"""
    import re
    import ast
    tests_dir = Path(__file__).resolve().parent
    header = (tests_dir.parent / "oracle" / "oracle.h").read_text()
    declared = set(re.findall(r"\b(oracle_[a-z0-9_]+)\s*\(", header))
    assert declared == set(PINNED_BY), (sorted(declared - set(PINNED_BY)), sorted(set(PINNED_BY) - declared))
    # wrapper methods of oracle_py.Oracle and the C functions each one calls
    wrappers = {}
    for node in ast.walk(ast.parse((tests_dir / "oracle_py.py").read_text())):
        if isinstance(node, ast.FunctionDef) and node.name != "__init__":
            wrappers[node.name] = set(re.findall(r"oracle_[a-z0-9_]+", ast.unparse(node)))
    for fn, (module, test) in PINNED_BY.items():
        tree = ast.parse((tests_dir / f"{module}.py").read_text())
        defs = {n.name: n for n in tree.body if isinstance(n, ast.FunctionDef)}
        assert test in defs, f"{fn}: {module}.{test} does not exist"
        body = ast.unparse(defs[test])
        called = set(re.findall(r"oracle_[a-z0-9_]+", body))
        for meth, cfns in wrappers.items():
            if re.search(r"\." + meth + r"\(", body):
                called |= cfns
        assert fn in called, f"{fn}: {module}.{test} does not call it"
