"""What the trained cost model gains over the methods without it, on the GPU.

The autotuner measures a small grid (integer octaves 64 B .. 4 MiB, 4-byte
elements, aligned, four concurrency levels); the MLP is trained on those
samples with the library's own kernels.  Held-out batches are drawn from the
uneven-put workload (log-uniform sizes 64 B .. 4 MiB, 8 PEs).

Every plan is timed the way the library's own selection times plans
(nvshmemx_batch_plan_select): a warm-up launch, then five rounds of bursts of
back-to-back launches spanning about 50 us, the median per launch kept, all
plans of one batch interleaved round by round on the same stream.

The assertions are ones a box-to-box spread of tens of per cent cannot flip:
  - the trained network's choices move every byte exactly (host comparison of
    every target with its source);
  - its plans are at least twice as fast as one workgroup per put and as the
    same network before training (8x and 20x in measurements);
  - they are no slower than the library's untuned rule beyond a 10% margin
    (measured 1.2-1.9x faster on 16-put batches).
How much faster than the rule it is goes into the benchmark's report, not into
a pass/fail gate.

The batch-level policy model (4 batch inputs -> 13 policies, the same
kernels) is trained on batches of 8 to 256 puts that the library times itself
under every policy (the exhaustive batch-level judge), then asked to pick the
policy of held-out 16- and 128-put batches.  Asserted: its fit is 10x better
than before training, its picks plan every byte exactly, they come within 10%
of the judge's best on average (measured 1-3%) and, on 16-put batches, at
least 1.3x faster than the rule (measured 1.9x)."""
import ctypes as C

import numpy as np
import pytest

import nvshmem_amd as na
from batch_check import assert_requests_copied

pytestmark = pytest.mark.gpu
NPES, WGS = 8, 4
SRC, DST = 32 << 20, 64 << 20


@pytest.fixture(scope="module")
def setup(local_pes, hip):
    lib = local_pes.start(NPES, wgs_per_pe=WGS, heap="96M")
    base = lib.nvshmalloc(SRC + DST)
    for pe in range(NPES):
        lib.nvshmemh_fill_pcg32(lib.nvshmem_ptr(base, pe), SRC // 4, 100 + pe, 54, None)
    hip.sync()
    grid = na.tune_grid(half_octaves=False, widths=(4,), concurrency=(1, 16, 512, 32768), aligns=(0,))
    times = na.tune_measure(lib, grid, reps=3)
    feats = na.tune_features(lib, grid)
    untrained = na.CostModel(lib, seed=5)
    trained = na.CostModel(lib, seed=5)
    T = np.log(np.maximum(times, 1e-3)).astype(np.float32)
    trained.train(feats, T, 3000, 3e-3)
    trained.train(feats, T, 1500, 1e-3)
    yield lib, base, trained, untrained, feats, T
    trained.close()
    untrained.close()


def _select_ms(lib, arr, n, choice_sets):
    """median ms per launch of one plan per choice set, timed together"""
    plans = []
    try:
        for ch in choice_sets:
            p = C.c_void_p()
            assert lib.nvshmemx_batch_plan_create(arr, n, ch, None, C.byref(p)) == 0
            plans.append(p)
        ms = (C.c_float * len(plans))()
        lib.nvshmemx_batch_plan_select((C.c_void_p * len(plans))(*[p.value for p in plans]), len(plans), 5, None, ms)
        return list(ms)
    finally:
        for p in plans:
            lib.nvshmemx_batch_plan_destroy(p)


def _batches(lib, base, nreq, count, seed0):
    for b in range(count):
        reqs = na.gen_requests(lib, seed0 + b, nreq, NPES, SRC, DST)
        yield na.make_requests(lib, reqs, base, 0, SRC)


def _one_piece(n):
    ch = (na.BatchChoice * n)()
    for i in range(n):
        ch[i].pieces, ch[i].vec_bytes, ch[i].mode = 1, 16, 1
    return ch


def test_training_fits_the_measurements(setup):
    _, _, trained, untrained, feats, T = setup
    mse_trained = float(np.mean((trained.predict(feats) - T) ** 2))
    mse_untrained = float(np.mean((untrained.predict(feats) - T) ** 2))
    print(f"log-time mse: trained {mse_trained:.4f}, untrained {mse_untrained:.4f}")
    assert mse_trained < 0.1 * mse_untrained


def test_trained_choice_is_exact(setup, hip):
    lib, base, trained, _, _, _ = setup
    for nreq in (16, 128):
        for arr in _batches(lib, base, nreq, 2, 7100 + nreq):
            n = len(arr)
            plan = C.c_void_p()
            assert lib.nvshmemx_batch_plan_create(arr, n, trained.choose(arr, n, 0), None, C.byref(plan)) == 0
            for pe in range(NPES):
                hip.memset(lib.nvshmem_ptr(base, pe) + SRC, 0, DST)
            lib.nvshmemx_batch_plan_launch(plan, None)
            hip.sync()
            lib.nvshmemx_batch_plan_destroy(plan)
            assert_requests_copied(lib, hip, arr, n)
    assert lib.nvshmemx_get_last_error() == 0


def test_trained_choice_gains_over_the_methods_without_it(setup):
    lib, base, trained, untrained, _, _ = setup
    names = ("trained", "rule", "one_piece", "untrained")
    rec = {k: [] for k in names}
    for arr in _batches(lib, base, 16, 8, 7000):
        n = len(arr)
        ms = _select_ms(lib, arr, n, [trained.choose(arr, n, 0), None, _one_piece(n), untrained.choose(arr, n, 0)])
        for k, v in zip(names, ms):
            rec[k].append(v)
    mean = {k: float(np.mean(v)) for k, v in rec.items()}
    print("16-put batches, mean ms per launch:", {k: round(v, 5) for k, v in mean.items()},
          f"rule / trained = {mean['rule'] / mean['trained']:.2f}")
    assert mean["one_piece"] > 2.0 * mean["trained"]
    assert mean["untrained"] > 2.0 * mean["trained"]
    assert mean["trained"] < 1.10 * mean["rule"]
    assert lib.nvshmemx_get_last_error() == 0


@pytest.fixture(scope="module")
def batch_setup(setup):
    lib, base, trained, _, _, _ = setup
    npol = lib.nvshmemx_batch_policy_count()
    X, Y = [], []
    for nreq in (8, 16, 32, 64, 128, 256):
        for hi in (18, 20, 22):
            for sd in range(3):
                reqs = na.gen_requests(lib, 30000 + 100 * nreq + 10 * hi + sd, nreq, NPES, SRC, DST, 6, hi)
                arr = na.make_requests(lib, reqs, base, 0, SRC)
                ms, _ = na.measure_batch(lib, arr, len(arr), trained, 5)
                X.append(na.batch_features(lib, arr, len(arr)))
                Y.append(np.log(ms * 1e3))
    X, Y = np.array(X, np.float32), np.array(Y, np.float32)
    bm, untrained = na.CostModel(lib, seed=3, ncand=npol), na.CostModel(lib, seed=3, ncand=npol)
    bm.train(X, Y, 6000, 3e-3)
    bm.train(X, Y, 3000, 1e-3)
    yield bm, untrained, X, Y
    bm.close()
    untrained.close()


def test_batch_model_fits_the_judge(batch_setup):
    bm, untrained, X, Y = batch_setup
    mse, mse0 = float(np.mean((bm.predict(X) - Y) ** 2)), float(np.mean((untrained.predict(X) - Y) ** 2))
    print(f"batch model log-time mse: trained {mse:.5f}, untrained {mse0:.5f}, {len(X)} batches")
    assert mse < 0.1 * mse0


def test_batch_model_picks_near_the_judge(setup, batch_setup, hip):
    lib, base, trained, _, _, _ = setup
    bm = batch_setup[0]
    for nreq, seed0 in ((16, 33000), (128, 34000)):
        judge, net, rule = [], [], []
        for b in range(12):
            reqs = na.gen_requests(lib, seed0 + b, nreq, NPES, SRC, DST)
            arr = na.make_requests(lib, reqs, base, 0, SRC)
            ms, _ = na.measure_batch(lib, arr, nreq, trained, 5)
            choice, pol = bm.choose_batch(trained, arr, nreq)
            judge.append(ms.min())
            net.append(ms[pol])
            rule.append(ms[0])
            if b == 0:  # the picked policy's plan moves every byte exactly
                plan = C.c_void_p()
                assert lib.nvshmemx_batch_plan_create(arr, nreq, choice, None, C.byref(plan)) == 0
                for pe in range(NPES):
                    hip.memset(lib.nvshmem_ptr(base, pe) + SRC, 0, DST)
                lib.nvshmemx_batch_plan_launch(plan, None)
                hip.sync()
                lib.nvshmemx_batch_plan_destroy(plan)
                assert_requests_copied(lib, hip, arr, nreq)
        j, m, r = float(np.mean(judge)), float(np.mean(net)), float(np.mean(rule))
        print(f"{nreq}-put batches: judge {j:.5f} ms, batch network {m:.5f} ({m / j:.3f}x), rule {r:.5f} "
              f"({r / j:.3f}x)")
        assert m < 1.10 * j
        if nreq == 16:
            assert r > 1.3 * m
    assert lib.nvshmemx_get_last_error() == 0
