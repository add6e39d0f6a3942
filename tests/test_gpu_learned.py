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
a pass/fail gate."""
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
