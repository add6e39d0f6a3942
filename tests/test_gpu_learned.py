"""What the trained cost model gains over the methods without it, on the GPU.

The autotuner measures a small grid (integer octaves 64 B .. 4 MiB, 4-byte
elements, aligned, four concurrency levels); the MLP is trained on those
samples with the library's own kernels.  Held-out batches are drawn from the
uneven-put workload (log-uniform sizes 64 B .. 4 MiB, 8 PEs).  Every plan is
timed as a captured hipGraph of back-to-back launches (HIP events), so launch
overhead does not blur microsecond-scale batches.

Compared with the trained network's choices:
  - the library's untuned rule (plans created without choices): the rule
    cannot see how many puts share the launch, so on a 16-put batch it leaves
    most of the device idle;
  - one workgroup per put (the plain fused batch with no splitting);
  - the same network before training (its random initial weights), reported
    beside them; its fit to the measurements is what is asserted.
Results are bit-exact in every case (checked once per plan kind)."""
import ctypes as C

import numpy as np
import pytest

import nvshmem_amd as na

pytestmark = pytest.mark.gpu
NPES, WGS = 8, 4
SRC, DST = 32 << 20, 64 << 20
REPLAYS = 20


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


def _graph_ms(lib, hip, arr, n, choices):
    plan = C.c_void_p()
    assert lib.nvshmemx_batch_plan_create(arr, n, choices, None, C.byref(plan)) == 0
    try:
        assert lib.nvshmemx_batch_plan_capture(plan, REPLAYS) == 0
        lib.nvshmemx_batch_plan_graph_launch(plan, None)   # warm-up
        e0, e1 = hip.event(), hip.event()
        hip.record(e0)
        lib.nvshmemx_batch_plan_graph_launch(plan, None)
        hip.record(e1)
        hip.sync()
        ms = hip.elapsed_ms(e0, e1) / REPLAYS
        hip.event_destroy(e0)
        hip.event_destroy(e1)
        assert int(lib.nvshmemh_verify_requests(arr, n)) == 0
        return ms
    finally:
        lib.nvshmemx_batch_plan_destroy(plan)


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


def test_trained_choice_beats_the_methods_without_it(setup, hip):
    lib, base, trained, untrained, _, _ = setup
    rec = {"trained": [], "rule": [], "one_piece": [], "untrained": []}
    for arr in _batches(lib, base, 16, 8, 7000):
        n = len(arr)
        rec["trained"].append(_graph_ms(lib, hip, arr, n, trained.choose(arr, n, 0)))
        rec["rule"].append(_graph_ms(lib, hip, arr, n, None))
        rec["one_piece"].append(_graph_ms(lib, hip, arr, n, _one_piece(n)))
        rec["untrained"].append(_graph_ms(lib, hip, arr, n, untrained.choose(arr, n, 0)))
    mean = {k: float(np.mean(v)) for k, v in rec.items()}
    print("16-put batches, mean ms:", {k: round(v, 4) for k, v in mean.items()})
    assert mean["rule"] > 1.3 * mean["trained"]
    assert mean["one_piece"] > 2.0 * mean["trained"]
    assert lib.nvshmemx_get_last_error() == 0
