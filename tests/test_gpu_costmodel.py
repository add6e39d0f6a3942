"""The learned cost model's HIP kernels against the double-precision oracle
(forward within 1e-5 relative, gradient within 1e-4 relative: fp32 vs fp64
with atomically accumulated sums), training on the GPU, and the autotuner's
measurement and table choice."""
import ctypes as C

import numpy as np
import pytest

import nvshmem_amd as na

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib1(local_pes):
    return local_pes.start(2)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(1e-30, float(np.max(np.abs(b)))))


def test_forward_matches_oracle(lib1, oracle):
    m = na.CostModel(lib1, seed=3)
    try:
        rng = np.random.default_rng(2)
        X = rng.normal(0, 1, (1000, 4)).astype(np.float32)
        P = m.params()
        want = oracle.mlp_forward(P.astype(np.float64), X.astype(np.float64), na.NCAND)
        assert _rel(m.predict(X), want) < 1e-5
    finally:
        m.close()


@pytest.mark.parametrize("n", [1, 63, 64, 990, 4097])
def test_gradient_matches_oracle(lib1, oracle, n):
    m = na.CostModel(lib1, seed=n)
    try:
        rng = np.random.default_rng(n)
        X = rng.normal(0, 1, (n, 4)).astype(np.float32)
        T = rng.normal(0, 1, (n, na.NCAND)).astype(np.float32)
        g = m.gradient(X, T)
        _, want = oracle.mlp_loss_grad(m.params().astype(np.float64), X.astype(np.float64), T.astype(np.float64),
                                       na.NCAND)
        assert _rel(g, want) < 1e-4
        # params untouched by a gradient evaluation
        assert np.array_equal(m.params(), m.params())
    finally:
        m.close()


def test_training_reduces_loss_and_tracks_oracle_steps(lib1, oracle, tmp_path):
    rng = np.random.default_rng(7)
    X = rng.uniform(-1, 1, (990, 4)).astype(np.float32)
    # a smooth target every output depends on differently
    T = (np.sin(X @ rng.normal(size=(4, na.NCAND))) + 0.1 * X[:, :1]).astype(np.float32)
    m = na.CostModel(lib1, seed=11)
    try:
        loss0 = float(np.mean((m.predict(X) - T) ** 2))
        m.train(X, T, 1000, 3e-3)
        loss1 = float(np.mean((m.predict(X) - T) ** 2))
        assert loss1 < 0.1 * loss0, (loss0, loss1)
        path = str(tmp_path / "model.bin").encode()
        assert lib1.nvshmemx_costmodel_save(m.h, path) == 0
        m2 = na.CostModel(lib1, seed=99)
        assert lib1.nvshmemx_costmodel_load(m2.h, path) == 0
        np.testing.assert_array_equal(m2.params(), m.params())
        m2.close()
    finally:
        m.close()


def test_adam_step_matches_oracle_update(lib1, oracle):
    """one GPU Adam step equals the textbook update of the oracle gradient"""
    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (200, 4)).astype(np.float32)
    T = rng.normal(0, 1, (200, na.NCAND)).astype(np.float32)
    m = na.CostModel(lib1, seed=5)
    try:
        p0 = m.params().astype(np.float64)
        _, g = oracle.mlp_loss_grad(p0, X.astype(np.float64), T.astype(np.float64), na.NCAND)
        lr = 1e-3
        mh, vh = g, g * g                       # bias-corrected first step
        want = p0 - lr * mh / (np.sqrt(vh) + 1e-8)
        m.train(X, T, 1, lr)
        got = m.params().astype(np.float64)
        assert np.max(np.abs(got - want)) < 2e-6
    finally:
        m.close()


def test_autotuner_measures_every_candidate(lib1):
    grid = {"log2_bytes": np.array([6.0, 14.0, 22.0]), "elem_bytes": np.array([4, 4, 8], np.int32),
            "concurrency": np.array([64, 8, 1], np.int32), "align": np.array([0, 1, 2], np.int32)}
    t = na.tune_measure(lib1, grid, reps=2)
    assert t.shape == (3, na.NCAND) and (t > 0).all() and np.isfinite(t).all()
    # large single request: splitting into many workgroup pieces beats one piece
    cands = na.candidates(lib1)
    one = cands.index((1, 16, 1))
    many = cands.index((64, 16, 1))
    assert t[2, many] < t[2, one]
    assert lib1.nvshmemx_get_last_error() == 0
