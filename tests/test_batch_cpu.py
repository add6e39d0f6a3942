"""CPU checks of the uneven-put workload: the generator the benchmark uses
(library harness) draws exactly the oracle's requests, and the requests have
the stated distribution (log-uniform 64 B..4 MiB, 4-byte multiples, src != dst,
256-byte aligned sources inside their region, disjoint packed destinations)."""
import numpy as np
import pytest

import nvshmem_amd as na

SRC, DST = 32 << 20, 64 << 20


@pytest.mark.parametrize("seed,npes", [(7, 64), (8, 8), (123456789, 2)])
def test_generator_matches_oracle(lib, oracle, seed, npes):
    dst = DST * max(1, 64 // npes)
    a = na.gen_requests(lib, seed, 4096, npes, SRC, dst)
    b = oracle.batch_generate(seed, 4096, npes, SRC, dst)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_workload_distribution(oracle):
    r = oracle.batch_generate(7, 4096, 64, SRC, DST)
    n = r["nbytes"].astype(np.int64)
    assert n.min() >= 64 and n.max() <= (4 << 20) and (n % 4 == 0).all()
    # log-uniform on [6, 22]: log2 sizes spread evenly; median near 2^14
    lg = np.log2(n)
    assert abs(np.median(lg) - 14) < 0.3
    hist, _ = np.histogram(lg, bins=8, range=(6, 22))
    assert hist.min() > 0.8 * 4096 / 8 and hist.max() < 1.2 * 4096 / 8
    assert (r["src_pe"] != r["dst_pe"]).all()
    assert (r["src_off"] % 256 == 0).all() and ((r["src_off"] + r["nbytes"]) <= SRC).all()
    assert abs(np.mean(r["dtype"]) - 0.5) < 0.05
    # destinations: disjoint within each PE and inside the region
    for pe in range(64):
        m = r["dst_pe"] == pe
        lo, hi = r["dst_off"][m].astype(np.int64), (r["dst_off"][m] + r["nbytes"][m]).astype(np.int64)
        order = np.argsort(lo)
        assert (lo[order][1:] >= hi[order][:-1]).all()
        assert (hi <= DST).all() and (lo % 256 == 0).all()
    # ~1.55 GB payload per step, as the workload states
    assert 1.2e9 < n.sum() < 1.9e9


def test_batch_apply_matches_sequential_puts(oracle):
    rng = np.random.default_rng(3)
    r = oracle.batch_generate(11, 300, 4, 1 << 16, 1 << 22, 6, 12)
    heaps = [rng.integers(0, 256, (1 << 16) + (1 << 22), dtype=np.uint8) for _ in range(4)]
    want = [h.copy() for h in heaps]
    for i in range(300):
        s, d = int(r["src_pe"][i]), int(r["dst_pe"][i])
        so, do, nb = int(r["src_off"][i]), (1 << 16) + int(r["dst_off"][i]), int(r["nbytes"][i])
        want[d][do:do + nb] = want[s][so:so + nb]
    oracle.batch_apply(heaps, r, 0, 1 << 16)
    for h, w in zip(heaps, want):
        np.testing.assert_array_equal(h, w)
