"""3-D 7-point Jacobi with halo exchange through the symmetric heap, 8 PEs in
a periodic 2x2x2 grid, bit-exact against the oracle's host replay of the whole
undecomposed domain.  Local mode (all 8 PEs in one process, kernel boundaries
as the barrier) at 16^3 per PE for 100 steps and at the full 256^3 per PE for
3 and 100 steps; process mode (8 processes, nvshmem_barrier_all_offload per
step) at 16^3 for 40 steps and at 256^3 for 100.  A step is normally one fused
kernel (halo read from the receive slots, faces written into the neighbours'
slots); the separate pack / unpack / stencil kernels are checked as well, and
so is the two-steps-per-launch kernel (cells past a PE's block read from the
neighbours' fields, no halo exchange; an odd last step fused)."""
import json
import os
import subprocess
import sys
import uuid
from pathlib import Path

import numpy as np
import pytest

import nvshmem_amd as na

pytestmark = pytest.mark.gpu
HERE = Path(__file__).resolve().parent
SEED = 42


def _run_local(lib, hip, n, steps, chunk, two_step=False):
    fb = lib.nvshmemh_jacobi_field_bytes(n)
    u, v = lib.nvshmalloc(fb), lib.nvshmalloc(fb)
    recv = lib.nvshmalloc(lib.nvshmemh_jacobi_recv_bytes(n))
    for pe in range(8):
        for p in (u, v):
            hip.memset(lib.nvshmem_ptr(p, pe), 0, fb)
    lib.nvshmemh_jacobi_init(u, n, 0, 8, SEED, None)
    chunk = 0 if chunk is None else chunk
    if two_step:
        out = lib.nvshmemh_jacobi_run2(u, v, recv, n, steps, chunk, None)
    else:
        out = lib.nvshmemh_jacobi_run(u, v, recv, n, steps, chunk, 0, None)
    hip.sync()
    assert lib.nvshmemx_get_last_error() == 0
    return out


def _check_blocks(hip, lib, out, want, n):
    for pe in range(8):
        got = na.jacobi_interior(hip, lib.nvshmem_ptr(out, pe), n, lib)
        blk = na.jacobi_block(want, pe, n)
        if not np.array_equal(got.view(np.uint64), blk.view(np.uint64)):
            bad = np.argwhere(got.view(np.uint64) != blk.view(np.uint64))
            raise AssertionError(f"PE {pe}: {len(bad)} cells differ, first {bad[0]}: {got[tuple(bad[0])]!r} vs "
                                 f"{blk[tuple(bad[0])]!r}")


_FULL = {}


def _full_domain_replay(oracle, n, steps):
    """the host replay of the 8 x 256^3 workload, computed once per session"""
    if (n, steps) not in _FULL:
        _FULL[(n, steps)] = oracle.jacobi_mt(2 * n, steps, SEED, max(1, min(16, os.cpu_count() or 1)))
    return _FULL[(n, steps)]


@pytest.mark.parametrize("n,steps,chunk", [(16, 100, None), (16, 7, 5), (48, 5, 7), (13, 6, 5), (130, 3, 64),
                                           (256, 3, 64)])
def test_jacobi_local_bitexact(local_pes, hip, oracle, n, steps, chunk):
    """n = 13 and 130 leave partial tiles in both j (8-row tiles) and k
    (128-column tiles) and odd row lengths"""
    heap = "16M" if n <= 16 else "320M"
    lib = local_pes.start(8, heap=heap)
    out = _run_local(lib, hip, n, steps, chunk)
    want = oracle.jacobi(2 * n, steps, SEED)
    for pe in range(8):
        got = na.jacobi_interior(hip, lib.nvshmem_ptr(out, pe), n, lib)
        blk = na.jacobi_block(want, pe, n)
        if not np.array_equal(got.view(np.uint64), blk.view(np.uint64)):
            bad = np.argwhere(got != blk)
            raise AssertionError(f"PE {pe}: {len(bad)} cells differ, first {bad[0]}: {got[tuple(bad[0])]!r} vs "
                                 f"{blk[tuple(bad[0])]!r}")


def test_jacobi_full_domain_100_steps(local_pes, hip, oracle):
    """the benchmarked workload run to the end: 8 PEs x 256^3 (a 512^3 global
    domain), 100 fused steps at the library's default chunk of planes, every
    cell bit-exact against the host replay of the whole domain (the replay's
    planes split over up to 16 host threads, same arithmetic per cell)"""
    n, steps = 256, 100
    lib = local_pes.start(8, heap="320M")
    out = _run_local(lib, hip, n, steps, None)
    _check_blocks(hip, lib, out, _full_domain_replay(oracle, n, steps), n)


@pytest.mark.parametrize("n,steps,chunk", [(16, 100, None), (16, 7, 5), (13, 6, 5), (48, 4, 7), (2, 6, 1), (3, 5, 2),
                                           (130, 4, 64), (128, 6, 127), (256, 4, None)])
def test_jacobi_two_step_local_bitexact(local_pes, hip, oracle, n, steps, chunk):
    """two steps per launch, every cell past a PE's block read from the
    neighbour's field: bit-exact against the host replay.  n = 2 and 3 fold
    two-cell halos onto blocks barely larger than them, 13 and 130 leave
    partial tiles and odd rows, odd step counts end with one fused step, and
    chunks run from one plane to the whole column"""
    heap = "16M" if n <= 16 else "320M"
    lib = local_pes.start(8, heap=heap)
    out = _run_local(lib, hip, n, steps, chunk, two_step=True)
    _check_blocks(hip, lib, out, oracle.jacobi(2 * n, steps, SEED), n)


def test_jacobi_two_step_full_domain_100_steps(local_pes, hip, oracle):
    """the benchmarked workload, 8 PEs x 256^3 for 100 steps as 50 two-step
    launches at the default chunk, bit-exact against the host replay"""
    n, steps = 256, 100
    lib = local_pes.start(8, heap="320M")
    out = _run_local(lib, hip, n, steps, None, two_step=True)
    _check_blocks(hip, lib, out, _full_domain_replay(oracle, n, steps), n)


@pytest.mark.parametrize("n,steps,chunk", [(128, 10, None), (128, 5, 2), (128, 4, 5), (128, 3, 127)])
def test_jacobi_chunks_bitexact(local_pes, hip, oracle, n, steps, chunk):
    """whole tiles in j and k with chunks of planes from 2 to 127 (a ragged
    last chunk) and the whole column"""
    lib = local_pes.start(8, heap="320M")
    out = _run_local(lib, hip, n, steps, chunk)
    want = oracle.jacobi(2 * n, steps, SEED)
    for pe in range(8):
        got = na.jacobi_interior(hip, lib.nvshmem_ptr(out, pe), n, lib)
        np.testing.assert_array_equal(got.view(np.uint64), na.jacobi_block(want, pe, n).view(np.uint64))


@pytest.mark.parametrize("n", [16, 48, 13, 130, 128])
@pytest.mark.parametrize("pack", ["direct", "api"])
def test_jacobi_separate_kernels_bitexact(local_pes, hip, oracle, n, pack):
    """the unfused step (pack+put of u's faces, unpack into u's halo, stencil
    reading the halo from u) gives the same bits as the fused step; n = 48 has
    partial stencil tiles in k, n = 13 and 130 in j and k.  The faces go out
    either by the direct-store pack kernel or through the library's own
    device puts (contiguous wavefront puts for the rows of the x and y faces,
    strided wavefront puts for the z faces)"""
    lib = local_pes.start(8, heap="16M" if n <= 48 else "96M")
    fb = lib.nvshmemh_jacobi_field_bytes(n)
    u, v = lib.nvshmalloc(fb), lib.nvshmalloc(fb)
    recv = lib.nvshmalloc(lib.nvshmemh_jacobi_recv_bytes(n))
    for pe in range(8):
        for p in (u, v):
            hip.memset(lib.nvshmem_ptr(p, pe), 0, fb)
    lib.nvshmemh_jacobi_init(u, n, 0, 8, SEED, None)
    steps = 6
    a = lib.nvshmemh_jacobi_run_separate(u, v, recv, n, steps, 7, 1 if pack == "api" else 0, None)
    hip.sync()
    assert lib.nvshmemx_get_last_error() == 0
    want = oracle.jacobi(2 * n, steps, SEED)
    for pe in range(8):
        got = na.jacobi_interior(hip, lib.nvshmem_ptr(a, pe), n, lib)
        np.testing.assert_array_equal(got.view(np.uint64), na.jacobi_block(want, pe, n).view(np.uint64))


@pytest.mark.parametrize("n,steps,how", [(16, 40, "fused"), (128, 6, "fused"), (16, 12, "api"), (48, 5, "api"),
                                         (16, 41, "two_step"), (128, 6, "two_step"), (256, 100, "fused"),
                                         (256, 100, "two_step")])
def test_jacobi_processes_bitexact(tmp_path, oracle, n, steps, how):
    """8 processes, one PE each: the fused step (faces stored straight into
    the neighbours' receive slots) and the unfused step whose faces go through
    the library's contiguous and strided device puts, the offload barrier
    between; the last case is the benchmarked workload, 256^3 per PE for 100
    steps"""
    name = f"nvshmem_jacobi_{os.getpid()}_{uuid.uuid4().hex[:8]}"
    heap = "16M" if n <= 16 else ("64M" if n <= 128 else "320M")
    procs = [subprocess.Popen([sys.executable, str(HERE / "pe_worker.py"), "jacobi", str(n), str(steps),
                               str(tmp_path), how],
                              env=na.shm_job_env(r, 8, name, {"NVSHMEM_HEAP_SIZE": heap,
                                                              "NVSHMEM_WAIT_TIMEOUT_MS": "5000",
                                                              "NVSHMEM_BOOTSTRAP_TIMEOUT_S": "60"}),
                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for r in range(8)]
    try:
        for p in procs:
            o, e = p.communicate(timeout=120)
            assert p.returncode == 0, e[-2000:]
            assert json.loads([l for l in o.splitlines() if l.startswith("{")][-1])["errors"] == 0
    finally:
        for p in procs:
            if p.poll() is None:
                p.kill()
        shm = Path("/dev/shm") / name
        if shm.exists():
            shm.unlink()
    want = _full_domain_replay(oracle, n, steps)
    for pe in range(8):
        got = np.load(tmp_path / f"jacobi_pe{pe}.npy")
        np.testing.assert_array_equal(got.view(np.uint64), na.jacobi_block(want, pe, n).view(np.uint64))
