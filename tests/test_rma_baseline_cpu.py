"""The CPU baseline (oracle/mpi_rma_baseline.c: the reference's semantics as
host MPI-3 RMA over a shared-memory window, run by MPICH on this host's
cores) is test and bench infrastructure; these CPU tests pin what it
reports: config 1 at its stated size (10,000 p + wait round trips, then a
1 MiB float put of PCG32(seed 42) uniform[-1,1) values, bit-exact), put and
get of every dtype bit-exact, the Jacobi application bit-exact against the
host replay, the uneven-put batch bit-exact, barrier episodes through the window, and a rank that never
reaches a barrier: the others' waits time out and the job ends with a
reported error instead of hanging (the reference's barrier would continue,
nvshmem_kernels.cu:65-68)."""
import json
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "oracle" / "build" / "rma_baseline"
MPIRUN = Path("/opt/conda/bin/mpirun")


def run(n, *args, env_extra=None, timeout=120):
    if not MPIRUN.exists() or not EXE.exists():
        pytest.skip("MPICH mpirun or the baseline program is not here")
    env = {k: v for k, v in os.environ.items() if not k.startswith("NVSHMEM_")}
    env.update(env_extra or {})
    return subprocess.run([str(MPIRUN), "-n", str(n), str(EXE)] + [str(a) for a in args], capture_output=True,
                          text=True, timeout=timeout, cwd="/tmp", env=env)


def records(p):
    return [json.loads(l) for l in p.stdout.splitlines() if l.startswith("{")]


def test_pingpong_and_1mib_put_config1():
    p = run(2, "pingpong", 10000)
    assert p.returncode == 0, p.stderr[-2000:]
    rec = records(p)[-1]
    assert rec["iters"] == 10000 and rec["put_1mib_bitexact"] is True, rec
    assert 0 < rec["rtt_p50_ns"] <= rec["rtt_p99_ns"]


@pytest.mark.parametrize("op", ["put", "get"])
@pytest.mark.parametrize("dtype", ["int", "float", "double"])
def test_putget_bitexact(op, dtype):
    p = run(2, "putget", op, dtype, 8, 1 << 16, 10, 1)
    assert p.returncode == 0, p.stderr[-2000:]
    recs = records(p)
    assert [r["bytes"] for r in recs] == [8 << k for k in range(14)]
    assert all(r["bitexact"] is True and r["mode"] == op and r["dtype"] == dtype for r in recs), recs


def test_jacobi_bitexact_against_host_replay():
    p = run(8, "jacobi", 8, 3, 1)
    assert p.returncode == 0, p.stderr[-2000:]
    assert records(p)[-1]["bitexact"] is True


def test_batch_bitexact():
    """the uneven-put workload's generator and sizes (seed 7), 64 logical PEs
    over 8 ranks, every destination checked"""
    p = run(8, "batch", 512, 7, 1)
    assert p.returncode == 0, p.stderr[-2000:]
    rec = records(p)[-1]
    assert rec["pes"] == 64 and rec["bitexact"] is True, rec


def test_barrier_episodes():
    p = run(4, "barrier", 200, -1)
    assert p.returncode == 0, p.stderr[-2000:]
    rec = records(p)[-1]
    assert rec["ranks"] == 4 and rec["rounds"] == 200


def test_barrier_skipped_by_one_rank_is_reported():
    """rank 2 leaves out the last of 20 episodes: the others' waits give up
    after the 300 ms budget, report it and end the job with status 3"""
    p = run(4, "barrier", 20, 2, env_extra={"NVSHMEM_WAIT_TIMEOUT_MS": "300"}, timeout=60)
    assert p.returncode != 0
    assert "barrier episode" in p.stderr and "timed out after 0.300 s" in p.stderr, p.stderr[-2000:]
