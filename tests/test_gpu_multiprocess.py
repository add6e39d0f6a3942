"""PEs as separate processes on one GPU: shared-memory bootstrap, heaps
exported with hipIpcGetMemHandle and mapped into every peer.  Each test starts
its PE processes (tests/pe_worker.py) and checks what they report."""
import json
import os
import subprocess
import sys
import uuid
from pathlib import Path

import pytest

import nvshmem_amd as na

pytestmark = pytest.mark.gpu
HERE = Path(__file__).resolve().parent


def run_job(npes, *args, heap="64M", timeout=90, extra=None):
    name = f"nvshmem_test_{os.getpid()}_{uuid.uuid4().hex[:8]}"
    env_extra = {"NVSHMEM_HEAP_SIZE": heap, "NVSHMEM_WAIT_TIMEOUT_MS": "5000",
                 "NVSHMEM_BOOTSTRAP_TIMEOUT_S": "60"}
    env_extra.update(extra or {})
    procs = [subprocess.Popen([sys.executable, str(HERE / "pe_worker.py"), *map(str, args)],
                              env=na.shm_job_env(r, npes, name, env_extra), stdout=subprocess.PIPE,
                              stderr=subprocess.PIPE, text=True) for r in range(npes)]
    outs = []
    try:
        for p in procs:
            o, e = p.communicate(timeout=timeout)
            assert p.returncode == 0, f"PE exited {p.returncode}: {e[-2000:]}"
            outs.append(json.loads([l for l in o.splitlines() if l.startswith("{")][-1]))
    finally:
        for p in procs:
            if p.poll() is None:
                p.kill()
        shm = Path("/dev/shm") / name
        if shm.exists():
            shm.unlink()
    return sorted(outs, key=lambda d: d["pe"])


@pytest.mark.parametrize("npes", [2, 4])
def test_multiprocess_putget_ring(npes):
    outs = run_job(npes, "putget")
    for o in outs:
        assert o["mode"] == na.__dict__.get("NVSHMEMX_MODE_PROCESS", 0) and o["bootstrap"] == "shm"
        assert o["put_ok"] and o["get_ok"] and o["errors"] == 0, o


def test_multiprocess_pingpong():
    iters = 10000
    outs = run_job(2, "pingpong", iters)
    assert all(o["errors"] == 0 for o in outs), outs
    assert all(o["final_flag"] == iters for o in outs), outs
    print(f"2-process p+wait RTT p50 {outs[0]['p50_ns']:.0f} ns p99 {outs[0]['p99_ns']:.0f} ns")


@pytest.mark.parametrize("npes", [2, 8])
def test_multiprocess_offload_barrier_ordering(npes):
    rounds = 200
    outs = run_job(npes, "barrier", rounds)
    for o in outs:
        assert o["errors"] == 0 and o["violations"] == 0, o
        assert o["slots"] == [rounds] * npes


@pytest.mark.parametrize("npes,discipline", [(2, na.BATCH_SOURCE_ORDER), (4, na.BATCH_SOURCE_ORDER),
                                             (4, na.BATCH_SHARED)])
def test_multiprocess_batch(npes, discipline):
    """batched puts between processes: each PE issues its share of one batch
    into the other PEs' IPC-mapped heaps; every PE's targets bit-exact"""
    outs = run_job(npes, "batch", 64 * npes, discipline)
    assert sum(o["requests_issued"] for o in outs) == 64 * npes
    assert all(o["batch_ok"] for o in outs), outs
    assert all(o["errors"] == 0 for o in outs)


@pytest.mark.parametrize("scenario,npes,args", [("putget", 2, ()), ("pingpong", 2, (2000,)), ("barrier", 4, (100,))])
def test_multiprocess_fine_grained_heap(scenario, npes, args):
    """the same exchanges with every PE's heap in fine-grained device memory
    (NVSHMEM_HEAP_KIND=fine), exported and opened through IPC"""
    outs = run_job(npes, scenario, *args, extra={"NVSHMEM_HEAP_KIND": "fine"})
    for o in outs:
        assert o["errors"] == 0 and o["heap_kind"] == 1, o
        if scenario == "putget":
            assert o["put_ok"] and o["get_ok"], o
        if scenario == "pingpong":
            assert o["final_flag"] == args[0], o
        if scenario == "barrier":
            assert o["violations"] == 0 and o["slots"] == [args[0]] * npes, o


def test_multiprocess_barrier_timeout_ends_the_process():
    """an offload barrier that only PE 0 enters gives up after its budget and
    PE 0's next nvshmem_barrier_all exits non-zero; PE 1, left in the host
    barrier, ends at the bootstrap's own time-out"""
    name = f"nvshmem_test_{os.getpid()}_{uuid.uuid4().hex[:8]}"
    env_extra = {"NVSHMEM_HEAP_SIZE": "16M", "NVSHMEM_WAIT_TIMEOUT_MS": "5000", "NVSHMEM_BOOTSTRAP_TIMEOUT_S": "5"}
    procs = [subprocess.Popen([sys.executable, str(HERE / "pe_worker.py"), "barrier_timeout"],
                              env=na.shm_job_env(r, 2, name, env_extra), stdout=subprocess.PIPE,
                              stderr=subprocess.PIPE, text=True) for r in range(2)]
    try:
        res = [p.communicate(timeout=90) for p in procs]
    finally:
        for p in procs:
            if p.poll() is None:
                p.kill()
        shm = Path("/dev/shm") / name
        if shm.exists():
            shm.unlink()
    (o0, e0), (o1, e1) = res
    assert "offload barrier returned" in o0, e0[-2000:]
    assert procs[0].returncode != 0 and "continued" not in o0
    assert "nvshmem_barrier_all: a device barrier outlived its 50 ms budget" in e0, e0[-2000:]
    assert procs[1].returncode != 0 and "continued" not in o1
