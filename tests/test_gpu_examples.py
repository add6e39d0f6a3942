"""The reference's caller pattern as drop-in MPI programs (examples/*.hip,
linked against libnvshmem.a with relocatable device code), run as 2 MPI
processes on the one GPU; and the same build of a program whose own kernel
posts non-blocking puts that the host then flushes."""
import re
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
EX = ROOT / "df-nvshmem-prototype_amd" / "build" / "examples"
MPIRUN = Path("/opt/conda/bin/mpirun")


def mpirun(exe, n=2, timeout=120):
    if not MPIRUN.exists():
        pytest.skip("MPICH mpirun not present")
    assert (EX / exe).exists(), f"{exe} not built"
    p = subprocess.run([str(MPIRUN), "-n", str(n), str(EX / exe)], capture_output=True, text=True, timeout=timeout,
                       env={"PATH": "/usr/bin:/bin", "HSA_ENABLE_IPC_MODE_LEGACY": "0",
                            "NVSHMEM_HEAP_SIZE": "64M", "NVSHMEM_WAIT_TIMEOUT_MS": "5000"})
    return p


def test_dropin_put_flag_latency():
    p = mpirun("shmem_latency")
    assert p.returncode == 0, p.stdout + p.stderr[-2000:]
    rows = [l.split() for l in p.stdout.splitlines() if re.match(r"^\s*\d+\s", l)]
    assert [int(r[0]) for r in rows] == [4 << k for k in range(15)]       # 4 B .. 64 KiB
    assert all(r[3] == "ok" for r in rows), p.stdout
    print(p.stdout)


def test_dropin_sync_latency():
    p = mpirun("shmem_sync_latency")
    assert p.returncode == 0, p.stdout + p.stderr[-2000:]
    m = re.search(r"sync latency: ([\d.]+) ns (\d+) cycles flags (\d+) (\d+)", p.stdout)
    assert m and int(m.group(3)) == int(m.group(4)) == 10100, p.stdout
    print(p.stdout)


def test_dropin_wrong_pe_count_is_reported():
    p = mpirun("shmem_sync_latency", n=1)
    assert p.returncode != 0 and "exactly two processes" in p.stderr


def test_dropin_posted_puts():
    """a user kernel, linked against libnvshmem.a, posts 903 puts into the
    other process's heap with nvshmemx_putmem_nbi; after the flush and the
    barrier each process finds every byte, and zeros past the last put"""
    p = mpirun("shmem_nbi_exchange")
    assert p.returncode == 0, p.stdout + p.stderr[-2000:]
    ok = [l for l in p.stdout.splitlines() if l.endswith(", ok")]
    assert len(ok) == 2 and all("903 puts, 1048512 bytes received" in l for l in ok), p.stdout
    print(p.stdout)
