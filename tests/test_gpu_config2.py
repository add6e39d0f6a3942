"""Two PEs as two processes on one MI355X, as the workload is configured:
{int32, float32, float64} x {put, get} x every power of two from 8 B to 1 GiB,
sources PCG32 words seeded by (PE, type, size) -- NaN bit patterns included,
so every copy has to be bitwise.  Both the host-launched typed calls
(nvshmemx_{T}_{put,get}_on_stream, the grid copy the benchmark times) and the
reference's device calls (nvshmem_{T}_{put,get}, one call per thread of 4096
elements) run at every size.  Each PE compares what landed in its own memory
with the oracle's put/get (reference semantics nvshmem_device.cu:487-540),
256 guard bytes on either side included.  The 1 GiB symmetric region keeps the
IPC-shared heap under 2 GiB; put sources and get targets are private memory."""
import pytest

from test_gpu_multiprocess import run_job

pytestmark = pytest.mark.gpu


def test_config2_every_type_direction_and_size():
    outs = run_job(2, "config2", 30, "host,thread", heap="1025M", timeout=900)
    for o in outs:
        assert o["errors"] == 0, o
        assert o["checked"] == 3 * 28 * 2 * 2, o
        assert not o["failures"], o["failures"][:20]
