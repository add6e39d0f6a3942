"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU.

The library under test is loaded from df-nvshmem-prototype_amd/build (C ABI via
ctypes); the oracle (test infrastructure) from oracle/build.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "df-nvshmem-prototype_amd"
sys.path.insert(0, str(PKG))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


# GPU files in the order they run: the device API and bulk put/get first (the
# hot path), then the processes sharing one GPU, tracing, the applications, and
# last whatever reports speed as well as results.  Files not listed keep their
# collection order after these; test_zz_* files always run last.
_GPU_ORDER = [
    "test_gpu_rma.py",
    "test_gpu_config2.py",
    "test_gpu_multiprocess.py",
    "test_gpu_trace.py",
    "test_gpu_examples.py",
    "test_gpu_idealize.py",
    "test_gpu_batch.py",
    "test_gpu_config4.py",
    "test_gpu_persistent.py",
    "test_gpu_nbi.py",
    "test_gpu_jacobi.py",
    "test_gpu_costmodel.py",
    "test_gpu_learned.py",
]


def pytest_collection_modifyitems(session, config, items):
    def rank(item):
        name = Path(str(item.fspath)).name
        if name.startswith("test_zz_"):
            return len(_GPU_ORDER) + 1
        return _GPU_ORDER.index(name) if name in _GPU_ORDER else len(_GPU_ORDER)

    items[:] = sorted(items, key=rank)  # stable: keeps the order inside each file


@pytest.fixture(scope="session")
def oracle():
    import oracle_py
    return oracle_py.Oracle()


@pytest.fixture(scope="session")
def lib():
    import nvshmem_amd
    return nvshmem_amd.load()


@pytest.fixture(scope="session")
def hip():
    import nvshmem_amd
    h = nvshmem_amd.Hip()
    if h.device_count() < 1:
        pytest.skip("no GPU visible")
    return h


class LocalPes:
    """(Re)starts the library in local mode with the requested layout."""

    def __init__(self, lib):
        self.lib = lib
        self.cfg = None

    def start(self, npes, wgs_per_pe=1, heap="256M", timeout_ms=5000):
        cfg = (npes, wgs_per_pe, heap, timeout_ms)
        if self.cfg == cfg and self.lib.nvshmemx_initialized():
            self.lib.nvshmcleanup()
            self.lib.nvshmemx_clear_errors()
            return self.lib
        if self.lib.nvshmemx_initialized():
            self.lib.nvstop_pes()
        os.environ["NVSHMEM_HEAP_SIZE"] = heap
        os.environ["NVSHMEM_WAIT_TIMEOUT_MS"] = str(timeout_ms)
        os.environ.pop("NVSHMEM_BOOTSTRAP", None)
        assert self.lib.nvshmemx_start_local_pes(npes, wgs_per_pe) == 0
        self.cfg = cfg
        return self.lib


@pytest.fixture(scope="session")
def local_pes(lib, hip):
    lp = LocalPes(lib)
    yield lp
    if lib.nvshmemx_initialized():
        lib.nvstop_pes()
