"""CPU checks of the drop-in boundary: the C-ABI library loads without a GPU,
exports every host entry point declared in include/*.h, the static library
carries them too, and a caller written against the reference's headers
compiles and links with relocatable device code."""
import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

import nvshmem_amd as na

ROOT = Path(__file__).resolve().parent.parent
BUILD = ROOT / "df-nvshmem-prototype_amd" / "build"

# every entry point of the reference's host header (nvshmem.h:44-52) and
# device header (nvshmem_device.h:38-71)
REFERENCE_HOST = ["nvstart_pes", "nvstop_pes", "nvshmem_my_pe", "nvshmem_n_pes", "nvshmalloc", "nvshmcleanup",
                  "nvshmem_barrier_all", "nvshmem_barrier_all_offload", "nvshmem_ptr"]
REFERENCE_DEVICE = ["nvshmem_trace_begin", "nvshmem_trace_end", "nvshmem_quiet"] + [
    f"nvshmem_{t}_{op}" for op in ("p", "g", "put", "get") for t in ("int", "float", "double")] + [
    "nvshmem_int_wait", "nvshmem_int_wait_until"]


def _defined_dynamic(path):
    out = subprocess.run(["nm", "-D", "--defined-only", str(path)], capture_output=True, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_headers_declare_the_reference_api():
    declared = na.declared_symbols()
    missing = [s for s in REFERENCE_HOST + REFERENCE_DEVICE if s not in declared]
    assert not missing


@pytest.mark.parametrize("trace", [False, True])
def test_shared_library_loads_and_exports_every_host_symbol(trace):
    lib = na.load(trace=trace)                    # loads on a machine without a GPU
    exported = _defined_dynamic(na.lib_path(trace))
    missing = sorted(s for s in na.declared_host_symbols() if s not in exported)
    assert not missing, missing
    for s in na.SIGNATURES:
        assert hasattr(lib, s), s


def test_static_library_has_every_host_symbol():
    out = subprocess.run(["nm", "--defined-only", str(BUILD / "libnvshmem.a")], capture_output=True, text=True,
                         check=True).stdout
    defined = {l.split()[-1] for l in out.splitlines() if re.search(r" T ", l)}
    missing = sorted(s for s in na.declared_host_symbols() if s not in defined)
    assert not missing, missing


def test_reference_style_caller_links(tmp_path):
    """a kernel calling every reference device entry point, compiled -fgpu-rdc
    and linked against libnvshmem.a: the device symbols resolve"""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    src = tmp_path / "caller.hip"
    src.write_text(r'''
#include <hip/hip_runtime.h>
#include "nvshmem.h"
#include "nvshmem_device.h"
__global__ void k(int *i, float *f, double *d, int pe) {
    nvshmem_trace_begin();
    nvshmem_int_p(i, 1, pe); nvshmem_float_p(f, 1.f, pe); nvshmem_double_p(d, 1.0, pe);
    i[1] = nvshmem_int_g(i, pe); f[1] = nvshmem_float_g(f, pe); d[1] = nvshmem_double_g(d, pe);
    nvshmem_int_put(i, i, 4, pe); nvshmem_float_put(f, f, 4, pe); nvshmem_double_put(d, d, 4, pe);
    nvshmem_int_get(i, i, 4, pe); nvshmem_float_get(f, f, 4, pe); nvshmem_double_get(d, d, 4, pe);
    nvshmem_quiet(); nvshmem_int_wait(i, 5); nvshmem_int_wait_until(i, 0, 1);
    nvshmem_trace_end();
}
int main() {
    nvstart_pes();
    int *i = (int *)nvshmalloc(64); float *f = (float *)nvshmalloc(64); double *d = (double *)nvshmalloc(64);
    (void)nvshmem_ptr(i, 0);
    hipLaunchKernelGGL(k, dim3(1), dim3(64), 0, 0, i, f, d, nvshmem_my_pe());
    nvshmem_barrier_all(); nvshmem_barrier_all_offload(); nvshmcleanup(); nvstop_pes();
    return nvshmem_n_pes() >= 0 ? 0 : 1;
}
''')
    exe = tmp_path / "caller"
    cmd = [hipcc, "--offload-arch=gfx950", "-O2", "-fgpu-rdc", f"-I{ROOT / 'include'}", str(src), "--hip-link",
           f"-L{BUILD}", "-lnvshmem", "-lpthread", "-lrt", "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]
    assert exe.exists()


def test_every_declared_device_function_links(tmp_path):
    """a kernel that takes the address of every __device__ function declared in
    include/*.h, compiled -fgpu-rdc and linked against libnvshmem.a: each one
    is defined in the library's device code (the link fails otherwise)"""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    names = sorted(na.declared_symbols() - na.declared_host_symbols())
    assert set(REFERENCE_DEVICE) <= set(names)
    refs = "\n".join(f"    sink[{i}] = (unsigned long long)(void *)&{n};" for i, n in enumerate(names))
    src = tmp_path / "every_device.hip"
    src.write_text(f"""
#include <hip/hip_runtime.h>
#include "nvshmem.h"
#include "nvshmem_device.h"
#include "nvshmemx_device.h"
__global__ void k(unsigned long long *sink) {{
{refs}
}}
int main() {{
    hipLaunchKernelGGL(k, dim3(1), dim3(1), 0, 0, (unsigned long long *)0);
    return 0;
}}
""")
    exe = tmp_path / "every_device"
    cmd = [hipcc, "--offload-arch=gfx950", "-O2", "-fgpu-rdc", f"-I{ROOT / 'include'}", str(src), "--hip-link",
           f"-L{BUILD}", "-lnvshmem", "-lpthread", "-lrt", "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]
    assert exe.exists()


_BEFORE_INIT = {
    "nvshmalloc": "lib.nvshmalloc(64)",
    "nvshmem_ptr": "lib.nvshmem_ptr(None, 0)",
    "nvshmem_barrier_all": "lib.nvshmem_barrier_all()",
    "nvshmem_barrier_all_offload": "lib.nvshmem_barrier_all_offload()",
    "nvshmemx_putmem_on_stream": "lib.nvshmemx_putmem_on_stream(None, None, 4, 0, None)",
    "nvshmemx_batch_plan_create": "lib.nvshmemx_batch_plan_create(None, 0, None, None, None)",
    "nvshmemx_nbi_flush_on_stream": "lib.nvshmemx_nbi_flush_on_stream(None)",
    "nvstop_pes": "lib.nvstop_pes()",
}


@pytest.mark.parametrize("fn", sorted(_BEFORE_INIT))
def test_calls_before_init_are_refused(fn):
    """library calls before nvstart_pes end the process with the reference's
    error convention ([file:line] [pe] message, non-zero exit) instead of
    touching uninitialised state; no GPU is needed to reach the check"""
    code = ("import sys; sys.path.insert(0, %r); import nvshmem_amd as na; lib = na.load(); %s; print('returned')"
            % (str(ROOT / "df-nvshmem-prototype_amd"), _BEFORE_INIT[fn]))
    env = {k: v for k, v in os.environ.items() if not k.startswith("NVSHMEM_")}
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, env=env)
    assert p.returncode != 0 and "returned" not in p.stdout
    assert f"{fn} called before nvstart_pes" in p.stderr, p.stderr[-1000:]
