"""ctypes loader for the MI355X OpenSHMEM-subset library.

The product is the C-ABI shared library ``build/libnvshmem_amd.so`` (and the
static ``build/libnvshmem.a`` for compiled callers).  This module only loads it,
declares argument types, and offers HIP runtime plumbing (device memory,
copies, streams, events) for test and benchmark drivers.  Nothing here computes
a result: if the library is missing the import of :func:`load` fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent.parent          # df-nvshmem-prototype_amd/
REPO_DIR = PKG_DIR.parent
BUILD_DIR = PKG_DIR / "build"
INCLUDE_DIR = REPO_DIR / "include"

DTYPES = {"int": 0, "float": 1, "double": 2}
DTYPE_BYTES = {"int": 4, "float": 4, "double": 8}

ERR_WAIT_TIMEOUT = 0x1
ERR_BARRIER_TIMEOUT = 0x2
ERR_BAD_PE = 0x4
ERR_TRACE_OVERFLOW = 0x8
ERR_TRACE_CONFLICT = 0x10
ERR_BAD_ADDRESS = 0x20


def lib_path(trace: bool = False, variant: str = "") -> Path:
    """build/libnvshmem_amd.so, or the tracing / idealized-communication build"""
    sub = "trace" if trace else variant
    return (BUILD_DIR / sub if sub else BUILD_DIR) / "libnvshmem_amd.so"


class LibraryMissing(RuntimeError):
    pass


_P = C.c_void_p
_SZ = C.c_size_t
_I = C.c_int
_U64 = C.c_ulonglong

# name -> (restype, argtypes); every symbol the loaders use
SIGNATURES = {
    # host API (include/nvshmem.h)
    "nvstart_pes": (None, []),
    "nvstop_pes": (None, []),
    "nvshmem_my_pe": (_I, []),
    "nvshmem_n_pes": (_I, []),
    "nvshmalloc": (_P, [_SZ]),
    "nvshmcleanup": (None, []),
    "nvshmem_barrier_all": (None, []),
    "nvshmem_barrier_all_offload": (None, []),
    "nvshmem_ptr": (_P, [_P, _I]),
    # extensions (include/nvshmemx.h)
    "nvshmemx_start_local_pes": (_I, [_I, _I]),
    "nvshmemx_initialized": (_I, []),
    "nvshmemx_mode": (_I, []),
    "nvshmemx_wgs_per_pe": (_I, []),
    "nvshmemx_heap_size": (_SZ, []),
    "nvshmemx_heap_kind": (_I, []),
    "nvshmemx_bootstrap_name": (C.c_char_p, []),
    "nvshmemx_wallclock_khz": (_I, []),
    "nvshmemx_get_last_error": (_I, []),
    "nvshmemx_clear_errors": (None, []),
    "nvshmemx_set_wait_timeout_ms": (None, [_U64]),
    "nvshmemx_set_stream": (None, [_P]),
    "nvshmemx_get_stream": (_P, []),
    "nvshmemx_putmem_on_stream": (None, [_P, _P, _SZ, _I, _P]),
    "nvshmemx_getmem_on_stream": (None, [_P, _P, _SZ, _I, _P]),
    "nvshmemx_int_put_on_stream": (None, [_P, _P, _SZ, _I, _P]),
    "nvshmemx_float_put_on_stream": (None, [_P, _P, _SZ, _I, _P]),
    "nvshmemx_double_put_on_stream": (None, [_P, _P, _SZ, _I, _P]),
    "nvshmemx_int_get_on_stream": (None, [_P, _P, _SZ, _I, _P]),
    "nvshmemx_float_get_on_stream": (None, [_P, _P, _SZ, _I, _P]),
    "nvshmemx_double_get_on_stream": (None, [_P, _P, _SZ, _I, _P]),
    "nvshmemx_putmem_config_on_stream": (None, [_P, _P, _SZ, _I, _P, _P]),
    "nvshmemx_default_copy_config": (None, [_SZ, _P]),
    "nvshmemx_putmem_2d_on_stream": (None, [_P, _SZ, _P, _SZ, _SZ, _SZ, _I, _P]),
    "nvshmemx_getmem_2d_on_stream": (None, [_P, _SZ, _P, _SZ, _SZ, _SZ, _I, _P]),
    "nvshmemx_int_iput_on_stream": (None, [_P, _P, C.c_ssize_t, C.c_ssize_t, _SZ, _I, _P]),
    "nvshmemx_float_iput_on_stream": (None, [_P, _P, C.c_ssize_t, C.c_ssize_t, _SZ, _I, _P]),
    "nvshmemx_double_iput_on_stream": (None, [_P, _P, C.c_ssize_t, C.c_ssize_t, _SZ, _I, _P]),
    "nvshmemx_int_iget_on_stream": (None, [_P, _P, C.c_ssize_t, C.c_ssize_t, _SZ, _I, _P]),
    "nvshmemx_float_iget_on_stream": (None, [_P, _P, C.c_ssize_t, C.c_ssize_t, _SZ, _I, _P]),
    "nvshmemx_double_iget_on_stream": (None, [_P, _P, C.c_ssize_t, C.c_ssize_t, _SZ, _I, _P]),
    "nvshmemh_strided_device": (None, [_I, _I, _P, _P, C.c_longlong, C.c_longlong, _SZ, _SZ, _SZ, _SZ, _SZ, _I, _I,
                                       _P]),
    "nvshmemx_barrier_all_on_stream": (None, [_P]),
    "nvshmemx_trace_flush": (None, []),
    "nvshmemx_trace_enabled": (_I, []),
    # caller harness (csrc/harness.hip)
    "nvshmemh_fill_pcg32": (None, [_P, _SZ, _U64, _U64, _P]),
    "nvshmemh_stream_read": (None, [_P, _SZ, _P, _P]),
    "nvshmemh_stream_write": (None, [_P, _SZ, _I, _P]),
    "nvshmemh_copy_stamped": (_I, [_P, _P, _SZ, _SZ, _I, _P, _SZ, _P]),
    "nvshmemh_post_nbi": (None, [_P, _P, _I, _P]),
    "nvshmemh_post_nbi_typed": (None, [_I, _I, _P, _P, _I, _I, _P]),
    "nvshmemh_post_nbi_mixed": (None, [_P, _P, _P, _I, _P]),
    "nvshmemh_pcg32_host": (None, [_P, _SZ, _U64, _U64]),
    "nvshmemh_thread_putget": (None, [_I, _I, _P, _P, _I, _I, _I, _P]),
    "nvshmemh_coop_putget": (None, [_I, _I, _P, _P, _SZ, _SZ, _I, _I, _P]),
    "nvshmemh_pg": (None, [_I, _I, _P, _P, _P, _I, _I, _P]),
    "nvshmemh_pingpong_process": (None, [_P, _I, _I, _I, _P, _I, _P, _P]),
    "nvshmemh_barrier_latency": (None, [_I, _P, _P]),
    "nvshmemh_pingpong_local": (None, [_P, _I, _I, _I, _P, _I, _P, _P]),
    "nvshmemh_order_local": (None, [_P, _P, _P, _P, _I, _I, _I, _I, _P, _P]),
    "nvshmemh_wg_barrier_check": (None, [_P, _I, _P, _I, _P]),
    "nvshmemh_publish_round": (None, [_P, _I, _P]),
    "nvshmemh_wait_once": (None, [_P, _I, _I, _P, _P]),
    "nvshmemh_traced_ops": (None, [_P, _P, _P, _P, _I, _I, _I, _I, _I, _P]),
    "nvshmemh_traced_lanes": (None, [_P, _I, _I, _I, _I, _I, _P]),
    "nvshmemh_check_round": (None, [_P, _I, _P, _P]),
}


class PutRequest(C.Structure):
    """nvshmemx_put_request_t"""
    _fields_ = [("target", _P), ("source", _P), ("nbytes", _SZ), ("pe", _I), ("elem_bytes", _I)]


class BatchChoice(C.Structure):
    """nvshmemx_batch_choice_t"""
    _fields_ = [("pieces", _I), ("vec_bytes", _I), ("mode", _I)]


class BatchLaunch(C.Structure):
    """nvshmemx_batch_launch_t"""
    _fields_ = [("workgroups", _I), ("block", _I), ("nontemporal", _I), ("discipline", _I)]


BATCH_SHARED, BATCH_OWNER, BATCH_SOURCE_ORDER = 0, 1, 2

SIGNATURES.update({
    "nvshmemx_batch_plan_create": (_I, [_P, _I, _P, _P, C.POINTER(_P)]),
    "nvshmemx_batch_plan_launch": (None, [_P, _P]),
    "nvshmemx_batch_plan_capture": (_I, [_P, _I]),
    "nvshmemx_batch_plan_graph_launch": (None, [_P, _P]),
    "nvshmemx_batch_plan_select": (_I, [_P, _I, _I, _P, _P]),
    "nvshmemx_batch_plan_pieces": (_SZ, [_P]),
    "nvshmemx_batch_plan_bytes": (_SZ, [_P]),
    "nvshmemx_batch_plan_destroy": (None, [_P]),
    "nvshmemx_batch_default_choice": (None, [_SZ, _I, _P]),
    "nvshmemx_put_requests_on_stream": (None, [_P, _I, _P]),
    "nvshmemh_gen_requests": (_I, [_U64, _I, _I, _U64, _U64, _I, _I, _P, _P, _P, _P, _P, _P]),
    "nvshmemh_verify_requests": (_U64, [_P, _I]),
    "nvshmemx_batch_put_persistent_on_stream": (_I, [_P, _I, _P]),
    "nvshmemx_batch_persistent_grid": (_I, []),
    "nvshmemx_batch_put_persistent_noalloc_on_stream": (_I, [_P, _I, _P]),
    "nvshmemx_batch_persistent_max_requests_noalloc": (_I, []),
    "nvshmemx_nbi_flush_on_stream": (_I, [_P]),
    "nvshmemx_nbi_capacity": (_I, []),
    "nvshmemx_tune_candidate_count": (_I, []),
    "nvshmemx_tune_candidate": (None, [_I, _P]),
    "nvshmemx_align_class": (_I, [_P, _P]),
    "nvshmemx_tune_features": (None, [C.c_double, _I, _I, _I, _P]),
    "nvshmemx_tune_measure": (_I, [_I, _P, _P, _P, _P, _I, _I, _I, _P]),
    "nvshmemx_tune_table_choose": (_I, [_P, _P, _I, _P, _I, _I, _P]),
    "nvshmemx_effective_concurrency": (_I, [_SZ, _SZ, _I]),
    "nvshmemx_costmodel_create": (_I, [_I, _U64, C.POINTER(_P)]),
    "nvshmemx_costmodel_num_params": (_I, [_P]),
    "nvshmemx_costmodel_get_params": (None, [_P, _P]),
    "nvshmemx_costmodel_set_params": (None, [_P, _P]),
    "nvshmemx_costmodel_gradient": (None, [_P, _P, _P, _I, _P]),
    "nvshmemx_costmodel_train": (None, [_P, _P, _P, _I, _I, C.c_float]),
    "nvshmemx_costmodel_predict": (None, [_P, _P, _I, _P]),
    "nvshmemx_costmodel_choose": (_I, [_P, _P, _I, _I, _P]),
    "nvshmemx_costmodel_save": (_I, [_P, C.c_char_p]),
    "nvshmemx_costmodel_load": (_I, [_P, C.c_char_p]),
    "nvshmemx_costmodel_destroy": (None, [_P]),
    "nvshmemx_costmodel_num_outputs": (_I, [_P]),
    "nvshmemx_batch_policy_count": (_I, []),
    "nvshmemx_batch_features": (None, [_P, _I, _P]),
    "nvshmemx_batch_policy_choices": (_I, [_P, _I, _I, _P, _P]),
    "nvshmemx_tune_measure_batch": (_I, [_P, _I, _P, _I, _P]),
    "nvshmemx_costmodel_choose_batch": (_I, [_P, _P, _P, _I, _P, _P]),
})

SIGNATURES.update({
    "nvshmemh_jacobi_field_bytes": (_SZ, [_I]),
    "nvshmemh_jacobi_layout": (None, [_I, _P, _P]),
    "nvshmemh_jacobi_recv_bytes": (_SZ, [_I]),
    "nvshmemh_jacobi_init": (None, [_P, _I, _I, _I, _U64, _P]),
    "nvshmemh_jacobi_pack_put": (None, [_P, _P, _I, _I, _I, _I, _P]),
    "nvshmemh_jacobi_unpack": (None, [_P, _P, _I, _I, _I, _I, _P]),
    "nvshmemh_jacobi_stencil": (None, [_P, _P, _I, _I, _I, _I, _P]),
    "nvshmemh_jacobi_stencil_fused": (None, [_P, _P, _P, _I, _I, _I, _I, _I, _I, _P]),
    "nvshmemh_jacobi_run": (_P, [_P, _P, _P, _I, _I, _I, _I, _P]),
    "nvshmemh_jacobi_pack_put_api": (None, [_P, _P, _I, _I, _I, _I, _P]),
    "nvshmemh_jacobi_run_separate": (_P, [_P, _P, _P, _I, _I, _I, _I, _P]),
    "nvshmemh_jacobi_stencil2": (None, [_P, _P, _I, _I, _I, _I, _P]),
    "nvshmemh_jacobi_run2": (_P, [_P, _P, _P, _I, _I, _I, _P]),
})


def jacobi_interior(hip, field_ptr, n, lib=None):
    """(n, n, n) interior of one PE's field (layout from nvshmemh_jacobi_layout)"""
    rs, ko = C.c_int(), C.c_int()
    (lib or _loaded_lib()).nvshmemh_jacobi_layout(n, C.byref(rs), C.byref(ko))
    a = hip.dtoh(field_ptr, (n + 2) * (n + 2) * rs.value * 8, "<f8").reshape(n + 2, n + 2, rs.value)
    return a[1:-1, 1:-1, ko.value + 1:ko.value + 1 + n].copy()


_LIBS = {}


def _loaded_lib():
    if "default" not in _LIBS:
        _LIBS["default"] = load()
    return _LIBS["default"]


def jacobi_block(global_field, pe, n):
    """PE pe's block of the global 2n^3 field (2x2x2 PE grid, pe = (x<<2)|(y<<1)|z)"""
    cx, cy, cz = pe >> 2, (pe >> 1) & 1, pe & 1
    return global_field[cx * n:(cx + 1) * n, cy * n:(cy + 1) * n, cz * n:(cz + 1) * n]


NCAND = 54


def candidates(lib):
    """list of (pieces, vec_bytes, mode) for every autotuner candidate"""
    out = []
    for i in range(lib.nvshmemx_tune_candidate_count()):
        c = BatchChoice()
        lib.nvshmemx_tune_candidate(i, C.byref(c))
        out.append((c.pieces, c.vec_bytes, c.mode))
    return out


def tune_grid(half_octaves=True, widths=(4, 8), concurrency=(1, 4, 16, 64, 256, 1024, 4096, 16384, 32768), aligns=(0, 1, 2),
              min_log2=6, max_log2=22):
    """feature points of the autotuner's exhaustive table"""
    import numpy as np
    step = 0.5 if half_octaves else 1.0
    sizes = np.arange(min_log2, max_log2 + 1e-9, step)
    pts = [(float(s), w, k, a) for s in sizes for w in widths for k in concurrency for a in aligns]
    return {"log2_bytes": np.array([p[0] for p in pts], np.float64), "elem_bytes": np.array([p[1] for p in pts],
                                                                                           np.int32),
            "concurrency": np.array([p[2] for p in pts], np.int32), "align": np.array([p[3] for p in pts], np.int32)}


def tune_measure(lib, grid, reps=3, workgroups=0, block=0):
    """times_us[npts, 54] from on-device measurement of every candidate"""
    import numpy as np
    n = len(grid["log2_bytes"])
    out = np.zeros((n, NCAND), np.float32)
    lib.nvshmemx_tune_measure(n, grid["log2_bytes"].ctypes.data, grid["elem_bytes"].ctypes.data,
                              grid["concurrency"].ctypes.data, grid["align"].ctypes.data, reps, workgroups, block,
                              out.ctypes.data)
    return out


def tune_features(lib, grid):
    import numpy as np
    n = len(grid["log2_bytes"])
    f = np.zeros((n, 4), np.float32)
    for i in range(n):
        lib.nvshmemx_tune_features(float(grid["log2_bytes"][i]), int(grid["elem_bytes"][i]),
                                   int(grid["concurrency"][i]), int(grid["align"][i]), f[i].ctypes.data)
    return f


class CostModel:
    """The library's MLP cost model (all compute in the library's HIP kernels)."""

    def __init__(self, lib, seed=1, ncand=NCAND):
        self.lib = lib
        self.h = C.c_void_p()
        lib.nvshmemx_costmodel_create(ncand, seed, C.byref(self.h))
        self.ncand = ncand
        self.nparams = lib.nvshmemx_costmodel_num_params(self.h)

    def params(self):
        import numpy as np
        p = np.zeros(self.nparams, np.float32)
        self.lib.nvshmemx_costmodel_get_params(self.h, p.ctypes.data)
        return p

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, np.float32)
        self.lib.nvshmemx_costmodel_set_params(self.h, p.ctypes.data)

    def gradient(self, X, T):
        import numpy as np
        X, T = np.ascontiguousarray(X, np.float32), np.ascontiguousarray(T, np.float32)
        g = np.zeros(self.nparams, np.float32)
        self.lib.nvshmemx_costmodel_gradient(self.h, X.ctypes.data, T.ctypes.data, len(X), g.ctypes.data)
        return g

    def train(self, X, T, steps, lr=3e-3):
        import numpy as np
        X, T = np.ascontiguousarray(X, np.float32), np.ascontiguousarray(T, np.float32)
        self.lib.nvshmemx_costmodel_train(self.h, X.ctypes.data, T.ctypes.data, len(X), steps, lr)

    def predict(self, X):
        import numpy as np
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((len(X), self.ncand), np.float32)
        self.lib.nvshmemx_costmodel_predict(self.h, X.ctypes.data, len(X), Y.ctypes.data)
        return Y

    def choose_batch(self, request_model, reqs_arr, n):
        """per-request choices of the policy this batch model predicts fastest, and the policy"""
        out = (BatchChoice * n)()
        pol = C.c_int()
        self.lib.nvshmemx_costmodel_choose_batch(self.h, request_model.h if request_model else None, reqs_arr, n,
                                                 out, C.byref(pol))
        return out, pol.value

    def choose(self, reqs_arr, n, concurrency):
        out = (BatchChoice * n)()
        self.lib.nvshmemx_costmodel_choose(self.h, reqs_arr, n, concurrency, out)
        return out

    def close(self):
        if self.h:
            self.lib.nvshmemx_costmodel_destroy(self.h)
            self.h = C.c_void_p()


def batch_features(lib, reqs_arr, n):
    import numpy as np
    f = np.zeros(4, np.float32)
    lib.nvshmemx_batch_features(reqs_arr, n, f.ctypes.data)
    return f


def measure_batch(lib, reqs_arr, n, request_model=None, reps=5):
    """every batch policy's plan timed on the device: (ms per policy, fastest policy)"""
    import numpy as np
    ms = np.zeros(lib.nvshmemx_batch_policy_count(), np.float32)
    best = lib.nvshmemx_tune_measure_batch(reqs_arr, n, request_model.h if request_model else None, reps,
                                           ms.ctypes.data)
    return ms, best


def gen_requests(lib, seed, nreq, npes, src_region, dst_region, min_log2=6, max_log2=22):
    """The uneven-put workload (harness generator); returns a dict of numpy arrays."""
    import numpy as np
    out = {k: np.zeros(nreq, dt) for k, dt in [("src_pe", np.int32), ("dst_pe", np.int32), ("src_off", np.uint64),
                                               ("dst_off", np.uint64), ("nbytes", np.uint64), ("dtype", np.int32)]}
    rc = lib.nvshmemh_gen_requests(seed, nreq, npes, src_region, dst_region, min_log2, max_log2,
                                   *[out[k].ctypes.data for k in ("src_pe", "dst_pe", "src_off", "dst_off", "nbytes",
                                                                  "dtype")])
    if rc != 0:
        raise ValueError(f"request generation failed ({rc}): destination region too small")
    return out


def make_requests(lib, reqs, sym_base, src_base, dst_base):
    """PutRequest array: sources at PE src_pe's heap + src_base + src_off (local
    addresses), targets at symmetric sym_base + dst_base + dst_off on dst_pe."""
    import numpy as np
    n = len(reqs["nbytes"])
    npes = lib.nvshmem_n_pes()
    src_bases = np.array([lib.nvshmem_ptr(sym_base + src_base, s) for s in range(npes)], np.uint64)
    rec = np.zeros(n, np.dtype([("target", "<u8"), ("source", "<u8"), ("nbytes", "<u8"), ("pe", "<i4"),
                                ("elem_bytes", "<i4")]))
    rec["target"] = np.uint64(sym_base + dst_base) + reqs["dst_off"].astype(np.uint64)
    rec["source"] = src_bases[reqs["src_pe"]] + reqs["src_off"].astype(np.uint64)
    rec["nbytes"] = reqs["nbytes"]
    rec["pe"] = reqs["dst_pe"]
    rec["elem_bytes"] = 4
    arr = (PutRequest * n).from_buffer(rec)
    arr._keepalive = rec
    return arr


class CopyConfig(C.Structure):
    """nvshmemx_copy_config_t (include/nvshmemx.h)."""
    _fields_ = [("workgroups", _I), ("block", _I), ("vec_bytes", _I), ("unroll", _I), ("nontemporal", _I),
                ("mode", _I), ("streams", _I)]


def load(trace: bool = False, variant: str = "") -> C.CDLL:
    path = lib_path(trace, variant)
    if not path.exists():
        raise LibraryMissing(f"{path} is not built; run make -C {PKG_DIR}")
    lib = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    return lib


def declared_symbols() -> set[str]:
    """Every function name declared with C linkage in include/*.h."""
    names: set[str] = set()
    pat = re.compile(r"^\s*(?:__device__\s+)?[A-Za-z_][\w\s\*]*?\b((?:nv|nvshmem)\w*)\s*\(", re.M)
    for h in sorted(INCLUDE_DIR.glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in pat.finditer(text):
            names.add(m.group(1))
    return names


def declared_host_symbols() -> set[str]:
    """Host (non-__device__) functions declared in include/*.h."""
    names: set[str] = set()
    pat = re.compile(r"^\s*(?!__device__)(?:const\s+)?[A-Za-z_]\w*[\s\*]+((?:nv|nvshmem)\w*)\s*\(", re.M)
    for h in sorted(INCLUDE_DIR.glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in pat.finditer(text):
            names.add(m.group(1))
    return names


# ---------------------------------------------------------------- HIP plumbing

class HipError(RuntimeError):
    pass


class Hip:
    """Minimal HIP runtime access for drivers: memory, copies, streams, events."""

    H2D, D2H, D2D = 1, 2, 3

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        L = self.lib
        for name, res, args in [
            ("hipMalloc", _I, [C.POINTER(_P), _SZ]),
            ("hipFree", _I, [_P]),
            ("hipMemcpy", _I, [_P, _P, _SZ, _I]),
            ("hipMemcpyAsync", _I, [_P, _P, _SZ, _I, _P]),
            ("hipMemset", _I, [_P, _I, _SZ]),
            ("hipMemsetAsync", _I, [_P, _I, _SZ, _P]),
            ("hipDeviceSynchronize", _I, []),
            ("hipStreamCreate", _I, [C.POINTER(_P)]),
            ("hipStreamDestroy", _I, [_P]),
            ("hipStreamSynchronize", _I, [_P]),
            ("hipEventCreate", _I, [C.POINTER(_P)]),
            ("hipEventRecord", _I, [_P, _P]),
            ("hipEventSynchronize", _I, [_P]),
            ("hipEventElapsedTime", _I, [C.POINTER(C.c_float), _P, _P]),
            ("hipEventDestroy", _I, [_P]),
            ("hipGetErrorString", C.c_char_p, [_I]),
            ("hipGetDeviceCount", _I, [C.POINTER(_I)]),
            ("hipSetDevice", _I, [_I]),
            ("hipStreamBeginCapture", _I, [_P, _I]),
            ("hipStreamEndCapture", _I, [_P, C.POINTER(_P)]),
            ("hipGraphInstantiate", _I, [C.POINTER(_P), _P, _P, _P, _SZ]),
            ("hipGraphLaunch", _I, [_P, _P]),
            ("hipGraphExecDestroy", _I, [_P]),
            ("hipGraphDestroy", _I, [_P]),
        ]:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args

    def check(self, rc, what=""):
        if rc != 0:
            raise HipError(f"{what}: {self.lib.hipGetErrorString(rc).decode()} ({rc})")

    def malloc(self, n):
        p = _P()
        self.check(self.lib.hipMalloc(C.byref(p), max(int(n), 1)), "hipMalloc")
        return p.value

    def free(self, p):
        if p:
            self.check(self.lib.hipFree(p), "hipFree")

    def htod(self, dst, arr):
        import numpy as np
        a = np.ascontiguousarray(arr)
        self.check(self.lib.hipMemcpy(dst, a.ctypes.data, a.nbytes, self.H2D), "hipMemcpy H2D")

    def dtoh(self, src, nbytes, dtype="u1"):
        import numpy as np
        out = np.empty(int(nbytes) // np.dtype(dtype).itemsize, dtype=dtype)
        if out.nbytes:
            self.check(self.lib.hipMemcpy(out.ctypes.data, src, out.nbytes, self.D2H), "hipMemcpy D2H")
        return out

    def dtod(self, dst, src, nbytes):
        self.check(self.lib.hipMemcpy(dst, src, nbytes, self.D2D), "hipMemcpy D2D")

    def memset(self, p, val, n):
        self.check(self.lib.hipMemset(p, val, n), "hipMemset")

    def sync(self):
        self.check(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def stream(self):
        s = _P()
        self.check(self.lib.hipStreamCreate(C.byref(s)), "hipStreamCreate")
        return s.value

    def stream_sync(self, s):
        self.check(self.lib.hipStreamSynchronize(s), "hipStreamSynchronize")

    def stream_destroy(self, s):
        self.check(self.lib.hipStreamDestroy(s), "hipStreamDestroy")

    def event(self):
        e = _P()
        self.check(self.lib.hipEventCreate(C.byref(e)), "hipEventCreate")
        return e.value

    def record(self, e, s=None):
        self.check(self.lib.hipEventRecord(e, s), "hipEventRecord")

    def elapsed_ms(self, a, b):
        self.check(self.lib.hipEventSynchronize(b), "hipEventSynchronize")
        ms = C.c_float()
        self.check(self.lib.hipEventElapsedTime(C.byref(ms), a, b), "hipEventElapsedTime")
        return float(ms.value)

    def event_destroy(self, e):
        self.check(self.lib.hipEventDestroy(e), "hipEventDestroy")

    def device_count(self):
        n = _I()
        rc = self.lib.hipGetDeviceCount(C.byref(n))
        return int(n.value) if rc == 0 else 0


def shm_job_env(rank: int, nranks: int, name: str, extra: dict | None = None) -> dict:
    """Environment for one PE process of a shared-memory-bootstrapped job."""
    env = dict(os.environ)
    env.update({
        "NVSHMEM_BOOTSTRAP": "shm",
        "NVSHMEM_BOOTSTRAP_RANK": str(rank),
        "NVSHMEM_BOOTSTRAP_NRANKS": str(nranks),
        "NVSHMEM_BOOTSTRAP_NAME": name,
    })
    if extra:
        env.update({k: str(v) for k, v in extra.items()})
    return env
