"""ctypes view of the CPU oracle (oracle/build/liboracle.so).  Test
infrastructure: imported only by tests, the smoke check and the benchmark's
CPU-baseline leg."""
import ctypes as C
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent.parent / "oracle"
_P, _SZ, _U64, _I = C.c_void_p, C.c_size_t, C.c_uint64, C.c_int


class OracleMissing(RuntimeError):
    pass


class Oracle:
    def __init__(self):
        path = ORACLE_DIR / "build" / "liboracle.so"
        if not path.exists():
            raise OracleMissing(f"{path} not built (make -C {ORACLE_DIR})")
        L = self.lib = C.CDLL(str(path))
        sigs = {
            "oracle_pcg32_fill": (None, [_P, _SZ, _U64, _U64]),
            "oracle_pcg32_uniform_pm1": (None, [_P, _SZ, _U64, _U64]),
            "oracle_pcg32_uniform01_f64": (None, [_P, _SZ, _U64, _U64]),
            "oracle_translate": (_U64, [_P, _U64, _U64, _U64, _I]),
            "oracle_put": (None, [_P, _SZ, _SZ, _P, _SZ, _I]),
            "oracle_get": (None, [_P, _SZ, _P, _SZ, _SZ, _I]),
            "oracle_p32": (None, [_P, _SZ, C.c_uint32, _I]),
            "oracle_p64": (None, [_P, _SZ, C.c_uint64, _I]),
            "oracle_g32": (C.c_uint32, [_P, _SZ, _I]),
            "oracle_g64": (C.c_uint64, [_P, _SZ, _I]),
            "oracle_int_wait_done": (_I, [_I, _I]),
            "oracle_int_wait_until_done": (_I, [_I, _I, _I]),
            "oracle_cmp_holds": (_I, [_I, _I, _I]),
            "oracle_ref_barrier_value": (_I, [C.c_longlong]),
            "oracle_ref_barrier_half": (_I, [C.c_longlong]),
            "oracle_nvshmalloc": (C.c_longlong, [_P, _SZ]),
            "oracle_heap_init": (None, [_P, _SZ, _SZ]),
            "oracle_nvshmcleanup": (None, [_P]),
            "oracle_gen_barrier_init": (None, [_P, _I]),
            "oracle_gen_barrier_read_gen": (_I, [_P]),
            "oracle_gen_barrier_arrive": (_I, [_P]),
            "oracle_gen_barrier_released": (_I, [_P, _I]),
            "oracle_trace_header": (C.c_char_p, []),
            "oracle_trace_op_name": (C.c_char_p, [_I]),
            "oracle_trace_row": (_I, [C.c_char_p, _SZ, C.c_longlong, C.c_ulonglong, _I, _I, _I, _U64, _U64, _I, _I,
                                      C.c_longlong]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype, fn.argtypes = res, args

    def pcg32(self, n, seed, seq=54):
        out = np.empty(int(n), dtype=np.uint32)
        if n:
            self.lib.oracle_pcg32_fill(out.ctypes.data, out.size, seed, seq)
        return out

    def uniform_pm1(self, n, seed, seq=54):
        out = np.empty(int(n), dtype=np.float32)
        self.lib.oracle_pcg32_uniform_pm1(out.ctypes.data, out.size, seed, seq)
        return out

    def uniform01_f64(self, n, seed, seq=54):
        out = np.empty(int(n), dtype=np.float64)
        self.lib.oracle_pcg32_uniform01_f64(out.ctypes.data, out.size, seed, seq)
        return out

    def put(self, heaps, target_off, source, pe):
        """heaps: list of writable uint8 numpy arrays (one per PE)."""
        ptrs = (C.c_void_p * len(heaps))(*[h.ctypes.data for h in heaps])
        src = np.ascontiguousarray(source).view(np.uint8)
        self.lib.oracle_put(ptrs, heaps[0].size, target_off, src.ctypes.data, src.size, pe)

    def get(self, heaps, target, source_off, pe):
        ptrs = (C.c_void_p * len(heaps))(*[h.ctypes.data for h in heaps])
        t = target.view(np.uint8)
        self.lib.oracle_get(ptrs, heaps[0].size, t.ctypes.data, source_off, t.size, pe)

    def heap(self, size, align):
        buf = (C.c_uint64 * 4)()
        self.lib.oracle_heap_init(buf, size, align)
        return buf

    def nvshmalloc(self, heap, size):
        return self.lib.oracle_nvshmalloc(heap, size)

    def translate(self, bases, anchor, mask, addr, pe):
        arr = (C.c_uint64 * len(bases))(*bases)
        return self.lib.oracle_translate(arr, anchor, mask, addr, pe)

    def gen_barrier(self, npes):
        buf = (C.c_int * 3)()
        self.lib.oracle_gen_barrier_init(buf, npes)
        return buf

    def trace_row(self, *fields):
        buf = C.create_string_buffer(512)
        self.lib.oracle_trace_row(buf, 512, *fields)
        return buf.value.decode()
