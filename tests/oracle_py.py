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
            "oracle_gen_barrier_init": (None, [_P, _I, C.c_uint32]),
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

    def gen_barrier(self, npes, generation=0):
        buf = (C.c_uint32 * 2)()
        self.lib.oracle_gen_barrier_init(buf, npes, generation)
        return buf

    def batch_generate(self, seed, nreq, npes, src_region, dst_region, min_log2=6, max_log2=22):
        f = self.lib.oracle_batch_generate
        f.restype = _I
        f.argtypes = [_U64, _I, _I, _U64, _U64, _I, _I, _P, _P, _P, _P, _P, _P]
        out = {k: np.zeros(nreq, dt) for k, dt in [("src_pe", np.int32), ("dst_pe", np.int32), ("src_off", np.uint64),
                                                   ("dst_off", np.uint64), ("nbytes", np.uint64), ("dtype", np.int32)]}
        rc = f(seed, nreq, npes, src_region, dst_region, min_log2, max_log2,
               *[out[k].ctypes.data for k in ("src_pe", "dst_pe", "src_off", "dst_off", "nbytes", "dtype")])
        if rc != 0:
            raise ValueError(f"oracle_batch_generate failed ({rc})")
        return out

    def batch_apply(self, heaps, reqs, src_base, dst_base):
        f = self.lib.oracle_batch_apply
        f.restype = None
        f.argtypes = [_P, _I, _P, _P, _P, _P, _P, _U64, _U64]
        ptrs = (C.c_void_p * len(heaps))(*[h.ctypes.data for h in heaps])
        f(ptrs, len(reqs["nbytes"]), *[reqs[k].ctypes.data for k in ("src_pe", "dst_pe", "src_off", "dst_off",
                                                                      "nbytes")], src_base, dst_base)

    def mlp_num_params(self, c):
        f = self.lib.oracle_mlp_num_params
        f.restype, f.argtypes = _I, [_I]
        return f(c)

    def mlp_forward(self, P, X, c):
        f = self.lib.oracle_mlp_forward
        f.restype, f.argtypes = None, [_P, _I, _P, _I, _P]
        P, X = np.ascontiguousarray(P, np.float64), np.ascontiguousarray(X, np.float64)
        Y = np.zeros((len(X), c), np.float64)
        f(P.ctypes.data, c, X.ctypes.data, len(X), Y.ctypes.data)
        return Y

    def mlp_loss_grad(self, P, X, T, c):
        f = self.lib.oracle_mlp_loss_grad
        f.restype, f.argtypes = C.c_double, [_P, _I, _P, _P, _I, _P]
        P = np.ascontiguousarray(P, np.float64)
        X, T = np.ascontiguousarray(X, np.float64), np.ascontiguousarray(T, np.float64)
        g = np.zeros_like(P)
        loss = f(P.ctypes.data, c, X.ctypes.data, T.ctypes.data, len(X), g.ctypes.data)
        return loss, g

    def jacobi(self, G, steps, seed):
        """global G^3 field after `steps` periodic Jacobi steps (float64)"""
        init, run = self.lib.oracle_jacobi_init, self.lib.oracle_jacobi_run
        init.restype, init.argtypes = None, [_P, _I, _U64]
        run.restype, run.argtypes = _I, [_P, _P, _I, _I]
        u = np.zeros(G ** 3, np.float64)
        v = np.zeros(G ** 3, np.float64)
        init(u.ctypes.data, G, seed)
        which = run(u.ctypes.data, v.ctypes.data, G, steps)
        return (v if which else u).reshape(G, G, G)

    def jacobi_mt(self, G, steps, seed, threads):
        """the same field, every step's planes split between `threads` threads"""
        init, run = self.lib.oracle_jacobi_init, self.lib.oracle_jacobi_run_mt
        init.restype, init.argtypes = None, [_P, _I, _U64]
        run.restype, run.argtypes = _I, [_P, _P, _I, _I, _I]
        u = np.zeros(G ** 3, np.float64)
        v = np.zeros(G ** 3, np.float64)
        init(u.ctypes.data, G, seed)
        which = run(u.ctypes.data, v.ctypes.data, G, steps, threads)
        return (v if which else u).reshape(G, G, G)

    def iput(self, heaps, target_off, tst, source, sst, nelems, elem_bytes, pe):
        f = self.lib.oracle_iput
        f.restype, f.argtypes = None, [_P, _SZ, C.c_ssize_t, _P, C.c_ssize_t, _SZ, _SZ, _I]
        ptrs = (C.c_void_p * len(heaps))(*[h.ctypes.data for h in heaps])
        src = np.ascontiguousarray(source)
        f(ptrs, target_off, tst, src.ctypes.data, sst, nelems, elem_bytes, pe)

    def iget(self, heaps, target, tst, source_off, sst, nelems, elem_bytes, pe):
        f = self.lib.oracle_iget
        f.restype, f.argtypes = None, [_P, _P, C.c_ssize_t, _SZ, C.c_ssize_t, _SZ, _SZ, _I]
        ptrs = (C.c_void_p * len(heaps))(*[h.ctypes.data for h in heaps])
        f(ptrs, target.ctypes.data, tst, source_off, sst, nelems, elem_bytes, pe)

    def put_2d(self, heaps, target_off, tpitch, source, spitch, width, height, pe):
        f = self.lib.oracle_put_2d
        f.restype, f.argtypes = None, [_P, _SZ, _SZ, _P, _SZ, _SZ, _SZ, _I]
        ptrs = (C.c_void_p * len(heaps))(*[h.ctypes.data for h in heaps])
        src = np.ascontiguousarray(source)
        f(ptrs, target_off, tpitch, src.ctypes.data, spitch, width, height, pe)

    def jacobi_step(self, u):
        """one periodic Jacobi step of a G^3 float64 field"""
        u = np.ascontiguousarray(u, np.float64)
        G = u.shape[0]
        v = np.empty_like(u)
        step = self.lib.oracle_jacobi_step
        step.restype, step.argtypes = None, [_P, _P, _I]
        step(u.ctypes.data, v.ctypes.data, G)
        return v

    def trace_row(self, *fields):
        buf = C.create_string_buffer(512)
        self.lib.oracle_trace_row(buf, 512, *fields)
        return buf.value.decode()
