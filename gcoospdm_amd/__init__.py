"""gcoospdm_amd -- MI355X-native GCOOSpDM (grouped-COO sparse x dense multiply).

Python binding of the C ABI in ``include/gcoo_spdm.h`` (libgcoospdm.so,
HIP kernels for gfx950).  The classes mirror the reference's plugin API
(reference src/ShGemm.h:7-24, src/SpGemmBuilder.h:12-23,
src/ShGCOOSpDM.h:8-33) so that code written against the reference reads the
same:

    algo = SpGemmBuilder.getShGemmAlgo(S, M, K, "gcoospdm")
    algo.mutiplyBy(D, K, N, C)          # convert S + multiply
    algo.build(); algo.runKernelTest(D, K, N, C)

Matrices are torch tensors on a HIP device (torch supplies device memory and
streams only; every computation runs in libgcoospdm.so).  There is no CPU
fallback: importing this package on a machine without the built library
raises, and every call fails loudly when the device path is unavailable.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgcoospdm.so")

GCOO_OK = 0
GCOO_ERR_INVALID = 1
GCOO_ERR_HIP = 2
GCOO_ERR_ALLOC = 3
GCOO_ERR_OVERFLOW = 4
GCOO_ERR_UNSUPPORTED = 5
GCOO_ERR_BUDGET = 6
GCOO_ERR_STATE = 7

# Every symbol include/gcoo_spdm.h declares: (name, restype, argtypes).
_c = ctypes
_P = _c.c_void_p
_I = _c.c_int
_I64 = _c.c_int64
_U64 = _c.c_uint64
_SZ = _c.c_size_t
_F = _c.c_float
_PP = _c.POINTER(_c.c_void_p)
_SIGNATURES = [
    ("gcoo_last_error", _c.c_char_p, []),
    ("gcoo_version", _c.c_char_p, []),
    ("gcoo_device_count", _I, [_c.POINTER(_I)]),
    ("gcoo_set_device", _I, [_I]),
    ("gcoo_dev_malloc", _I, [_PP, _SZ]),
    ("gcoo_dev_free", _I, [_P]),
    ("gcoo_memcpy_h2d", _I, [_P, _P, _SZ, _P]),
    ("gcoo_memcpy_d2h", _I, [_P, _P, _SZ, _P]),
    ("gcoo_memset_async", _I, [_P, _I, _SZ, _P]),
    ("gcoo_stream_create", _I, [_PP]),
    ("gcoo_stream_destroy", _I, [_P]),
    ("gcoo_stream_sync", _I, [_P]),
    ("gcoo_device_sync", _I, []),
    ("gcoo_timer_create", _I, [_PP]),
    ("gcoo_timer_destroy", _I, [_P]),
    ("gcoo_timer_start", _I, [_P, _P]),
    ("gcoo_timer_stop", _I, [_P, _P]),
    ("gcoo_timer_elapsed_ms", _I, [_P, _c.POINTER(_F)]),
    ("gcoo_fill_uniform_async", _I, [_P, _I64, _U64, _U64, _F, _P]),
    ("gcoo_fill_sparse_async", _I, [_P, _I64, _U64, _U64, _F, _U64, _P]),
    ("gcoo_create", _I, [_I, _I, _I, _PP]),
    ("gcoo_destroy", _I, [_P]),
    ("gcoo_convert_dense", _I, [_P, _P, _P]),
    ("gcoo_convert_coo", _I, [_P, _P, _P, _P, _I64, _P]),
    ("gcoo_get_info", _I, [_P, _c.POINTER(_I), _c.POINTER(_I), _c.POINTER(_I),
                           _c.POINTER(_I), _c.POINTER(_I64)]),
    ("gcoo_get_arrays", _I, [_P, _PP, _PP, _PP, _PP, _PP]),
    ("gcoo_spdm_f32_async", _I, [_P, _P, _I, _I, _P, _P]),
    ("gcoo_spdm_f32", _I, [_P, _P, _I, _I, _P, _P]),
    ("gcoo_spdm_f32_noalloc_async", _I, [_P, _P, _I, _I, _P, _P]),
    ("gcoo_multiply_by", _I, [_P, _P, _P, _I, _I, _P, _P]),
    ("gcoo_test_set_conversion_abort", _I, [_P, _P]),
    ("gcoo_set_spdm_config", _I, [_P, _P]),
    ("gcoo_get_spdm_config", _I, [_P, _P]),
    ("gcoo_get_spdm_choice", _I, [_P, _I, _P, _I, _P, _c.POINTER(_I), _c.POINTER(_I),
                                  _c.POINTER(_I), _c.POINTER(_I)]),
    ("gcoo_f32_to_bf16_async", _I, [_P, _P, _I64, _P]),
    ("gcoo_spdm_bf16_async", _I, [_P, _P, _I, _I, _P, _P]),
    ("gcoo_spdm_bf16", _I, [_P, _P, _I, _I, _P, _P]),
    ("gcoo_batch_create", _I, [_PP]),
    ("gcoo_batch_destroy", _I, [_P]),
    ("gcoo_batch_clear", _I, [_P]),
    ("gcoo_batch_add", _I, [_P, _P, _P, _I, _I, _P, _I, _I, _c.POINTER(_I)]),
    ("gcoo_batch_commit", _I, [_P, _P]),
    ("gcoo_batch_run_async", _I, [_P, _P]),
    ("gcoo_batch_graph_capture", _I, [_P, _P]),
    ("gcoo_batch_graph_launch", _I, [_P, _P]),
    ("gcoo_batch_info", _I, [_P, _c.POINTER(_I), _c.POINTER(_I64), _c.POINTER(_I)]),
    ("gcoo_batch_job_tile", _I, [_P, _I, _c.POINTER(_I), _c.POINTER(_I), _c.POINTER(_I)]),
    ("gcoo_batch_set_panel_mode", _I, [_P, _I, _c.c_double, _c.c_double]),
    ("gcoo_set_panel_balance", _I, [_P, _I]),
    ("gcoo_csr_create", _I, [_I, _I, _PP]),
    ("gcoo_csr_destroy", _I, [_P]),
    ("gcoo_csr_convert_dense", _I, [_P, _P, _P]),
    ("gcoo_csr_get_arrays", _I, [_P, _PP, _PP, _PP, _c.POINTER(_I64)]),
    ("gcoo_csr_spmm_f32", _I, [_P, _P, _I, _P, _P]),
    ("gcoo_csr_spmm_f32_async", _I, [_P, _P, _I, _P, _P]),
    ("gcoo_batch_panel_info", _I, [_P, _c.POINTER(_I), _c.POINTER(_I)]),
    ("gcoo_batch_job_is_panel", _I, [_P, _I, _c.POINTER(_I)]),
    ("gcoo_server_create", _I, [_P, _PP]),
    ("gcoo_server_destroy", _I, [_P]),
    ("gcoo_server_register", _I, [_P, _P, _P, _I, _I, _I, _I, _c.POINTER(_I)]),
    ("gcoo_server_submit", _I, [_P, _I, _c.POINTER(_U64)]),
    ("gcoo_server_wait", _I, [_P, _U64, _c.POINTER(_c.c_double)]),
    ("gcoo_server_result", _I, [_P, _U64, _PP, _c.POINTER(_I), _c.POINTER(_I)]),
    ("gcoo_server_release", _I, [_P, _U64]),
    ("gcoo_server_stats", _I, [_P, _P]),
    ("gcoo_server_reset_stats", _I, [_P]),
    ("gcoo_test_server_inject_failure", _I, [_P, _I]),
]

_lib: Optional[ctypes.CDLL] = None


class SpdmConfig(ctypes.Structure):
    """gcoo_spdm_config (include/gcoo_spdm.h): variant, vec, waves, chunk,
    unroll; zero fields mean automatic."""
    _fields_ = [("variant", ctypes.c_int), ("vec", ctypes.c_int),
                ("waves", ctypes.c_int), ("chunk", ctypes.c_int),
                ("unroll", ctypes.c_int)]


class GCOOError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


def lib() -> ctypes.CDLL:
    """Load libgcoospdm.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `make -C gcoospdm_amd` "
                "(or __graft_entry__.build()) first")
        handle = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGNATURES:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def declared_symbols():
    return [s[0] for s in _SIGNATURES]


def _check(rc: int, where: str):
    if rc != GCOO_OK:
        raise GCOOError(rc, where, lib().gcoo_last_error().decode())


def _ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


def _stream(stream=None) -> int:
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return int(stream.cuda_stream)


def _require_dev(t, name, dtype=None):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a HIP device tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous (row-major)")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")


class GCOOMatrix:
    """A sparse S (M x K) in grouped COO, owned by libgcoospdm.so."""

    def __init__(self, M: int, K: int, cpg: int = 4):
        h = ctypes.c_void_p()
        _check(lib().gcoo_create(M, K, cpg, ctypes.byref(h)), "gcoo_create")
        self._h = h
        self.M, self.K, self.cpg = M, K, cpg

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().gcoo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def convert_dense(self, S, stream=None):
        import torch
        _require_dev(S, "S", torch.float32)
        if tuple(S.shape) != (self.M, self.K):
            raise ValueError(f"S shape {tuple(S.shape)} != {(self.M, self.K)}")
        _check(lib().gcoo_convert_dense(self._h, _ptr(S), _stream(stream)),
               "gcoo_convert_dense")
        return self

    def convert_coo(self, rows, cols, vals, stream=None):
        import torch
        _require_dev(rows, "rows", torch.int32)
        _require_dev(cols, "cols", torch.int32)
        _require_dev(vals, "vals", torch.float32)
        n = rows.numel()
        if cols.numel() != n or vals.numel() != n:
            raise ValueError("rows, cols, vals must have the same length")
        _check(lib().gcoo_convert_coo(self._h, _ptr(rows), _ptr(cols), _ptr(vals), n,
                                      _stream(stream)), "gcoo_convert_coo")
        return self

    def info(self):
        M, K, cpg, ng = (ctypes.c_int() for _ in range(4))
        nnz = ctypes.c_int64()
        _check(lib().gcoo_get_info(self._h, ctypes.byref(M), ctypes.byref(K),
                                   ctypes.byref(cpg), ctypes.byref(ng),
                                   ctypes.byref(nnz)), "gcoo_get_info")
        return dict(M=M.value, K=K.value, cpg=cpg.value, nGroup=ng.value, nnz=nnz.value)

    @property
    def nnz(self) -> int:
        return self.info()["nnz"]

    @property
    def nGroup(self) -> int:
        return self.info()["nGroup"]

    def arrays_host(self):
        """Copy the GCOO arrays to host numpy arrays (for parity checks).

        Returns dict(vals f32[nnz], rows i32[nnz], cols i32[nnz],
        groupIndex i32[nGroup+1], nnzPerGroup i32[nGroup+1])."""
        import numpy as np
        ptrs = [ctypes.c_void_p() for _ in range(5)]
        _check(lib().gcoo_get_arrays(self._h, *[ctypes.byref(p) for p in ptrs]),
               "gcoo_get_arrays")
        inf = self.info()
        nnz, ng = inf["nnz"], inf["nGroup"]
        out = {}
        for name, p, n, dt in zip(
                ["vals", "rows", "cols", "groupIndex", "nnzPerGroup"], ptrs,
                [nnz, nnz, nnz, ng + 1, ng + 1],
                [np.float32, np.int32, np.int32, np.int32, np.int32]):
            a = np.empty(n, dtype=dt)
            if n:
                _check(lib().gcoo_memcpy_d2h(a.ctypes.data, p, a.nbytes, _stream()),
                       "gcoo_memcpy_d2h")
            out[name] = a
        return out

    def spdm(self, D, C, stream=None, sync=True):
        """C = S * D (FP32, bit-identical to the reference kernel)."""
        import torch
        _require_dev(D, "D", torch.float32)
        _require_dev(C, "C", torch.float32)
        K, N = D.shape
        if tuple(C.shape) != (self.M, N):
            raise ValueError(f"C shape {tuple(C.shape)} != {(self.M, N)}")
        fn = lib().gcoo_spdm_f32 if sync else lib().gcoo_spdm_f32_async
        _check(fn(self._h, _ptr(D), K, N, _ptr(C), _stream(stream)), "gcoo_spdm_f32")
        return C

    def spdm_bf16(self, D16, C, stream=None, sync=True):
        """C = S * D with BF16 storage of S's values and of D (a torch.int16
        or torch.bfloat16 tensor of BF16 bits), FP32 accumulation."""
        import torch
        if D16.dtype not in (torch.int16, torch.bfloat16):
            raise ValueError("D16 must hold BF16 bits (torch.bfloat16 or torch.int16)")
        _require_dev(D16, "D16")
        _require_dev(C, "C", torch.float32)
        K, N = D16.shape
        if tuple(C.shape) != (self.M, N):
            raise ValueError(f"C shape {tuple(C.shape)} != {(self.M, N)}")
        fn = lib().gcoo_spdm_bf16 if sync else lib().gcoo_spdm_bf16_async
        _check(fn(self._h, _ptr(D16), K, N, _ptr(C), _stream(stream)), "gcoo_spdm_bf16")
        return C

    def set_panel_balance(self, mode: int):
        """Row order of the panel-staged multiply's lists (gcoo_set_panel_balance):
        0 in order (default), 1 balanced where the lists outlive the call, 2
        always balanced."""
        _check(lib().gcoo_set_panel_balance(self._h, mode), "gcoo_set_panel_balance")
        return self

    def set_config(self, variant=0, vec=0, waves=0, chunk=0, unroll=0):
        """Multiply-kernel variant / tile shape (0 = automatic)."""
        cfg = SpdmConfig(variant, vec, waves, chunk, unroll)
        _check(lib().gcoo_set_spdm_config(self._h, ctypes.byref(cfg)), "gcoo_set_spdm_config")
        return self

    def get_config(self):
        cfg = SpdmConfig()
        _check(lib().gcoo_get_spdm_config(self._h, ctypes.byref(cfg)), "gcoo_get_spdm_config")
        return dict(variant=cfg.variant, vec=cfg.vec, waves=cfg.waves, chunk=cfg.chunk,
                    unroll=cfg.unroll)

    def spdm_choice(self, D, C):
        """(variant, vec, waves, rows per wave) of the multiply kernel a launch
        with D (FP32 or BF16 bits) and C would run (gcoo_get_spdm_choice);
        waves and rows are 0 unless the variant is 6."""
        import torch
        dtype = 0 if D.dtype == torch.float32 else 1
        v, vec, w, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().gcoo_get_spdm_choice(self._h, dtype, _ptr(D), D.shape[1], _ptr(C),
                                          ctypes.byref(v), ctypes.byref(vec), ctypes.byref(w),
                                          ctypes.byref(r)),
               "gcoo_get_spdm_choice")
        return v.value, vec.value, w.value, r.value

    def multiply_by(self, S, D, C, stream=None):
        """Convert S then C = S * D (the reference's mutiplyBy).  Returns once
        the conversion is checked; the multiply may still run on `stream`
        (torch's current stream by default), so work queued on that stream
        sees C complete."""
        import torch
        _require_dev(S, "S", torch.float32)
        _require_dev(D, "D", torch.float32)
        _require_dev(C, "C", torch.float32)
        K, N = D.shape
        _check(lib().gcoo_multiply_by(self._h, _ptr(S), _ptr(D), K, N, _ptr(C),
                                      _stream(stream)), "gcoo_multiply_by")
        return C


class Batch:
    """Many independent multiplies in one persistent launch (gcoo_batch_*).

    add() keeps references to the handle and tensors so they outlive the
    batch's runs."""

    def __init__(self):
        b = ctypes.c_void_p()
        _check(lib().gcoo_batch_create(ctypes.byref(b)), "gcoo_batch_create")
        self._b = b
        self._keep = []

    def close(self):
        if getattr(self, "_b", None) is not None and self._b.value:
            lib().gcoo_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self):
        _check(lib().gcoo_batch_clear(self._b), "gcoo_batch_clear")
        self._keep = []

    def add(self, m: "GCOOMatrix", D, C, rows=0, vec=0) -> int:
        import torch
        _require_dev(C, "C", torch.float32)
        _require_dev(D, "D")
        if D.dtype == torch.float32:
            dtype = 0
        elif D.dtype in (torch.bfloat16, torch.int16):
            dtype = 1
        else:
            raise ValueError("D must be float32 or BF16 bits")
        K, N = D.shape
        if K != m.K or tuple(C.shape) != (m.M, N):
            raise ValueError("shape mismatch")
        idx = ctypes.c_int()
        _check(lib().gcoo_batch_add(self._b, m.handle, _ptr(D), dtype, N, _ptr(C), rows, vec,
                                    ctypes.byref(idx)), "gcoo_batch_add")
        self._keep.append((m, D, C))
        return idx.value

    def commit(self, stream=None):
        _check(lib().gcoo_batch_commit(self._b, _stream(stream)), "gcoo_batch_commit")
        return self

    def run(self, stream=None):
        _check(lib().gcoo_batch_run_async(self._b, _stream(stream)), "gcoo_batch_run_async")

    def capture(self, stream):
        _check(lib().gcoo_batch_graph_capture(self._b, _stream(stream)),
               "gcoo_batch_graph_capture")

    def replay(self, stream=None):
        _check(lib().gcoo_batch_graph_launch(self._b, _stream(stream)),
               "gcoo_batch_graph_launch")

    def info(self):
        j, g = ctypes.c_int(), ctypes.c_int()
        it = ctypes.c_int64()
        _check(lib().gcoo_batch_info(self._b, ctypes.byref(j), ctypes.byref(it),
                                     ctypes.byref(g)), "gcoo_batch_info")
        return dict(jobs=j.value, items=it.value, grid_blocks=g.value)

    def set_panel_mode(self, mode: int, ratio_f32: float = 0.0, ratio_bf16: float = 0.0):
        """0: no panel jobs, 1: the cost model's choice (default; a job is a
        panel job when its panel estimate is below ratio x its persistent
        estimate, per element type, 0 = the built-in ratio), 2: every job
        whose operands allow the panel kernel.  Takes effect at commit."""
        _check(lib().gcoo_batch_set_panel_mode(self._b, mode, float(ratio_f32),
                                               float(ratio_bf16)),
               "gcoo_batch_set_panel_mode")
        return self

    def panel_info(self):
        pj, wg = ctypes.c_int(), ctypes.c_int()
        _check(lib().gcoo_batch_panel_info(self._b, ctypes.byref(pj), ctypes.byref(wg)),
               "gcoo_batch_panel_info")
        return dict(panel_jobs=pj.value, workgroups=wg.value)

    def job_is_panel(self, job) -> bool:
        p = ctypes.c_int()
        _check(lib().gcoo_batch_job_is_panel(self._b, job, ctypes.byref(p)),
               "gcoo_batch_job_is_panel")
        return bool(p.value)

    def job_tile(self, job):
        r, v, ln = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().gcoo_batch_job_tile(self._b, job, ctypes.byref(r), ctypes.byref(v),
                                         ctypes.byref(ln)), "gcoo_batch_job_tile")
        return dict(rows=r.value, vec=v.value, lanes=ln.value)


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("slots", ctypes.c_int),
                ("max_batch_jobs", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("wall_s", ctypes.c_double), ("peak_bytes", ctypes.c_size_t),
                ("budget_bytes", ctypes.c_size_t)]


class Server:
    """Continuous-batching request loop (gcoo_server_*), operands resident."""

    def __init__(self, hbm_budget_bytes: int, slots: int = 2, max_batch_jobs: int = 64):
        cfg = ServerConfig(hbm_budget_bytes, slots, max_batch_jobs)
        s = ctypes.c_void_p()
        _check(lib().gcoo_server_create(ctypes.byref(cfg), ctypes.byref(s)), "gcoo_server_create")
        self._s = s
        self._keep = []

    def close(self):
        if getattr(self, "_s", None) is not None and self._s.value:
            lib().gcoo_server_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def register(self, m: "GCOOMatrix", D, rows=0, vec=0) -> int:
        import torch
        _require_dev(D, "D")
        dtype = 0 if D.dtype == torch.float32 else 1
        op = ctypes.c_int()
        _check(lib().gcoo_server_register(self._s, m.handle, _ptr(D), dtype, D.shape[1], rows, vec,
                                          ctypes.byref(op)), "gcoo_server_register")
        self._keep.append((m, D))
        return op.value

    def submit(self, op: int) -> int:
        rid = ctypes.c_uint64()
        _check(lib().gcoo_server_submit(self._s, op, ctypes.byref(rid)), "gcoo_server_submit")
        return rid.value

    def wait(self, rid: int) -> float:
        lat = ctypes.c_double()
        _check(lib().gcoo_server_wait(self._s, rid, ctypes.byref(lat)), "gcoo_server_wait")
        return lat.value

    def result_host(self, rid: int):
        """Copy a completed request's C to a host numpy array."""
        import numpy as np
        p = ctypes.c_void_p()
        M, N = ctypes.c_int(), ctypes.c_int()
        _check(lib().gcoo_server_result(self._s, rid, ctypes.byref(p), ctypes.byref(M),
                                        ctypes.byref(N)), "gcoo_server_result")
        out = np.empty((M.value, N.value), np.float32)
        if out.size:
            _check(lib().gcoo_memcpy_d2h(out.ctypes.data, p, out.nbytes, None), "gcoo_memcpy_d2h")
        return out

    def release(self, rid: int):
        _check(lib().gcoo_server_release(self._s, rid), "gcoo_server_release")

    def stats(self):
        st = ServerStats()
        _check(lib().gcoo_server_stats(self._s, ctypes.byref(st)), "gcoo_server_stats")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def reset_stats(self):
        _check(lib().gcoo_server_reset_stats(self._s), "gcoo_server_reset_stats")


class ShGemm:
    """Reference src/ShGemm.h:7-24 (abstract)."""

    def __init__(self, nRows: int, nCols: int):
        self.m_nRows = nRows
        self.m_nCols = nCols
        self.m_pDenseMatrix = None

    def mutiplyBy(self, A, nRows, nCols, C):
        raise NotImplementedError

    def mutiply(self, B, nRows, nCols, C):
        raise NotImplementedError

    def build(self):
        raise NotImplementedError

    def runKernelTest(self, A, nRows, nCols, C):
        raise NotImplementedError


class ShGCOOSpDM(ShGemm):
    """Reference src/ShGCOOSpDM.h:8-33 over the C ABI.

    denseMatrix: device tensor S (nRows x nCols).  mutiplyBy / runKernelTest
    take the dense operand A (nRows_A x nCols_A with nRows_A == nCols) and
    write C (nRows x nCols_A)."""

    def __init__(self, denseMatrix, nRows: int, nCols: int, cpg: int = 4):
        super().__init__(nRows, nCols)
        self.m_pDenseMatrix = denseMatrix
        self._m = GCOOMatrix(nRows, nCols, cpg)
        self.m_nGroup = (nRows + cpg - 1) // cpg

    def _shape_check(self, A, nRows, nCols):
        if tuple(A.shape) != (nRows, nCols):
            raise ValueError(f"A shape {tuple(A.shape)} != {(nRows, nCols)}")

    def mutiplyBy(self, A, nRows, nCols, C):
        self._shape_check(A, nRows, nCols)
        self._m.multiply_by(self.m_pDenseMatrix, A, C)
        return 0.0

    def mutiply(self, B, nRows, nCols, C):
        return self.mutiplyBy(B, nRows, nCols, C)

    def build(self):
        self._m.convert_dense(self.m_pDenseMatrix)

    def runKernelTest(self, A, nRows, nCols, C):
        self._shape_check(A, nRows, nCols)
        self._m.spdm(A, C)
        return 0.0

    @property
    def matrix(self) -> GCOOMatrix:
        return self._m


class CSRMatrix:
    """A CSR matrix on the device (gcoo_csr_*): rowPtr, colInd, vals."""

    def __init__(self, M: int, K: int):
        c = ctypes.c_void_p()
        _check(lib().gcoo_csr_create(M, K, ctypes.byref(c)), "gcoo_csr_create")
        self._c = c
        self.M, self.K = M, K

    def close(self):
        if getattr(self, "_c", None) is not None and self._c.value:
            lib().gcoo_csr_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def convert_dense(self, S, stream=None):
        import torch
        _require_dev(S, "S", torch.float32)
        if tuple(S.shape) != (self.M, self.K):
            raise ValueError("S shape mismatch")
        _check(lib().gcoo_csr_convert_dense(self._c, _ptr(S), _stream(stream)),
               "gcoo_csr_convert_dense")
        return self

    def spmm(self, D, C, stream=None, sync=True):
        import torch
        _require_dev(D, "D", torch.float32)
        _require_dev(C, "C", torch.float32)
        K, N = D.shape
        if K != self.K or tuple(C.shape) != (self.M, N):
            raise ValueError("shape mismatch")
        fn = lib().gcoo_csr_spmm_f32 if sync else lib().gcoo_csr_spmm_f32_async
        _check(fn(self._c, _ptr(D), N, _ptr(C), _stream(stream)), "gcoo_csr_spmm_f32")
        return C

    def arrays_host(self):
        """(rowPtr, colInd, vals) copied to the host as numpy arrays."""
        import numpy as np
        rp, ci, va = (ctypes.c_void_p() for _ in range(3))
        n = ctypes.c_int64()
        _check(lib().gcoo_csr_get_arrays(self._c, ctypes.byref(rp), ctypes.byref(ci),
                                         ctypes.byref(va), ctypes.byref(n)), "gcoo_csr_get_arrays")
        out = {}
        for key, ptr, count, dt in (("rowPtr", rp, self.M + 1, np.int32),
                                    ("colInd", ci, n.value, np.int32),
                                    ("vals", va, n.value, np.float32)):
            host = np.empty(count, dt)
            if count:
                _check(lib().gcoo_memcpy_d2h(host.ctypes.data_as(ctypes.c_void_p), ptr,
                                             host.nbytes, None), "gcoo_memcpy_d2h")
            out[key] = host
        out["nnz"] = n.value
        return out


class ShCsrSpGemm(ShGemm):
    """Factory key "csr": the reference's row-wise CSR SpDM ablation
    (src/ShCsrSpGemm.h, src/ShCsrSpGemm.cu:9-146) over gcoo_csr_*: the
    three-array CSR of S and the reference kernel's structure (a workgroup
    per row and column tile, the row's entries staged in LDS)."""

    def __init__(self, denseMatrix, nRows: int, nCols: int):
        super().__init__(nRows, nCols)
        self.m_pDenseMatrix = denseMatrix
        self._c = CSRMatrix(nRows, nCols)

    def _shape_check(self, A, nRows, nCols):
        if tuple(A.shape) != (nRows, nCols):
            raise ValueError(f"A shape {tuple(A.shape)} != {(nRows, nCols)}")

    def mutiplyBy(self, A, nRows, nCols, C):
        self._shape_check(A, nRows, nCols)
        self._c.convert_dense(self.m_pDenseMatrix)
        self._c.spmm(A, C)
        return 0.0

    def mutiply(self, B, nRows, nCols, C):
        return self.mutiplyBy(B, nRows, nCols, C)

    def build(self):
        self._c.convert_dense(self.m_pDenseMatrix)

    def runKernelTest(self, A, nRows, nCols, C):
        self._shape_check(A, nRows, nCols)
        self._c.spmm(A, C)
        return 0.0

    @property
    def matrix(self) -> CSRMatrix:
        return self._c


class SpGemmBuilder:
    """Reference src/SpGemmBuilder.h:12-23: "gcoospdm", "csr", and the vendor
    baselines "cusparse" (rocSPARSE) and "cublas" (rocBLAS); None otherwise."""

    @staticmethod
    def getShGemmAlgo(denseMatrix, nRows: int, nCols: int, desc: str):
        if desc == "gcoospdm":
            return ShGCOOSpDM(denseMatrix, nRows, nCols)
        if desc == "csr":
            return ShCsrSpGemm(denseMatrix, nRows, nCols)
        if desc == "cusparse":
            from .vendor import ShRocsparseGemm
            return ShRocsparseGemm(denseMatrix, nRows, nCols)
        if desc == "cublas":
            from .vendor import ShRocblasGemm
            return ShRocblasGemm(denseMatrix, nRows, nCols)
        return None


def keep_threshold(sparsity: float) -> int:
    """floor(sparsity * 2^40): the Bernoulli cut of gcoo_fill_sparse_async."""
    import math
    if sparsity <= 0.0:
        return 0
    if sparsity >= 1.0:
        return 1 << 40
    return int(math.floor(sparsity * 1099511627776.0))


def to_bf16(x, out=None, stream=None):
    """RNE FP32 -> BF16 on the device (gcoo_f32_to_bf16_async); returns a
    torch.bfloat16 tensor of the same shape."""
    import torch
    _require_dev(x, "x", torch.float32)
    if out is None:
        out = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device)
    _require_dev(out, "out")
    _check(lib().gcoo_f32_to_bf16_async(_ptr(x), _ptr(out), x.numel(), _stream(stream)),
           "gcoo_f32_to_bf16_async")
    return out


def fill_uniform(t, seed: int, stream_id: int, scale: float = 1.0, stream=None):
    import torch
    _require_dev(t, "t", torch.float32)
    _check(lib().gcoo_fill_uniform_async(_ptr(t), t.numel(), seed, stream_id, scale,
                                         _stream(stream)), "gcoo_fill_uniform_async")
    return t


def fill_sparse(t, seed: int, stream_id: int, sparsity: float, scale: float = 1.0,
                stream=None):
    import torch
    _require_dev(t, "t", torch.float32)
    _check(lib().gcoo_fill_sparse_async(_ptr(t), t.numel(), seed, stream_id, scale,
                                        keep_threshold(sparsity), _stream(stream)),
           "gcoo_fill_sparse_async")
    return t
