"""Vendor baselines of the reference (cuSPARSE CSR SpMM, cuBLAS SGEMM) on
rocSPARSE / rocBLAS, through libgcoo_vendor.so (include/gcoo_vendor.h).

ShRocsparseGemm is the factory key "cusparse" (reference
src/ShCusparseGemm.cpp:23-118), ShRocblasGemm the key "cublas" (reference
src/ShCublasGemm.cpp:16-40).  Both compute C = S * A like the GCOO backend;
their summation order is the library's, so they match the FP64 product to
the FP32 tolerance rather than bit for bit."""
import ctypes
import os

from . import GCOOError, ShGemm, _ptr, _require_dev, _stream

VENDOR_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgcoo_vendor.so")

_c = ctypes
_VENDOR_SIGNATURES = [
    ("gcoo_vendor_last_error", _c.c_char_p, []),
    ("gcoo_vendor_csr_create", _c.c_int, [_c.c_int, _c.c_int, _c.POINTER(_c.c_void_p)]),
    ("gcoo_vendor_csr_destroy", _c.c_int, [_c.c_void_p]),
    ("gcoo_vendor_csr_convert", _c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p]),
    ("gcoo_vendor_csr_nnz", _c.c_int, [_c.c_void_p, _c.POINTER(_c.c_int64)]),
    ("gcoo_vendor_csr_spmm", _c.c_int,
     [_c.c_void_p, _c.c_void_p, _c.c_int, _c.c_int, _c.c_void_p, _c.c_void_p]),
    ("gcoo_vendor_blas_create", _c.c_int, [_c.POINTER(_c.c_void_p)]),
    ("gcoo_vendor_blas_destroy", _c.c_int, [_c.c_void_p]),
    ("gcoo_vendor_sgemm", _c.c_int,
     [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_int, _c.c_int, _c.c_int, _c.c_void_p,
      _c.c_void_p]),
]

_vlib = None


def vendor_lib() -> ctypes.CDLL:
    global _vlib
    if _vlib is None:
        if not os.path.exists(VENDOR_LIB_PATH):
            raise GCOOError(-1, f"{VENDOR_LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(VENDOR_LIB_PATH)
        for name, res, args in _VENDOR_SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _vlib = L
    return _vlib


def declared_vendor_symbols():
    return [name for name, _, _ in _VENDOR_SIGNATURES]


def _vcheck(rc: int, where: str):
    if rc != 0:
        raise GCOOError(rc, f"{where}: {vendor_lib().gcoo_vendor_last_error().decode()}")


class ShRocsparseGemm(ShGemm):
    """Factory key "cusparse": dense S -> CSR on every mutiplyBy, then CSR SpMM."""

    def __init__(self, denseMatrix, nRows: int, nCols: int):
        super().__init__(nRows, nCols)
        self.m_pDenseMatrix = denseMatrix
        self._h = ctypes.c_void_p()
        _vcheck(vendor_lib().gcoo_vendor_csr_create(nRows, nCols, ctypes.byref(self._h)),
                "gcoo_vendor_csr_create")

    def __del__(self):
        if getattr(self, "_h", None) and _vlib is not None:
            _vlib.gcoo_vendor_csr_destroy(self._h)
            self._h = None

    def build(self):
        _require_dev(self.m_pDenseMatrix, "S")
        _vcheck(vendor_lib().gcoo_vendor_csr_convert(self._h, _ptr(self.m_pDenseMatrix),
                                                     _stream()), "gcoo_vendor_csr_convert")

    def runKernelTest(self, A, nRows, nCols, C):
        import torch
        _require_dev(A, "A", torch.float32)
        _require_dev(C, "C", torch.float32)
        _vcheck(vendor_lib().gcoo_vendor_csr_spmm(self._h, _ptr(A), nRows, nCols, _ptr(C),
                                                  _stream()), "gcoo_vendor_csr_spmm")
        torch.cuda.current_stream().synchronize()
        return 0.0

    def mutiplyBy(self, A, nRows, nCols, C):
        self.build()
        return self.runKernelTest(A, nRows, nCols, C)

    def mutiply(self, B, nRows, nCols, C):
        return self.mutiplyBy(B, nRows, nCols, C)

    @property
    def nnz(self) -> int:
        n = ctypes.c_int64()
        _vcheck(vendor_lib().gcoo_vendor_csr_nnz(self._h, ctypes.byref(n)), "gcoo_vendor_csr_nnz")
        return n.value


class ShRocblasGemm(ShGemm):
    """Factory key "cublas": dense SGEMM of S and A."""

    def __init__(self, denseMatrix, nRows: int, nCols: int):
        super().__init__(nRows, nCols)
        self.m_pDenseMatrix = denseMatrix
        self._h = ctypes.c_void_p()
        _vcheck(vendor_lib().gcoo_vendor_blas_create(ctypes.byref(self._h)),
                "gcoo_vendor_blas_create")

    def __del__(self):
        if getattr(self, "_h", None) and _vlib is not None:
            _vlib.gcoo_vendor_blas_destroy(self._h)
            self._h = None

    def build(self):
        pass

    def mutiplyBy(self, A, nRows, nCols, C):
        import torch
        _require_dev(self.m_pDenseMatrix, "S", torch.float32)
        _require_dev(A, "A", torch.float32)
        _require_dev(C, "C", torch.float32)
        if nRows != self.m_nCols:
            raise ValueError("rows of A must equal columns of S")
        _vcheck(vendor_lib().gcoo_vendor_sgemm(self._h, _ptr(self.m_pDenseMatrix), _ptr(A),
                                               self.m_nRows, self.m_nCols, nCols, _ptr(C),
                                               _stream()), "gcoo_vendor_sgemm")
        torch.cuda.current_stream().synchronize()
        return 0.0

    def mutiply(self, B, nRows, nCols, C):
        return self.mutiplyBy(B, nRows, nCols, C)

    def runKernelTest(self, A, nRows, nCols, C):
        return self.mutiplyBy(A, nRows, nCols, C)
