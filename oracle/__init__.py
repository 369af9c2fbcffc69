"""CPU oracle for the GCOOSpDM hot path -- TEST INFRASTRUCTURE ONLY.

ctypes binding of liboracle.so (gcoo_oracle.c, a C restatement of the
reference's conversion and multiply; every function cites the reference
file:line it follows) and of oracle/_ref/libref_utils.so (the reference's own
src/utils.cpp compiled from its sources, when that was built).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, always as the checker or the CPU baseline, never as the product.

Parity status of each function (how it is pinned; tests/test_oracle.py):
  convert / count / scan   brute force (numpy sort by (group, k, r)) on tiny
                           inputs incl. -0.0 / NaN / denormals / ragged M,
                           committed golden fixtures, invariants.
  spdm_f32                 exact-rational fmaf emulation (fractions.Fraction)
                           on tiny inputs; S = I gives D; single nonzero.
  spdm_bf16                exact-rational emulation after the textbook RNE
                           bf16 rounding (itself checked against hand values).
  spdm_fp64 / dense_fp64   bit-exact vs the reference's matrixMulCPU
                           (oracle/_ref) and vs numpy float64 dot.
  rel_frobenius            closed form on hand-made vectors.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libref_utils.so")

_lib = None
_ref = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        i, i64, vp = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p
        L.orc_gcoo_count.restype = i64
        L.orc_gcoo_count.argtypes = [_f32p, i, i, i, _i32p]
        L.orc_exclusive_scan.restype = None
        L.orc_exclusive_scan.argtypes = [_i32p, _i32p, i]
        L.orc_gcoo_convert.restype = i64
        L.orc_gcoo_convert.argtypes = [_f32p, i, i, i, _f32p, _i32p, _i32p, _i32p, _i32p]
        L.orc_gcoo_spdm_f32.restype = None
        L.orc_gcoo_spdm_f32.argtypes = [_f32p, _i32p, _i32p, _i32p, _i32p, i, i, i,
                                        _f32p, i, _f32p, i]
        L.orc_gcoo_spdm_bf16.restype = None
        L.orc_gcoo_spdm_bf16.argtypes = [_u16p, _i32p, _i32p, _i32p, _i32p, i, i, i,
                                         _u16p, i, _f32p, i]
        L.orc_f32_to_bf16.restype = None
        L.orc_f32_to_bf16.argtypes = [_f32p, _u16p, i64]
        L.orc_spdm_fp64.restype = None
        L.orc_spdm_fp64.argtypes = [_f32p, _i32p, _i32p, _i32p, i, i, i, _f32p, i,
                                    _f64p, vp, i]
        L.orc_dense_fp64.restype = None
        L.orc_dense_fp64.argtypes = [_f32p, _f32p, i, i, i, _f64p, vp]
        L.orc_rel_frobenius.restype = ctypes.c_double
        L.orc_rel_frobenius.argtypes = [_f32p, _f64p, i64]
        L.orc_max_threads.restype = i
        L.orc_max_threads.argtypes = []
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref():
    """The reference's own utils.cpp (randomInit, matrixMulCPU)."""
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_LIB)
        R.ref_srand.restype = None
        R.ref_srand.argtypes = [ctypes.c_uint]
        R.ref_random_init.restype = ctypes.c_int
        R.ref_random_init.argtypes = [_f32p, ctypes.c_size_t, ctypes.c_float]
        R.ref_matrix_mul_cpu.restype = None
        R.ref_matrix_mul_cpu.argtypes = [_f32p, _f32p, _f32p, ctypes.c_uint,
                                         ctypes.c_uint, ctypes.c_uint]
        _ref = R
    return _ref


def max_threads() -> int:
    return lib().orc_max_threads()


def count(S, cpg=4):
    S = np.ascontiguousarray(S, dtype=np.float32)
    M, K = S.shape
    ng = (M + cpg - 1) // cpg
    npg = np.zeros(ng + 1, np.int32)
    tot = lib().orc_gcoo_count(S, M, K, cpg, npg)
    return npg, tot


def exclusive_scan(src):
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.zeros_like(src)
    lib().orc_exclusive_scan(src, dst, src.size)
    return dst


def convert(S, cpg=4):
    """Dense -> GCOO.  Returns dict of numpy arrays (reference layout)."""
    S = np.ascontiguousarray(S, dtype=np.float32)
    M, K = S.shape
    ng = (M + cpg - 1) // cpg
    npg, nnz = count(S, cpg)
    vals = np.zeros(max(nnz, 1), np.float32)
    rows = np.zeros(max(nnz, 1), np.int32)
    cols = np.zeros(max(nnz, 1), np.int32)
    gi = np.zeros(ng + 1, np.int32)
    npg2 = np.zeros(ng + 1, np.int32)
    n2 = lib().orc_gcoo_convert(S, M, K, cpg, vals, rows, cols, gi, npg2)
    assert n2 == nnz
    return dict(vals=vals[:nnz], rows=rows[:nnz], cols=cols[:nnz], groupIndex=gi,
                nnzPerGroup=npg2, nGroup=ng, nnz=int(nnz), M=M, K=K, cpg=cpg)


def _pad(a, dt):
    a = np.ascontiguousarray(a, dtype=dt)
    return a if a.size else np.zeros(1, dt)


def spdm_f32(g, D, threads=1):
    """C = S * D with the reference's per-output fmaf chain (bit-exact)."""
    D = np.ascontiguousarray(D, dtype=np.float32)
    K, N = D.shape
    C = np.zeros((g["M"], N), np.float32)
    if C.size == 0:
        return C
    lib().orc_gcoo_spdm_f32(_pad(g["vals"], np.float32), _pad(g["rows"], np.int32),
                            _pad(g["cols"], np.int32), g["groupIndex"], g["nnzPerGroup"],
                            g["nGroup"], g["cpg"], g["M"], _pad(D, np.float32), N, C,
                            threads)
    return C


def to_bf16(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros(x.shape, np.uint16)
    if x.size:
        lib().orc_f32_to_bf16(x.reshape(-1), y.reshape(-1), x.size)
    return y


def bf16_to_f32(y):
    return (np.asarray(y, np.uint16).astype(np.uint32) << 16).view(np.float32)


def spdm_bf16(g, vals_bf16, D_bf16, threads=1):
    D_bf16 = np.ascontiguousarray(D_bf16, dtype=np.uint16)
    K, N = D_bf16.shape
    C = np.zeros((g["M"], N), np.float32)
    if C.size == 0:
        return C
    lib().orc_gcoo_spdm_bf16(_pad(vals_bf16, np.uint16), _pad(g["rows"], np.int32),
                             _pad(g["cols"], np.int32), g["groupIndex"], g["nnzPerGroup"],
                             g["nGroup"], g["cpg"], g["M"], _pad(D_bf16, np.uint16), N,
                             C, threads)
    return C


def spdm_fp64(g, D, threads=1):
    """FP64 reference: returns (Cd float64, C float32 = (float)Cd)."""
    D = np.ascontiguousarray(D, dtype=np.float32)
    K, N = D.shape
    Cd = np.zeros((g["M"], N), np.float64)
    C = np.zeros((g["M"], N), np.float32)
    if Cd.size:
        lib().orc_spdm_fp64(_pad(g["vals"], np.float32), _pad(g["rows"], np.int32),
                            _pad(g["cols"], np.int32), g["groupIndex"], g["nGroup"],
                            g["cpg"], g["M"], _pad(D, np.float32), N, Cd,
                            C.ctypes.data, threads)
    return Cd, C


def dense_fp64(S, D):
    S = np.ascontiguousarray(S, dtype=np.float32)
    D = np.ascontiguousarray(D, dtype=np.float32)
    M, K = S.shape
    _, N = D.shape
    Cd = np.zeros((M, N), np.float64)
    C = np.zeros((M, N), np.float32)
    if Cd.size:
        lib().orc_dense_fp64(_pad(S, np.float32), _pad(D, np.float32), M, K, N, Cd,
                             C.ctypes.data)
    return Cd, C


def rel_frobenius(C, R):
    C = np.ascontiguousarray(C, dtype=np.float32).reshape(-1)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1)
    return lib().orc_rel_frobenius(_pad(C, np.float32), _pad(R, np.float64), C.size)
