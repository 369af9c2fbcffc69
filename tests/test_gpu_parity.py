"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars: GCOO arrays bit-exact; FP32 product bit-exact with the reference's
fmaf chain (NaN payload included); relative Frobenius error vs FP64 <= 1e-5
(the reference's sdkCompareL2fe tolerance, src/main.cu:136)."""
import os
import subprocess

import numpy as np
import pytest

import gcoospdm_amd as G
import oracle
from tests import helpers as H

pytestmark = pytest.mark.gpu

FP32_REL_TOL = 1e-5


def to_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def check_arrays(m, g_ref):
    got = m.arrays_host()
    assert m.nnz == g_ref["nnz"]
    for key in ("rows", "cols", "groupIndex", "nnzPerGroup"):
        np.testing.assert_array_equal(got[key], g_ref[key], err_msg=key)
    assert H.bits_equal(got["vals"], g_ref["vals"])


EDGE_SHAPES = [
    (13, 7, 5, 0.5),      # M % 4 != 0, K % 4 != 0 (scalar loads)
    (41, 8, 3, 0.05),     # empty groups, N odd
    (9, 16, 64, 1.0),     # dense
    (8, 12, 6, 0.0),      # all zero
    (1, 17, 1, 0.4),
    (64, 1000, 300, 0.02),
    (257, 515, 129, 0.1),
    (8, 4000, 16, 0.5),   # rows longer than the single-pass conversion's LDS list
    (6, 0, 5, 0.5),       # K = 0: no entries, C all +0.0
    (0, 8, 3, 0.5),       # M = 0: no groups
    (5, 8, 0, 0.5),       # N = 0: nothing to write
]


@pytest.mark.parametrize("cpg", [1, 2, 4])
@pytest.mark.parametrize("M,K,N,dens", EDGE_SHAPES)
def test_convert_and_spdm_bit_exact(dev, M, K, N, dens, cpg):
    import torch
    print(f"case M={M} K={K} N={N} dens={dens} cpg={cpg}", flush=True)
    rng = np.random.default_rng(M * 7 + K * 3 + N + cpg)
    S = H.rng_sparse(rng, M, K, dens)
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    g = oracle.convert(S, cpg)
    m = G.GCOOMatrix(M, K, cpg).convert_dense(to_dev(S, dev))
    check_arrays(m, g)
    C = torch.full((M, N), float("nan"), device=dev)
    m.spdm(to_dev(D, dev), C)
    assert H.bits_equal(C.cpu().numpy(), oracle.spdm_f32(g, D))


def test_edge_values(dev):
    import torch
    S = np.zeros((7, 9), np.float32)
    S[0, 1] = -0.0
    S[0, 2] = 1.5
    S[1, 0] = np.nan
    S[2, 8] = np.float32(1e-40)
    S[3, 3] = -np.float32(1e-45)
    S[4, 4] = np.inf
    S[6, 8] = 3.0
    D = np.random.default_rng(0).uniform(-1, 1, (9, 5)).astype(np.float32)
    D[8, 0] = np.float32(1e-39)  # denormal * normal products
    g = oracle.convert(S, 4)
    m = G.GCOOMatrix(7, 9, 4).convert_dense(to_dev(S, dev))
    check_arrays(m, g)
    C = torch.zeros((7, 5), device=dev)
    m.spdm(to_dev(D, dev), C)
    assert H.bits_equal(C.cpu().numpy(), oracle.spdm_f32(g, D))


def test_unaligned_operands(dev):
    """Sub-tensor views that are not 16-byte aligned take the scalar paths."""
    import torch
    rng = np.random.default_rng(11)
    M, K, N = 33, 20, 12
    S = H.rng_sparse(rng, M, K, 0.2)
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    Sbuf = torch.zeros(M * K + 1, device=dev)
    Sbuf[1:] = to_dev(S.reshape(-1), dev)
    Sv = Sbuf[1:].view(M, K)
    Dbuf = torch.zeros(K * N + 1, device=dev)
    Dbuf[1:] = to_dev(D.reshape(-1), dev)
    Dv = Dbuf[1:].view(K, N)
    Cbuf = torch.zeros(M * N + 1, device=dev)
    Cv = Cbuf[1:].view(M, N)
    g = oracle.convert(S, 4)
    m = G.GCOOMatrix(M, K, 4).convert_dense(Sv)
    check_arrays(m, g)
    m.spdm(Dv, Cv)
    assert H.bits_equal(Cv.cpu().numpy(), oracle.spdm_f32(g, D))


def test_config1_800_s09_vs_fp64(dev):
    """Plain method, small: 800 x 800, p(nonzero) = 0.1, U[-1,1], fixed seed;
    checked bit-exact vs the fmaf chain and against FP64 (rel <= 1e-5)."""
    import torch
    n = 800
    S = H.counter_fill(n * n, 1, 1, 1.0, G.keep_threshold(0.9), True).reshape(n, n)
    D = H.counter_fill(n * n, 1, 0, 1.0).reshape(n, n)
    Sd = torch.empty((n, n), device=dev)
    Dd = torch.empty((n, n), device=dev)
    G.fill_sparse(Sd, 1, 1, 0.9)
    G.fill_uniform(Dd, 1, 0)
    # the device generator is bit-identical to its host definition
    assert H.bits_equal(Sd.cpu().numpy(), S) and H.bits_equal(Dd.cpu().numpy(), D)
    algo = G.SpGemmBuilder.getShGemmAlgo(Sd, n, n, "gcoospdm")
    C = torch.empty((n, n), device=dev)
    algo.mutiplyBy(Dd, n, n, C)
    g = oracle.convert(S, 4)
    check_arrays(algo.matrix, g)
    Cn = C.cpu().numpy()
    assert H.bits_equal(Cn, oracle.spdm_f32(g, D, threads=8))
    Cd, _ = oracle.spdm_fp64(g, D, threads=8)
    err = oracle.rel_frobenius(Cn, Cd)
    print(f"config1 rel_frobenius={err:.3e}")
    assert err <= FP32_REL_TOL


@pytest.mark.parametrize("variant,waves,rows", [(0, 0, 0), (6, 16, 16), (6, 16, 21), (6, 12, 0),
                                               (6, 8, 0)])
@pytest.mark.parametrize("sparsity", [0.99, 0.995, 0.999])
def test_config2_n8000_full_size(dev, sparsity, variant, waves, rows):
    """n = 8000: conversion bit-exact at full size, and all of C bit-exact
    against the oracle's fmaf chain plus FP64 rel-Frobenius <= 1e-5 over all
    of C (the check of src/main.cu:128-142 at this size)."""
    import torch
    n = 8000
    print(f"config2 n={n} s={sparsity} variant={variant} waves={waves} rows={rows}", flush=True)
    Sd = torch.empty((n, n), device=dev)
    Dd = torch.empty((n, n), device=dev)
    G.fill_sparse(Sd, 7, 1, sparsity)
    G.fill_uniform(Dd, 7, 0)
    C = torch.empty((n, n), device=dev)
    m = G.GCOOMatrix(n, n, 4).set_config(variant=variant, waves=waves, chunk=rows)
    m.multiply_by(Sd, Dd, C)
    S = Sd.cpu().numpy()
    D = Dd.cpu().numpy()
    g = oracle.convert(S, 4)
    check_arrays(m, g)
    Cn = C.cpu().numpy()
    Cref = oracle.spdm_f32(g, D, threads=oracle.max_threads())
    assert H.bits_equal(Cn, Cref)
    Cd, _ = oracle.spdm_fp64(g, D, threads=oracle.max_threads())
    err = oracle.rel_frobenius(Cn, Cd)
    print(f"config2 s={sparsity} rel_frobenius={err:.3e}", flush=True)
    assert err <= FP32_REL_TOL


@pytest.mark.parametrize("variant", [0, 6])
def test_repeat_and_capacity_growth(dev, variant):
    """mutiplyBy on a handle re-used with denser inputs (capacity grows) and
    then sparser ones (capacity kept) matches a fresh conversion each time."""
    import torch
    M, K, N = 200, 300, 96
    print(f"capacity growth variant={variant}", flush=True)
    m = G.GCOOMatrix(M, K, 4).set_config(variant=variant)
    rng = np.random.default_rng(3)
    for dens in (0.01, 0.2, 0.5, 0.05, 0.5):
        S = H.rng_sparse(rng, M, K, dens)
        D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        C = torch.empty((M, N), device=dev)
        m.multiply_by(to_dev(S, dev), to_dev(D, dev), C)
        g = oracle.convert(S, 4)
        check_arrays(m, g)
        assert H.bits_equal(C.cpu().numpy(), oracle.spdm_f32(g, D))


def test_reference_api_mirror(dev):
    """build() + runKernelTest == mutiplyBy == mutiply (reference
    src/main.cu:53-75 call sequence)."""
    import torch
    M, K, N = 96, 64, 80
    rng = np.random.default_rng(9)
    S = to_dev(H.rng_sparse(rng, M, K, 0.1), dev)
    D = to_dev(rng.uniform(-1, 1, (K, N)).astype(np.float32), dev)
    algo = G.SpGemmBuilder.getShGemmAlgo(S, M, K, "gcoospdm")
    C1, C2, C3 = (torch.empty((M, N), device=dev) for _ in range(3))
    algo.build()
    algo.runKernelTest(D, K, N, C1)
    algo.mutiplyBy(D, K, N, C2)
    algo.mutiply(D, K, N, C3)
    assert torch.equal(C1.view(torch.int32), C2.view(torch.int32))
    assert torch.equal(C1.view(torch.int32), C3.view(torch.int32))
    with pytest.raises(G.GCOOError):
        algo.runKernelTest(D[:-1], K - 1, N, C1)   # K mismatch -> error, no crash


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("vec", [1, 2, 4])
@pytest.mark.parametrize("M,K,N,dens", [(130, 257, 520, 0.05), (61, 1030, 200, 0.3),
                                        (1000, 900, 64, 0.01)])
def test_kernel_variants_bit_exact(dev, variant, vec, M, K, N, dens):
    """Every multiply variant / tile width gives the reference's bits."""
    import torch
    print(f"variant={variant} vec={vec} M={M} K={K} N={N}", flush=True)
    rng = np.random.default_rng(M + K + N)
    S = H.rng_sparse(rng, M, K, dens)
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    g = oracle.convert(S, 4)
    m = G.GCOOMatrix(M, K, 4).convert_dense(to_dev(S, dev))
    m.set_config(variant=variant, vec=vec)
    C = torch.full((M, N), float("nan"), device=dev)
    m.spdm(to_dev(D, dev), C)
    assert H.bits_equal(C.cpu().numpy(), oracle.spdm_f32(g, D))


@pytest.mark.parametrize("rows_per_wave", [1, 4])
@pytest.mark.parametrize("unroll", [8, 16, 32])
@pytest.mark.parametrize("M,K,N,dens", [(37, 200, 70, 0.5), (130, 3000, 130, 0.02),
                                        (9, 5000, 64, 0.3)])
def test_row_segment_rounds_bit_exact(dev, rows_per_wave, unroll, M, K, N, dens):
    """Variant 2 with one row or a 4-row group per wave and 8, 16 or 32
    entries per round (then rounds of 8, then the guarded last 1..7), FP32
    and BF16: rows of every length modulo the round sizes."""
    import torch
    print(f"rowseg rpw={rows_per_wave} unroll={unroll} M={M} K={K} N={N}", flush=True)
    rng = np.random.default_rng(M * 3 + K + N + unroll)
    S = H.rng_sparse(rng, M, K, dens)
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    g = oracle.convert(S, 4)
    m = G.GCOOMatrix(M, K, 4).convert_dense(to_dev(S, dev))
    m.set_config(variant=2, chunk=rows_per_wave, unroll=unroll)
    C = torch.full((M, N), float("nan"), device=dev)
    m.spdm(to_dev(D, dev), C)
    assert H.bits_equal(C.cpu().numpy(), oracle.spdm_f32(g, D))
    D16 = G.to_bf16(to_dev(D, dev))
    C.fill_(float("nan"))
    m.spdm_bf16(D16, C)
    ref16 = oracle.spdm_bf16(g, oracle.to_bf16(g["vals"]), oracle.to_bf16(D))
    assert H.bits_equal(C.cpu().numpy(), ref16)


@pytest.mark.parametrize("vec", [1, 2, 4, 8])
def test_row_segment_tails_every_width(dev, vec):
    """Variant 2, one row per wave, at every column width: the guarded last
    1..7 entries of a row load 2, 4, 8 or 16 bytes per lane (BF16) and are
    widened next to their FMAs; rows of 0..40 entries cover every tail
    length after rounds of 16 and of 8.  FP32 runs up to 4 columns per lane."""
    import torch
    M, K, N = 41, 300, 64
    print(f"rowseg tails vec={vec}", flush=True)
    rng = np.random.default_rng(4100 + vec)
    S = np.zeros((M, K), np.float32)
    for r in range(M):
        cols = rng.choice(K, size=r, replace=False)
        S[r, cols] = rng.uniform(-1, 1, r).astype(np.float32)
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    g = oracle.convert(S, 4)
    m = G.GCOOMatrix(M, K, 4).convert_dense(to_dev(S, dev))
    m.set_config(variant=2, chunk=1, unroll=16, vec=vec)
    Dd = to_dev(D, dev)
    C = torch.full((M, N), float("nan"), device=dev)
    D16 = G.to_bf16(Dd)
    ch = m.spdm_choice(D16, C)
    assert ch[:2] == (2, vec), ch
    m.spdm_bf16(D16, C)
    ref16 = oracle.spdm_bf16(g, oracle.to_bf16(g["vals"]), oracle.to_bf16(D))
    assert H.bits_equal(C.cpu().numpy(), ref16)
    if vec <= 4:
        C.fill_(float("nan"))
        ch = m.spdm_choice(Dd, C)
        assert ch[:2] == (2, vec), ch
        m.spdm(Dd, C)
        assert H.bits_equal(C.cpu().numpy(), oracle.spdm_f32(g, D))


@pytest.mark.parametrize("M,K,N,dens", [
    (300, 1000, 512, 0.02),    # 2 panels (one partial), 16 chunks
    (257, 129, 260, 0.3),      # dense chunks: the 8-entry batches and every tail size
    (513, 70, 1024, 0.0),      # no entries
    (5, 3, 8, 0.5),            # one partial panel, one partial chunk, one partial tile
    (1000, 5000, 256, 0.005),  # most (row, chunk) lists empty
    (256, 10000, 512, 0.01),   # 134 chunks: three 64-chunk windows
    (300, 80000, 256, 0.001),  # 1067 chunks: list counts beyond LDS
    (64, 2000, 1030, 0.05),    # N * 4 not a multiple of 16: row-segment fallback
    (300, 301, 256, 0.05),     # odd K: BF16's last staged row pair straddles K
])
@pytest.mark.parametrize("dtype,waves,rows", [("f32", 16, 0), ("f32", 16, 21), ("f32", 12, 0),
                                              ("f32", 8, 0), ("bf16", 16, 0), ("bf16", 16, 21),
                                              ("bf16", 12, 0)])
def test_panel_variant_shapes(dev, M, K, N, dens, dtype, waves, rows):
    """The panel-staged multiply (variant 6; workgroups of 16 waves x 16
    rows, 16 x 21, 12 x 28 or 8 x 42; FP32 D in 75-row chunks, BF16 D in
    150-row chunks) across panel, chunk and tile edges gives the reference's
    bits (BF16: the chain over BF16-widened operands), and follows
    re-conversions of the handle (its panel lists are rebuilt)."""
    import torch
    print(f"panel {dtype} waves={waves} rows={rows} M={M} K={K} N={N} dens={dens}", flush=True)
    rng = np.random.default_rng(M * 7 + K + N)
    m = G.GCOOMatrix(M, K, 4).set_config(variant=6, waves=waves, chunk=rows)
    for d in (dens, dens / 2):
        S = H.rng_sparse(rng, M, K, d)
        S[M // 2, : min(K, 700)] = rng.uniform(-1, 1, min(K, 700))  # a long row
        D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        g = oracle.convert(S, 4)
        m.convert_dense(to_dev(S, dev))
        C = torch.full((M, N), float("nan"), device=dev)
        if dtype == "f32":
            m.spdm(to_dev(D, dev), C)
            want = oracle.spdm_f32(g, D)
        else:
            m.spdm_bf16(G.to_bf16(to_dev(D, dev)), C)
            want = oracle.spdm_bf16(g, oracle.to_bf16(g["vals"]), oracle.to_bf16(D))
        assert H.bits_equal(C.cpu().numpy(), want)


def test_row_segments_match_gcoo(dev):
    """The per-row segments the conversion emits hold exactly the GCOO
    entries of each row, k ascending (checked through the product of
    S = permutation-like sparse matrices)."""
    import torch
    M, K = 50, 40
    S = np.zeros((M, K), np.float32)
    rng = np.random.default_rng(12)
    for r in range(M):
        ks = rng.choice(K, size=rng.integers(0, 6), replace=False)
        S[r, ks] = rng.uniform(-1, 1, ks.size)
    D = np.eye(K, dtype=np.float32)   # C = S exactly
    m = G.GCOOMatrix(M, K, 4).convert_dense(to_dev(S, dev))
    m.set_config(variant=2)
    C = torch.empty((M, K), device=dev)
    m.spdm(to_dev(D, dev), C)
    assert H.bits_equal(C.cpu().numpy(), oracle.spdm_f32(oracle.convert(S, 4), D))


def _coo_of(S, rng, extra_zeros=0, row_sorted=False):
    """COO arrays of S in random order, or (row_sorted) sorted by row with
    the columns of a row in random order."""
    r, k = np.nonzero(S != 0.0)
    v = S[r, k]
    if extra_zeros:   # explicit zeros in the COO input are dropped
        zr = rng.integers(0, S.shape[0], extra_zeros)
        zk = rng.integers(0, S.shape[1], extra_zeros)
        keep = S[zr, zk] == 0.0
        zr, zk = zr[keep], zk[keep]
        key = np.unique(zr.astype(np.int64) * S.shape[1] + zk)
        zr, zk = (key // S.shape[1]).astype(np.int64), (key % S.shape[1]).astype(np.int64)
        r = np.concatenate([r, zr])
        k = np.concatenate([k, zk])
        v = np.concatenate([v, np.zeros(zr.size, np.float32)])
    perm = rng.permutation(r.size)
    if row_sorted:
        perm = perm[np.argsort(r[perm], kind="stable")]
    return (r[perm].astype(np.int32), k[perm].astype(np.int32), v[perm].astype(np.float32))


@pytest.mark.parametrize("row_sorted", [False, True])
@pytest.mark.parametrize("cpg", [1, 2, 4])
@pytest.mark.parametrize("M,K,dens", [(37, 50, 0.1), (200, 3000, 0.05), (9, 20000, 0.5),
                                      (5, 40000, 0.5), (64, 64, 0.0)])
def test_coo_conversion_matches_dense(dev, M, K, dens, cpg, row_sorted):
    """COO input (with explicit zeros), shuffled or sorted by row, gives the
    dense path's arrays bit for bit; rows of > 64, > 8192 and > 16384
    entries take the LDS sort (beyond 64 KB of LDS from 8192) and the
    global-memory sort."""
    import torch
    print(f"coo M={M} K={K} dens={dens} cpg={cpg} row_sorted={row_sorted}", flush=True)
    rng = np.random.default_rng(M * 31 + K + cpg)
    S = H.rng_sparse(rng, M, K, dens)
    r, k, v = _coo_of(S, rng, extra_zeros=17, row_sorted=row_sorted)
    g = oracle.convert(S, cpg)
    m = G.GCOOMatrix(M, K, cpg).convert_coo(to_dev(r, dev), to_dev(k, dev), to_dev(v, dev))
    check_arrays(m, g)
    N = 70
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    C = torch.empty((M, N), device=dev)
    m.spdm(to_dev(D, dev), C)
    assert H.bits_equal(C.cpu().numpy(), oracle.spdm_f32(g, D))


def test_coo_rejects_bad_input(dev):
    r = np.array([0, 1, 1], np.int32)
    k = np.array([2, 3, 3], np.int32)
    v = np.array([1.0, 2.0, 3.0], np.float32)
    m = G.GCOOMatrix(4, 5, 4)
    with pytest.raises(G.GCOOError) as e:
        m.convert_coo(to_dev(r, dev), to_dev(k, dev), to_dev(v, dev))
    assert e.value.code == G.GCOO_ERR_INVALID
    k2 = np.array([2, 3, 5], np.int32)   # column 5 outside K = 5
    with pytest.raises(G.GCOOError) as e:
        m.convert_coo(to_dev(r, dev), to_dev(k2, dev), to_dev(v, dev))
    assert e.value.code == G.GCOO_ERR_INVALID
    # the handle is usable after a rejected input
    r3 = np.array([0, 3], np.int32)
    k3 = np.array([4, 0], np.int32)
    v3 = np.array([1.5, -2.0], np.float32)
    m.convert_coo(to_dev(r3, dev), to_dev(k3, dev), to_dev(v3, dev))
    assert m.nnz == 2


MAIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "gcoospdm_amd", "host", "main")


@pytest.mark.parametrize("kernelonly", [0, 1])
def test_main_driver_check(dev, kernelonly):
    """`main -algo=gcoospdm` on the config-1 size with its FP64 check."""
    r = subprocess.run([MAIN, "-algo=gcoospdm", "-nRowsA=800", "-nColsA=800",
                        "-nRowsB=800", "-nColsB=800", "-sparsity=0.9", "-dist=uniform",
                        "-check=1", "-iters=5", f"-kernelonly={kernelonly}"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    assert "Test passed" in r.stdout and "Time:" in r.stdout


def test_main_driver_reference_inputs(dev):
    """The reference's own generator (glibc rand, values in [0,1])."""
    r = subprocess.run([MAIN, "-algo=gcoospdm", "-nRowsA=512", "-nColsA=512",
                        "-nRowsB=512", "-nColsB=512", "-sparsity=0.95", "-check=1",
                        "-iters=3"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    assert "Test passed" in r.stdout


@pytest.mark.parametrize("M,K,N,dens,variant", [(257, 516, 96, 0.1, 0), (2000, 2048, 256, 0.01, 0),
                                                (2000, 2048, 256, 0.01, 6)])
def test_conversion_stall_fallback(dev, M, K, N, dens, variant):
    """The single-pass conversion's abort word is raised from the host before
    the launch (as a group whose bounded look-back wait gave up would raise
    it): every kernel still exits normally, the multiply queued behind the
    conversion skips, and the library converts again with the count / scan /
    scatter kernels.  Arrays and C must still equal the oracle's, and the
    next conversion takes the single-pass kernel again.  With the panel
    kernel (variant 6) the status word reaches the host through its
    panel-list kernel."""
    import torch
    print(f"stall fallback M={M} K={K} N={N} dens={dens} variant={variant}", flush=True)
    rng = np.random.default_rng(M + K)
    m = G.GCOOMatrix(M, K, 4).set_config(variant=variant)
    for step in range(3):
        S = H.rng_sparse(rng, M, K, dens * (1 + step))
        D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        Sd, Dd = to_dev(S, dev), to_dev(D, dev)
        C = torch.full((M, N), float("nan"), device=dev)
        if step < 2:
            rc = G.lib().gcoo_test_set_conversion_abort(m.handle, G._stream())
            assert rc == 0
        if step == 1:
            m.convert_dense(Sd)
            m.spdm(Dd, C)
        else:
            m.multiply_by(Sd, Dd, C)
        g = oracle.convert(S, 4)
        check_arrays(m, g)
        assert H.bits_equal(C.cpu().numpy(), oracle.spdm_f32(g, D)), step


@pytest.mark.parametrize("M,K,N,s", [(8000, 8000, 512, 0.99), (700, 900, 260, 0.97),
                                      (50, 1500, 256, 0.98), (1000, 75, 256, 0.9),
                                      (400, 12000, 256, 0.99)])
def test_panel_balanced_order_same_bits(dev, M, K, N, s):
    """The panel kernel's lists with each panel's rows dealt to the waves in
    the balanced order (gcoo_set_panel_balance 2) give the same bits as in
    order, for FP32 and BF16: ragged and partial panels, one chunk, and K past
    the balanced order's 128-chunk limit (which keeps the plain order)."""
    import torch
    print(f"balanced order M={M} K={K} N={N} s={s}", flush=True)
    S = torch.empty((M, K), device=dev)
    D = torch.empty((K, N), device=dev)
    G.fill_sparse(S, 31, 1, s)
    G.fill_uniform(D, 31, 0)
    m = G.GCOOMatrix(M, K, 4).convert_dense(S)
    m.set_config(variant=6, waves=16, chunk=21)
    D16 = G.to_bf16(D)
    out = {}
    for mode in (0, 2):
        m.set_panel_balance(mode)
        C = torch.full((M, N), float("nan"), device=dev)
        m.spdm(D, C)
        C16 = torch.full((M, N), float("nan"), device=dev)
        m.spdm_bf16(D16, C16)
        out[mode] = (C, C16)
    assert torch.equal(out[0][0].view(torch.int32), out[2][0].view(torch.int32))
    assert torch.equal(out[0][1].view(torch.int32), out[2][1].view(torch.int32))
    if M * K <= 1_000_000:
        Sh, Dh = S.cpu().numpy(), D.cpu().numpy()
        assert H.bits_equal(out[2][0].cpu().numpy(), oracle.spdm_f32(oracle.convert(Sh, 4), Dh))
