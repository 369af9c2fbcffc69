"""GPU parity on skewed matrices built from COO: the 65536 x 65536 matrix
with Zipf(1.5) row lengths (mean 16), N = 256; and small skewed matrices
that force the long-row kernel with odd shapes.  The reference GCOO arrays
are built on the host by sorting the COO entries by (group, k, r); the
product is checked bit for bit against the oracle's fmaf chain run on those
arrays, and against FP64."""
import numpy as np
import pytest

import gcoospdm_amd as G
import oracle
from gcoospdm_amd import workloads as W
from tests import helpers as H

pytestmark = pytest.mark.gpu


def host_gcoo(rows, cols, vals, M, cpg=4):
    keep = vals != 0.0
    r, k, v = rows[keep].astype(np.int64), cols[keep].astype(np.int64), vals[keep]
    g = r // cpg
    order = np.lexsort((r, k, g))
    r, k, v, g = r[order], k[order], v[order], g[order]
    ng = (M + cpg - 1) // cpg
    npg = np.zeros(ng + 1, np.int32)
    np.add.at(npg, g, 1)
    gi = np.zeros(ng + 1, np.int32)
    gi[1:] = np.cumsum(npg[:-1])
    return dict(vals=v.astype(np.float32), rows=r.astype(np.int32), cols=k.astype(np.int32),
                groupIndex=gi, nnzPerGroup=npg, nGroup=ng, nnz=int(r.size), M=M, cpg=cpg)


def to_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def check_arrays(m, g):
    got = m.arrays_host()
    assert m.nnz == g["nnz"]
    for key in ("rows", "cols", "groupIndex", "nnzPerGroup"):
        np.testing.assert_array_equal(got[key], g[key], err_msg=key)
    assert H.bits_equal(got["vals"], g["vals"])


def test_zipf_65536_n256(dev):
    import torch
    M = K = 65536
    N = 256
    rows, cols, vals = W.zipf_coo(M, K, seed=11)
    g = host_gcoo(rows, cols, vals, M)
    m = G.GCOOMatrix(M, K, 4).convert_coo(to_dev(rows, dev), to_dev(cols, dev), to_dev(vals, dev))
    check_arrays(m, g)
    rng = np.random.default_rng(12)
    D = rng.standard_normal((K, N)).astype(np.float32)
    Dd = to_dev(D, dev)
    ref = oracle.spdm_f32(g, D, threads=oracle.max_threads())
    for variant in (0, 2, 4):
        print(f"zipf variant={variant}", flush=True)
        m.set_config(variant=variant)
        C = torch.full((M, N), float("nan"), device=dev)
        m.spdm(Dd, C)
        assert H.bits_equal(C.cpu().numpy(), ref), variant
    Cd, _ = oracle.spdm_fp64(g, D, threads=oracle.max_threads())
    assert oracle.rel_frobenius(ref, Cd) <= 1e-5
    # BF16 form on the same matrix
    C = torch.empty((M, N), device=dev)
    m.set_config()
    m.spdm_bf16(G.to_bf16(Dd), C)
    refb = oracle.spdm_bf16(g, oracle.to_bf16(g["vals"]), oracle.to_bf16(D),
                            threads=oracle.max_threads())
    assert H.bits_equal(C.cpu().numpy(), refb)
    assert oracle.rel_frobenius(C.cpu().numpy(), Cd) <= 1e-2


@pytest.mark.parametrize("bf16", [False, True])
@pytest.mark.parametrize("N", [8, 64, 256, 264, 520])
def test_long_row_kernel_shapes(dev, bf16, N):
    """Rows of 0 .. 300 entries with threshold 16: the long-row kernel takes
    a large share of the rows, with partial chunks and partial column tiles."""
    import torch
    M, K = 97, 700
    rng = np.random.default_rng(N + (7 if bf16 else 0))
    lens = rng.integers(0, 300, M)
    lens[::7] = 0
    S = np.zeros((M, K), np.float32)
    for r in range(M):
        ks = rng.choice(K, size=lens[r], replace=False)
        S[r, ks] = rng.uniform(-1, 1, lens[r])
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    g = oracle.convert(S, 4)
    m = G.GCOOMatrix(M, K, 4).convert_dense(to_dev(S, dev))
    m.set_config(variant=4, chunk=16)
    C = torch.full((M, N), float("nan"), device=dev)
    if bf16:
        m.spdm_bf16(G.to_bf16(to_dev(D, dev)), C)
        ref = oracle.spdm_bf16(g, oracle.to_bf16(g["vals"]), oracle.to_bf16(D))
    else:
        m.spdm(to_dev(D, dev), C)
        ref = oracle.spdm_f32(g, D)
    assert H.bits_equal(C.cpu().numpy(), ref)


@pytest.mark.parametrize("bf16", [False, True])
@pytest.mark.parametrize("N", [64, 256, 520])
def test_long_rows_first_order(dev, bf16, N):
    """Variant 2 with its longest rows walked first (rows with >= 2^waves
    entries listed on the device and taken by the first workgroups of each
    column tile), against the natural order (waves = 1) and the oracle:
    thresholds that list most rows, a few, or none; one row per wave and the
    group's rows per wave; column tiles that divide the 8 XCDs (the
    interleaved placement) and that do not."""
    import torch
    M, K = 301, 900
    rng = np.random.default_rng(N + (5 if bf16 else 0))
    lens = np.minimum(rng.zipf(1.6, M), 600)
    lens[::11] = 0
    S = np.zeros((M, K), np.float32)
    for r in range(M):
        ks = rng.choice(K, size=lens[r], replace=False)
        S[r, ks] = rng.uniform(-1, 1, lens[r])
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    g = oracle.convert(S, 4)
    m = G.GCOOMatrix(M, K, 4).convert_dense(to_dev(S, dev))
    Dd = G.to_bf16(to_dev(D, dev)) if bf16 else to_dev(D, dev)
    ref = (oracle.spdm_bf16(g, oracle.to_bf16(g["vals"]), oracle.to_bf16(D)) if bf16
           else oracle.spdm_f32(g, D))
    for vec in ((1, 2, 4, 8) if bf16 else (1, 2, 4)):
        for waves in (1, 2, 4, 6, 9, 0):
            for chunk in (1, 2):
                print(f"long-first bf16={bf16} N={N} vec={vec} waves={waves} chunk={chunk}",
                      flush=True)
                m.set_config(variant=2, vec=vec, waves=waves, chunk=chunk)
                C = torch.full((M, N), float("nan"), device=dev)
                (m.spdm_bf16 if bf16 else m.spdm)(Dd, C)
                assert H.bits_equal(C.cpu().numpy(), ref), (vec, waves, chunk)


def test_multiply_by_with_long_rows(dev):
    """mutiplyBy (conversion + multiply enqueued together) on a skewed dense
    matrix, twice (the second call reuses capacity)."""
    import torch
    M, K, N = 256, 4096, 256
    rng = np.random.default_rng(3)
    m = G.GCOOMatrix(M, K, 4)
    for rep in range(2):
        S = H.rng_sparse(rng, M, K, 0.002)
        S[5] = rng.uniform(-1, 1, K).astype(np.float32)   # one dense row
        D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        C = torch.empty((M, N), device=dev)
        m.multiply_by(to_dev(S, dev), to_dev(D, dev), C)
        assert H.bits_equal(C.cpu().numpy(), oracle.spdm_f32(oracle.convert(S, 4), D))
