"""GPU parity of the BF16 storage form.

Bars: the device RNE conversion is bit-exact with the oracle's; the BF16
product is bit-exact with the oracle's BF16 fmaf chain (same order, operands
widened exactly); its relative Frobenius error against the FP64 product of
the original FP32 operands is <= 1e-2 (the stated BF16 tolerance)."""
import numpy as np
import pytest

import gcoospdm_amd as G
import oracle
from tests import helpers as H

pytestmark = pytest.mark.gpu

BF16_REL_TOL = 1e-2


def to_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def test_bf16_rounding_matches_oracle(dev):
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-4, 4, 10000).astype(np.float32),
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -0.0, np.inf, -np.inf,
                                  1e-40, 3.4e38, np.nan], np.float32)])
    y = G.to_bf16(to_dev(x, dev)).view(torch.int16).cpu().numpy().view(np.uint16)
    ref = oracle.to_bf16(x)
    finite = ~np.isnan(x)
    np.testing.assert_array_equal(y[finite], ref[finite])
    assert (y[~finite] & 0x7F80 == 0x7F80).all() and (y[~finite] & 0x7F != 0).all()


@pytest.mark.parametrize("variant", [2, 3, 5, 6])
@pytest.mark.parametrize("vec", [1, 2, 4, 8])
@pytest.mark.parametrize("M,K,N,dens", [(130, 257, 512, 0.05), (61, 1030, 96, 0.3),
                                        (300, 77, 1030, 0.2)])
def test_bf16_bit_exact(dev, variant, vec, M, K, N, dens):
    import torch
    print(f"bf16 variant={variant} vec={vec} M={M} K={K} N={N}", flush=True)
    rng = np.random.default_rng(M * 3 + K + N)
    S = H.rng_sparse(rng, M, K, dens)
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    g = oracle.convert(S, 4)
    m = G.GCOOMatrix(M, K, 4).convert_dense(to_dev(S, dev))
    m.set_config(variant=variant, vec=vec)
    D16 = G.to_bf16(to_dev(D, dev))
    C = torch.full((M, N), float("nan"), device=dev)
    m.spdm_bf16(D16, C)
    Cn = C.cpu().numpy()
    ref = oracle.spdm_bf16(g, oracle.to_bf16(g["vals"]), oracle.to_bf16(D))
    assert H.bits_equal(Cn, ref)
    Cd, _ = oracle.spdm_fp64(g, D)
    if np.any(Cd):
        assert oracle.rel_frobenius(Cn, Cd) <= BF16_REL_TOL


def test_bf16_after_reconversion(dev):
    """The BF16 values follow a re-conversion of the handle."""
    import torch
    M, K, N = 64, 80, 128
    rng = np.random.default_rng(5)
    m = G.GCOOMatrix(M, K, 4)
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    D16 = G.to_bf16(to_dev(D, dev))
    for dens in (0.1, 0.3):
        S = H.rng_sparse(rng, M, K, dens)
        m.convert_dense(to_dev(S, dev))
        C = torch.empty((M, N), device=dev)
        m.spdm_bf16(D16, C)
        g = oracle.convert(S, 4)
        ref = oracle.spdm_bf16(g, oracle.to_bf16(g["vals"]), oracle.to_bf16(D))
        assert H.bits_equal(C.cpu().numpy(), ref)


def test_bf16_n8000_error_bound(dev):
    """n = 8000, s = 0.99: BF16 error against FP64 on a sample of row groups."""
    import torch
    n = 8000
    Sd = torch.empty((n, n), device=dev)
    Dd = torch.empty((n, n), device=dev)
    G.fill_sparse(Sd, 3, 1, 0.99)
    G.fill_uniform(Dd, 3, 0)
    m = G.GCOOMatrix(n, n, 4).convert_dense(Sd)
    C = torch.empty((n, n), device=dev)
    m.spdm_bf16(G.to_bf16(Dd), C)
    S = Sd.cpu().numpy()
    D = Dd.cpu().numpy()
    rows = np.concatenate([np.arange(4 * q, 4 * q + 4) for q in range(0, n // 4, 97)])
    sub = oracle.convert(S[rows], 4)
    Cd, _ = oracle.spdm_fp64(sub, D, threads=oracle.max_threads())
    err = oracle.rel_frobenius(C.cpu().numpy()[rows], Cd)
    print(f"bf16 n=8000 rel_frobenius={err:.3e}")
    assert err <= BF16_REL_TOL
