"""The no-allocation multiply (gcoo_spdm_f32_noalloc_async) through the C
ABI against the existing entry point (gcoo_spdm_f32): bit-equal C on two
calls in a row, the second on a smaller input, for the panel-staged kernel
(panel lists in the library's static scratch), the long-row split and
variant 2's long-rows-first order (their row list and counter there); a
multiply whose lists exceed the scratch is refused before anything is
enqueued."""
import numpy as np
import pytest

import gcoospdm_amd as G
import oracle
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _noalloc(m, D, C):
    import torch
    K, N = D.shape
    rc = G.lib().gcoo_spdm_f32_noalloc_async(m.handle, D.data_ptr(), K, N, C.data_ptr(),
                                             G._stream())
    G._check(rc, "gcoo_spdm_f32_noalloc_async")
    torch.cuda.synchronize()


def _case(dev, M, K, N, s, seed, variant=0):
    import torch
    S = torch.empty((M, K), device=dev)
    G.fill_sparse(S, seed, 1, s)
    D = torch.empty((K, N), device=dev)
    G.fill_uniform(D, seed, 0)
    m = G.GCOOMatrix(M, K, 4).convert_dense(S)
    m.set_config(variant=variant)
    return S, D, m


@pytest.mark.parametrize("variant", [0, 4])
def test_noalloc_matches_plain_two_calls(dev, variant):
    import torch
    # both calls take the panel kernel (lists in the static scratch) by
    # default: 16 x 21 panels at 1 % and at 0.6 %
    shapes = [(8000, 8000, 8000, 0.99), (6000, 6000, 6000, 0.994)]
    for call, (M, K, N, s) in enumerate(shapes):
        print(f"noalloc variant={variant} call={call} M={M} K={K} N={N} s={s}", flush=True)
        S, D, m = _case(dev, M, K, N, s, 31 + call, variant)
        want = torch.empty((M, N), device=dev)
        m.spdm(D, want)
        if variant == 0:
            assert m.spdm_choice(D, want)[0] == 6
        got = torch.full((M, N), float("nan"), device=dev)
        _noalloc(m, D, got)
        assert torch.equal(got.view(torch.int32), want.view(torch.int32)), (call, M, s)
        if call == 1:
            # the smaller input also matches the oracle's fmaf chain
            g = oracle.convert(S.cpu().numpy(), 4)
            assert H.bits_equal(got.cpu().numpy(),
                                oracle.spdm_f32(g, D.cpu().numpy(), threads=oracle.max_threads()))
        del S, D, m, want, got


def test_noalloc_long_rows_first(dev):
    """Variant 2 with its long rows first: the row list and its counter in
    the static scratch, two calls in a row (the second smaller), bit-equal
    to the existing entry point and to the oracle."""
    import torch
    for call, (M, K, N) in enumerate([(4096, 3000, 256), (1500, 2000, 128)]):
        print(f"noalloc long-first call={call} M={M} K={K} N={N}", flush=True)
        rng = np.random.default_rng(40 + call)
        lens = np.minimum(rng.zipf(1.5, M), K // 2)
        S = np.zeros((M, K), np.float32)
        for r in range(M):
            S[r, rng.choice(K, size=lens[r], replace=False)] = rng.uniform(-1, 1, lens[r])
        D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        Sd = torch.from_numpy(S).to(dev)
        Dd = torch.from_numpy(D).to(dev)
        m = G.GCOOMatrix(M, K, 4).convert_dense(Sd)
        m.set_config(variant=2, vec=1, waves=6, chunk=1)
        want = torch.empty((M, N), device=dev)
        m.spdm(Dd, want)
        got = torch.full((M, N), float("nan"), device=dev)
        _noalloc(m, Dd, got)
        assert torch.equal(got.view(torch.int32), want.view(torch.int32)), call
        g = oracle.convert(S, 4)
        assert H.bits_equal(got.cpu().numpy(), oracle.spdm_f32(g, D, threads=oracle.max_threads()))
        del Sd, Dd, m, want, got


def test_noalloc_rejects_oversized_lists(dev):
    """At n = 8000, s = 0.95 the panel lists need ~26 MB of scratch: refused
    with GCOO_ERR_BUDGET; the same handle with variant 2 (no scratch) runs."""
    import torch
    S, D, m = _case(dev, 8000, 8000, 8000, 0.95, 5)
    assert m.nnz * 8 > 16 << 20
    C = torch.empty((8000, 8000), device=dev)
    K, N = D.shape
    rc = G.lib().gcoo_spdm_f32_noalloc_async(m.handle, D.data_ptr(), K, N, C.data_ptr(),
                                             G._stream())
    assert rc == G.GCOO_ERR_BUDGET
    assert b"scratch" in G.lib().gcoo_last_error()
    m.set_config(variant=2)
    _noalloc(m, D, C)
    want = torch.empty_like(C)
    m.spdm(D, want)
    assert torch.equal(C.view(torch.int32), want.view(torch.int32))
