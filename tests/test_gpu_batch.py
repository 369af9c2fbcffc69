"""GPU tests of the batched persistent path: every job's C in a batch is
bit-identical to the same job run alone (FP32 and BF16, any tile), and HIP
graph replays reproduce the eager run, with eager work between replays."""
import numpy as np
import pytest

import gcoospdm_amd as G
import oracle
from gcoospdm_amd import workloads as W
from tests import helpers as H

pytestmark = pytest.mark.gpu


def small_specs(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        M = int(rng.integers(1, 700))
        K = int(rng.integers(1, 900))
        N = int(rng.choice([1, 3, 16, 32, 48, 64, 100, 128, 256, 260]))
        out.append(W.JobSpec(i, M, K, N, float(rng.uniform(0.8, 0.999)), seed * 1000 + i))
    return out


def alone(m, D, N, bf16):
    import torch
    C = torch.empty((m.M, N), device=D.device)
    if bf16:
        m.spdm_bf16(D, C)
    else:
        m.spdm(D, C)
    return C


@pytest.mark.parametrize("bf16", [False, True])
@pytest.mark.parametrize("rows,vec", [(0, 0), (1, 1), (8, 2), (16, 4)])
def test_batch_matches_alone(dev, bf16, rows, vec):
    import torch
    print(f"batch bf16={bf16} rows={rows} vec={vec}", flush=True)
    specs = small_specs(40, 7 + rows + vec)
    jobs = [W.materialize(sp, dev, bf16) for sp in specs]
    b = G.Batch()
    for m, D, C in jobs:
        C.fill_(float("nan"))
        b.add(m, D, C, rows=rows, vec=vec)
    b.commit()
    b.run()
    torch.cuda.synchronize()
    for (m, D, C), sp in zip(jobs, specs):
        ref = alone(m, D, sp.N, bf16)
        assert torch.equal(C.view(torch.int32), ref.view(torch.int32)), sp


def test_batch_vs_oracle(dev):
    """A few jobs checked directly against the oracle's fmaf chain."""
    import torch
    specs = small_specs(6, 99)
    b = G.Batch()
    jobs = []
    for sp in specs:
        S = torch.empty((sp.M, sp.K), device=dev)
        G.fill_sparse(S, sp.seed, 1, sp.sparsity)
        m = G.GCOOMatrix(sp.M, sp.K, 4).convert_dense(S)
        D = torch.empty((sp.K, sp.N), device=dev)
        G.fill_uniform(D, sp.seed, 0)
        C = torch.empty((sp.M, sp.N), device=dev)
        b.add(m, D, C)
        jobs.append((S.cpu().numpy(), D.cpu().numpy(), C))
    b.commit()
    b.run()
    torch.cuda.synchronize()
    for S, D, C in jobs:
        assert H.bits_equal(C.cpu().numpy(), oracle.spdm_f32(oracle.convert(S, 4), D))


def test_batch_graph_replays(dev):
    """Capture once, replay several times with eager work in between; each
    replay reproduces the eager result."""
    import torch
    specs = small_specs(24, 5)
    jobs = [W.materialize(sp, dev, False) for sp in specs]
    b = G.Batch()
    for m, D, C in jobs:
        b.add(m, D, C)
    b.commit()
    b.run()
    torch.cuda.synchronize()
    eager = [C.clone() for _, _, C in jobs]
    s = torch.cuda.Stream()
    b.capture(s)
    for rep in range(4):
        for _, _, C in jobs:
            C.zero_()                          # eager work between replays
        torch.cuda.synchronize()
        b.replay(s)
        s.synchronize()
        for (_, _, C), ref in zip(jobs, eager):
            assert torch.equal(C.view(torch.int32), ref.view(torch.int32)), rep


def test_batch_config4_sample(dev):
    """64 jobs drawn like the 512-job workload (M, K in [256, 8192], N in
    {32..256}, s in [0.95, 0.999]): batch == alone, FP32 and BF16."""
    import torch
    specs = W.batch_jobs(64, 2024)
    for bf16 in (False, True):
        jobs = [W.materialize(sp, dev, bf16) for sp in specs]
        b = G.Batch()
        for m, D, C in jobs:
            b.add(m, D, C)
        b.commit()
        b.run()
        torch.cuda.synchronize()
        for (m, D, C), sp in zip(jobs, specs):
            ref = alone(m, D, sp.N, bf16)
            assert torch.equal(C.view(torch.int32), ref.view(torch.int32)), sp
        del jobs
