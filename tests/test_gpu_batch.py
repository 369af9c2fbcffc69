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


def _host_job(sp, dev):
    """The job's S and D on the host, generated by the same device kernels."""
    import torch
    S = torch.empty((sp.M, sp.K), device=dev)
    G.fill_sparse(S, sp.seed, 1, sp.sparsity)
    D = torch.empty((sp.K, sp.N), device=dev)
    G.fill_uniform(D, sp.seed, 0)
    return S.cpu().numpy(), D.cpu().numpy()


@pytest.mark.parametrize("bf16", [False, True])
def test_batch_config4_512_graph(dev, bf16):
    """The full batching workload: 512 jobs (M, K log-uniform in [256, 8192],
    N in {32, 64, 128, 256}, s in [0.95, 0.999]) with tuned tiles, captured
    into one HIP graph and replayed once.  Every C is bit-equal to the job
    run alone, and 32 jobs spread over the batch are bit-equal to the
    oracle's fmaf chain (BF16: the chain over BF16-widened operands), with
    the relative Frobenius error against FP64 of the FP32 operands within
    1e-5 (FP32) / 1e-2 (BF16), the check of src/main.cu:128-142."""
    import torch
    from gcoospdm_amd import autotune
    print(f"batch 512 graph bf16={bf16}", flush=True)
    specs = W.batch_jobs(512, 2024)
    jobs = [W.materialize(sp, dev, bf16) for sp in specs]
    b = G.Batch()
    for (m, D, C), sp in zip(jobs, specs):
        C.fill_(float("nan"))
        rows, vec = autotune.batch_tile(sp.N, sp.K * (1 - sp.sparsity), bf16)
        b.add(m, D, C, rows=rows, vec=vec)
    b.commit()
    s = torch.cuda.Stream()
    b.capture(s)
    b.replay(s)
    s.synchronize()
    assert b.info()["jobs"] == 512
    for (m, D, C), sp in zip(jobs, specs):
        ref = alone(m, D, sp.N, bf16)
        assert torch.equal(C.view(torch.int32), ref.view(torch.int32)), sp
    tol = 1e-2 if bf16 else 1e-5
    worst = 0.0
    for i in np.linspace(0, 511, 32).astype(int):
        sp = specs[i]
        S, Dh = _host_job(sp, dev)
        g = oracle.convert(S, 4)
        Cn = jobs[i][2].cpu().numpy()
        if bf16:
            want = oracle.spdm_bf16(g, oracle.to_bf16(g["vals"]), oracle.to_bf16(Dh),
                                    threads=oracle.max_threads())
        else:
            want = oracle.spdm_f32(g, Dh, threads=oracle.max_threads())
        assert H.bits_equal(Cn, want), sp
        Cd, _ = oracle.spdm_fp64(g, Dh, threads=oracle.max_threads())
        err = oracle.rel_frobenius(Cn, Cd)
        worst = max(worst, err)
        assert err <= tol, (sp, err)
    print(f"batch 512 bf16={bf16}: 32 jobs oracle-exact, worst rel_frobenius {worst:.3e}",
          flush=True)


def test_batch_refuses_after_reconvert(dev):
    """Converting a job's matrix again after commit (here with growth, so its
    arrays are reallocated) makes run, capture and replay refuse with
    GCOO_ERR_STATE, launching nothing; a new commit binds the new arrays."""
    import torch
    specs = small_specs(6, 21)
    jobs = [W.materialize(sp, dev, bf16) for sp, bf16 in zip(specs, [False, True] * 3)]
    b = G.Batch()
    for m, D, C in jobs:
        b.add(m, D, C)
    b.commit()
    s = torch.cuda.Stream()
    b.capture(s)
    b.replay(s)
    s.synchronize()
    # job 1 (BF16) gets a much denser S: its entry arrays are reallocated
    m, D, C = jobs[1]
    S = torch.empty((m.M, m.K), device=dev)
    G.fill_sparse(S, 99, 1, 0.3)
    m.convert_dense(S)
    for call in (b.run, lambda: b.capture(s), lambda: b.replay(s)):
        with pytest.raises(G.GCOOError) as e:
            call()
        assert e.value.code == G.GCOO_ERR_STATE
    for _, _, C in jobs:
        C.fill_(float("nan"))
    b.commit()
    b.run()
    torch.cuda.synchronize()
    for (m, D, C), sp in zip(jobs, specs):
        ref = alone(m, D, sp.N, D.dtype != torch.float32)
        assert torch.equal(C.view(torch.int32), ref.view(torch.int32)), sp
