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


@pytest.mark.parametrize("bf16", [False, True])
def test_batch_empty_rows_narrow_streams(dev, bf16):
    """Jobs with no entries at all, and jobs whose first rows are empty, on
    16- and 32-lane row streams (N = 16, 32): the streams' loads past a
    row's end read entries they do not use; every C is bit-equal to the job
    alone (+0.0 where a row has no entries)."""
    import torch
    print(f"batch empty rows bf16={bf16}", flush=True)
    rng = np.random.default_rng(5)
    b = G.Batch().set_panel_mode(0)
    jobs = []
    for M, K, N, kind in [(37, 50, 16, "zero"), (64, 90, 32, "zero"), (45, 70, 32, "head"),
                          (9, 20, 16, "head"), (130, 40, 32, "all")]:
        S = np.zeros((M, K), np.float32)
        if kind == "head":
            S[M // 2:] = H.rng_sparse(rng, M - M // 2, K, 0.3)
        elif kind == "all":
            S = H.rng_sparse(rng, M, K, 0.4)
        D = torch.from_numpy(rng.uniform(-1, 1, (K, N)).astype(np.float32)).to(dev)
        if bf16:
            D = G.to_bf16(D)
        m = G.GCOOMatrix(M, K, 4).convert_dense(torch.from_numpy(S).to(dev))
        C = torch.full((M, N), float("nan"), device=dev)
        b.add(m, D, C)
        jobs.append((m, D, C, N))
    b.commit()
    b.run()
    torch.cuda.synchronize()
    for m, D, C, N in jobs:
        ref = alone(m, D, N, bf16)
        assert torch.equal(C.view(torch.int32), ref.view(torch.int32)), (m.M, m.K, N)
    b.close()


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


def _panel_specs():
    """Jobs that the panel kernel takes (dense enough, N a multiple of 4) next
    to jobs it does not: ragged M (not multiples of the 336-row panel), one
    and two 256-column tiles (N = 256, 260, 512), short K, FP32 and BF16."""
    shapes = [(700, 900, 256, 0.97), (336, 400, 260, 0.95), (1000, 75, 512, 0.9),
              (50, 1500, 256, 0.98), (1200, 2000, 128, 0.96), (400, 600, 48, 0.95),
              (673, 151, 256, 0.99), (9, 5, 256, 0.5), (800, 800, 100, 0.97)]
    return [W.JobSpec(i, M, K, N, s, 7000 + i) for i, (M, K, N, s) in enumerate(shapes)]


@pytest.mark.parametrize("mode", [1, 2])
def test_batch_panel_jobs_match_gather(dev, mode):
    """The same jobs with no panel jobs (mode 0) and with panel jobs (mode 1:
    the cost model's choice, 2: every job the operands allow), FP32 and BF16
    mixed in one batch: identical bits, and the panel jobs bit-equal to the
    oracle's chain."""
    import torch
    specs = _panel_specs()
    jobs = [W.materialize(sp, dev, i % 2 == 1) for i, sp in enumerate(specs)]
    results = {}
    for md in (0, mode):
        b = G.Batch().set_panel_mode(md)
        for m, D, C in jobs:
            C.fill_(float("nan"))
            b.add(m, D, C)
        b.commit()
        b.run()
        torch.cuda.synchronize()
        info = b.panel_info()
        print(f"panel mode {md}: {info}", flush=True)
        if md == 0:
            assert info["panel_jobs"] == 0
        else:
            assert info["panel_jobs"] >= 1
        if md == 2:
            # every job with entries whose D rows are whole 16-byte pieces
            # (the operands are aligned)
            assert info["panel_jobs"] == sum(
                1 for i, (sp, (m, _, _)) in enumerate(zip(specs, jobs))
                if (sp.N * (2 if i % 2 else 4)) % 16 == 0 and m.nnz > 0)
        results[md] = ([C.clone() for _, _, C in jobs], [b.job_is_panel(i) for i in range(len(jobs))])
    for i, sp in enumerate(specs):
        assert torch.equal(results[0][0][i].view(torch.int32),
                           results[mode][0][i].view(torch.int32)), sp
    for i, sp in enumerate(specs):
        if not results[mode][1][i]:
            continue
        S, Dh = _host_job(sp, dev)
        g = oracle.convert(S, 4)
        if i % 2 == 1:
            want = oracle.spdm_bf16(g, oracle.to_bf16(g["vals"]), oracle.to_bf16(Dh))
        else:
            want = oracle.spdm_f32(g, Dh)
        assert H.bits_equal(results[mode][0][i].cpu().numpy(), want), sp


def test_batch_panel_same_matrix_two_types(dev):
    """One matrix in an FP32 and a BF16 job of the same batch: the panel lists
    are per element type, so the second job takes the persistent kernel; both
    results equal the jobs run alone."""
    import torch
    sp = W.JobSpec(0, 900, 700, 256, 0.97, 4242)
    m, D, C = W.materialize(sp, dev, False)
    D16 = G.to_bf16(D)
    C16 = torch.empty_like(C)
    b = G.Batch().set_panel_mode(2)
    b.add(m, D, C)
    b.add(m, D16, C16)
    b.commit()
    b.run()
    torch.cuda.synchronize()
    assert b.job_is_panel(0) and not b.job_is_panel(1)
    assert torch.equal(C.view(torch.int32), alone(m, D, 256, False).view(torch.int32))
    assert torch.equal(C16.view(torch.int32), alone(m, D16, 256, True).view(torch.int32))


def test_batch_panel_lists_rebuilt_refuses(dev):
    """After commit, a plain multiply of a panel job's matrix with another
    panel shape rebuilds the handle's lists: run, capture and replay refuse
    with GCOO_ERR_STATE, launching nothing; a new commit runs again."""
    import torch
    sp = W.JobSpec(0, 1000, 1200, 256, 0.97, 4343)
    m, D, C = W.materialize(sp, dev, False)
    b = G.Batch().set_panel_mode(2)
    b.add(m, D, C)
    b.commit()
    assert b.job_is_panel(0)
    s = torch.cuda.Stream()
    b.capture(s)
    b.replay(s)
    s.synchronize()
    ref = C.clone()
    m.set_config(variant=6, waves=16, chunk=16)   # 256-row panels: other lists
    C2 = torch.empty_like(C)
    m.spdm(D, C2)
    m.set_config()
    assert torch.equal(C2.view(torch.int32), ref.view(torch.int32))
    for call in (b.run, lambda: b.capture(s), lambda: b.replay(s)):
        with pytest.raises(G.GCOOError) as e:
            call()
        assert e.value.code == G.GCOO_ERR_STATE
    C.fill_(float("nan"))
    b.commit()
    b.run()
    torch.cuda.synchronize()
    assert torch.equal(C.view(torch.int32), ref.view(torch.int32))
