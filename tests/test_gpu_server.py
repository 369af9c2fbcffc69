"""GPU tests of the continuous-batching server: results in the mix equal
the same multiply alone (bit for bit, FP32 and BF16), admission never
exceeds the HBM budget, oversized requests are rejected, requests
submitted while batches run are batched and complete, served results at the
batching workload's sizes match the oracle and the FP64 bound, and a failed
launch fails its requests instead of reporting them done."""
import threading

import numpy as np
import pytest

import gcoospdm_amd as G
import oracle
from gcoospdm_amd import workloads as W

pytestmark = pytest.mark.gpu


def make_ops(dev, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        sp = W.JobSpec(i, int(rng.integers(16, 600)), int(rng.integers(16, 700)),
                       int(rng.choice([32, 64, 128, 256])), float(rng.uniform(0.9, 0.999)),
                       seed * 100 + i)
        bf16 = bool(i % 2)
        m, D, _ = W.materialize(sp, dev, bf16)
        out.append((sp, m, D, bf16))
    return out


def alone(m, D, bf16):
    import torch
    C = torch.empty((m.M, D.shape[1]), device=D.device)
    if bf16:
        m.spdm_bf16(D, C)
    else:
        m.spdm(D, C)
    return C.cpu().numpy()


def test_server_mix_equals_alone(dev):
    ops = make_ops(dev, 30, 3)
    srv = G.Server(hbm_budget_bytes=64 << 20, slots=3, max_batch_jobs=8)
    ids = [srv.register(m, D) for _, m, D, _ in ops]
    # two submitting threads, interleaved, twice over
    reqs = []
    lock = threading.Lock()

    def client(part):
        for rep in range(2):
            for k in range(part, len(ids), 2):
                r = srv.submit(ids[k])
                with lock:
                    reqs.append((k, r))

    th = [threading.Thread(target=client, args=(p,)) for p in (0, 1)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    refs = {k: alone(m, D, bf16) for k, (_, m, D, bf16) in enumerate(ops)}
    for k, r in reqs:
        lat = srv.wait(r)
        assert lat >= 0
        got = srv.result_host(r)
        assert np.array_equal(got.view(np.uint32), refs[k].view(np.uint32)), k
        srv.release(r)
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["peak_bytes"] <= st["budget_bytes"]
    assert st["batches"] >= 2
    srv.close()


def test_server_budget_admission(dev):
    """A budget of ~2 outputs forces queueing; everything still completes,
    never above the budget; a request larger than the budget is refused."""
    ops = make_ops(dev, 12, 4)
    biggest = max(m.M * D.shape[1] * 4 for _, m, D, _ in ops)
    budget = 2 * ((biggest + 255) // 256 * 256)
    srv = G.Server(hbm_budget_bytes=budget, slots=2, max_batch_jobs=4)
    ids = [srv.register(m, D) for _, m, D, _ in ops]
    reqs = [(k, srv.submit(i)) for k, i in enumerate(ids)]
    for k, r in reqs:
        srv.wait(r)
        got = srv.result_host(r)
        _, m, D, bf16 = ops[k]
        assert np.array_equal(got.view(np.uint32), alone(m, D, bf16).view(np.uint32))
        srv.release(r)
    st = srv.stats()
    assert st["peak_bytes"] <= budget
    # an operand whose output alone exceeds the budget
    import torch
    sp = W.JobSpec(0, 2048, 64, 256, 0.99, 1)
    m, D, _ = W.materialize(sp, dev, False)
    big = srv.register(m, D)
    with pytest.raises(G.GCOOError) as e:
        srv.submit(big)
    assert e.value.code == G.GCOO_ERR_BUDGET
    srv.close()


def test_server_continuous_arrivals(dev):
    """Requests keep arriving while earlier batches run (no waiting between
    submissions); each completes with the alone result."""
    import time
    ops = make_ops(dev, 16, 5)
    srv = G.Server(hbm_budget_bytes=256 << 20, slots=4, max_batch_jobs=64)
    ids = [srv.register(m, D) for _, m, D, _ in ops]
    refs = [alone(m, D, bf16) for _, m, D, bf16 in ops]
    reqs = []
    for wave in range(5):
        for k, i in enumerate(ids):
            reqs.append((k, srv.submit(i)))
        time.sleep(0.0005)
    for k, r in reqs:
        srv.wait(r)
        assert np.array_equal(srv.result_host(r).view(np.uint32), refs[k].view(np.uint32))
        srv.release(r)
    assert srv.stats()["completed"] == len(reqs)
    srv.close()


def _config4_ops(dev, count):
    """`count` jobs drawn like the 512-job workload (the first `count` of
    it), each with its FP32 D and the BF16 form of the same D."""
    specs = W.batch_jobs(count, 2024)
    out = []
    for sp in specs:
        m, D, _ = W.materialize(sp, dev, False)
        out.append((sp, m, D, G.to_bf16(D)))
    return out


def _poisson_submit(srv, seq, rate, seed):
    """Submit op ids in `seq` at Poisson arrival times of the given rate
    (busy-waiting on the clock between arrivals)."""
    import time
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, len(seq))
    t = time.perf_counter()
    out = []
    for op, gap in zip(seq, gaps):
        t += gap
        while time.perf_counter() < t:
            pass
        out.append(srv.submit(op))
    return out


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_server_config4_poisson_vs_oracle(dev, mode):
    """Serving at the batching workload's sizes: 48 jobs (M, K in [256,
    8192], N in {32..256}, s in [0.95, 0.999]) registered in FP32 and BF16
    form; 200 requests over both forms arrive as a Poisson stream.  Every
    result equals the job run alone bit for bit, and 16 of the jobs (both
    forms) equal the oracle's fmaf chains with the relative Frobenius error
    against FP64 of the FP32 operands within 1e-5 (FP32) / 1e-2 (BF16)."""
    import torch
    print(f"server config4 poisson mode={mode}", flush=True)
    ops = _config4_ops(dev, 48)
    if mode == "batched":
        srv = G.Server(hbm_budget_bytes=1 << 30, slots=4, max_batch_jobs=64)
    else:
        srv = G.Server(hbm_budget_bytes=1 << 30, slots=1, max_batch_jobs=1)
    ids = {}
    for k, (sp, m, D, D16) in enumerate(ops):
        ids[(k, False)] = srv.register(m, D)
        ids[(k, True)] = srv.register(m, D16)
    rng = np.random.default_rng(17)
    keys = [(int(k), bool(b)) for k, b in zip(rng.integers(0, 48, 200), rng.integers(0, 2, 200))]
    # every sampled job appears in both forms
    sample = list(range(0, 48, 3))
    keys += [(k, b) for k in sample for b in (False, True)]
    rng.shuffle(keys)
    rids = _poisson_submit(srv, [ids[k] for k in keys], 20000.0, 3)
    got = {}
    for key, r in zip(keys, rids):
        assert srv.wait(r) >= 0
        got.setdefault(key, []).append(srv.result_host(r))
        srv.release(r)
    for (k, bf16), results in got.items():
        sp, m, D, D16 = ops[k]
        ref = alone(m, D16 if bf16 else D, bf16)
        for C in results:
            assert np.array_equal(C.view(np.uint32), ref.view(np.uint32)), (sp, bf16)
    worst = {False: 0.0, True: 0.0}
    for k in sample:
        sp = ops[k][0]
        S = torch.empty((sp.M, sp.K), device=dev)
        G.fill_sparse(S, sp.seed, 1, sp.sparsity)
        Sh, Dh = S.cpu().numpy(), ops[k][2].cpu().numpy()
        g = oracle.convert(Sh, 4)
        Cd, _ = oracle.spdm_fp64(g, Dh, threads=oracle.max_threads())
        for bf16 in (False, True):
            C = got[(k, bf16)][0]
            if bf16:
                want = oracle.spdm_bf16(g, oracle.to_bf16(g["vals"]), oracle.to_bf16(Dh),
                                        threads=oracle.max_threads())
            else:
                want = oracle.spdm_f32(g, Dh, threads=oracle.max_threads())
            assert np.array_equal(C.view(np.uint32), want.view(np.uint32)), (sp, bf16)
            err = oracle.rel_frobenius(C, Cd)
            worst[bf16] = max(worst[bf16], err)
            assert err <= (1e-2 if bf16 else 1e-5), (sp, bf16, err)
    st = srv.stats()
    assert st["completed"] == len(keys)
    assert st["peak_bytes"] <= st["budget_bytes"]
    print(f"server {mode}: {len(keys)} requests, p50 {st['p50_ms']:.3f} ms, "
          f"p99 {st['p99_ms']:.3f} ms, worst rel_frobenius f32 {worst[False]:.2e} "
          f"bf16 {worst[True]:.2e}", flush=True)
    srv.close()


def test_server_injected_failure(dev):
    """A launch that fails (forced from the host before anything is
    enqueued) fails its request and every request not yet launched with the
    same error code; later submits are refused with it; failed requests can
    be released.  No kernel is made to fail."""
    ops = make_ops(dev, 4, 8)
    srv = G.Server(hbm_budget_bytes=64 << 20, slots=1, max_batch_jobs=1)
    ids = [srv.register(m, D) for _, m, D, _ in ops]
    ok = srv.submit(ids[0])
    assert srv.wait(ok) >= 0
    srv.release(ok)
    assert lib_inject(srv, G.GCOO_ERR_UNSUPPORTED) == 0
    # the first submit is accepted; once its launch has failed, submits are
    # refused with the launch's error code
    rids = []
    for i in ids:
        try:
            rids.append(srv.submit(i))
        except G.GCOOError as e:
            assert e.code == G.GCOO_ERR_UNSUPPORTED
            break
    assert rids
    for r in rids:
        with pytest.raises(G.GCOOError) as e:
            srv.wait(r)
        assert e.value.code == G.GCOO_ERR_UNSUPPORTED
        with pytest.raises(G.GCOOError):
            srv.result_host(r)
        srv.release(r)
    with pytest.raises(G.GCOOError) as e:
        srv.submit(ids[0])
    assert e.value.code == G.GCOO_ERR_UNSUPPORTED
    srv.close()


def lib_inject(srv, code):
    return G.lib().gcoo_test_server_inject_failure(srv._s, code)


def test_server_wait_while_submitting(dev):
    """One thread waits on requests while another keeps submitting (so the
    server's request table grows and rehashes under the waiter)."""
    ops = make_ops(dev, 6, 9)
    srv = G.Server(hbm_budget_bytes=256 << 20, slots=2, max_batch_jobs=32)
    ids = [srv.register(m, D) for _, m, D, _ in ops]
    refs = [alone(m, D, bf16) for _, m, D, bf16 in ops]
    first = [srv.submit(ids[k]) for k in range(len(ids))]
    more = []

    def submitter():
        for rep in range(300):
            k = rep % len(ids)
            more.append((k, srv.submit(ids[k])))

    th = threading.Thread(target=submitter)
    th.start()
    for k, r in enumerate(first):
        srv.wait(r)
        assert np.array_equal(srv.result_host(r).view(np.uint32), refs[k].view(np.uint32))
        srv.release(r)
    th.join()
    for k, r in more:
        srv.wait(r)
        assert np.array_equal(srv.result_host(r).view(np.uint32), refs[k].view(np.uint32))
        srv.release(r)
    srv.close()
