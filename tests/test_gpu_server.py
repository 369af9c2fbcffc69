"""GPU tests of the continuous-batching server: results in the mix equal
the same multiply alone (bit for bit, FP32 and BF16), admission never
exceeds the HBM budget, oversized requests are rejected, and requests
submitted while batches run are batched and complete."""
import threading

import numpy as np
import pytest

import gcoospdm_amd as G
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
