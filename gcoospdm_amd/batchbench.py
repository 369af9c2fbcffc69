"""Measurement of the batched path on the many-jobs workload.

512 independent jobs (M, K log-uniform in [256, 8192], N in {32..256},
sparsity in [0.95, 0.999], see workloads.batch_jobs), FP32 or BF16:
  sequential  one plain multiply launch per job, back to back on one stream;
  batched     one HIP-graph replay of the batch with the autotuned tiles:
              the panel-staged launch of the jobs the batch's cost model
              gives it, then the persistent kernel for the others
              (panel_mode 0: every job on the persistent kernel).
Both are timed with HIP events on the stream they run on; GFLOP/s counts
2 * nnz * N per job.  Used by bench.py; also runnable alone:
    python -m gcoospdm_amd.batchbench [--jobs 512] [--bf16]
"""
from __future__ import annotations

import argparse
import json
import time


def run(count=512, bf16=False, reps=10, seed=2024, table=None, panel_mode=None, panel_ratio=0.0):
    import torch
    import gcoospdm_amd as G
    from . import autotune, workloads as W

    specs = W.batch_jobs(count, seed)
    jobs = [W.materialize(sp, bf16=bf16) for sp in specs]
    nnz = [m.nnz for m, _, _ in jobs]
    flops = sum(2.0 * z * sp.N for z, sp in zip(nnz, specs))
    algo_bytes = sum(W.spdm_bytes(z, sp.M, sp.K, sp.N, bf16) for z, sp in zip(nnz, specs))
    stream = torch.cuda.Stream()

    def seq():
        for m, D, C in jobs:
            if bf16:
                m.spdm_bf16(D, C, stream=stream, sync=False)
            else:
                m.spdm(D, C, stream=stream, sync=False)

    # plain tiles from the table as well, so both sides are tuned
    for (m, D, C), sp in zip(jobs, specs):
        v, vec = autotune.plain_config(sp.N, sp.K * (1 - sp.sparsity), bf16, table)
        m.set_config(variant=v, vec=vec)

    def timed(fn):
        # untimed runs first, about 40 ms of GPU work: after the host-side
        # set-up the GPU clock needs that much load to come back up
        t0 = time.perf_counter()
        while True:
            with torch.cuda.stream(stream):
                fn()
            stream.synchronize()
            if time.perf_counter() - t0 > 0.04:
                break
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3

    t_seq = timed(seq)
    ref = [C.clone() for _, _, C in jobs]

    if panel_mode is None:  # the tuned choice
        panel_mode, panel_ratio = autotune.batch_panel(bf16, table)
    b = G.Batch().set_panel_mode(panel_mode, panel_ratio, panel_ratio)
    for (m, D, C), sp in zip(jobs, specs):
        rows, vec = autotune.batch_tile(sp.N, sp.K * (1 - sp.sparsity), bf16, table)
        b.add(m, D, C, rows=rows, vec=vec)
    b.commit()
    b.capture(stream)
    t_batch = timed(lambda: b.replay(stream))
    same = all(torch.equal(C.view(torch.int32), r.view(torch.int32))
               for (_, _, C), r in zip(jobs, ref))
    info = b.info()
    pinfo = b.panel_info()
    pflops = sum(2.0 * z * sp.N for i, (z, sp) in enumerate(zip(nnz, specs)) if b.job_is_panel(i))
    return {
        "workload": f"batch_{count}_jobs", "dtype": "bf16" if bf16 else "f32",
        "jobs": count, "nnz": int(sum(nnz)), "gflop": round(flops / 1e9, 3),
        "algorithmic_bytes": int(algo_bytes),
        "sequential_ms": round(t_seq * 1e3, 4), "batched_ms": round(t_batch * 1e3, 4),
        "sequential_gflops": round(flops / t_seq / 1e9, 2),
        "batched_gflops": round(flops / t_batch / 1e9, 2),
        "speedup": round(t_seq / t_batch, 3),
        "batched_hbm_frac": round(algo_bytes / t_batch / 8e12, 4),
        "work_items": info["items"], "grid_blocks": info["grid_blocks"],
        "panel_mode": panel_mode, "panel_ratio": panel_ratio, "panel_jobs": pinfo["panel_jobs"],
        "panel_workgroups": pinfo["workgroups"],
        "panel_flop_share": round(pflops / flops, 3) if flops else 0.0,
        "bit_identical_to_alone": bool(same),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=512)
    ap.add_argument("--bf16", action="store_true")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--panel-mode", type=int, default=None,
                    help="0: no panel jobs, 1: the cost model's, 2: all eligible "
                         "(default: the tuned choice)")
    ap.add_argument("--panel-ratio", type=float, default=0.0,
                    help="mode 1: panel when its estimate is below ratio x the other (0: 1.0)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also copy a 256 MiB tensor 3 times (known bytes for counter runs)")
    args = ap.parse_args()
    print(json.dumps(run(args.jobs, args.bf16, args.reps, panel_mode=args.panel_mode,
                         panel_ratio=args.panel_ratio)), flush=True)
    if args.calibrate:
        import torch
        X = torch.empty(64 << 20, device="cuda:0")   # 256 MiB
        Y = torch.empty_like(X)
        for _ in range(3):
            Y.copy_(X)
        torch.cuda.synchronize()
        print("calibration: copied 256 MiB x3", flush=True)


if __name__ == "__main__":
    main()
