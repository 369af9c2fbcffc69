"""Autotuner for the multiply tiles.

The reference grid-searches its compile-time COLUMN_PER_GROUP / BS by
substituting them into templates, recompiling and timing `main`
(src/autotune.py:21-86, src/constants.h.t:4-5).  Here the tile parameters
are runtime arguments, so the search needs no rebuild: each candidate is
timed with HIP events on the device and the fastest is written to a table.

Four tables:
  batch[dtype][class] = {rows, vec}   for gcoo_batch_add, where
      class = (N bucket, entries-per-row bucket) of a job (searched with
      every job on the persistent kernel, the kernel these tiles are for);
  batch_panel[dtype] = {mode, ratio}  for gcoo_batch_set_panel_mode: which
      jobs of a batch run the panel-staged kernel instead, searched on the
      whole many-jobs workload with the tuned tiles;
  plain[dtype][class] = {variant, vec} for gcoo_set_spdm_config;
  skew[dtype][class]  = {cpg, variant, vec, unroll, chunk, waves} for matrices
      with skewed rows, class = (N bucket, entries-per-row bucket, skew
      bucket = longest row over the mean row): the group width (the
      reference's COLUMN_PER_GROUP, src/constants.h.t:4, grid-searched by
      src/autotune.py:89-97), the variant and its tile, searched on the
      skewed workload itself (65536 x 65536, Zipf(1.5) rows of mean 16,
      N = 256).

Run:  python -m gcoospdm_amd.autotune [--out gcoospdm_amd/tuning.json]
The table shipped in the package was produced by that command on an MI355X.
"""
from __future__ import annotations

import argparse
import json
import os
import time

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(_HERE, "tuning.json")

N_BUCKETS = [32, 64, 128, 256]            # upper bounds (last open-ended)
ROW_BUCKETS = [8, 32, 128]                # entries per row: <8, <32, <128, >=128


def n_class(N: int) -> int:
    for i, b in enumerate(N_BUCKETS):
        if N <= b:
            return i
    return len(N_BUCKETS) - 1


def row_class(per_row: float) -> int:
    for i, b in enumerate(ROW_BUCKETS):
        if per_row < b:
            return i
    return len(ROW_BUCKETS)


def class_key(N: int, per_row: float) -> str:
    return f"n{n_class(N)}r{row_class(per_row)}"


SKEW_BUCKETS = [4, 32]                    # longest row / mean row: <4, <32, >=32


def skew_class(max_row: float, per_row: float) -> int:
    ratio = max_row / max(per_row, 1e-9)
    for i, b in enumerate(SKEW_BUCKETS):
        if ratio < b:
            return i
    return len(SKEW_BUCKETS)


def skew_key(N: int, per_row: float, max_row: float) -> str:
    return f"n{n_class(N)}r{row_class(per_row)}s{skew_class(max_row, per_row)}"


def skew_config(N: int, per_row: float, max_row: float, bf16: bool, table: dict | None = None):
    """Tuned {cpg, variant, vec, unroll, chunk, waves} for a skewed matrix, or None."""
    t = (table if table is not None else load()).get("skew", {}).get("bf16" if bf16 else "f32", {})
    return t.get(skew_key(N, per_row, max_row))


_table_cache = {}


def load(path: str = DEFAULT_TABLE) -> dict:
    if path not in _table_cache:
        if os.path.exists(path):
            with open(path) as f:
                _table_cache[path] = json.load(f)
        else:
            _table_cache[path] = {}
    return _table_cache[path]


def batch_tile(N: int, per_row: float, bf16: bool, table: dict | None = None):
    """(rows, vec) for a job of width N with `per_row` entries per row;
    (0, 0) = the library's built-in choice."""
    t = (table if table is not None else load()).get("batch", {}).get("bf16" if bf16 else "f32", {})
    e = t.get(class_key(N, per_row))
    return (e["rows"], e["vec"]) if e else (0, 0)


def batch_panel(bf16: bool, table: dict | None = None):
    """(mode, ratio) for Batch.set_panel_mode; (1, 0.0) = the library's
    built-in choice."""
    t = (table if table is not None else load()).get("batch_panel", {}).get(
        "bf16" if bf16 else "f32")
    return (t["mode"], t["ratio"]) if t else (1, 0.0)


def plain_config(N: int, per_row: float, bf16: bool, table: dict | None = None):
    t = (table if table is not None else load()).get("plain", {}).get("bf16" if bf16 else "f32", {})
    e = t.get(class_key(N, per_row))
    return (e["variant"], e["vec"]) if e else (0, 0)


def _time(fn, reps=5, trials=3, warm=5):
    """Median over trials of the mean of reps launches, after warm untimed
    launches (the clock runs slow for the first launches after an idle gap)."""
    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    ts = []
    for _ in range(trials):
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_time(e1) / reps)
    return sorted(ts)[len(ts) // 2]


def tune_batch(jobs_by_class, bf16: bool, log):
    """jobs_by_class: {key: [(m, D, C, N), ...]} -> {key: {rows, vec}}"""
    import gcoospdm_amd as G
    out = {}
    vecs = [2, 4, 8] if bf16 else [1, 2, 4]
    for key, jobs in sorted(jobs_by_class.items()):
        best = None
        for rows in (1, 2, 4, 8, 16):
            for vec in vecs:
                b = G.Batch().set_panel_mode(0)
                for m, D, C, N in jobs:
                    b.add(m, D, C, rows=rows, vec=vec)
                b.commit()
                ms = _time(b.run)
                log(dict(kind="batch", dtype="bf16" if bf16 else "f32", cls=key, rows=rows,
                         vec=vec, ms=round(ms, 4), jobs=len(jobs)))
                if best is None or ms < best[0]:
                    best = (ms, rows, vec)
                b.close()
        out[key] = {"rows": best[1], "vec": best[2], "ms": round(best[0], 4)}
    return out


def tune_batch_panel(specs, bf16: bool, log, table: dict):
    """The panel-job choice of a batch, searched on the whole workload
    (every job materialised, tuned tiles): mode 0 (none), mode 1 at several
    ratios, mode 2 (every eligible job); -> {mode, ratio, ms}."""
    import torch
    import gcoospdm_amd as G
    from . import workloads as W
    jobs = [W.materialize(sp, bf16=bf16) for sp in specs]
    cands = [(0, 0.0)] + [(1, r) for r in (0.5, 0.7, 1.0, 1.4, 2.0, 4.0, 8.0, 16.0)] + [(2, 0.0)]
    best, ref = None, None
    stream = torch.cuda.Stream()
    for mode, ratio in cands:
        b = G.Batch().set_panel_mode(mode, ratio, ratio)
        for (m, D, C), sp in zip(jobs, specs):
            rows, vec = batch_tile(sp.N, sp.K * (1 - sp.sparsity), bf16, table)
            b.add(m, D, C, rows=rows, vec=vec)
        b.commit()
        b.capture(stream)
        with torch.cuda.stream(stream):
            ms = _time(lambda: b.replay(stream), reps=5)
        out = [C.clone() for _, _, C in jobs]
        if ref is None:
            ref = out
        same = all(torch.equal(x.view(torch.int32), y.view(torch.int32)) for x, y in zip(out, ref))
        info = b.panel_info()
        log(dict(kind="batch_panel", dtype="bf16" if bf16 else "f32", mode=mode, ratio=ratio,
                 ms=round(ms, 4), panel_jobs=info["panel_jobs"], same=same))
        if not same:
            raise RuntimeError(f"batch panel mode {mode} ratio {ratio} changed the bits")
        if best is None or ms < best[0]:
            best = (ms, mode, ratio)
        b.close()
    return {"mode": best[1], "ratio": best[2], "ms": round(best[0], 4)}


def tune_plain(bf16: bool, log, device="cuda:0"):
    """Representative square-ish shapes per class; variant 2 vs 3, vec."""
    import torch
    import gcoospdm_amd as G
    out = {}
    shapes = {  # class key -> (M, K, N, sparsity)
        class_key(32, 4): (4096, 4096, 32, 0.999), class_key(256, 4): (4096, 4096, 256, 0.999),
        class_key(256, 20): (4096, 4096, 256, 0.995), class_key(256, 80): (8000, 8000, 256, 0.99),
        class_key(64, 20): (4096, 4096, 64, 0.995), class_key(128, 20): (4096, 4096, 128, 0.995),
        class_key(256, 400): (4096, 4096, 256, 0.9),
    }
    vecs = [2, 4, 8] if bf16 else [1, 2, 4]
    for key, (M, K, N, s) in shapes.items():
        S = torch.empty((M, K), device=device)
        G.fill_sparse(S, 5, 1, s)
        m = G.GCOOMatrix(M, K, 4).convert_dense(S)
        del S
        D = torch.empty((K, N), device=device)
        G.fill_uniform(D, 5, 0)
        if bf16:
            D = G.to_bf16(D)
        C = torch.empty((M, N), device=device)
        best = None
        for variant in (2, 3):
            for vec in vecs:
                m.set_config(variant=variant, vec=vec)
                fn = (lambda: m.spdm_bf16(D, C, sync=False)) if bf16 else \
                     (lambda: m.spdm(D, C, sync=False))
                ms = _time(fn)
                log(dict(kind="plain", dtype="bf16" if bf16 else "f32", cls=key,
                         variant=variant, vec=vec, ms=round(ms, 4)))
                if best is None or ms < best[0]:
                    best = (ms, variant, vec)
        out[key] = {"variant": best[1], "vec": best[2], "ms": round(best[0], 4)}
    return out


def tune_skew(bf16: bool, log, device="cuda:0", seed=11):
    """Grid search on the skewed workload: cpg in {1, 2, 4} x variants
    {2, 4, 6} x their tile parameters; every candidate's C is checked
    bit-equal to the first one's."""
    import numpy as np
    import torch
    import gcoospdm_amd as G
    from . import workloads as W
    M = K = 65536
    N = 256
    rows, cols, vals = W.zipf_coo(M, K, seed=seed)
    counts = np.bincount(rows, minlength=M)
    per_row, max_row = float(counts.mean()), float(counts.max())
    r, c, v = (torch.from_numpy(x).to(device) for x in (rows, cols, vals))
    D = torch.from_numpy(np.random.default_rng(seed + 1).standard_normal((K, N))
                         .astype(np.float32)).to(device)
    if bf16:
        D = G.to_bf16(D)
    C = torch.empty((M, N), device=device)
    cands = []
    vecs = [2, 4, 8] if bf16 else [1, 2, 4]
    for vec in vecs:
        for unroll in (8, 16, 32):
            for chunk in (0, 1):
                # waves: the long-rows-first threshold of variant 2 (0
                # automatic, 1 off, w: rows of 2^w entries or more first)
                for waves in (0, 1, 8):
                    cands.append(dict(variant=2, vec=vec, unroll=unroll, chunk=chunk,
                                      waves=waves))
        for chunk in (64, 256, 1024):
            cands.append(dict(variant=4, vec=vec, unroll=16, chunk=chunk, waves=0))
    cands.append(dict(variant=6, vec=4, unroll=0, chunk=0, waves=0))
    best, ref = None, None
    for cpg in (1, 2, 4):
        m = G.GCOOMatrix(M, K, cpg).convert_coo(r, c, v)
        for cfg in cands:
            m.set_config(**cfg)
            fn = (lambda: m.spdm_bf16(D, C, sync=False)) if bf16 else \
                 (lambda: m.spdm(D, C, sync=False))
            ms = _time(fn, reps=5)
            if ref is None:
                ref = C.clone()
            same = bool(torch.equal(C.view(torch.int32), ref.view(torch.int32)))
            log(dict(kind="skew", dtype="bf16" if bf16 else "f32", cpg=cpg, ms=round(ms, 4),
                     same=same, **cfg))
            if not same:
                raise RuntimeError(f"skew candidate cpg={cpg} {cfg} changed the bits")
            if best is None or ms < best[0]:
                best = (ms, cpg, cfg)
        m.close()
    key = skew_key(N, per_row, max_row)
    return {key: dict(cpg=best[1], ms=round(best[0], 4), **best[2])}


def _jobs_by_class(specs, bf16: bool, per_class: int):
    """Up to per_class materialised jobs of each (N, entries-per-row) class."""
    from . import workloads as W
    by_class = {}
    for sp in specs:
        key = class_key(sp.N, sp.K * (1.0 - sp.sparsity))
        lst = by_class.setdefault(key, [])
        if len(lst) < per_class:
            m, D, C = W.materialize(sp, bf16=bf16)
            lst.append((m, D, C, sp.N))
    return by_class


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--jobs", type=int, default=512)
    ap.add_argument("--per-class", type=int, default=24)
    ap.add_argument("--log", default=None)
    ap.add_argument("--skew-only", action="store_true",
                    help="rerun only the skew search, keeping the other tables")
    ap.add_argument("--batch-only", action="store_true",
                    help="rerun only the batch tile search, keeping the other tables")
    ap.add_argument("--batch-panel-only", action="store_true",
                    help="rerun only the batch panel-job search, keeping the other tables")
    args = ap.parse_args()
    from . import workloads as W

    logf = open(args.log, "a") if args.log else None

    def log(rec):
        line = json.dumps(rec)
        print(line, flush=True)
        if logf:
            logf.write(line + "\n")

    t0 = time.time()
    if args.skew_only:
        table = dict(load(args.out)) if os.path.exists(args.out) else {}
        table["skew"] = {d: tune_skew(d == "bf16", log) for d in ("f32", "bf16")}
        with open(args.out, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        print(f"wrote {args.out}", flush=True)
        return
    specs = W.batch_jobs(args.jobs, 2024)
    if args.batch_panel_only:
        table = dict(load(args.out)) if os.path.exists(args.out) else {}
        table["batch_panel"] = {d: tune_batch_panel(specs, d == "bf16", log, table)
                                for d in ("f32", "bf16")}
        with open(args.out, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        print(f"wrote {args.out}", flush=True)
        return
    if args.batch_only:
        table = dict(load(args.out)) if os.path.exists(args.out) else {}
        table["batch"] = {}
        for bf16 in (False, True):
            table["batch"]["bf16" if bf16 else "f32"] = tune_batch(
                _jobs_by_class(specs, bf16, args.per_class), bf16, log)
        with open(args.out, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        print(f"wrote {args.out}", flush=True)
        return
    table = {"device": "MI355X (gfx950)", "batch": {}, "batch_panel": {}, "plain": {}, "skew": {}}
    for bf16 in (False, True):
        by_class = _jobs_by_class(specs, bf16, args.per_class)
        table["batch"]["bf16" if bf16 else "f32"] = tune_batch(by_class, bf16, log)
        del by_class
        table["batch_panel"]["bf16" if bf16 else "f32"] = tune_batch_panel(specs, bf16, log, table)
        table["plain"]["bf16" if bf16 else "f32"] = tune_plain(bf16, log)
        table["skew"]["bf16" if bf16 else "f32"] = tune_skew(bf16, log)
    table["seconds"] = round(time.time() - t0, 1)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
