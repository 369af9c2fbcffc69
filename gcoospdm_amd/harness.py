"""The reference's timing harness (src/bench.py:12-20,52-133 with
src/perf_compare.py:40-54), driving this build's `main`: one run of
`main -algo=<algo>` per (n, sparsity) or per Matrix Market file, the
"Time:" line (microseconds per call) read back as milliseconds, one
tab-separated row per case with the reference's columns:

    N  sparsity  gcoospdm  cusparse  cublas            (synthetic)
    N  sparsity  filename  gcoospdm  cusparse  cublas  (files)

Synthetic rows end with two more columns: the reference's roofline model of
the GCOO multiply restated for MI355X (perf_model.time_of_s, in ms) and the
HBM bound of its algorithmic bytes (perf_model.hbm_bound_time), beside the
measured time.  A failed run is written as 10000000.0, as the reference
does.  Usage:

    python -m gcoospdm_amd.harness --ns 8000 --sparsities 0.99,0.995,0.999
    python -m gcoospdm_amd.harness --files <dir with .mtx files> --kernel-only
"""
import argparse
import fnmatch
import os
import subprocess
import sys

MAIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host", "main")
FAILED = 10000000.0


def run_main(args, timeout=1800):
    """Milliseconds per call from `main`'s "Time:" line, or FAILED."""
    try:
        r = subprocess.run([MAIN] + args, capture_output=True, text=True, timeout=timeout)
    except (OSError, subprocess.TimeoutExpired):
        return FAILED
    if r.returncode != 0:
        return FAILED
    for line in r.stdout.splitlines():
        if line.startswith("Time:"):
            return float(line.split(":", 1)[1]) / 1000.0
    return FAILED


def synthetic(ns, sparsities, algos, kernel_only=False, iters=None, dist="ref", out=sys.stdout):
    from . import perf_model
    out.write("N\tsparsity\t" + "\t".join(algos) + "\tgcoospdm_model\thbm_bound\n")
    for n in ns:
        for s in sparsities:
            ms = []
            for algo in algos:
                a = [f"-sparsity={s:f}", f"-nRowsA={n}", f"-nRowsB={n}", f"-nColsA={n}",
                     f"-nColsB={n}", f"-algo={algo}", f"-dist={dist}"]
                if kernel_only:
                    a.append("-kernelonly=1")
                if iters:
                    a.append(f"-iters={iters}")
                ms.append(run_main(a))
            model = perf_model.time_of_s(s, n)[0] * 1e3
            bound = perf_model.hbm_bound_time(s, n) * 1e3
            out.write(f"{n}\t{s:.4f}\t" + "\t".join(f"{t:.6f}" for t in ms)
                      + f"\t{model:.6f}\t{bound:.6f}\n")
            out.flush()


def mtx_size(path):
    with open(path) as f:
        for line in f:
            if not line.startswith("%"):
                return [int(x) for x in line.split()]
    return [0, 0, 0]


def public_dataset(paths, algos, kernel_only=False, out=sys.stdout):
    out.write("N\tsparsity\tfilename\t" + "\t".join(algos) + "\n")
    for p in paths:
        size = mtx_size(p)
        if size[0] != size[1] or size[0] > 50000 or size[0] <= 1:
            continue
        ms = []
        for algo in algos:
            a = [f"-file={p}", f"-algo={algo}"] + (["-kernelonly=1"] if kernel_only else [])
            ms.append(run_main(a))
        out.write(f"{size[0]}\t{size[2] / (size[0] * size[1]):.9f}\t{os.path.basename(p)}\t"
                  + "\t".join(str(t) for t in ms) + "\n")
        out.flush()


def find_mtx(root):
    found = []
    for d, _, files in os.walk(root):
        found += [os.path.join(d, f) for f in fnmatch.filter(files, "*.mtx")]
    return sorted(found)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--ns", default="8000")
    ap.add_argument("--sparsities", default="0.99,0.995,0.999")
    ap.add_argument("--algos", default="gcoospdm,cusparse,cublas")
    ap.add_argument("--files", default=None, help="directory searched for .mtx files")
    ap.add_argument("--kernel-only", action="store_true")
    ap.add_argument("--iters", type=int, default=None)
    ap.add_argument("--dist", default="ref", choices=["ref", "uniform"])
    a = ap.parse_args(argv)
    algos = a.algos.split(",")
    if a.files:
        public_dataset(find_mtx(a.files), algos, a.kernel_only)
    else:
        synthetic([int(x) for x in a.ns.split(",")], [float(x) for x in a.sparsities.split(",")],
                  algos, a.kernel_only, a.iters, a.dist)


if __name__ == "__main__":
    main()
