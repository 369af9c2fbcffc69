"""Roofline performance model of the GCOO multiply, restated for MI355X.

The reference models its kernels with a plain roofline
(scripts/perf_model.py:206-218, `predict_perf`): with peak bandwidth BW
(GB/s) and peak single-precision rate P (GFLOP/s), the ridge is rc = P / BW;
a kernel of operational intensity r FLOP/B runs at 0.85 P when r >= rc and
at BW * r otherwise.  Its intensities (scripts/perf_model.py:368-381) are

    GCOO SpDM   r = 2 N (1 - s) / 4
    cuSPARSE    r = 2 N (1 - s) / ((1 - s) N) = 2
    cuBLAS      r = N / 6

for an N x N x N product at sparsity s, and the predicted time is the
product's flop count (2 (1 - s) N^3 for the sparse ones, 2 N^3 for cuBLAS)
over the predicted rate.  The device table keeps the reference's three GPUs
(scripts/perf_model.py:206-207) and adds MI355X (8000 GB/s HBM3E, 157.3
TFLOP/s FP32 vector).

`hbm_bound_time` is the bound this build measures against instead: the
algorithmic bytes of the multiply (12 B per entry, 8 B per group, the K x N
operand and the M x N result, all in FP32) over the HBM peak.
"""
from __future__ import annotations

# (peak bandwidth GB/s, peak FP32 GFLOP/s)
DEVICES = {
    "980": (224.0, 4981.0),
    "titanx": (433.0, 10970.0),
    "p100": (732.0, 9500.0),
    "mi355x": (8000.0, 157300.0),
}
EFFICIENCY_AT_PEAK = 0.85


def predict_perf(r: float, device: str = "mi355x") -> float:
    """Predicted rate in FLOP/s at operational intensity r (FLOP/B)."""
    bw, flops = DEVICES[device]
    rc = flops / bw
    if r >= rc:
        return flops * 1e9 * EFFICIENCY_AT_PEAK
    return bw * r * 1e9


def gcoo_intensity(s: float, N: int) -> float:
    return 2.0 * N * (1.0 - s) / 4.0


def time_of_s(s: float, N: int, device: str = "mi355x"):
    """(seconds, intensity) of the GCOO multiply of an N^3 product at sparsity s."""
    r = gcoo_intensity(s, N)
    return 2.0 * (1.0 - s) * N ** 3 / predict_perf(r, device), r


def time_of_s_cusparse(s: float, N: int, device: str = "mi355x"):
    r = 2.0 * N * (1.0 - s) / ((1.0 - s) * N)
    return 2.0 * (1.0 - s) * N ** 3 / predict_perf(r, device), r


def time_of_cublas(N: int, device: str = "mi355x"):
    r = N / 6.0
    return 2.0 * N ** 3 / predict_perf(r, device), r


def spdm_bytes(nnz: float, M: int, K: int, N: int, cpg: int = 4) -> float:
    ng = (M + cpg - 1) // cpg
    return 12.0 * nnz + 8.0 * (ng + 1) + 4.0 * K * N + 4.0 * M * N


def hbm_bound_time(s: float, N: int, device: str = "mi355x") -> float:
    """Seconds to move the multiply's algorithmic bytes at peak bandwidth
    (expected nnz = (1 - s) N^2)."""
    bw, _ = DEVICES[device]
    return spdm_bytes((1.0 - s) * N * N, N, N, N) / (bw * 1e9)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="predicted GCOO / cuSPARSE / cuBLAS times")
    ap.add_argument("--ns", default="8000")
    ap.add_argument("--sparsities", default="0.99,0.995,0.999")
    ap.add_argument("--device", default="mi355x", choices=sorted(DEVICES))
    a = ap.parse_args(argv)
    print("N\tsparsity\tr\tgcoospdm_ms\tcusparse_ms\tcublas_ms\thbm_bound_ms")
    for n in (int(x) for x in a.ns.split(",")):
        for s in (float(x) for x in a.sparsities.split(",")):
            t, r = time_of_s(s, n, a.device)
            tc, _ = time_of_s_cusparse(s, n, a.device)
            tb, _ = time_of_cublas(n, a.device)
            print(f"{n}\t{s:.4f}\t{r:.2f}\t{t * 1e3:.6f}\t{tc * 1e3:.6f}\t{tb * 1e3:.6f}\t"
                  f"{hbm_bound_time(s, n, a.device) * 1e3:.6f}")


if __name__ == "__main__":
    main()
