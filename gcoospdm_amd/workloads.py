"""Synthetic workloads of the benchmark configurations.

The reference draws its inputs with glibc rand() on the host
(src/utils.cpp:5-19, src/main.cu:103-121); these generators produce the
configurations this library is measured on, seeded and reproducible:

  square(n, s)   n x n S with i.i.d. Bernoulli(1 - s) nonzeros, U(-1,1)
                 values, and a dense n x n U(-1,1) D, generated in HBM by the
                 library's counter-based kernels (gcoo_fill_*).
  zipf_coo(...)  a large S built from COO: per-row nnz from a Zipf(a) law
                 rescaled to a target mean and capped, uniform-random distinct
                 columns, N(0,1) values; D dense N(0,1).
  batch_jobs(..) the specs of many independent multiplies: M, K log-uniform
                 in [256, 8192], N from {32, 64, 128, 256}, sparsity uniform
                 in [0.95, 0.999], one seed per job (master ^ index).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import GCOOMatrix, fill_sparse, fill_uniform, to_bf16


def square(n: int, sparsity: float, seed: int = 1, device="cuda:0"):
    import torch
    S = torch.empty((n, n), device=device)
    D = torch.empty((n, n), device=device)
    fill_sparse(S, seed, 1, sparsity)
    fill_uniform(D, seed, 0)
    return S, D


def zipf_row_counts(M: int, a: float = 1.5, mean: float = 16.0, cap: int = 4096, seed: int = 1):
    """Per-row nonzero counts: x ~ Zipf(a) truncated at `cap`, rescaled so the
    truncated law's mean becomes `mean`: clamp(round(x * mean / mu), 1, cap)
    with mu the mean of the truncated law (49.03 for a = 1.5, cap = 4096)."""
    ks = np.arange(1, cap + 1, dtype=np.float64)
    w = ks ** (-a)
    w /= w.sum()
    mu = float((ks * w).sum())
    rng = np.random.default_rng(seed)
    x = rng.choice(ks, size=M, p=w)
    return np.clip(np.rint(x * mean / mu), 1, cap).astype(np.int64), mu


def zipf_coo(M: int, K: int, a: float = 1.5, mean: float = 16.0, cap: int = 4096, seed: int = 1):
    """Host COO arrays (rows, cols, vals) of the skewed matrix, rows sorted,
    columns distinct within a row (uniform without replacement)."""
    counts, _ = zipf_row_counts(M, a, mean, cap, seed)
    rng = np.random.default_rng(seed + 1)
    rows = np.repeat(np.arange(M, dtype=np.int32), counts)
    cols = np.empty(rows.size, np.int32)
    pos = 0
    for r, c in enumerate(counts):
        c = int(c)
        if c * 4 < K:
            # rejection-free enough: draw, dedupe, top up
            got = np.unique(rng.integers(0, K, size=c + 8))
            while got.size < c:
                got = np.unique(np.concatenate([got, rng.integers(0, K, size=c)]))
            sel = rng.permutation(got)[:c]
        else:
            sel = rng.choice(K, size=c, replace=False)
        cols[pos:pos + c] = sel
        pos += c
    vals = rng.standard_normal(rows.size).astype(np.float32)
    vals[vals == 0.0] = 1.0   # a drawn exact zero would be dropped by the format
    return rows, cols, vals


@dataclass
class JobSpec:
    index: int
    M: int
    K: int
    N: int
    sparsity: float
    seed: int

    @property
    def expected_nnz(self) -> float:
        return self.M * self.K * (1.0 - self.sparsity)


def batch_jobs(count: int = 512, master_seed: int = 2024):
    rng = np.random.default_rng(master_seed)
    lo, hi = math.log(256), math.log(8192)
    specs = []
    for i in range(count):
        M = int(round(math.exp(rng.uniform(lo, hi))))
        K = int(round(math.exp(rng.uniform(lo, hi))))
        N = int(rng.choice([32, 64, 128, 256]))
        s = float(rng.uniform(0.95, 0.999))
        specs.append(JobSpec(i, M, K, N, s, master_seed ^ i))
    return specs


def materialize(spec: JobSpec, device="cuda:0", bf16: bool = False):
    """Device operands of one job: (GCOOMatrix converted from a dense S, D
    (FP32 or BF16), C).  S is generated densely then dropped."""
    import torch
    S = torch.empty((spec.M, spec.K), device=device)
    fill_sparse(S, spec.seed, 1, spec.sparsity)
    m = GCOOMatrix(spec.M, spec.K, 4).convert_dense(S)
    del S
    D = torch.empty((spec.K, spec.N), device=device)
    fill_uniform(D, spec.seed, 0)
    if bf16:
        D = to_bf16(D)
    C = torch.empty((spec.M, spec.N), device=device)
    return m, D, C


def spdm_bytes(nnz, M, K, N, bf16=False, cpg=4):
    """Algorithmic HBM bytes of one multiply (see the roofline notes)."""
    ng = (M + cpg - 1) // cpg
    if bf16:
        return 10 * nnz + 8 * (ng + 1) + 2 * K * N + 4 * M * N
    return 12 * nnz + 8 * (ng + 1) + 4 * K * N + 4 * M * N
