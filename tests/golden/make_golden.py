"""Writes the small golden fixtures of this directory.

The reference ships no golden vectors for its GCOO path, so these fixtures
are derived from the definition of the format and of the arithmetic,
independently of the C oracle: enumeration + sort for the GCOO arrays,
exact-rational fmaf chains for the FP32 product, numpy float64 for the FP64
product (tests/helpers.py).  Inputs are seeded numpy draws plus hand-placed
edge values (-0.0, NaN, denormals).  Re-run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from tests import helpers as H  # noqa: E402


def write(name, S, D, cpg):
    g = H.brute_gcoo(S, cpg)
    C = H.spdm_fmaf_chain(S, D)
    Sd = S.astype(np.float64)
    Sd[np.isnan(Sd)] = 0.0
    C64 = np.zeros((S.shape[0], D.shape[1]), np.float64)
    for r in range(S.shape[0]):
        for k in range(S.shape[1]):
            if S[r, k] != 0.0:
                C64[r] += np.float64(S[r, k]) * D[k].astype(np.float64)
    np.savez(os.path.join(HERE, name), S=S, D=D, cpg=np.int32(cpg), vals=g["vals"],
             rows=g["rows"], cols=g["cols"], groupIndex=g["groupIndex"],
             nnzPerGroup=g["nnzPerGroup"], C=C, C64=C64)


def main():
    rng = np.random.default_rng(20171101)
    S = H.rng_sparse(rng, 10, 12, 0.3)
    D = rng.uniform(-1, 1, (12, 6)).astype(np.float32)
    write("g_10x12x6_cpg4.npz", S, D, 4)
    S = H.rng_sparse(rng, 9, 8, 0.5)
    D = rng.uniform(-1, 1, (8, 5)).astype(np.float32)
    write("g_9x8x5_cpg2.npz", S, D, 2)
    S = H.rng_sparse(rng, 6, 7, 0.4)
    S[0, 0] = -0.0
    S[1, 2] = np.float32(3e-39)   # denormal
    S[5, 6] = np.float32(-7e-42)
    D = rng.uniform(-1, 1, (7, 4)).astype(np.float32)
    write("g_6x7x4_edges_cpg4.npz", S, D, 4)


if __name__ == "__main__":
    main()
