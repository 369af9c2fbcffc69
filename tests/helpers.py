"""Independent restatements used to pin the oracle (tests only).

Nothing here calls the oracle: these are brute-force / exact-arithmetic
re-derivations written in numpy and fractions.Fraction, so that a mistake in
the C oracle (a dropped term, a wrong order, a transposed operand) shows up
as a mismatch."""
from __future__ import annotations

from fractions import Fraction

import numpy as np

GAMMA = 0x9E3779B97F4A7C15
MASK64 = (1 << 64) - 1


def brute_gcoo(S, cpg):
    """GCOO arrays by enumeration + sort: entries of rows [cpg*g, cpg*g+cpg)
    with S[r][k] != 0.0, ordered by (g, k, r)."""
    S = np.asarray(S, np.float32)
    M, K = S.shape
    ng = (M + cpg - 1) // cpg
    r, k = np.nonzero(S != 0.0)   # NaN != 0 is True, -0.0 != 0 is False
    g = r // cpg
    order = np.lexsort((r, k, g))
    r, k, g = r[order], k[order], g[order]
    npg = np.zeros(ng + 1, np.int32)
    np.add.at(npg, g, 1)
    gi = np.zeros(ng + 1, np.int32)
    gi[1:] = np.cumsum(npg[:-1])
    return dict(vals=S[r, k].astype(np.float32), rows=r.astype(np.int32),
                cols=k.astype(np.int32), groupIndex=gi, nnzPerGroup=npg,
                nGroup=ng, nnz=int(r.size), M=M, K=K, cpg=cpg)


def _round_f32(x: Fraction) -> np.float32:
    """Exact rational -> nearest float32, ties to even (finite range only)."""
    f = np.float32(float(x))
    best = f
    for cand in (np.nextafter(f, np.float32(-np.inf)), np.nextafter(f, np.float32(np.inf))):
        if not np.isfinite(cand):
            continue
        dc, db = abs(Fraction(float(cand)) - x), abs(Fraction(float(best)) - x)
        if dc < db or (dc == db and (int(cand.view(np.uint32)) & 1) == 0
                       and (int(best.view(np.uint32)) & 1) == 1):
            best = cand
    return np.float32(best)


def fmaf_exact(a, b, c) -> np.float32:
    """IEEE fusedMultiplyAdd in float32 from exact rational arithmetic."""
    a, b, c = np.float32(a), np.float32(b), np.float32(c)
    x = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    if x == 0:
        # exact zero: sign per IEEE 754 (round to nearest): +0 unless both
        # addends are -0 (a*b is -0 when signs differ and one factor is 0)
        prod_neg = (np.signbit(a) != np.signbit(b))
        prod_zero = float(a) == 0.0 or float(b) == 0.0
        if prod_zero and prod_neg and np.signbit(c) and float(c) == 0.0:
            return np.float32(-0.0)
        return np.float32(0.0)
    return _round_f32(x)


def spdm_fmaf_chain(S, D):
    """C[r][j] = fmaf chain from +0.0f over row r's nonzeros, ascending k,
    evaluated with exact rational fmaf.  Tiny inputs only."""
    S = np.asarray(S, np.float32)
    D = np.asarray(D, np.float32)
    M, K = S.shape
    _, N = D.shape
    C = np.zeros((M, N), np.float32)
    for r in range(M):
        ks = [k for k in range(K) if S[r, k] != 0.0]
        for j in range(N):
            acc = np.float32(0.0)
            for k in ks:
                acc = fmaf_exact(S[r, k], D[k, j], acc)
            C[r, j] = acc
    return C


def mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def counter_fill(n, seed, stream_id, scale=1.0, keep_thr=0, sparse=False):
    """numpy restatement of gcoo_fill_*: see include/gcoo_spdm.h."""
    key = mix64((seed * GAMMA + stream_id) & MASK64)
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(key) + i * np.uint64(GAMMA))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        h = z ^ (z >> np.uint64(31))
    num = 2 * (h >> np.uint64(40)).astype(np.int64) - (1 << 24) + 1
    v = (num.astype(np.float32) * np.float32(2.0 ** -24)) * np.float32(scale)
    if sparse:
        low = h & np.uint64((1 << 40) - 1)
        v = np.where(low < np.uint64(keep_thr), np.float32(0.0), v).astype(np.float32)
    return v.astype(np.float32)


def bits_equal(a, b):
    """Bitwise equality of float32 arrays (NaN payloads included)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def rng_sparse(rng, M, K, density, lo=-1.0, hi=1.0):
    S = rng.uniform(lo, hi, size=(M, K)).astype(np.float32)
    S[rng.random((M, K)) >= density] = 0.0
    return S
