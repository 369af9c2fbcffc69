"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against an independent derivation
(tests/helpers.py), the committed golden fixtures (tests/golden/), the
reference's own matrixMulCPU compiled from its sources (oracle/_ref), or a
mathematical identity -- chosen so that a plausible mistake fails one of
them.  CPU only."""
import os

import numpy as np
import pytest

import oracle
from tests import helpers as H

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _edge_matrix():
    """Rows covering -0.0 (dropped), NaN (kept), denormals (kept), inf."""
    S = np.zeros((7, 9), np.float32)
    S[0, 1] = -0.0
    S[0, 2] = 1.5
    S[1, 0] = np.nan
    S[2, 8] = np.float32(1e-40)         # denormal
    S[3, 3] = -np.float32(1e-45)        # smallest denormal
    S[4, 4] = np.inf
    S[6, 0] = -2.0
    S[6, 8] = 3.0
    return S


CASES = [
    ("small", 12, 10, 0.3),
    ("ragged_M", 13, 7, 0.5),     # M % 4 != 0
    ("tall_sparse", 41, 5, 0.05),  # empty groups
    ("dense", 9, 11, 1.0),
    ("empty", 8, 6, 0.0),
    ("one_row", 1, 17, 0.4),
    ("one_col", 15, 1, 0.6),
]


@pytest.mark.parametrize("cpg", [1, 2, 4])
@pytest.mark.parametrize("name,M,K,dens", CASES)
def test_convert_matches_brute_force(name, M, K, dens, cpg):
    rng = np.random.default_rng(hash((name, cpg)) & 0xFFFF)
    S = H.rng_sparse(rng, M, K, dens)
    got = oracle.convert(S, cpg)
    exp = H.brute_gcoo(S, cpg)
    for key in ("rows", "cols", "groupIndex", "nnzPerGroup"):
        np.testing.assert_array_equal(got[key], exp[key], err_msg=key)
    assert H.bits_equal(got["vals"], exp["vals"])
    assert got["nnz"] == exp["nnz"] == int(np.count_nonzero(S != 0.0))


@pytest.mark.parametrize("cpg", [1, 2, 4])
def test_convert_edge_values(cpg):
    S = _edge_matrix()
    got = oracle.convert(S, cpg)
    exp = H.brute_gcoo(S, cpg)
    for key in ("rows", "cols", "groupIndex", "nnzPerGroup"):
        np.testing.assert_array_equal(got[key], exp[key], err_msg=key)
    assert H.bits_equal(got["vals"], exp["vals"])
    # -0.0 dropped, NaN / denormals / inf kept
    assert got["nnz"] == 7
    assert not np.any((got["rows"] == 0) & (got["cols"] == 1))


def test_convert_invariants():
    rng = np.random.default_rng(5)
    S = H.rng_sparse(rng, 37, 29, 0.2)
    g = oracle.convert(S, 4)
    gi, npg = g["groupIndex"], g["nnzPerGroup"]
    assert npg[-1] == 0 and gi[0] == 0 and gi[-1] == g["nnz"]
    assert np.all(np.diff(gi) == npg[:-1])
    for grp in range(g["nGroup"]):
        a, b = gi[grp], gi[grp + 1]
        r, k = g["rows"][a:b], g["cols"][a:b]
        assert np.all(r // 4 == grp)
        key = k.astype(np.int64) * 64 + r
        assert np.all(np.diff(key) > 0)   # strictly increasing (k, r)
        assert np.all(S[r, k] != 0.0)


def test_scan_closed_form():
    src = np.array([3, 0, 5, 1, 0], np.int32)
    np.testing.assert_array_equal(oracle.exclusive_scan(src), [0, 3, 3, 8, 9])
    rng = np.random.default_rng(1)
    src = rng.integers(0, 100, 1000).astype(np.int32)
    exp = np.concatenate([[0], np.cumsum(src)[:-1]]).astype(np.int32)
    np.testing.assert_array_equal(oracle.exclusive_scan(src), exp)


@pytest.mark.parametrize("cpg", [1, 2, 4])
@pytest.mark.parametrize("M,K,N,dens", [(6, 9, 5, 0.4), (9, 4, 3, 0.8), (5, 12, 7, 0.25)])
def test_spdm_f32_bit_exact_vs_exact_fmaf(M, K, N, dens, cpg):
    rng = np.random.default_rng(M * 100 + K * 10 + N)
    S = H.rng_sparse(rng, M, K, dens)
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    g = oracle.convert(S, cpg)
    got = oracle.spdm_f32(g, D)
    exp = H.spdm_fmaf_chain(S, D)
    assert H.bits_equal(got, exp)


def test_spdm_f32_fma_not_mul_add():
    """A chain where fmaf and round(round(a*b) + c) differ: a mul+add oracle
    fails it.  c1 = -(1 + 2^-11); fma(1 + 2^-12, 1 + 2^-12, c1) = 2^-24
    exactly, while round(a*b) ties to 1 + 2^-11 and mul+add gives 0."""
    S = np.array([[-(1.0 + 2.0 ** -11), 1.0 + 2.0 ** -12]], np.float32)
    D = np.array([[1.0], [1.0 + 2.0 ** -12]], np.float32)
    got = oracle.spdm_f32(oracle.convert(S, 4), D)
    assert got[0, 0] == np.float32(2.0 ** -24)
    assert H.bits_equal(got, H.spdm_fmaf_chain(S, D))


def test_spdm_f32_identity_and_single_nonzero():
    rng = np.random.default_rng(2)
    n, N = 10, 6
    D = rng.uniform(-1, 1, (n, N)).astype(np.float32)
    g = oracle.convert(np.eye(n, dtype=np.float32), 4)
    assert H.bits_equal(oracle.spdm_f32(g, D), D)
    S = np.zeros((n, n), np.float32)
    S[3, 7] = -2.5
    C = oracle.spdm_f32(oracle.convert(S, 4), D)
    exp = np.zeros((n, N), np.float32)
    exp[3] = np.float32(-2.5) * D[7]
    assert H.bits_equal(C, exp)
    # all-zero S gives +0.0 everywhere
    C0 = oracle.spdm_f32(oracle.convert(np.zeros((n, n), np.float32), 4), D)
    assert np.all(C0.view(np.uint32) == 0)


def test_spdm_f32_threads_identical():
    rng = np.random.default_rng(3)
    S = H.rng_sparse(rng, 64, 48, 0.1)
    D = rng.uniform(-1, 1, (48, 33)).astype(np.float32)
    g = oracle.convert(S, 4)
    assert H.bits_equal(oracle.spdm_f32(g, D, threads=1), oracle.spdm_f32(g, D, threads=4))


def test_fp64_matches_numpy_and_dense():
    rng = np.random.default_rng(4)
    S = H.rng_sparse(rng, 30, 40, 0.2)
    D = rng.uniform(-1, 1, (40, 20)).astype(np.float32)
    g = oracle.convert(S, 4)
    Cd, C = oracle.spdm_fp64(g, D)
    Dd, Df = oracle.dense_fp64(S, D)
    # f32*f32 is exact in double and zero terms add exactly 0: the sparse and
    # dense accumulations agree bit for bit
    np.testing.assert_array_equal(Cd, Dd)
    assert H.bits_equal(C, Df)
    ref = S.astype(np.float64) @ D.astype(np.float64)
    np.testing.assert_allclose(Cd, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.skipif(not oracle.ref_available(),
                    reason="oracle/_ref not built (reference sources absent)")
def test_fp64_matches_reference_matrixMulCPU():
    """The reference's own matrixMulCPU (src/utils.cpp:30-46), compiled from
    its sources, against the oracle's FP64 paths."""
    rng = np.random.default_rng(6)
    M, K, N = 23, 31, 17
    S = H.rng_sparse(rng, M, K, 0.3)
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    Cref = np.zeros((M, N), np.float32)
    oracle.ref().ref_matrix_mul_cpu(Cref, S, D, M, K, N)
    _, Cdense = oracle.dense_fp64(S, D)
    _, Csparse = oracle.spdm_fp64(oracle.convert(S, 4), D)
    assert H.bits_equal(Cdense, Cref)
    assert H.bits_equal(Csparse, Cref)


@pytest.mark.skipif(not oracle.ref_available(),
                    reason="oracle/_ref not built (reference sources absent)")
def test_reference_generator_kat():
    """The reference's randomInit (src/utils.cpp:5-19) on the default glibc
    seed: the drawn matrix converts to the same GCOO in the oracle and by
    brute force, and its zero pattern follows r % 1e6 < s * 1e6."""
    R = oracle.ref()
    R.ref_srand(1)
    A = np.zeros(64 * 64, np.float32)
    nnz = R.ref_random_init(A, A.size, np.float32(0.9))
    S = A.reshape(64, 64)
    assert nnz == int(np.count_nonzero(S))
    assert np.all((S >= 0) & (S <= 1))
    g = oracle.convert(S, 4)
    e = H.brute_gcoo(S, 4)
    np.testing.assert_array_equal(g["cols"], e["cols"])
    np.testing.assert_array_equal(g["rows"], e["rows"])


def test_bf16_rounding_hand_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.0, 0.0, -0.0,
                  np.inf, 1.0 + 2 ** -7 + 2 ** -9], np.float32)
    y = oracle.to_bf16(x)
    # bf16 keeps 7 fraction bits: ulp(1) = 2^-7.  1 + 2^-8 is a tie -> even
    # (0x3F80); 1 + 3*2^-9 is above half -> 0x3F81; 1 + 2^-9 below -> 0x3F80;
    # 1 + 2^-7 + 2^-9 -> 0x3F81.
    exp = np.array([0x3F80, 0x3F80, 0x3F81, 0x3F80, 0xC000, 0x0000, 0x8000, 0x7F80,
                    0x3F81], np.uint16)
    np.testing.assert_array_equal(y, exp)
    nan = oracle.to_bf16(np.array([np.nan], np.float32))
    assert (nan[0] & 0x7F80) == 0x7F80 and (nan[0] & 0x7F) != 0


def test_spdm_bf16_vs_exact():
    rng = np.random.default_rng(8)
    S = H.rng_sparse(rng, 6, 8, 0.5)
    D = rng.uniform(-1, 1, (8, 5)).astype(np.float32)
    g = oracle.convert(S, 4)
    vb = oracle.to_bf16(g["vals"])
    Db = oracle.to_bf16(D)
    got = oracle.spdm_bf16(g, vb, Db)
    Sb = np.zeros_like(S)
    Sb[g["rows"], g["cols"]] = oracle.bf16_to_f32(vb)
    exp = H.spdm_fmaf_chain(Sb, oracle.bf16_to_f32(Db).reshape(D.shape))
    assert H.bits_equal(got, exp)


def test_rel_frobenius_closed_form():
    R = np.array([3.0, 4.0], np.float64)
    C = np.array([3.0, 4.0], np.float32)
    assert oracle.rel_frobenius(C, R) == 0.0
    C = np.array([0.0, 0.0], np.float32)
    assert oracle.rel_frobenius(C, R) == pytest.approx(1.0)
    C = np.array([3.0, 5.0], np.float32)
    assert oracle.rel_frobenius(C, R) == pytest.approx(0.2)


@pytest.mark.parametrize("fname", sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))
                         if os.path.isdir(GOLDEN) else [])
def test_golden_fixtures(fname):
    z = np.load(os.path.join(GOLDEN, fname))   # allow_pickle=False (default)
    S, D, cpg = z["S"], z["D"], int(z["cpg"])
    g = oracle.convert(S, cpg)
    for key in ("rows", "cols", "groupIndex", "nnzPerGroup"):
        np.testing.assert_array_equal(g[key], z[key], err_msg=key)
    assert H.bits_equal(g["vals"], z["vals"])
    assert H.bits_equal(oracle.spdm_f32(g, D), z["C"])
    np.testing.assert_array_equal(oracle.spdm_fp64(g, D)[0], z["C64"])
