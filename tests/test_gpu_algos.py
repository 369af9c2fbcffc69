"""The reference's SpGemmBuilder algorithms beside "gcoospdm": "csr" (the
row-wise CSR ablation: three-array CSR and the reference kernel's
structure) gives the reference's bits; the vendor baselines "cusparse" (rocSPARSE) and "cublas"
(rocBLAS) sum in their own order, so they are held to the FP32 tolerance
against the FP64 product (relative Frobenius <= 1e-5, src/main.cu:136)."""
import os
import subprocess

import numpy as np
import pytest

import gcoospdm_amd as G
import oracle
from tests import helpers as H

pytestmark = pytest.mark.gpu

FP32_REL_TOL = 1e-5
MAIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "gcoospdm_amd", "host", "main")


def to_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("algo", ["gcoospdm", "csr", "cusparse", "cublas"])
@pytest.mark.parametrize("M,K,N,dens", [(130, 257, 520, 0.05), (61, 1030, 200, 0.3),
                                        (300, 77, 96, 0.2), (9, 5, 3, 0.0)])
def test_algorithms_match(dev, algo, M, K, N, dens):
    import torch
    print(f"algo={algo} M={M} K={K} N={N} dens={dens}", flush=True)
    rng = np.random.default_rng(M * 5 + K + N)
    S = H.rng_sparse(rng, M, K, dens)
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    Sd = to_dev(S, dev)
    Dd = to_dev(D, dev)
    a = G.SpGemmBuilder.getShGemmAlgo(Sd, M, K, algo)
    assert a is not None
    C = torch.full((M, N), float("nan"), device=dev)
    a.mutiplyBy(Dd, K, N, C)
    C2 = torch.full((M, N), float("nan"), device=dev)
    a.build()
    a.runKernelTest(Dd, K, N, C2)
    Cn = C.cpu().numpy()
    assert H.bits_equal(Cn, C2.cpu().numpy())
    g = oracle.convert(S, 4)
    if algo in ("gcoospdm", "csr"):
        assert H.bits_equal(Cn, oracle.spdm_f32(g, D))
    Cd, _ = oracle.spdm_fp64(g, D)
    if np.any(Cd):
        assert oracle.rel_frobenius(Cn, Cd) <= FP32_REL_TOL
    else:
        assert not np.any(Cn)
    if algo == "cusparse":
        assert a.nnz == int((S != 0).sum())


@pytest.mark.parametrize("algo", ["csr", "cusparse", "cublas"])
def test_main_driver_algorithms(dev, algo):
    """`main -algo=<algo>` on the 800 x 800 case with its FP64 check."""
    r = subprocess.run([MAIN, f"-algo={algo}", "-nRowsA=800", "-nColsA=800", "-nRowsB=800",
                        "-nColsB=800", "-sparsity=0.9", "-dist=uniform", "-iters=3", "-check=1"],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-600:], r.stderr[-600:])
    assert r.returncode == 0
    assert f"Algorithm: {algo}" in r.stdout and "Time:" in r.stdout and "Test passed" in r.stdout


def _read_mtx_like_reference(text):
    """The reference MatrixReader's parse (src/MatrixReader.cpp:11-84),
    restated: 'general' in the header or mirrored; a value-less entry takes
    glibc rand() (default seed), and rand() is drawn for every entry line."""
    import ctypes
    import ctypes.util
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.srand(1)
    lines = text.splitlines()
    general = "general" in lines[0]
    i = 1
    while lines[i].startswith("%"):
        i += 1
    r, c, _ = (int(x) for x in lines[i].split())
    A = np.zeros((r, c), np.float32)
    nnz = 0
    for ln in lines[i + 1:]:
        if not ln.strip():
            continue
        parts = ln.split(" ")
        row, col = int(parts[0]) - 1, int(parts[1]) - 1
        val = np.float32(libc.rand())
        if len(parts) > 2:
            val = np.float32(float(parts[2]))
        A[row, col] = val
        nnz += 1
        if not general:
            A[col, row] = val
            nnz += col != row
    return A, nnz


@pytest.mark.parametrize("kind", ["general", "symmetric", "pattern"])
def test_main_matrix_market_file(dev, tmp_path, kind):
    """`main -file=` (the reference's timeFromFile, src/main.cu:149-178): the
    file is read as the reference reads it and C = S * S is the reference's
    fmaf chain on that matrix."""
    rng = np.random.default_rng({"general": 1, "symmetric": 2, "pattern": 3}[kind])
    n = 70
    S = H.rng_sparse(rng, n, n, 0.08)
    if kind == "symmetric":
        S = np.tril(S)
    rr, cc = np.nonzero(S)
    field = "pattern" if kind == "pattern" else "real"
    sym = "symmetric" if kind == "symmetric" else "general"
    body = [f"%%MatrixMarket matrix coordinate {field} {sym}", "% written by the test",
            f"{n} {n} {rr.size}"]
    for r_, c_ in zip(rr, cc):
        body.append(f"{r_ + 1} {c_ + 1}" if kind == "pattern" else f"{r_ + 1} {c_ + 1} {float(S[r_, c_])!r}")
    text = "\n".join(body) + "\n\n"
    path = tmp_path / f"{kind}.mtx"
    path.write_text(text)
    A, nnz = _read_mtx_like_reference(text)
    out = tmp_path / "C.bin"
    r = subprocess.run([MAIN, f"-file={path}", "-algo=gcoospdm", "-iters=2", "-check=1",
                        f"-dump={out}"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-500:], r.stderr[-500:])
    assert r.returncode == 0 and "Test passed" in r.stdout
    assert f"nnz cpu: {nnz}," in r.stdout
    C = np.fromfile(out, np.float32).reshape(n, n)
    assert H.bits_equal(C, oracle.spdm_f32(oracle.convert(A, 4), A))


def test_harness_tsv(dev, tmp_path):
    """The reference-format TSV harness: one row per case, one column per
    algorithm, milliseconds per call, no failed runs."""
    import io
    from gcoospdm_amd import harness
    buf = io.StringIO()
    harness.synthetic([256], [0.9, 0.99], ["gcoospdm", "csr", "cusparse", "cublas"],
                      iters=3, dist="uniform", out=buf)
    rows = [ln.split("\t") for ln in buf.getvalue().strip().splitlines()]
    print(buf.getvalue())
    assert rows[0] == ["N", "sparsity", "gcoospdm", "csr", "cusparse", "cublas",
                       "gcoospdm_model", "hbm_bound"]
    assert len(rows) == 3
    from gcoospdm_amd import perf_model
    for r in rows[1:]:
        assert r[0] == "256" and all(0.0 < float(t) < harness.FAILED for t in r[2:])
        s = float(r[1])
        assert abs(float(r[6]) - perf_model.time_of_s(s, 256)[0] * 1e3) < 1e-5
    n = 40
    rng = np.random.default_rng(9)
    S = H.rng_sparse(rng, n, n, 0.1)
    rr, cc = np.nonzero(S)
    (tmp_path / "m.mtx").write_text(
        "%%MatrixMarket matrix coordinate real general\n" + f"{n} {n} {rr.size}\n"
        + "".join(f"{a + 1} {b + 1} {float(S[a, b])!r}\n" for a, b in zip(rr, cc)))
    buf = io.StringIO()
    harness.public_dataset(harness.find_mtx(str(tmp_path)), ["gcoospdm", "cusparse"], out=buf)
    rows = [ln.split("\t") for ln in buf.getvalue().strip().splitlines()]
    assert rows[1][0] == str(n) and rows[1][2] == "m.mtx"
    assert all(0.0 < float(t) < harness.FAILED for t in rows[1][3:])


@pytest.mark.parametrize("M,K,N,dens", [(130, 257, 520, 0.05), (1, 1, 1, 1.0), (7, 300, 1, 0.5),
                                        (40, 33, 256, 0.0), (700, 900, 300, 0.02)])
def test_csr_format_and_kernel(dev, M, K, N, dens):
    """The CSR arrays equal the oracle's conversion with one row per group
    (whose groups are rows: groupIndex is rowPtr, the (k, r) order within a
    one-row group is ascending k), edge values included (-0.0 dropped, NaN,
    denormals and inf kept); the CSR kernel's C is bit-equal to the oracle's
    fmaf chain; the product of a converted matrix can be re-run."""
    import torch
    rng = np.random.default_rng(M + 3 * K + 7 * N)
    S = H.rng_sparse(rng, M, K, dens)
    if M * K >= 8:
        S.flat[rng.integers(0, M * K, 4)] = np.array([-0.0, np.nan, 1e-40, np.inf], np.float32)
    D = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    c = G.CSRMatrix(M, K).convert_dense(to_dev(S, dev))
    got = c.arrays_host()
    g = oracle.convert(S, 1)
    assert np.array_equal(got["rowPtr"], g["groupIndex"])
    assert got["nnz"] == g["nnz"]
    assert np.array_equal(got["colInd"], g["cols"])
    assert np.array_equal(got["vals"].view(np.uint32), g["vals"].view(np.uint32))
    want = oracle.spdm_f32(oracle.convert(S, 4), D)
    for rep in range(2):
        C = torch.full((M, N), float("nan"), device=dev)
        c.spmm(to_dev(D, dev), C)
        assert H.bits_equal(C.cpu().numpy(), want), rep
