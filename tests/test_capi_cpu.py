"""Host-side checks of the C ABI and the `main` driver (no GPU needed):
the library loads, exports every function include/gcoo_spdm.h declares, and
argument validation fails with a status code instead of exiting."""
import ctypes
import os
import re
import subprocess

import pytest

import gcoospdm_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gcoo_spdm.h")
MAIN = os.path.join(ROOT, "gcoospdm_amd", "host", "main")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gcoo_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(G.declared_symbols())


def test_library_exports_every_declared_symbol():
    lib = G.lib()
    for name in header_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (gcoo_\w+)", out))
    assert set(header_functions()) <= exported


def test_version_and_error_paths():
    lib = G.lib()
    assert b"gfx950" in lib.gcoo_version()
    h = ctypes.c_void_p()
    assert lib.gcoo_create(8, 8, 3, ctypes.byref(h)) == G.GCOO_ERR_INVALID
    assert b"cpg" in lib.gcoo_last_error()
    assert lib.gcoo_create(-1, 8, 4, ctypes.byref(h)) == G.GCOO_ERR_INVALID
    assert lib.gcoo_create(8, 8, 4, None) == G.GCOO_ERR_INVALID
    assert lib.gcoo_spdm_f32(None, None, 8, 8, None, None) == G.GCOO_ERR_INVALID
    assert lib.gcoo_convert_dense(None, None, None) == G.GCOO_ERR_INVALID
    assert lib.gcoo_get_arrays(None, None, None, None, None, None) == G.GCOO_ERR_INVALID
    assert lib.gcoo_destroy(None) == G.GCOO_OK
    assert lib.gcoo_fill_uniform_async(None, -1, 0, 0, 1.0, None) == G.GCOO_ERR_INVALID


def test_python_mirror_factory_keys():
    # unknown keys return None like the reference's NULL (the known ones
    # build device state and are exercised by the GPU tests)
    for key in ("nope", "groupspgemm", ""):
        assert G.SpGemmBuilder.getShGemmAlgo(None, 4, 4, key) is None


VENDOR_HEADER = os.path.join(ROOT, "include", "gcoo_vendor.h")


def vendor_header_functions():
    txt = open(VENDOR_HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gcoo_vendor_[a-z0-9_]+)\s*\(", txt)))


def test_vendor_header_binding_and_exports():
    from gcoospdm_amd import vendor as V
    assert vendor_header_functions() == sorted(V.declared_vendor_symbols())
    lib = V.vendor_lib()
    out = subprocess.run(["nm", "-D", "--defined-only", V.VENDOR_LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (gcoo_vendor_\w+)", out))
    assert set(vendor_header_functions()) <= exported
    h = ctypes.c_void_p()
    assert lib.gcoo_vendor_csr_create(-1, 4, ctypes.byref(h)) == G.GCOO_ERR_INVALID
    assert lib.gcoo_vendor_csr_create(4, 4, None) == G.GCOO_ERR_INVALID
    assert lib.gcoo_vendor_csr_destroy(None) == G.GCOO_OK
    assert lib.gcoo_vendor_blas_destroy(None) == G.GCOO_OK
    assert lib.gcoo_vendor_sgemm(None, None, None, 1, 1, 1, None, None) == G.GCOO_ERR_INVALID
    assert b"argument" in lib.gcoo_vendor_last_error()


@pytest.mark.skipif(not os.path.exists(MAIN), reason="main driver not built")
def test_main_driver_flag_validation():
    r = subprocess.run([MAIN, "-help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "-algo" in r.stdout
    r = subprocess.run([MAIN, "-nRowsA=8"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "algo" in r.stderr
    r = subprocess.run([MAIN, "-algo=gcoospdm", "-nRowsA=8", "-nColsB=9"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "nColsB" in r.stderr
    r = subprocess.run([MAIN, "-algo=gcoospdm", "-file=x.mtx"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and a message,
    before anything touches a device."""
    r = subprocess.run(["python3", os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_bench_traffic_from_request_sizes(monkeypatch):
    """bench.py's `traffic`: read bytes from the L2's request counts by size,
    write bytes from WRITE_SIZE scaled by the 256 MiB copy, and FETCH_SIZE
    scaled by the copy beside them.  Counter readings are stubbed with the
    shape of a real pass: the copy's reads are all 128-byte requests, which
    FETCH_SIZE counts as 64 bytes (factor 2); a kernel with 64-byte requests
    as well is overstated by the scaled FETCH_SIZE and exact from the counts."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    copy, k = "__amd_rocclr_copyBuffer", "spdm_panel_kernel<F32, 75, 16, 21>"
    n128, n64 = 3_000_000, 2_000_000
    counts = {
        "FETCH_SIZE": {copy: [131072.0, 10.0], k: [(n128 + n64) * 64 / 1024]},
        "WRITE_SIZE": {copy: [262144.0], k: [250000.0]},
        "TCC_EA0_RDREQ_32B_sum": {copy: [0.0], k: [0.0]},
        "TCC_EA0_RDREQ_64B_sum": {copy: [0.0], k: [float(n64)]},
        "TCC_EA0_RDREQ_128B_sum": {copy: [2097152.0], k: [float(n128)]},
        "TCC_EA0_RDREQ_sum": {copy: [2097152.0], k: [float(n128 + n64)]},
    }

    def fake_pass(counters, n, s, seed, timeout_s):
        return {c: counts[c] for c in counters}

    monkeypatch.setattr(bench, "_counter_pass", fake_pass)
    out = bench.measure_traffic(8000, 0.99, 1, (k,))
    assert out["fetch_factor"] == 2.0 and out["write_factor"] == 1.0
    assert out["copy_read_check"] == 1.0
    r = out[k]
    assert r["fetch_bytes"] == 128 * n128 + 64 * n64
    assert r["write_bytes"] == 250000 * 1024
    assert r["bytes"] == r["fetch_bytes"] + r["write_bytes"]
    assert r["fetch_bytes_copy_scaled"] == 2 * 64 * (n128 + n64)
    # request counts that do not add up to the total are an error, reported
    counts["TCC_EA0_RDREQ_sum"][k] = [float(n128 + 2 * n64)]
    assert "do not add up" in bench.measure_traffic(8000, 0.99, 1, (k,))["error"]


@pytest.mark.skipif(not os.path.exists(MAIN), reason="main not built")
@pytest.mark.parametrize("seed,sparsity", [(None, 0.7), (7, 0.9), (12345, 0.0)])
def test_main_ref_inputs_match_reference_randomInit(tmp_path, seed, sparsity):
    """`main -dist=ref` generates the reference's own inputs (src/main.cu:103-121:
    the dense operand first with sparsity 0, then the sparse one, both by
    randomInit, src/utils.cpp:5-19, from glibc's rand()).  Its operands are
    compared bit for bit with the reference's randomInit compiled from its
    sources into oracle/_ref, called in the same order from the same seed
    (glibc's default seed is 1)."""
    import numpy as np
    import oracle
    if not oracle.ref_available():
        pytest.skip("oracle/_ref (the reference's utils.cpp) not built")
    nRowsA, nColsA, nRowsB = 48, 20, 33
    prefix = str(tmp_path / "in")
    cmd = [MAIN, "-algo=gcoospdm", "-dist=ref", f"-nRowsA={nRowsA}", f"-nColsA={nColsA}",
           f"-nRowsB={nRowsB}", f"-nColsB={nRowsA}", f"-sparsity={sparsity}",
           f"-inputs={prefix}"]
    if seed is not None:
        cmd.append(f"-seed={seed}")
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=60, check=True).stdout
    D = np.fromfile(prefix + ".D.bin", np.float32)
    S = np.fromfile(prefix + ".S.bin", np.float32)
    R = oracle.ref()
    R.ref_srand(1 if seed is None else seed)
    Dr = np.empty(nRowsA * nColsA, np.float32)
    Sr = np.empty(nRowsB * nRowsA, np.float32)
    assert R.ref_random_init(Dr, Dr.size, 0.0) == Dr.size
    nnz = R.ref_random_init(Sr, Sr.size, sparsity)
    assert np.array_equal(D.view(np.uint32), Dr.view(np.uint32))
    assert np.array_equal(S.view(np.uint32), Sr.view(np.uint32))
    assert f"nnz cpu: {nnz}," in out
    if sparsity > 0:
        assert 0 < nnz < Sr.size
