// main.cpp -- the `main -algo=gcoospdm` driver (reference src/main.cu:1-223).
//
// Flags kept from the reference (src/main.cu:183-215):
//   -nRowsA -nColsA  dense operand D (the reference's "A")
//   -nRowsB -nColsB  sparse operand S (the reference's "B")
//   -sparsity -algo -kernelonly -help
// Added:
//   -dist=ref|uniform  ref: the reference's glibc rand() inputs (default);
//                      uniform: values U(-1,1), S kept with p = 1 - sparsity
//   -seed=N            srand(N) for ref, counter seed for uniform
//   -iters=N           timed repeats (reference: 100, 20 when nRowsA > 8000)
//   -cpg=N             rows per group (1, 2, 4)
//   -check=1           FP64 check, relative L2 <= 1e-5 (the reference's
//                      disabled check, src/main.cu:128-142)
//   -file=path.mtx     the reference's timeFromFile (src/main.cu:149-178): the
//                      Matrix Market matrix is both the sparse S and the dense
//                      operand (square matrices only)
//   -dump=path         write C (row-major float32, no header) after the run
//   -inputs=prefix     write the generated operands (prefix.D.bin, then
//                      prefix.S.bin, row-major float32) and exit before any
//                      device work
// Output keeps the reference's lines ("nnz cpu:", "A[..] * B[..]",
// "Algorithm:", "Time:" in microseconds) so its harness can parse it, plus a
// GFLOPS line (2 * nnz * N / t).
//
// The product computed is C = S * D with S = nRowsB x nColsB and
// D = nRowsA x nColsA, so nColsB must equal nRowsA; C is nRowsB x nColsA.
// (The reference allocates C as nRowsA x nColsB, consistent only for square
// inputs, and checks neither shape.)
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "SpGemmBuilder.h"
#include "gcoo_spdm.h"
#include "host_utils.h"

using gcoo_host::Flags;

namespace {

#define CHECK_GCOO(expr)                                                          \
    do {                                                                          \
        int rc_ = (expr);                                                         \
        if (rc_ != GCOO_OK) {                                                     \
            std::fprintf(stderr, "%s failed (%d): %s\n", #expr, rc_,              \
                         gcoo_last_error());                                      \
            std::exit(1);                                                         \
        }                                                                         \
    } while (0)

void usage()
{
    std::printf("Usage -nRowsA=nRowsA   (rows of dense A = columns of sparse B)\n");
    std::printf("      -nColsA=nColsA\n");
    std::printf("      -nRowsB=nRowsB\n");
    std::printf("      -nColsB=nColsB\n");
    std::printf("      -sparsity=sparsity (Sparsity of matrix B)\n");
    std::printf("      -algo=gcoospdm|csr|cusparse|cublas\n");
    std::printf("      -kernelonly=0 (Benchmark kernel only)\n");
    std::printf("      -dist=ref|uniform -seed=N -iters=N -cpg=4 -check=0\n");
    std::printf("      -file=matrix.mtx (S and the dense operand) -dump=C.bin -inputs=prefix\n");
    std::printf("C = B * A with B sparse: nColsB must equal nRowsA\n");
}

struct Result {
    double us_per_iter = 0;
    long long nnz = 0;
};

// Reference benchSpGemm (src/main.cu:29-95).
Result bench(const float *A, size_t nRowsA, size_t nColsA, const float *B, size_t nRowsB,
             size_t nColsB, float *C, const std::string &algo, bool kernelonly, int iters,
             int cpg)
{
    std::printf("A[%zu,%zu] * B[%zu,%zu]\n", nRowsA, nColsA, nRowsB, nColsB);
    std::printf("Algorithm: %s\n", algo.c_str());
    const size_t nA = nRowsA * nColsA, nB = nRowsB * nColsB, nC = nRowsB * nColsA;
    float *dA = nullptr, *dB = nullptr, *dC = nullptr;
    CHECK_GCOO(gcoo_dev_malloc((void **)&dA, nA * sizeof(float)));
    CHECK_GCOO(gcoo_dev_malloc((void **)&dB, nB * sizeof(float)));
    CHECK_GCOO(gcoo_dev_malloc((void **)&dC, nC * sizeof(float)));
    CHECK_GCOO(gcoo_memcpy_h2d(dA, A, nA * sizeof(float), nullptr));
    CHECK_GCOO(gcoo_memcpy_h2d(dB, B, nB * sizeof(float), nullptr));

    std::unique_ptr<ShGemm> algoObj;
    if (algo == "gcoospdm" && cpg != 4)
        algoObj.reset(new ShGCOOSpDM(dB, (int)nRowsB, (int)nColsB, cpg));
    else
        algoObj.reset(SpGemmBuilder::getShGemmAlgo(dB, (int)nRowsB, (int)nColsB, algo));
    if (!algoObj) {
        std::fprintf(stderr,
                     "unknown algorithm '%s' (available: gcoospdm, csr, cusparse, cublas)\n",
                     algo.c_str());
        std::exit(2);
    }
    const int n = (int)nRowsA, m = (int)nColsA;
    if (kernelonly) algoObj->build();
    // warm-up
    if (kernelonly)
        algoObj->runKernelTest(dA, n, m, dC);
    else
        algoObj->mutiplyBy(dA, n, m, dC);
    gcoo_timer_t t;
    CHECK_GCOO(gcoo_timer_create(&t));
    CHECK_GCOO(gcoo_timer_start(t, nullptr));
    for (int i = 0; i < iters; ++i) {
        if (kernelonly)
            algoObj->runKernelTest(dA, n, m, dC);
        else
            algoObj->mutiplyBy(dA, n, m, dC);
    }
    CHECK_GCOO(gcoo_device_sync());
    CHECK_GCOO(gcoo_timer_stop(t, nullptr));
    float ms = 0;
    CHECK_GCOO(gcoo_timer_elapsed_ms(t, &ms));
    gcoo_timer_destroy(t);
    Result res;
    res.us_per_iter = iters > 0 ? ms * 1000.0 / iters : 0.0;
    if (auto *g = dynamic_cast<ShGCOOSpDM *>(algoObj.get())) res.nnz = g->nnz();
    if (auto *c = dynamic_cast<ShCsrSpGemm *>(algoObj.get())) {
        int64_t n = 0;
        if (gcoo_csr_get_arrays(c->csr(), nullptr, nullptr, nullptr, &n) == GCOO_OK) res.nnz = n;
    }
    CHECK_GCOO(gcoo_memcpy_d2h(C, dC, nC * sizeof(float), nullptr));
    std::printf("Time:%f\n", res.us_per_iter);
    algoObj.reset();
    gcoo_dev_free(dA);
    gcoo_dev_free(dB);
    gcoo_dev_free(dC);
    return res;
}

bool dump_matrix(const std::string &path, const std::vector<float> &C)
{
    if (path.empty()) return true;
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    const bool ok = std::fwrite(C.data(), sizeof(float), C.size(), fp) == C.size();
    return std::fclose(fp) == 0 && ok;
}

}  // namespace

int main(int argc, char **argv)
{
    std::printf("[Benchmark sparse gemm on GPU (sparse B multiply dense A)] - Starting...\n");
    Flags f(argc, argv);
    if (f.has("help") || f.has("?")) {
        usage();
        return 0;
    }
    const size_t nRowsA = (size_t)f.get_int("nRowsA", 1024);
    const size_t nColsA = (size_t)f.get_int("nColsA", 1024);
    const size_t nRowsB = (size_t)f.get_int("nRowsB", 1024);
    const size_t nColsB = (size_t)f.get_int("nColsB", 1024);
    const double sparsity = f.get_float("sparsity", 0.5);
    const bool kernelonly = f.get_int("kernelonly", 0) == 1;
    const std::string algo = f.get_str("algo", "");
    const std::string dist = f.get_str("dist", "ref");
    const int cpg = (int)f.get_int("cpg", 4);
    const bool check = f.get_int("check", 0) == 1;
    int iters = (int)f.get_int("iters", nRowsA > 8000 ? 20 : 100);
    if (algo.empty()) {
        std::fprintf(stderr, "missing -algo= (available: gcoospdm, csr, cusparse, cublas)\n");
        usage();
        return 2;
    }
    if (f.has("file")) {
        const std::string path = f.get_str("file", "");
        gcoo_host::MtxMatrix mm;
        std::string err;
        if (!gcoo_host::read_matrix_market(path, mm, err)) {
            std::fprintf(stderr, "%s\n", err.c_str());
            return 2;
        }
        std::printf("Read mtx file: %s\n", path.c_str());
        if (mm.nRows != mm.nCols) {
            std::fprintf(stderr, "%s: %d x %d is not square (S is also the dense operand)\n",
                         path.c_str(), mm.nRows, mm.nCols);
            return 2;
        }
        const size_t n = (size_t)mm.nRows;
        std::printf("nnz cpu: %lld, density: %.4f\n", mm.nnz,
                    n ? mm.nnz * 1.0 / ((double)n * (double)n) : 0.0);
        long long nzv = 0;  // nonzero values, the multiply's work (duplicates collapse)
        for (float v : mm.dense) nzv += v != 0.0f;
        std::vector<float> C(n * n);
        Result r;
        try {
            r = bench(mm.dense.data(), n, n, mm.dense.data(), n, n, C.data(), algo, kernelonly,
                      iters, cpg);
        } catch (const std::exception &e) {
            std::fprintf(stderr, "error: %s\n", e.what());
            return 1;
        }
        if (r.us_per_iter > 0)
            std::printf("GFLOPS:%f\n", 2.0 * (double)nzv * (double)n / (r.us_per_iter * 1e3));
        if (!dump_matrix(f.get_str("dump", ""), C)) {
            std::fprintf(stderr, "cannot write -dump file\n");
            return 1;
        }
        if (check) {
            const int threads = (int)std::max(1u, std::thread::hardware_concurrency());
            const double e = gcoo_host::rel_l2_vs_fp64(mm.dense.data(), mm.dense.data(), C.data(),
                                                       (int)n, (int)n, (int)n, threads);
            std::printf("RelL2:%.3e\n", e);
            std::printf(e <= 1e-5 ? "Test passed\n" : "Test failed\n");
            if (e > 1e-5) return 1;
        }
        return 0;
    }
    if (nColsB != nRowsA) {
        std::fprintf(stderr, "nColsB (%zu) must equal nRowsA (%zu)\n", nColsB, nRowsA);
        return 2;
    }
    if (nRowsA > 2147483647u || nColsA > 2147483647u || nRowsB > 2147483647u ||
        nColsB > 2147483647u) {
        std::fprintf(stderr, "dimensions must fit in int\n");
        return 2;
    }
    std::vector<float> A(nRowsA * nColsA), B(nRowsB * nColsB), C(nRowsB * nColsA);
    long long nnz = 0;
    if (dist == "ref") {
        if (f.has("seed")) srand((unsigned)f.get_int("seed", 1));
        gcoo_host::random_init_ref(A.data(), A.size(), 0.0f);
        nnz = gcoo_host::random_init_ref(B.data(), B.size(), (float)sparsity);
    } else if (dist == "uniform") {
        const uint64_t seed = (uint64_t)f.get_int("seed", 1);
        gcoo_host::fill_counter(A.data(), A.size(), seed, 0, 1.0f, 0, false);
        nnz = gcoo_host::fill_counter(B.data(), B.size(), seed, 1, 1.0f,
                                      gcoo_host::keep_threshold(sparsity), true);
    } else {
        std::fprintf(stderr, "unknown -dist=%s (ref|uniform)\n", dist.c_str());
        return 2;
    }
    std::printf("nnz cpu: %lld, density: %.4f\n", nnz,
                B.empty() ? 0.0 : nnz * 1.0 / (double)B.size());
    if (f.has("inputs")) {
        const std::string prefix = f.get_str("inputs", "");
        if (!dump_matrix(prefix + ".D.bin", A) || !dump_matrix(prefix + ".S.bin", B)) {
            std::fprintf(stderr, "cannot write -inputs files\n");
            return 1;
        }
        return 0;
    }
    Result r;
    try {
        r = bench(A.data(), nRowsA, nColsA, B.data(), nRowsB, nColsB, C.data(), algo, kernelonly,
                  iters, cpg);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    if (r.us_per_iter > 0)
        std::printf("GFLOPS:%f\n", 2.0 * (double)nnz * (double)nColsA / (r.us_per_iter * 1e3));
    if (!dump_matrix(f.get_str("dump", ""), C)) {
        std::fprintf(stderr, "cannot write -dump file\n");
        return 1;
    }
    if (check) {
        const int threads = (int)std::max(1u, std::thread::hardware_concurrency());
        const double err = gcoo_host::rel_l2_vs_fp64(B.data(), A.data(), C.data(), (int)nRowsB,
                                                     (int)nColsB, (int)nColsA, threads);
        std::printf("RelL2:%.3e\n", err);
        if (err <= 1e-5) {
            std::printf("Test passed\n");
        } else {
            std::printf("Test failed\n");
            return 1;
        }
    }
    return 0;
}
