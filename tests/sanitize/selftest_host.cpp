// Sanitizer self-test of the `main` driver's host utilities
// (gcoospdm_amd/host/host_utils.cpp: the reference-input and counter
// generators, the FP64 check, the Matrix Market reader and the flag parser),
// built with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/sanitize/Makefile) and run by tests/test_sanitizers.py.  The Matrix
// Market files include the malformed cases the reader must reject (the
// reference's reader writes out of bounds on them, src/MatrixReader.cpp:
// 45-75), so an out-of-bounds access stops the run with a report.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../gcoospdm_amd/host/host_utils.h"

using namespace gcoo_host;

#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            std::fprintf(stderr, "check failed: %s (%s:%d)\n", #c, __FILE__, __LINE__); \
            std::exit(1);                                                         \
        }                                                                         \
    } while (0)

static std::string write_file(const char *name, const char *text)
{
    const char *dir = std::getenv("SANITIZE_TMP");
    std::string path = std::string(dir ? dir : "/tmp") + "/" + name;
    FILE *f = std::fopen(path.c_str(), "w");
    CHECK(f != nullptr);
    std::fputs(text, f);
    std::fclose(f);
    return path;
}

int main()
{
    std::vector<float> a(1000);
    const long long nz = random_init_ref(a.data(), a.size(), 0.9f);
    long long cnt = 0;
    for (float v : a) cnt += v != 0.0f;
    CHECK(nz == cnt);
    std::vector<float> b(777);
    const long long nz2 = fill_counter(b.data(), b.size(), 3, 1, 1.0f, keep_threshold(0.5), true);
    cnt = 0;
    for (float v : b) cnt += v != 0.0f;
    CHECK(nz2 == cnt);
    CHECK(keep_threshold(0.0) == 0 && keep_threshold(1.0) == (1ull << 40));

    // FP64 check of an exact product: S = I (3 x 3), C = D
    const int M = 3, K = 3, N = 4;
    std::vector<float> S(M * K, 0.0f), D(K * N), C(M * N);
    for (int i = 0; i < 3; ++i) S[i * K + i] = 1.0f;
    for (int i = 0; i < K * N; ++i) D[i] = 0.25f * (float)(i + 1);
    C = D;
    CHECK(rel_l2_vs_fp64(S.data(), D.data(), C.data(), M, K, N, 2) == 0.0);
    C[5] += 1.0f;
    CHECK(rel_l2_vs_fp64(S.data(), D.data(), C.data(), M, K, N, 1) > 0.0);

    MtxMatrix m;
    std::string err;
    CHECK(read_matrix_market(write_file("san_general.mtx",
                                        "%%MatrixMarket matrix coordinate real general\n"
                                        "% comment\n3 4 3\n1 1 2.5\n3 4 -1\n2 2 0.5\n"),
                             m, err));
    CHECK(m.nRows == 3 && m.nCols == 4 && m.dense[0] == 2.5f && m.dense[2 * 4 + 3] == -1.0f);
    CHECK(read_matrix_market(write_file("san_sym.mtx",
                                        "%%MatrixMarket matrix coordinate real symmetric\n"
                                        "3 3 2\n2 1 4\n3 3 1\n"),
                             m, err));
    CHECK(m.dense[1 * 3 + 0] == 4.0f && m.dense[0 * 3 + 1] == 4.0f);
    CHECK(read_matrix_market(write_file("san_pattern.mtx",
                                        "%%MatrixMarket matrix coordinate pattern general\n"
                                        "2 2 2\n1 1\n2 2\n\n"),
                             m, err));
    CHECK(m.dense[0] != 0.0f && m.dense[3] != 0.0f);
    CHECK(!read_matrix_market(write_file("san_badidx.mtx",
                                         "%%MatrixMarket matrix coordinate real general\n"
                                         "2 2 1\n3 1 1.0\n"),
                              m, err));
    CHECK(!read_matrix_market(write_file("san_zero.mtx",
                                         "%%MatrixMarket matrix coordinate real general\n"
                                         "2 2 1\n0 1 1.0\n"),
                              m, err));
    CHECK(!read_matrix_market(write_file("san_trunc.mtx",
                                         "%%MatrixMarket matrix coordinate real general\n2 2\n"),
                              m, err));
    CHECK(!read_matrix_market("/nonexistent/file.mtx", m, err));

    const char *argv[] = {"main", "-nRowsA=12", "-sparsity=0.25", "-algo=csr", "-kernelonly"};
    Flags f(5, const_cast<char **>(argv));
    CHECK(f.get_int("nRowsA", 0) == 12 && std::fabs(f.get_float("sparsity", 0) - 0.25) < 1e-12);
    CHECK(f.get_str("algo", "") == "csr" && f.has("kernelonly") && !f.has("missing"));
    CHECK(f.get_int("nColsA", 7) == 7);
    std::printf("host utilities sanitizer self-test ok\n");
    return 0;
}
