// host_utils.h -- host-side helpers of the `main` driver: input generators,
// the FP64 result check and the -name=value flag parser.
#ifndef GCOO_HOST_UTILS_H
#define GCOO_HOST_UTILS_H

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace gcoo_host {

// The reference generator, reference src/utils.cpp:5-19: one glibc rand()
// draw r per element; the element is 0 when r % 1000000 < sparsity * 1e6,
// else r / RAND_MAX.  Returns the number of nonzeros.  Reproduces the
// reference's own inputs (same libc, same seed: glibc's default of 1 unless
// srand was called).
long long random_init_ref(float *data, size_t size, float sparsity);

// Counter-based generator, the host twin of gcoo_fill_sparse_async
// (include/gcoo_spdm.h): bit-identical values for the same seed/stream.
uint64_t keep_threshold(double sparsity);
long long fill_counter(float *data, size_t size, uint64_t seed, uint64_t stream_id,
                       float scale, uint64_t keep_thr, bool sparse);

// FP64 check of C = S * D (S: M x K, D: K x N, row-major): the reference's
// matrixMulCPU (src/utils.cpp:30-46) accumulated in double over S's
// nonzeros, then the relative L2 error ||C - ref|| / ||ref|| of
// sdkCompareL2fe (src/main.cu:136).  Multi-threaded over rows with
// std::thread.
double rel_l2_vs_fp64(const float *S, const float *D, const float *C, int M, int K,
                      int N, int threads);

// A Matrix Market file read the way the reference's MatrixReader does
// (src/MatrixReader.cpp:11-84): the header's "general" keyword selects an
// unsymmetric matrix, anything else is mirrored (a(c, r) = a(r, c), an
// off-diagonal entry counted twice); comment lines start with '%'; each entry
// line is "row col [value]" (1-based, fields split at single spaces), and an
// entry with no value takes rand() as its value -- rand() is drawn for every
// entry line, as in the reference.  Returns false with a message on a
// missing or malformed file or an index outside the stated size (the
// reference writes out of bounds there).
struct MtxMatrix {
    int nRows = 0, nCols = 0;
    long long nnz = 0;           // the reference's count (m_nnz)
    std::vector<float> dense;    // row-major nRows x nCols
};
bool read_matrix_market(const std::string &path, MtxMatrix &out, std::string &err);

class Flags
{
public:
    Flags(int argc, char **argv) : argc_(argc), argv_(argv) {}
    bool has(const char *name) const;
    long long get_int(const char *name, long long def) const;
    double get_float(const char *name, double def) const;
    std::string get_str(const char *name, const std::string &def) const;

private:
    const char *find(const char *name) const;
    int argc_;
    char **argv_;
};

}  // namespace gcoo_host

#endif
