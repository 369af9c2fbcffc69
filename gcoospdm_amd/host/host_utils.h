// host_utils.h -- host-side helpers of the `main` driver: input generators,
// the FP64 result check and the -name=value flag parser.
#ifndef GCOO_HOST_UTILS_H
#define GCOO_HOST_UTILS_H

#include <cstddef>
#include <cstdint>
#include <string>

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
