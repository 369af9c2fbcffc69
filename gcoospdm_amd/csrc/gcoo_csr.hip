// gcoo_csr.hip -- the "csr" algorithm: the reference's row-wise CSR SpDM
// ablation (src/ShCsrSpGemm.cu:9-146), beside the grouped-COO path.
//
// Format: standard three-array CSR of S (M x K): rowPtr[M + 1], colInd[nnz]
// and vals[nnz], a row's entries in ascending column order.  The reference
// builds it with cusparseSnnz + cusparseSdense2csr (src/ShCsrSpGemm.cu:
// 67-110); here the dense -> sparse conversion is this library's own (one
// row per group, whose row offsets are rowPtr), followed by one pass that
// splits its (column, value) pairs into colInd and vals.
//
// Multiply: the reference kernel's structure
// (column_major_dense_sparse_csc_mat_mul_kernel_v2, src/ShCsrSpGemm.cu:9-43):
// one workgroup per (row of S, tile of BS output columns), consecutive
// workgroups taking consecutive rows of the same column tile; the row's
// entries are staged in LDS BS at a time (one per thread), then every
// thread runs its column's chain over them, c = fmaf(val, D[k][j], c) from
// +0.0f in ascending k.  The reference works on transposed copies of D and C
// (cublasSgeam before and after, :119-137) because its kernel reads the
// dense operand column-major; with row-major D and C the same chain reads
// D[k][j] directly, so no transposes are needed.  (The reference passes
// rowPtr and colInd to the kernel in swapped order, :128 against :9; the
// kernel's evident meaning -- BI = row offsets, BJ = column indices -- is
// what is built here.)  The bits equal every GCOO kernel's: one fmaf chain
// per output over the row's entries in ascending k.
#include "gcoo_internal.h"

#include <new>

struct gcoo_csr {
    int M = 0, K = 0;
    gcoo_handle g = nullptr;   // the conversion (one row per group)
    int *colInd = nullptr;     // [cap]
    float *vals = nullptr;     // [cap]
    int64_t cap = 0;
    bool converted = false;
};

namespace gcoo {
namespace {

constexpr int kCsrBS = 256;  // threads (output columns) per workgroup

__global__ __launch_bounds__(256) void csr_split_kernel(const int2 *__restrict__ seg, int64_t n,
                                                        int *__restrict__ colInd,
                                                        float *__restrict__ vals)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int2 e = seg[i];
        colInd[i] = e.x;
        vals[i] = __int_as_float(e.y);
    }
}

template <int BS>
__global__ __launch_bounds__(BS) void csr_spmm_kernel(const int *__restrict__ rowPtr,
                                                      const int *__restrict__ colInd,
                                                      const float *__restrict__ vals,
                                                      const float *__restrict__ D, int N,
                                                      float *__restrict__ C)
{
    __shared__ int sCol[BS];
    __shared__ float sVal[BS];
    const int r = blockIdx.x;
    const int t = threadIdx.x;
    const int j = blockIdx.y * BS + t;
    const int b = rowPtr[r], e = rowPtr[r + 1];
    float c = 0.0f;
    for (int k0 = b; k0 < e; k0 += BS) {
        const int n = e - k0 < BS ? e - k0 : BS;
        if (t < n) {
            sCol[t] = colInd[k0 + t];
            sVal[t] = vals[k0 + t];
        }
        __syncthreads();
        if (j < N) {
            const float *dcol = D + j;
            for (int i = 0; i < n; ++i) c = __builtin_fmaf(sVal[i], dcol[(size_t)sCol[i] * N], c);
        }
        __syncthreads();
    }
    if (j < N) C[(size_t)r * N + j] = c;
}

}  // namespace
}  // namespace gcoo

using namespace gcoo;

extern "C" {

int gcoo_csr_create(int M, int K, gcoo_csr_t *out)
{
    clear_error();
    if (!out) return set_error(GCOO_ERR_INVALID, "null csr out");
    if (M < 0 || K < 0) return set_error(GCOO_ERR_INVALID, "bad shape %d x %d", M, K);
    auto *c = new (std::nothrow) gcoo_csr;
    if (!c) return set_error(GCOO_ERR_ALLOC, "csr");
    c->M = M;
    c->K = K;
    int rc = gcoo_create(M, K, 1, &c->g);
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return GCOO_OK;
}

int gcoo_csr_destroy(gcoo_csr_t c)
{
    if (!c) return GCOO_OK;
    if (c->colInd) (void)hipFree(c->colInd);
    if (c->vals) (void)hipFree(c->vals);
    gcoo_destroy(c->g);
    delete c;
    return GCOO_OK;
}

int gcoo_csr_convert_dense(gcoo_csr_t c, const float *S, gcoo_stream stream)
{
    clear_error();
    if (!c) return set_error(GCOO_ERR_INVALID, "null csr");
    c->converted = false;
    if (int rc = gcoo_convert_dense(c->g, S, stream)) return rc;
    const int64_t n = c->g->nnz;
    if (n > c->cap) {
        if (c->colInd) GCOO_HIP_TRY(hipFree(c->colInd));
        if (c->vals) GCOO_HIP_TRY(hipFree(c->vals));
        c->colInd = nullptr;
        c->vals = nullptr;
        c->cap = 0;
        const int64_t cap = n + n / 16 + 1024;
        if (hipMalloc(&c->colInd, sizeof(int) * (size_t)cap) != hipSuccess ||
            hipMalloc(&c->vals, sizeof(float) * (size_t)cap) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(GCOO_ERR_ALLOC, "cannot allocate CSR arrays of %lld entries",
                             (long long)cap);
        }
        c->cap = cap;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (n > 0) {
        const int64_t blocks64 = (n + 255) / 256;
        const unsigned blocks = (unsigned)(blocks64 < 65536 ? blocks64 : 65536);
        hipLaunchKernelGGL(csr_split_kernel, dim3(blocks), dim3(256), 0, s, c->g->seg, n,
                           c->colInd, c->vals);
        GCOO_CHECK_LAUNCH("csr_split_kernel");
    }
    GCOO_HIP_TRY(hipStreamSynchronize(s));
    c->converted = true;
    return GCOO_OK;
}

int gcoo_csr_get_arrays(gcoo_csr_t c, const int **rowPtr, const int **colInd, const float **vals,
                        int64_t *nnz)
{
    clear_error();
    if (!c) return set_error(GCOO_ERR_INVALID, "null csr");
    if (!c->converted) return set_error(GCOO_ERR_STATE, "CSR matrix not converted");
    if (rowPtr) *rowPtr = c->g->rowptr;
    if (colInd) *colInd = c->colInd;
    if (vals) *vals = c->vals;
    if (nnz) *nnz = c->g->nnz;
    return GCOO_OK;
}

int gcoo_csr_spmm_f32_async(gcoo_csr_t c, const float *D, int N, float *C, gcoo_stream stream)
{
    clear_error();
    if (!c) return set_error(GCOO_ERR_INVALID, "null csr");
    if (!c->converted) return set_error(GCOO_ERR_STATE, "CSR matrix not converted");
    if (N < 0) return set_error(GCOO_ERR_INVALID, "N = %d", N);
    if (N == 0 || c->M == 0) return GCOO_OK;
    if (!C || (!D && c->K > 0)) return set_error(GCOO_ERR_INVALID, "null D or C");
    const long long tiles = (N + kCsrBS - 1) / kCsrBS;
    if (tiles > 65535) return set_error(GCOO_ERR_UNSUPPORTED, "N = %d too wide", N);
    hipLaunchKernelGGL((csr_spmm_kernel<kCsrBS>), dim3((unsigned)c->M, (unsigned)tiles),
                       dim3(kCsrBS), 0, reinterpret_cast<hipStream_t>(stream), c->g->rowptr,
                       c->colInd, c->vals, D, N, C);
    GCOO_CHECK_LAUNCH("csr_spmm_kernel");
    return GCOO_OK;
}

int gcoo_csr_spmm_f32(gcoo_csr_t c, const float *D, int N, float *C, gcoo_stream stream)
{
    if (int rc = gcoo_csr_spmm_f32_async(c, D, N, C, stream)) return rc;
    GCOO_HIP_TRY(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
    return GCOO_OK;
}

}  // extern "C"
