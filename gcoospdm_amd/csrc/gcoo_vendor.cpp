// gcoo_vendor.cpp -- libgcoo_vendor.so: the reference's vendor baselines
// (cuSPARSE CSR SpMM and cuBLAS SGEMM, src/ShCusparseGemm.cpp:23-118 and
// src/ShCublasGemm.cpp:16-26) on ROCm's rocSPARSE and rocBLAS, for
// comparison runs beside the GCOO path.  Host code only; the device work is
// the libraries'.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsparse/rocsparse.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/gcoo_vendor.h"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                      \
            return fail(GCOO_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

#define SPARSE_TRY(expr)                                                           \
    do {                                                                           \
        rocsparse_status e_ = (expr);                                              \
        if (e_ != rocsparse_status_success)                                        \
            return fail(GCOO_ERR_HIP, "%s: rocsparse status %d", #expr, (int)e_);  \
    } while (0)

#define BLAS_TRY(expr)                                                             \
    do {                                                                           \
        rocblas_status e_ = (expr);                                                \
        if (e_ != rocblas_status_success)                                          \
            return fail(GCOO_ERR_HIP, "%s: rocblas status %d", #expr, (int)e_);    \
    } while (0)

hipStream_t as_stream(gcoo_stream s) { return reinterpret_cast<hipStream_t>(s); }

// Grow a device buffer to at least `bytes` (contents discarded).
int grow(void **p, size_t *have, size_t bytes)
{
    if (bytes <= *have) return GCOO_OK;
    if (*p) HIP_TRY(hipFree(*p));
    *p = nullptr;
    *have = 0;
    const size_t want = bytes + bytes / 8 + 256;
    if (hipMalloc(p, want) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(GCOO_ERR_ALLOC, "cannot allocate %zu bytes", want);
    }
    *have = want;
    return GCOO_OK;
}

}  // namespace

struct gcoo_vendor_csr {
    int M = 0, K = 0;
    int64_t nnz = -1;  // -1: not converted
    rocsparse_handle handle = nullptr;
    rocsparse_mat_descr descr = nullptr;
    int *rowptr = nullptr;  // [M + 1]
    int *perRow = nullptr;  // [M]
    void *colind = nullptr;
    size_t colindBytes = 0;
    void *vals = nullptr;
    size_t valsBytes = 0;
    void *work = nullptr;
    size_t workBytes = 0;
};

struct gcoo_vendor_blas {
    rocblas_handle handle = nullptr;
};

extern "C" {

const char *gcoo_vendor_last_error(void) { return g_err; }

int gcoo_vendor_csr_create(int M, int K, gcoo_vendor_csr_t *out)
{
    if (!out || M < 0 || K < 0) return fail(GCOO_ERR_INVALID, "bad csr_create arguments");
    *out = nullptr;
    auto *c = new gcoo_vendor_csr;
    c->M = M;
    c->K = K;
    if (rocsparse_create_handle(&c->handle) != rocsparse_status_success ||
        rocsparse_create_mat_descr(&c->descr) != rocsparse_status_success ||
        hipMalloc(&c->rowptr, sizeof(int) * ((size_t)M + 1)) != hipSuccess ||
        hipMalloc(&c->perRow, sizeof(int) * ((size_t)M + 1)) != hipSuccess) {
        (void)hipGetLastError();
        gcoo_vendor_csr_destroy(c);
        return fail(GCOO_ERR_ALLOC, "cannot create the rocSPARSE CSR state");
    }
    *out = c;
    return GCOO_OK;
}

int gcoo_vendor_csr_destroy(gcoo_vendor_csr_t c)
{
    if (!c) return GCOO_OK;
    if (c->descr) rocsparse_destroy_mat_descr(c->descr);
    if (c->handle) rocsparse_destroy_handle(c->handle);
    if (c->rowptr) (void)hipFree(c->rowptr);
    if (c->perRow) (void)hipFree(c->perRow);
    if (c->colind) (void)hipFree(c->colind);
    if (c->vals) (void)hipFree(c->vals);
    if (c->work) (void)hipFree(c->work);
    delete c;
    return GCOO_OK;
}

int gcoo_vendor_csr_convert(gcoo_vendor_csr_t c, const float *S, gcoo_stream s)
{
    if (!c || (!S && (size_t)c->M * c->K > 0)) return fail(GCOO_ERR_INVALID, "null argument");
    SPARSE_TRY(rocsparse_set_stream(c->handle, as_stream(s)));
    if (c->M == 0 || c->K == 0) {
        HIP_TRY(hipMemsetAsync(c->rowptr, 0, sizeof(int) * ((size_t)c->M + 1), as_stream(s)));
        HIP_TRY(hipStreamSynchronize(as_stream(s)));
        c->nnz = 0;
        return GCOO_OK;
    }
    // Row-major S (M x K) is column-major S^T (K x M): its per-column counts
    // are S's per-row counts and its CSC is S's CSR.
    rocsparse_int total = 0;
    SPARSE_TRY(rocsparse_set_pointer_mode(c->handle, rocsparse_pointer_mode_host));
    SPARSE_TRY(rocsparse_snnz(c->handle, rocsparse_direction_column, c->K, c->M, c->descr, S,
                              c->K, c->perRow, &total));
    int rc;
    if ((rc = grow(&c->colind, &c->colindBytes, sizeof(int) * ((size_t)total + 1)))) return rc;
    if ((rc = grow(&c->vals, &c->valsBytes, sizeof(float) * ((size_t)total + 1)))) return rc;
    SPARSE_TRY(rocsparse_sdense2csc(c->handle, c->K, c->M, c->descr, S, c->K, c->perRow,
                                    static_cast<float *>(c->vals), c->rowptr,
                                    static_cast<int *>(c->colind)));
    HIP_TRY(hipStreamSynchronize(as_stream(s)));
    c->nnz = total;
    return GCOO_OK;
}

int gcoo_vendor_csr_nnz(gcoo_vendor_csr_t c, int64_t *nnz)
{
    if (!c || !nnz) return fail(GCOO_ERR_INVALID, "null argument");
    if (c->nnz < 0) return fail(GCOO_ERR_STATE, "not converted");
    *nnz = c->nnz;
    return GCOO_OK;
}

int gcoo_vendor_csr_spmm(gcoo_vendor_csr_t c, const float *D, int K, int N, float *C,
                         gcoo_stream s)
{
    if (!c || (!D && (size_t)K * N > 0) || (!C && (size_t)c->M * N > 0))
        return fail(GCOO_ERR_INVALID, "null argument");
    if (c->nnz < 0) return fail(GCOO_ERR_STATE, "not converted");
    if (K != c->K || N < 0) return fail(GCOO_ERR_INVALID, "K %d (expected %d), N %d", K, c->K, N);
    if (c->M == 0 || N == 0) return GCOO_OK;
    SPARSE_TRY(rocsparse_set_stream(c->handle, as_stream(s)));
    SPARSE_TRY(rocsparse_set_pointer_mode(c->handle, rocsparse_pointer_mode_host));
    rocsparse_spmat_descr A = nullptr;
    rocsparse_dnmat_descr B = nullptr, Cd = nullptr;
    int rc = GCOO_OK;
    const float alpha = 1.0f, beta = 0.0f;
    size_t bytes = 0;
    if (rocsparse_create_csr_descr(&A, c->M, c->K, c->nnz, c->rowptr, c->colind, c->vals,
                                   rocsparse_indextype_i32, rocsparse_indextype_i32,
                                   rocsparse_index_base_zero,
                                   rocsparse_datatype_f32_r) != rocsparse_status_success ||
        rocsparse_create_dnmat_descr(&B, K, N, N, const_cast<float *>(D),
                                     rocsparse_datatype_f32_r,
                                     rocsparse_order_row) != rocsparse_status_success ||
        rocsparse_create_dnmat_descr(&Cd, c->M, N, N, C, rocsparse_datatype_f32_r,
                                     rocsparse_order_row) != rocsparse_status_success) {
        rc = fail(GCOO_ERR_HIP, "cannot create the rocSPARSE descriptors");
    }
    if (!rc &&
        rocsparse_spmm(c->handle, rocsparse_operation_none, rocsparse_operation_none, &alpha, A,
                       B, &beta, Cd, rocsparse_datatype_f32_r, rocsparse_spmm_alg_default,
                       rocsparse_spmm_stage_buffer_size, &bytes,
                       nullptr) != rocsparse_status_success)
        rc = fail(GCOO_ERR_HIP, "rocsparse_spmm (buffer size) failed");
    if (!rc) rc = grow(&c->work, &c->workBytes, bytes ? bytes : 1);
    if (!rc &&
        rocsparse_spmm(c->handle, rocsparse_operation_none, rocsparse_operation_none, &alpha, A,
                       B, &beta, Cd, rocsparse_datatype_f32_r, rocsparse_spmm_alg_default,
                       rocsparse_spmm_stage_preprocess, &bytes,
                       c->work) != rocsparse_status_success)
        rc = fail(GCOO_ERR_HIP, "rocsparse_spmm (preprocess) failed");
    if (!rc &&
        rocsparse_spmm(c->handle, rocsparse_operation_none, rocsparse_operation_none, &alpha, A,
                       B, &beta, Cd, rocsparse_datatype_f32_r, rocsparse_spmm_alg_default,
                       rocsparse_spmm_stage_compute, &bytes,
                       c->work) != rocsparse_status_success)
        rc = fail(GCOO_ERR_HIP, "rocsparse_spmm (compute) failed");
    if (Cd) rocsparse_destroy_dnmat_descr(Cd);
    if (B) rocsparse_destroy_dnmat_descr(B);
    if (A) rocsparse_destroy_spmat_descr(A);
    return rc;
}

int gcoo_vendor_blas_create(gcoo_vendor_blas_t *out)
{
    if (!out) return fail(GCOO_ERR_INVALID, "null argument");
    *out = nullptr;
    auto *b = new gcoo_vendor_blas;
    if (rocblas_create_handle(&b->handle) != rocblas_status_success) {
        delete b;
        return fail(GCOO_ERR_HIP, "rocblas_create_handle failed");
    }
    *out = b;
    return GCOO_OK;
}

int gcoo_vendor_blas_destroy(gcoo_vendor_blas_t b)
{
    if (!b) return GCOO_OK;
    if (b->handle) rocblas_destroy_handle(b->handle);
    delete b;
    return GCOO_OK;
}

int gcoo_vendor_sgemm(gcoo_vendor_blas_t b, const float *S, const float *D, int M, int K,
                      int N, float *C, gcoo_stream s)
{
    if (!b || M < 0 || K < 0 || N < 0) return fail(GCOO_ERR_INVALID, "bad sgemm arguments");
    if ((size_t)M * N == 0) return GCOO_OK;
    if ((!S || !D) && K > 0) return fail(GCOO_ERR_INVALID, "null argument");
    if (!C) return fail(GCOO_ERR_INVALID, "null argument");
    BLAS_TRY(rocblas_set_stream(b->handle, as_stream(s)));
    BLAS_TRY(rocblas_set_pointer_mode(b->handle, rocblas_pointer_mode_host));
    const float alpha = 1.0f, beta = 0.0f;
    // Row-major C = S * D is column-major C^T = D^T * S^T.
    BLAS_TRY(rocblas_sgemm(b->handle, rocblas_operation_none, rocblas_operation_none, N, M, K,
                           &alpha, D, N, S, K, &beta, C, N));
    return GCOO_OK;
}

}  // extern "C"
