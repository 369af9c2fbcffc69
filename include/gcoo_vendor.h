/* gcoo_vendor.h -- C ABI of libgcoo_vendor.so: the reference's vendor
 * baselines on ROCm, for comparison runs beside the GCOO path.
 *
 *   "cusparse"  reference src/ShCusparseGemm.cpp:23-118 (cusparseSnnz +
 *               cusparseSdense2csr, then cusparseScsrmm2 on every
 *               mutiplyBy): here rocSPARSE -- rocsparse_snnz +
 *               rocsparse_sdense2csc on the column-major view of S (whose
 *               CSC is S's CSR), then rocsparse_spmm.
 *   "cublas"    reference src/ShCublasGemm.cpp:16-26 (cublasSgemm on the
 *               dense S): here rocblas_sgemm.
 *
 * Every product is C = S * D with S (M x K), D (K x N) and C (M x N)
 * row-major FP32 device arrays -- the GCOO path's convention, so the three
 * algorithms compute the same thing.  (The reference's vendor calls are
 * written in column-major terms and agree with it only for square inputs.)
 * The summation order is the library's: results agree with the FP64 product
 * to the FP32 tolerance, not bit for bit.  Status codes are those of
 * gcoo_spdm.h; on failure gcoo_vendor_last_error() describes it.  These
 * baselines have no counterpart in the GCOO path's own ABI and are not used
 * by it. */
#ifndef GCOO_VENDOR_H
#define GCOO_VENDOR_H

#include <stdint.h>

#include "gcoo_spdm.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *gcoo_vendor_last_error(void);

/* rocSPARSE CSR form of an M x K sparse S. */
typedef struct gcoo_vendor_csr *gcoo_vendor_csr_t;
int gcoo_vendor_csr_create(int M, int K, gcoo_vendor_csr_t *out);
int gcoo_vendor_csr_destroy(gcoo_vendor_csr_t c);
/* Dense S (device, row-major M x K) -> CSR; synchronous (reads nnz back). */
int gcoo_vendor_csr_convert(gcoo_vendor_csr_t c, const float *S, gcoo_stream s);
int gcoo_vendor_csr_nnz(gcoo_vendor_csr_t c, int64_t *nnz);
/* C = S * D with the converted S (rocsparse_spmm); enqueued on s. */
int gcoo_vendor_csr_spmm(gcoo_vendor_csr_t c, const float *D, int K, int N, float *C,
                         gcoo_stream s);

/* rocBLAS dense SGEMM. */
typedef struct gcoo_vendor_blas *gcoo_vendor_blas_t;
int gcoo_vendor_blas_create(gcoo_vendor_blas_t *out);
int gcoo_vendor_blas_destroy(gcoo_vendor_blas_t b);
/* C = S * D, all dense; enqueued on s. */
int gcoo_vendor_sgemm(gcoo_vendor_blas_t b, const float *S, const float *D, int M, int K,
                      int N, float *C, gcoo_stream s);

#ifdef __cplusplus
}
#endif

#endif
