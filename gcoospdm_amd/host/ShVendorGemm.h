// ShVendorGemm.h -- the reference's vendor baselines behind its ShGemm
// interface, on ROCm libraries through libgcoo_vendor.so (include/gcoo_vendor.h):
//   ShRocsparseGemm  factory key "cusparse" (reference src/ShCusparseGemm.cpp:23-118):
//                    dense S -> CSR on every mutiplyBy, then a CSR SpMM;
//   ShRocblasGemm    factory key "cublas" (reference src/ShCublasGemm.cpp:16-40):
//                    dense SGEMM of S and A.
// Both compute C = S * A (row-major, S = m_nRows x m_nCols), like the GCOO
// backend, and return the call's wall time in milliseconds (the reference
// returns 0.0).  Failures throw std::runtime_error.
#ifndef GCOO_SHVENDORGEMM_H
#define GCOO_SHVENDORGEMM_H

#include "ShGemm.h"
#include "gcoo_vendor.h"

class ShRocsparseGemm : public ShGemm
{
public:
    ShRocsparseGemm(float *denseMatrix, int nRows, int nCols, gcoo_stream stream = nullptr);
    ~ShRocsparseGemm() override;
    ShRocsparseGemm(const ShRocsparseGemm &) = delete;
    ShRocsparseGemm &operator=(const ShRocsparseGemm &) = delete;

    double mutiplyBy(float *A, int nRows, int nCols, float *C) override;
    double mutiply(float *B, int nRows, int nCols, float *C) override;
    void build() override;
    double runKernelTest(float *A, int nRows, int nCols, float *C) override;
    long long nnz() const;

private:
    gcoo_vendor_csr_t m_csr = nullptr;
    gcoo_stream m_stream = nullptr;
};

class ShRocblasGemm : public ShGemm
{
public:
    ShRocblasGemm(float *denseMatrix, int nRows, int nCols, gcoo_stream stream = nullptr);
    ~ShRocblasGemm() override;
    ShRocblasGemm(const ShRocblasGemm &) = delete;
    ShRocblasGemm &operator=(const ShRocblasGemm &) = delete;

    double mutiplyBy(float *A, int nRows, int nCols, float *C) override;
    double mutiply(float *B, int nRows, int nCols, float *C) override;
    void build() override {}
    double runKernelTest(float *A, int nRows, int nCols, float *C) override;

private:
    gcoo_vendor_blas_t m_blas = nullptr;
    gcoo_stream m_stream = nullptr;
};

#endif
