// ShVendorGemm.cpp -- rocSPARSE / rocBLAS baselines over libgcoo_vendor.so.
#include "ShVendorGemm.h"

#include <chrono>
#include <stdexcept>
#include <string>

namespace {

void check(int rc, const char *what)
{
    if (rc != GCOO_OK)
        throw std::runtime_error(std::string(what) + ": " + gcoo_vendor_last_error());
}

void check_gcoo(int rc, const char *what)
{
    if (rc != GCOO_OK) throw std::runtime_error(std::string(what) + ": " + gcoo_last_error());
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
        .count();
}

}  // namespace

ShRocsparseGemm::ShRocsparseGemm(float *denseMatrix, int nRows, int nCols, gcoo_stream stream)
    : ShGemm(nRows, nCols), m_stream(stream)
{
    m_pDenseMatrix = denseMatrix;
    check(gcoo_vendor_csr_create(nRows, nCols, &m_csr), "gcoo_vendor_csr_create");
}

ShRocsparseGemm::~ShRocsparseGemm() { gcoo_vendor_csr_destroy(m_csr); }

double ShRocsparseGemm::mutiplyBy(float *A, int nRows, int nCols, float *C)
{
    auto t0 = std::chrono::steady_clock::now();
    check(gcoo_vendor_csr_convert(m_csr, m_pDenseMatrix, m_stream), "gcoo_vendor_csr_convert");
    check(gcoo_vendor_csr_spmm(m_csr, A, nRows, nCols, C, m_stream), "gcoo_vendor_csr_spmm");
    check_gcoo(gcoo_stream_sync(m_stream), "gcoo_stream_sync");
    return ms_since(t0);
}

double ShRocsparseGemm::mutiply(float *B, int nRows, int nCols, float *C)
{
    return mutiplyBy(B, nRows, nCols, C);
}

void ShRocsparseGemm::build()
{
    check(gcoo_vendor_csr_convert(m_csr, m_pDenseMatrix, m_stream), "gcoo_vendor_csr_convert");
}

double ShRocsparseGemm::runKernelTest(float *A, int nRows, int nCols, float *C)
{
    auto t0 = std::chrono::steady_clock::now();
    check(gcoo_vendor_csr_spmm(m_csr, A, nRows, nCols, C, m_stream), "gcoo_vendor_csr_spmm");
    check_gcoo(gcoo_stream_sync(m_stream), "gcoo_stream_sync");
    return ms_since(t0);
}

long long ShRocsparseGemm::nnz() const
{
    int64_t n = 0;
    return gcoo_vendor_csr_nnz(m_csr, &n) == GCOO_OK ? (long long)n : -1;
}

ShRocblasGemm::ShRocblasGemm(float *denseMatrix, int nRows, int nCols, gcoo_stream stream)
    : ShGemm(nRows, nCols), m_stream(stream)
{
    m_pDenseMatrix = denseMatrix;
    check(gcoo_vendor_blas_create(&m_blas), "gcoo_vendor_blas_create");
}

ShRocblasGemm::~ShRocblasGemm() { gcoo_vendor_blas_destroy(m_blas); }

double ShRocblasGemm::mutiplyBy(float *A, int nRows, int nCols, float *C)
{
    if (nRows != m_nCols)
        throw std::invalid_argument("ShRocblasGemm: rows of A must equal columns of S");
    auto t0 = std::chrono::steady_clock::now();
    check(gcoo_vendor_sgemm(m_blas, m_pDenseMatrix, A, m_nRows, m_nCols, nCols, C, m_stream),
          "gcoo_vendor_sgemm");
    check_gcoo(gcoo_stream_sync(m_stream), "gcoo_stream_sync");
    return ms_since(t0);
}

double ShRocblasGemm::mutiply(float *B, int nRows, int nCols, float *C)
{
    return mutiplyBy(B, nRows, nCols, C);
}

double ShRocblasGemm::runKernelTest(float *A, int nRows, int nCols, float *C)
{
    return mutiplyBy(A, nRows, nCols, C);
}
