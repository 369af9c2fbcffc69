// ShCsrSpGemm.h -- factory key "csr": the reference's row-wise CSR SpDM
// ablation (reference src/ShCsrSpGemm.h, src/ShCsrSpGemm.cu:9-146) over the
// C ABI's gcoo_csr_* entry points: S converted to the standard three-array
// CSR (rowPtr, colInd, vals; the reference uses cusparseSnnz +
// cusparseSdense2csr, src/ShCsrSpGemm.cu:67-110) and multiplied by a kernel
// of the reference kernel's structure (one workgroup per row and column
// tile, the row's entries staged in LDS, one output column per thread).
// Each C element is the fmaf chain over the row's entries in ascending k.
#ifndef GCOO_SHCSRSPGEMM_H
#define GCOO_SHCSRSPGEMM_H

#include <chrono>
#include <stdexcept>
#include <string>

#include "ShGemm.h"
#include "gcoo_spdm.h"

class ShCsrSpGemm : public ShGemm
{
public:
    // denseMatrix: device pointer to S (nRows x nCols), not owned.
    ShCsrSpGemm(float *denseMatrix, int nRows, int nCols, gcoo_stream stream = nullptr)
        : ShGemm(nRows, nCols), m_stream(stream)
    {
        m_pDenseMatrix = denseMatrix;
        check(gcoo_csr_create(nRows, nCols, &m_c), "gcoo_csr_create");
    }
    ~ShCsrSpGemm() override { gcoo_csr_destroy(m_c); }
    ShCsrSpGemm(const ShCsrSpGemm &) = delete;
    ShCsrSpGemm &operator=(const ShCsrSpGemm &) = delete;

    // Reference src/ShCsrSpGemm.cu:56-146: convert, multiply (synchronous).
    double mutiplyBy(float *A, int nRows, int nCols, float *C) override
    {
        const auto t0 = std::chrono::steady_clock::now();
        shape(nRows);
        check(gcoo_csr_convert_dense(m_c, m_pDenseMatrix, m_stream), "gcoo_csr_convert_dense");
        check(gcoo_csr_spmm_f32(m_c, A, nCols, C, m_stream), "gcoo_csr_spmm_f32");
        return ms_since(t0);
    }
    double mutiply(float *B, int nRows, int nCols, float *C) override
    {
        return mutiplyBy(B, nRows, nCols, C);
    }
    void build() override
    {
        check(gcoo_csr_convert_dense(m_c, m_pDenseMatrix, m_stream), "gcoo_csr_convert_dense");
    }
    double runKernelTest(float *A, int nRows, int nCols, float *C) override
    {
        const auto t0 = std::chrono::steady_clock::now();
        shape(nRows);
        check(gcoo_csr_spmm_f32(m_c, A, nCols, C, m_stream), "gcoo_csr_spmm_f32");
        return ms_since(t0);
    }
    gcoo_csr_t csr() const { return m_c; }

private:
    static void check(int rc, const char *what)
    {
        if (rc != GCOO_OK)
            throw std::runtime_error(std::string(what) + ": " + gcoo_last_error());
    }
    static double ms_since(std::chrono::steady_clock::time_point t0)
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
            .count();
    }
    void shape(int nRowsA) const
    {
        if (nRowsA != m_nCols)
            throw std::invalid_argument("ShCsrSpGemm: A has " + std::to_string(nRowsA) +
                                        " rows, S has " + std::to_string(m_nCols) + " columns");
    }
    gcoo_csr_t m_c = nullptr;
    gcoo_stream m_stream = nullptr;
};

#endif
