// ShGCOOSpDM.cpp -- reference src/ShGCOOSpDM.cu:13-218 restated over the
// C ABI: conversion and multiply are HIP kernels inside libgcoospdm.so.
#include "ShGCOOSpDM.h"

#include <chrono>
#include <stdexcept>
#include <string>

namespace {

void check(int rc, const char *what)
{
    if (rc != GCOO_OK)
        throw std::runtime_error(std::string(what) + ": " + gcoo_last_error());
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
        .count();
}

}  // namespace

ShGCOOSpDM::ShGCOOSpDM(float *denseMatrix, int nRows, int nCols, int cpg, gcoo_stream stream)
    : ShGemm(nRows, nCols), m_stream(stream), m_cpg(cpg)
{
    m_pDenseMatrix = denseMatrix;
    check(gcoo_create(nRows, nCols, cpg, &m_h), "gcoo_create");
    m_nGroup = (nRows + cpg - 1) / cpg;
}

ShGCOOSpDM::~ShGCOOSpDM() { gcoo_destroy(m_h); }

void ShGCOOSpDM::refresh_members()
{
    const float *v;
    const int *r, *c, *gi, *np;
    check(gcoo_get_arrays(m_h, &v, &r, &c, &gi, &np), "gcoo_get_arrays");
    m_pVals = const_cast<float *>(v);
    m_pRows = const_cast<int *>(r);
    m_pCols = const_cast<int *>(c);
    m_pGroupIndex = const_cast<int *>(gi);
    m_pNnzPerGroup = const_cast<int *>(np);
}

void ShGCOOSpDM::convertToGroupCOOFormat(float *denseMatrix, int nRows, int nCols,
                                         float *&pVals, int *&pCols, int *&pRows,
                                         int *&pGroupIndex, int *&pNnzPerGroup, int nGroup)
{
    if (nRows != m_nRows || nCols != m_nCols || nGroup != m_nGroup)
        throw std::invalid_argument("convertToGroupCOOFormat: shape differs from the object's");
    check(gcoo_convert_dense(m_h, denseMatrix, m_stream), "gcoo_convert_dense");
    refresh_members();
    pVals = m_pVals;
    pCols = m_pCols;
    pRows = m_pRows;
    pGroupIndex = m_pGroupIndex;
    pNnzPerGroup = m_pNnzPerGroup;
}

double ShGCOOSpDM::mutiplyBy(float *A, int nRows, int nCols, float *C)
{
    auto t0 = std::chrono::steady_clock::now();
    check(gcoo_multiply_by(m_h, m_pDenseMatrix, A, nRows, nCols, C, m_stream),
          "gcoo_multiply_by");
    return ms_since(t0);
}

double ShGCOOSpDM::mutiply(float *B, int nRows, int nCols, float *C)
{
    return mutiplyBy(B, nRows, nCols, C);
}

void ShGCOOSpDM::build()
{
    check(gcoo_convert_dense(m_h, m_pDenseMatrix, m_stream), "gcoo_convert_dense");
    refresh_members();
}

double ShGCOOSpDM::runKernelTest(float *A, int nRows, int nCols, float *C)
{
    auto t0 = std::chrono::steady_clock::now();
    check(gcoo_spdm_f32(m_h, A, nRows, nCols, C, m_stream), "gcoo_spdm_f32");
    return ms_since(t0);
}

long long ShGCOOSpDM::nnz() const
{
    int64_t n = 0;
    gcoo_get_info(m_h, nullptr, nullptr, nullptr, nullptr, &n);
    return (long long)n;
}
