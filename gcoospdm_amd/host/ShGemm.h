// ShGemm.h -- abstract sparse GEMM algorithm, the reference's plugin
// interface (reference src/ShGemm.h:7-24), kept signature for signature,
// including the spelling of mutiplyBy / mutiply.
//
// All matrix pointers are caller-owned DEVICE pointers; row-major FP32.
// The object wraps a sparse S (m_nRows x m_nCols) given as a dense device
// matrix (m_pDenseMatrix, not owned).  Differences from the reference, by
// design: the destructor is virtual (the reference deletes through a base
// pointer with a non-virtual destructor, src/main.cu:90) and failures throw
// std::runtime_error instead of exiting.
#ifndef GCOO_SHGEMM_H
#define GCOO_SHGEMM_H

class ShGemm
{
public:
    ShGemm() {}
    ShGemm(int nRows, int nCols) : m_nRows(nRows), m_nCols(nCols) {}
    virtual ~ShGemm() {}
    // C = S * A, A is nRows x nCols dense (nRows must equal m_nCols).
    // Converts S and multiplies on every call.  Returns the call's wall
    // time in milliseconds.
    virtual double mutiplyBy(float *A, int nRows, int nCols, float *C) = 0;
    // C = S * B (same product; the reference left it unimplemented).
    virtual double mutiply(float *B, int nRows, int nCols, float *C) = 0;
    // Convert S once and keep the sparse form.
    virtual void build() = 0;
    // C = S * A with the form kept by build(): the multiply kernel only.
    virtual double runKernelTest(float *A, int nRows, int nCols, float *C) = 0;

public:
    int m_nRows = 0;
    int m_nCols = 0;
    float *m_pDenseMatrix = nullptr;
};

#endif
