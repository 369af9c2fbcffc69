// ShGCOOSpDM.h -- the GCOO backend behind the reference's ShGemm interface
// (reference src/ShGCOOSpDM.h:1-35).  All device work goes through the C ABI
// of libgcoospdm.so (include/gcoo_spdm.h); this class holds no HIP types.
#ifndef GCOO_SHGCOOSPDM_H
#define GCOO_SHGCOOSPDM_H

#include "ShGemm.h"
#include "gcoo_spdm.h"

class ShGCOOSpDM : public ShGemm
{
public:
    // denseMatrix: device pointer to S (nRows x nCols), not owned.
    // cpg: rows per group (1, 2 or 4; the reference's COLUMN_PER_GROUP = 4).
    ShGCOOSpDM(float *denseMatrix, int nRows, int nCols, int cpg = 4,
               gcoo_stream stream = nullptr);
    ~ShGCOOSpDM() override;
    ShGCOOSpDM(const ShGCOOSpDM &) = delete;
    ShGCOOSpDM &operator=(const ShGCOOSpDM &) = delete;

    // Reference src/ShGCOOSpDM.cu:19-126.  The arrays stay owned by this
    // object (the reference hands out fresh allocations and leaks them).
    void convertToGroupCOOFormat(float *denseMatrix, int nRows, int nCols,
                                 float *&pVals, int *&pCols, int *&pRows,
                                 int *&pGroupIndex, int *&pNnzPerGroup,
                                 int nGroup);

    // Returns once the conversion is checked, with the multiply possibly
    // still running on the object's stream (as the reference's does):
    // synchronise that stream before reading C.
    double mutiplyBy(float *A, int nRows, int nCols, float *C) override;
    double mutiply(float *B, int nRows, int nCols, float *C) override;
    void build() override;
    double runKernelTest(float *A, int nRows, int nCols, float *C) override;

    long long nnz() const;
    gcoo_handle handle() const { return m_h; }

public:
    // Public like the reference's members (src/ShGCOOSpDM.h:25-32); valid
    // after build() or convertToGroupCOOFormat(), device pointers.
    float *m_pVals = nullptr;
    int *m_pCols = nullptr;
    int *m_pRows = nullptr;
    int *m_pGroupIndex = nullptr;
    int *m_pNnzPerGroup = nullptr;
    int m_nGroup = 0;

private:
    void refresh_members();
    gcoo_handle m_h = nullptr;
    gcoo_stream m_stream = nullptr;
    int m_cpg = 4;
};

#endif
