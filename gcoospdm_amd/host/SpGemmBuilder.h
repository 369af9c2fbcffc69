// SpGemmBuilder.h -- string factory of ShGemm algorithms (reference
// src/SpGemmBuilder.h:12-23).  "gcoospdm" is the GCOO backend; "csr" the
// row-wise CSR ablation (the same library, one row per group); "cusparse"
// and "cublas" the vendor baselines on rocSPARSE and rocBLAS
// (libgcoo_vendor.so).  Unknown keys return NULL, as in the reference.
#ifndef GCOO_SPGEMMBUILDER_H
#define GCOO_SPGEMMBUILDER_H

#include <string>

#include "ShCsrSpGemm.h"
#include "ShGCOOSpDM.h"
#include "ShGemm.h"
#include "ShVendorGemm.h"

class SpGemmBuilder
{
public:
    static ShGemm *getShGemmAlgo(float *denseMatrix, int nRows, int nCols,
                                 const std::string &desc)
    {
        if (desc == "gcoospdm") return new ShGCOOSpDM(denseMatrix, nRows, nCols);
        if (desc == "csr") return new ShCsrSpGemm(denseMatrix, nRows, nCols);
        if (desc == "cusparse") return new ShRocsparseGemm(denseMatrix, nRows, nCols);
        if (desc == "cublas") return new ShRocblasGemm(denseMatrix, nRows, nCols);
        return nullptr;
    }
};

#endif
