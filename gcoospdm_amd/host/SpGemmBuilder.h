// SpGemmBuilder.h -- string factory of ShGemm algorithms (reference
// src/SpGemmBuilder.h:12-23).  "gcoospdm" is the GCOO backend.  The
// reference's vendor baselines ("cusparse", "cublas") and its "csr" kernel
// are not part of this library; their keys return NULL exactly like an
// unknown key does in the reference.
#ifndef GCOO_SPGEMMBUILDER_H
#define GCOO_SPGEMMBUILDER_H

#include <string>

#include "ShGCOOSpDM.h"
#include "ShGemm.h"

class SpGemmBuilder
{
public:
    static ShGemm *getShGemmAlgo(float *denseMatrix, int nRows, int nCols,
                                 const std::string &desc)
    {
        if (desc == "gcoospdm") return new ShGCOOSpDM(denseMatrix, nRows, nCols);
        return nullptr;
    }
};

#endif
