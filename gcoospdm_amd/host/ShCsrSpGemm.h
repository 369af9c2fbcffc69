// ShCsrSpGemm.h -- factory key "csr": the reference's row-wise CSR SpDM
// ablation (src/ShCsrSpGemm.cu:9-146: one block per row of S, each thread
// one output column, accumulating the row's entries in CSR order).  Here it
// is the GCOO path with one row per group, whose row segments are exactly
// CSR, multiplied by the row-segment kernel: each output is one fmaf chain
// over the row's entries in ascending k, as in the reference kernel.
#ifndef GCOO_SHCSRSPGEMM_H
#define GCOO_SHCSRSPGEMM_H

#include <stdexcept>
#include <string>

#include "ShGCOOSpDM.h"

class ShCsrSpGemm : public ShGCOOSpDM
{
public:
    ShCsrSpGemm(float *denseMatrix, int nRows, int nCols, gcoo_stream stream = nullptr)
        : ShGCOOSpDM(denseMatrix, nRows, nCols, 1, stream)
    {
        const gcoo_spdm_config cfg = {2, 0, 0, 0, 0};  // row segments
        if (gcoo_set_spdm_config(handle(), &cfg) != GCOO_OK)
            throw std::runtime_error(std::string("gcoo_set_spdm_config: ") + gcoo_last_error());
    }
};

#endif
