// gcoo_coo.hip -- COO -> grouped COO (placeholder until the device path lands).
#include "gcoo_internal.h"

namespace gcoo {

int convert_coo(gcoo_matrix *h, const int *rows, const int *cols,
                const float *vals, int64_t nnz, hipStream_t s)
{
    (void)h; (void)rows; (void)cols; (void)vals; (void)nnz; (void)s;
    return set_error(GCOO_ERR_UNSUPPORTED, "COO conversion not built yet");
}

}  // namespace gcoo
