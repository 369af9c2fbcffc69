// C-linkage entry points onto the reference's own C++ host utilities
// (src/utils.cpp), compiled from the reference sources into oracle/_ref/.
// Test infrastructure only: lets tests call the reference's matrixMulCPU and
// randomInit through ctypes.  It declares the reference functions; it does
// not restate them.
#include <stddef.h>
#include <stdlib.h>

int randomInit(float *data, size_t size, float sparsity);
void matrixMulCPU(float *C, const float *A, const float *B, unsigned int hA,
                  unsigned int wA, unsigned int wB);

extern "C" {
void ref_srand(unsigned int seed) { srand(seed); }
int ref_random_init(float *data, size_t size, float sparsity)
{
    return randomInit(data, size, sparsity);
}
void ref_matrix_mul_cpu(float *C, const float *A, const float *B,
                        unsigned int hA, unsigned int wA, unsigned int wB)
{
    matrixMulCPU(C, A, B, hA, wA, wB);
}
}
