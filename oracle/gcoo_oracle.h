/*
 * gcoo_oracle.h -- CPU restatement of the reference GCOOSpDM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and by nothing else.  The product (gcoospdm_amd/) never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Storage is row-major FP32 throughout: S is the sparse operand (M x K),
 * D the dense operand (K x N), C = S * D (M x N), exactly as the reference
 * computes it (reference src/main.cu:131, src/utils.cpp:30-46).
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   orc_gcoo_count / orc_exclusive_scan / orc_gcoo_convert
 *       brute-force enumeration + sort on tiny inputs, closed-form invariants,
 *       committed golden fixtures (tests/golden/).
 *   orc_gcoo_spdm_f32
 *       bit-exact against an independent numpy float32 fmaf emulation on tiny
 *       inputs, S = I  =>  C = D, single nonzero => scaled row, and the FP64
 *       bound below.
 *   orc_spdm_fp64 / orc_dense_fp64
 *       bit-exact against the reference's own matrixMulCPU compiled from
 *       /root/reference/src/utils.cpp into oracle/_ref/ (when that builds),
 *       and against numpy float64 matmul within 1 ulp-of-double bounds.
 *   orc_rel_frobenius
 *       closed form on hand-made vectors.
 */
#ifndef GCOO_ORACLE_H
#define GCOO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* K1: nnzPerGroup[g] = #{(r,k): cpg*g <= r < min(cpg*g+cpg, M), S[r][k] != 0.0}
 * for g < nGroup, and nnzPerGroup[nGroup] = 0 (the memset of the shared
 * allocation).  Returns the total.
 * Reference: src/gcoo_spdm_kernels.cu:4-28, src/ShGCOOSpDM.cu:33-51. */
int64_t orc_gcoo_count(const float *S, int M, int K, int cpg, int *nnzPerGroup);

/* K2: dst[0] = 0; dst[i] = dst[i-1] + src[i-1] for 0 < i < n.
 * Reference: src/gcoo_spdm_kernels.cu:160-168, launched src/ShGCOOSpDM.cu:65. */
void orc_exclusive_scan(const int *src, int *dst, int n);

/* A2..A6: full dense -> GCOO conversion.  Within group g the entries are the
 * nonzeros of rows [cpg*g, cpg*g+cpg) sorted by (k, r), stored at
 * [groupIndex[g], groupIndex[g+1]).  vals keeps the bits of S[r][k].
 * groupIndex and nnzPerGroup have nGroup+1 entries.  Returns nnz.
 * Arrays vals/rows/cols must hold at least nnz entries (call orc_gcoo_count
 * first to size them).
 * Reference: src/gcoo_spdm_kernels.cu:30-90 (K3), src/ShGCOOSpDM.cu:19-126. */
int64_t orc_gcoo_convert(const float *S, int M, int K, int cpg,
                         float *vals, int *rows, int *cols,
                         int *groupIndex, int *nnzPerGroup);

/* K4: C = S * D from the GCOO arrays.  Each C[r][j] is one FP32 fmaf chain
 * starting from +0.0f over row r's nonzeros in ascending k (the group's
 * (k, r) order), exactly the reference kernel's arithmetic.  Rows of the
 * tail group past M are not written.  nthreads > 1 uses OpenMP over groups
 * (each group is independent, so the result is identical).
 * Reference: src/gcoo_spdm_kernels.cu:92-158, src/ShGCOOSpDM.cu:153-167. */
void orc_gcoo_spdm_f32(const float *vals, const int *rows, const int *cols,
                       const int *groupIndex, const int *nnzPerGroup,
                       int nGroup, int cpg, int M, const float *D, int N,
                       float *C, int nthreads);

/* Same arithmetic with BF16 storage: vals_bf16 and D_bf16 are the upper 16
 * bits of FP32 values (round-to-nearest-even from FP32 by orc_f32_to_bf16),
 * widened exactly to FP32 and accumulated with fmaf in the same order. */
void orc_gcoo_spdm_bf16(const uint16_t *vals_bf16, const int *rows,
                        const int *cols, const int *groupIndex,
                        const int *nnzPerGroup, int nGroup, int cpg, int M,
                        const uint16_t *D_bf16, int N, float *C, int nthreads);

/* IEEE RNE FP32 -> BF16 (NaN kept quiet).  Textbook definition. */
void orc_f32_to_bf16(const float *x, uint16_t *y, int64_t n);

/* FP64 oracle, sparse form: Cd[i][j] = sum_k (double)S[i][k]*(double)D[k][j],
 * k ascending, over the GCOO arrays.  Equal to matrixMulCPU for finite
 * inputs because zero terms add exactly 0 and f32*f32 is exact in double.
 * C (optional) receives (float)Cd.
 * Reference: src/utils.cpp:30-46. */
void orc_spdm_fp64(const float *vals, const int *rows, const int *cols,
                   const int *groupIndex, int nGroup, int cpg, int M,
                   const float *D, int N, double *Cd, float *C, int nthreads);

/* matrixMulCPU semantics on the dense S (single thread, k ascending, double
 * accumulator).  Reference: src/utils.cpp:30-46. */
void orc_dense_fp64(const float *S, const float *D, int M, int K, int N,
                    double *Cd, float *C);

/* ||C - R||_F / ||R||_F in double (the relative-L2 of sdkCompareL2fe,
 * restated from its mathematical definition; reference src/main.cu:136). */
double orc_rel_frobenius(const float *C, const double *R, int64_t n);

/* Number of threads the OpenMP runtime would use (for baselines). */
int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
