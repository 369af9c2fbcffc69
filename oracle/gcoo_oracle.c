/*
 * gcoo_oracle.c -- CPU restatement of the reference GCOOSpDM hot path.
 * TEST INFRASTRUCTURE ONLY (see gcoo_oracle.h): used by tests/, smoke() and
 * bench.py's cpu_baseline leg as the checker and the CPU baseline; never by
 * the product.
 */
#include "gcoo_oracle.h"
#include <math.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* The reference tests `v != 0.0` (float promoted to double): -0.0 is dropped,
 * NaN and denormals are kept.  src/gcoo_spdm_kernels.cu:17,59,76. */
static inline int is_nz(float v) { return v != 0.0; }

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* src/gcoo_spdm_kernels.cu:4-28 (count per group), src/ShGCOOSpDM.cu:37
 * (nnzPerGroup[nGroup] stays 0 from the memset). */
int64_t orc_gcoo_count(const float *S, int M, int K, int cpg, int *nnzPerGroup)
{
    int nGroup = (M + cpg - 1) / cpg;
    int64_t total = 0;
    for (int g = 0; g < nGroup; g++) {
        int c = 0;
        for (int j = 0; j < cpg; j++) {
            int64_t r = (int64_t)g * cpg + j;
            if (r >= M) break;
            const float *row = S + r * (int64_t)K;
            for (int k = 0; k < K; k++) c += is_nz(row[k]);
        }
        nnzPerGroup[g] = c;
        total += c;
    }
    nnzPerGroup[nGroup] = 0;
    return total;
}

/* src/gcoo_spdm_kernels.cu:160-168 */
void orc_exclusive_scan(const int *src, int *dst, int n)
{
    if (n <= 0) return;
    dst[0] = 0;
    for (int i = 1; i < n; i++) dst[i] = dst[i - 1] + src[i - 1];
}

/* src/gcoo_spdm_kernels.cu:30-90 and src/ShGCOOSpDM.cu:19-126.
 * The reference walks each group in 512-wide k chunks and, inside a chunk,
 * emits thread t's (k = chunk + t) nonzeros in row order after a block scan;
 * the net order within a group is therefore (k ascending, then r ascending). */
int64_t orc_gcoo_convert(const float *S, int M, int K, int cpg,
                         float *vals, int *rows, int *cols,
                         int *groupIndex, int *nnzPerGroup)
{
    int nGroup = (M + cpg - 1) / cpg;
    orc_gcoo_count(S, M, K, cpg, nnzPerGroup);
    orc_exclusive_scan(nnzPerGroup, groupIndex, nGroup + 1);
    for (int g = 0; g < nGroup; g++) {
        int64_t p = groupIndex[g];
        int r0 = g * cpg;
        int r1 = r0 + cpg < M ? r0 + cpg : M;
        for (int k = 0; k < K; k++) {
            for (int r = r0; r < r1; r++) {
                float v = S[(int64_t)r * K + k];
                if (is_nz(v)) {
                    vals[p] = v;
                    rows[p] = r;
                    cols[p] = k;
                    p++;
                }
            }
        }
    }
    return groupIndex[nGroup];
}

/* src/gcoo_spdm_kernels.cu:92-158.  Per block (group g, column tile) the
 * reference keeps one accumulator per group row, initialised to 0.0
 * (:97), walks the group's entries in stored order and applies
 * c[r & (CPG-1)] = fmaf(a, D[k][j], c[...]) (:133-142); rows < M are
 * written at the end (:148-157).  Group-major and per-row order is all
 * that matters for the bits, so a row-indexed accumulator restates it. */
void orc_gcoo_spdm_f32(const float *vals, const int *rows, const int *cols,
                       const int *groupIndex, const int *nnzPerGroup,
                       int nGroup, int cpg, int M, const float *D, int N,
                       float *C, int nthreads)
{
    (void)nnzPerGroup;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int g = 0; g < nGroup; g++) {
        int r0 = g * cpg;
        int r1 = r0 + cpg < M ? r0 + cpg : M;
        for (int r = r0; r < r1; r++)
            memset(C + (int64_t)r * N, 0, sizeof(float) * (size_t)N);
        for (int64_t p = groupIndex[g]; p < groupIndex[g + 1]; p++) {
            float a = vals[p];
            int r = rows[p];
            const float *drow = D + (int64_t)cols[p] * N;
            float *crow = C + (int64_t)r * N;
            for (int j = 0; j < N; j++) crow[j] = fmaf(a, drow[j], crow[j]);
        }
    }
}

static inline float bf16_to_f32(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

void orc_f32_to_bf16(const float *x, uint16_t *y, int64_t n)
{
    for (int64_t i = 0; i < n; i++) {
        uint32_t u;
        memcpy(&u, &x[i], 4);
        if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) {
            y[i] = (uint16_t)((u >> 16) | 0x0040u); /* quiet NaN */
        } else {
            uint32_t lsb = (u >> 16) & 1u;
            y[i] = (uint16_t)((u + 0x7fffu + lsb) >> 16);
        }
    }
}

void orc_gcoo_spdm_bf16(const uint16_t *vals_bf16, const int *rows,
                        const int *cols, const int *groupIndex,
                        const int *nnzPerGroup, int nGroup, int cpg, int M,
                        const uint16_t *D_bf16, int N, float *C, int nthreads)
{
    (void)nnzPerGroup;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int g = 0; g < nGroup; g++) {
        int r0 = g * cpg;
        int r1 = r0 + cpg < M ? r0 + cpg : M;
        for (int r = r0; r < r1; r++)
            memset(C + (int64_t)r * N, 0, sizeof(float) * (size_t)N);
        for (int64_t p = groupIndex[g]; p < groupIndex[g + 1]; p++) {
            float a = bf16_to_f32(vals_bf16[p]);
            const uint16_t *drow = D_bf16 + (int64_t)cols[p] * N;
            float *crow = C + (int64_t)rows[p] * N;
            for (int j = 0; j < N; j++)
                crow[j] = fmaf(a, bf16_to_f32(drow[j]), crow[j]);
        }
    }
}

/* src/utils.cpp:30-46 restated over the nonzeros only. */
void orc_spdm_fp64(const float *vals, const int *rows, const int *cols,
                   const int *groupIndex, int nGroup, int cpg, int M,
                   const float *D, int N, double *Cd, float *C, int nthreads)
{
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int g = 0; g < nGroup; g++) {
        int r0 = g * cpg;
        int r1 = r0 + cpg < M ? r0 + cpg : M;
        for (int r = r0; r < r1; r++)
            memset(Cd + (int64_t)r * N, 0, sizeof(double) * (size_t)N);
        for (int64_t p = groupIndex[g]; p < groupIndex[g + 1]; p++) {
            double a = (double)vals[p];
            const float *drow = D + (int64_t)cols[p] * N;
            double *crow = Cd + (int64_t)rows[p] * N;
            for (int j = 0; j < N; j++) crow[j] += a * (double)drow[j];
        }
        if (C) {
            for (int r = r0; r < r1; r++)
                for (int j = 0; j < N; j++)
                    C[(int64_t)r * N + j] = (float)Cd[(int64_t)r * N + j];
        }
    }
}

/* src/utils.cpp:30-46 (matrixMulCPU(reference, S, D, hA=M, wA=K, wB=N)). */
void orc_dense_fp64(const float *S, const float *D, int M, int K, int N,
                    double *Cd, float *C)
{
    for (int64_t i = 0; i < M; i++)
        for (int64_t j = 0; j < N; j++) {
            double sum = 0;
            for (int64_t k = 0; k < K; k++) {
                double a = S[i * K + k];
                double b = D[k * N + j];
                sum += a * b;
            }
            if (Cd) Cd[i * N + j] = sum;
            if (C) C[i * N + j] = (float)sum;
        }
}

double orc_rel_frobenius(const float *C, const double *R, int64_t n)
{
    double num = 0.0, den = 0.0;
    for (int64_t i = 0; i < n; i++) {
        double d = (double)C[i] - R[i];
        num += d * d;
        den += R[i] * R[i];
    }
    if (den == 0.0) return num == 0.0 ? 0.0 : INFINITY;
    return sqrt(num / den);
}
