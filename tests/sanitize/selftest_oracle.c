/* Sanitizer self-test of the CPU oracle (oracle/gcoo_oracle.c), built with
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/Makefile)
 * and run by tests/test_sanitizers.py.  It drives every oracle function on
 * small seeded inputs -- ragged shapes, every group width, empty rows and
 * matrices, -0.0 / NaN / denormal / inf values -- and checks the invariants
 * the conversion and the chains must keep, so that an out-of-bounds access or
 * undefined operation in any of them stops the run with a report. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/gcoo_oracle.h"

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint64_t next(void)
{
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static float unif(void) { return (float)((double)(next() >> 11) / 9007199254740992.0 * 2.0 - 1.0); }

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            fprintf(stderr, "check failed: %s (%s:%d)\n", #c, __FILE__, __LINE__); \
            exit(1);                                                    \
        }                                                               \
    } while (0)

static void one_case(int M, int K, int N, int cpg, double density, int edge)
{
    float *S = calloc((size_t)M * K + 1, sizeof(float));
    float *D = malloc(((size_t)K * N + 1) * sizeof(float));
    for (size_t i = 0; i < (size_t)M * K; ++i)
        if ((double)(next() >> 11) / 9007199254740992.0 < density) S[i] = unif();
    for (size_t i = 0; i < (size_t)K * N; ++i) D[i] = unif();
    if (edge && M * K >= 4) {
        S[0] = -0.0f;
        S[1] = NAN;
        S[2] = 1e-40f;
        S[3] = INFINITY;
    }
    const int nGroup = (M + cpg - 1) / cpg;
    int *npg = malloc(((size_t)nGroup + 1) * sizeof(int));
    const int64_t nnz = orc_gcoo_count(S, M, K, cpg, npg);
    CHECK(npg[nGroup] == 0);
    int64_t sum = 0;
    for (int g = 0; g < nGroup; ++g) sum += npg[g];
    CHECK(sum == nnz);
    float *vals = malloc(((size_t)nnz + 1) * sizeof(float));
    int *rows = malloc(((size_t)nnz + 1) * sizeof(int));
    int *cols = malloc(((size_t)nnz + 1) * sizeof(int));
    int *gi = malloc(((size_t)nGroup + 1) * sizeof(int));
    int *npg2 = malloc(((size_t)nGroup + 1) * sizeof(int));
    CHECK(orc_gcoo_convert(S, M, K, cpg, vals, rows, cols, gi, npg2) == nnz);
    CHECK(gi[0] == 0 && gi[nGroup] == nnz);
    for (int g = 0; g < nGroup; ++g) {
        CHECK(npg2[g] == npg[g] && gi[g + 1] - gi[g] == npg[g]);
        for (int i = gi[g]; i < gi[g + 1]; ++i) {
            CHECK(rows[i] / cpg == g && cols[i] >= 0 && cols[i] < K);
            CHECK(S[(size_t)rows[i] * K + cols[i]] != 0.0f || isnan(vals[i]));
            if (i > gi[g]) CHECK(cols[i] > cols[i - 1] || (cols[i] == cols[i - 1] && rows[i] > rows[i - 1]));
        }
    }
    int *scan = malloc(((size_t)nGroup + 1) * sizeof(int));
    orc_exclusive_scan(npg, scan, nGroup + 1);
    for (int g = 0; g <= nGroup; ++g) CHECK(scan[g] == gi[g]);

    float *C = malloc(((size_t)M * N + 1) * sizeof(float));
    float *C2 = malloc(((size_t)M * N + 1) * sizeof(float));
    orc_gcoo_spdm_f32(vals, rows, cols, gi, npg, nGroup, cpg, M, D, N, C, 1);
    orc_gcoo_spdm_f32(vals, rows, cols, gi, npg, nGroup, cpg, M, D, N, C2, 3);
    CHECK(memcmp(C, C2, (size_t)M * N * sizeof(float)) == 0);

    double *Cd = malloc(((size_t)M * N + 1) * sizeof(double));
    double *Cd2 = malloc(((size_t)M * N + 1) * sizeof(double));
    float *Cf = malloc(((size_t)M * N + 1) * sizeof(float));
    orc_spdm_fp64(vals, rows, cols, gi, nGroup, cpg, M, D, N, Cd, Cf, 2);
    orc_dense_fp64(S, D, M, K, N, Cd2, NULL);
    if (!edge) {
        for (size_t i = 0; i < (size_t)M * N; ++i) CHECK(Cd[i] == Cd2[i]);
        const double e = orc_rel_frobenius(C, Cd, (int64_t)M * N);
        CHECK(e <= 1e-5 || M * N == 0 || nnz == 0);
    }

    uint16_t *v16 = malloc(((size_t)nnz + 1) * sizeof(uint16_t));
    uint16_t *d16 = malloc(((size_t)K * N + 1) * sizeof(uint16_t));
    orc_f32_to_bf16(vals, v16, nnz);
    orc_f32_to_bf16(D, d16, (int64_t)K * N);
    orc_gcoo_spdm_bf16(v16, rows, cols, gi, npg, nGroup, cpg, M, d16, N, C2, 2);

    free(S); free(D); free(npg); free(vals); free(rows); free(cols); free(gi); free(npg2);
    free(scan); free(C); free(C2); free(Cd); free(Cd2); free(Cf); free(v16); free(d16);
}

int main(void)
{
    const int shapes[][3] = {{1, 1, 1}, {7, 9, 5}, {13, 64, 3}, {33, 20, 12}, {4, 1, 7},
                             {1, 300, 2}, {65, 70, 33}, {0, 5, 3}, {5, 0, 3}, {6, 6, 0}};
    int cases = 0;
    for (size_t s = 0; s < sizeof(shapes) / sizeof(shapes[0]); ++s)
        for (int cpg = 1; cpg <= 4; cpg *= 2)
            for (int edge = 0; edge < 2; ++edge) {
                one_case(shapes[s][0], shapes[s][1], shapes[s][2], cpg, 0.3, edge);
                one_case(shapes[s][0], shapes[s][1], shapes[s][2], cpg, 1.0, edge);
                cases += 2;
            }
    float x[4] = {1.0f, -0.0f, NAN, 3.00390625f};
    uint16_t y[4];
    orc_f32_to_bf16(x, y, 4);
    CHECK(y[0] == 0x3f80 && y[1] == 0x8000 && (y[2] & 0x7fc0) == 0x7fc0);
    printf("oracle sanitizer self-test: %d cases ok\n", cases);
    return 0;
}
