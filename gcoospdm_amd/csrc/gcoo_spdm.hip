// gcoo_spdm.hip -- grouped-COO sparse x dense multiply (FP32) on gfx950.
//
// Replaces the reference kernel sparse_dense_groupcoo_mat_mul_kernel
// (src/gcoo_spdm_kernels.cu:92-158, launched src/ShGCOOSpDM.cu:153-167 and
// :197-211).  Arithmetic contract, kept bit for bit: C[r][j] is one FP32 fmaf
// chain from +0.0f over row r's nonzeros in ascending k, i.e. the group's
// stored (k, r) order.
//
// Mapping (wave64-first, not the reference's 128-thread block per group):
//   * one wave owns one group (cpg rows) and a 64*VEC-column tile of C; lane l
//     owns columns [tile + VEC*l, tile + VEC*l + VEC) and keeps cpg*VEC
//     accumulators in VGPRs;
//   * the group's (val, row, col) triples are wave-uniform: they are read with
//     scalar loads, no LDS staging and no barriers are needed;
//   * per nonzero every lane issues one VEC-wide load of D[k][cols] -- a
//     coalesced 64*VEC*4-byte row segment per wave; U nonzeros are issued
//     before their FMAs to keep U loads in flight per wave;
//   * workgroups are remapped so each XCD sweeps whole column tiles of C:
//     the D column panel of a tile (K x 64*VEC floats) then stays in that
//     XCD's L2 while all groups stream through it.
#include "gcoo_internal.h"

namespace gcoo {
namespace {

constexpr int kSpdmWaves = 4;  // waves per workgroup (groups per workgroup)
constexpr int kSpdmThreads = 64 * kSpdmWaves;

template <int VEC>
__device__ __forceinline__ void load_vec(float (&d)[VEC], const float *p)
{
    if constexpr (VEC == 1) {
        d[0] = *p;
    } else if constexpr (VEC == 2) {
        f32x2 x = *reinterpret_cast<const f32x2 *>(p);
        d[0] = x.x; d[1] = x.y;
    } else {
        f32x4 x = *reinterpret_cast<const f32x4 *>(p);
        d[0] = x.x; d[1] = x.y; d[2] = x.z; d[3] = x.w;
    }
}

template <int VEC>
__device__ __forceinline__ void store_vec(float *p, const float (&c)[VEC])
{
    if constexpr (VEC == 1) {
        __builtin_nontemporal_store(c[0], p);
    } else if constexpr (VEC == 2) {
        f32x2 x = {c[0], c[1]};
        __builtin_nontemporal_store(x, reinterpret_cast<f32x2 *>(p));
    } else {
        f32x4 x = {c[0], c[1], c[2], c[3]};
        __builtin_nontemporal_store(x, reinterpret_cast<f32x4 *>(p));
    }
}

// acc = fma(a, d, acc), IEEE fused multiply-add (the reference's fmaf).  Written
// as v_fmac_f32 so that the compiler cannot merge the per-row-slot copies of
// this code into one FMA on a selected accumulator: merged, the accumulators
// become a dynamically indexed private array, which lives in scratch memory.
__device__ __forceinline__ void fmac(float &acc, float a, float d)
{
    asm("v_fmac_f32 %0, %1, %2" : "+v"(acc) : "s"(a), "v"(d));
}

template <int VEC>
__device__ __forceinline__ void fma_vec(float (&acc)[VEC], float a, const float (&d)[VEC])
{
#pragma unroll
    for (int v = 0; v < VEC; ++v) fmac(acc[v], a, d[v]);
}

// Accumulators of one group, one named array per row slot.
template <int CPG, int VEC>
struct GroupAcc {
    float a0[VEC], a1[VEC], a2[VEC], a3[VEC];

    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int v = 0; v < VEC; ++v) a0[v] = a1[v] = a2[v] = a3[v] = 0.0f;
    }

    // acc[r][:] = fmaf(a, d[:], acc[r][:]) for a wave-uniform row slot r
    // (uniform, so the switch is a scalar branch).
    __device__ __forceinline__ void fma(int r, float a, const float (&d)[VEC])
    {
        if constexpr (CPG == 1) {
            fma_vec<VEC>(a0, a, d);
        } else if constexpr (CPG == 2) {
            if (r == 0) fma_vec<VEC>(a0, a, d);
            else fma_vec<VEC>(a1, a, d);
        } else {
            switch (r) {
            case 0: fma_vec<VEC>(a0, a, d); break;
            case 1: fma_vec<VEC>(a1, a, d); break;
            case 2: fma_vec<VEC>(a2, a, d); break;
            default: fma_vec<VEC>(a3, a, d); break;
            }
        }
    }

    __device__ __forceinline__ const float (&row(int q) const)[VEC]
    {
        return q == 0 ? a0 : q == 1 ? a1 : q == 2 ? a2 : a3;
    }
};

template <int CPG, int VEC, int U>
__global__ __launch_bounds__(kSpdmThreads) void spdm_f32_wave_kernel(
    const float *__restrict__ vals, const int *__restrict__ rows,
    const int *__restrict__ cols, const int *__restrict__ groupIndex,
    int nGroup, int M, const float *__restrict__ D, int N,
    float *__restrict__ C, int nGroupBlocks,
    const ConvStatus *__restrict__ guard, long long cap)
{
    // When launched right behind an enqueued conversion, skip the work if
    // that conversion overflowed the entry capacity (the host re-runs it).
    if (guard != nullptr && guard->total > cap) return;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int ct = L / nGroupBlocks;
    const int gb = L - ct * nGroupBlocks;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = __lane_id();
    const int g = gb * kSpdmWaves + wave;
    if (g >= nGroup) return;
    const int j = ct * (64 * VEC) + lane * VEC;
    const bool active = j < N;
    const int r0 = g * CPG;

    GroupAcc<CPG, VEC> acc;
    acc.zero();

    const int p0 = groupIndex[g];
    const int p1 = groupIndex[g + 1];
    const float *Dj = D + j;
    int p = p0;
    for (; p + U <= p1; p += U) {
        float a[U];
        int rr[U];
        float d[U][VEC];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a[u] = vals[p + u];
            rr[u] = rows[p + u] - r0;
            const int k = cols[p + u];
            if (active) {
                load_vec<VEC>(d[u], Dj + (size_t)k * N);
            } else {
#pragma unroll
                for (int v = 0; v < VEC; ++v) d[u][v] = 0.0f;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc.fma(rr[u], a[u], d[u]);
    }
    for (; p < p1; ++p) {
        const float a = vals[p];
        const int r = rows[p] - r0;
        const int k = cols[p];
        float d[VEC];
        if (active) {
            load_vec<VEC>(d, Dj + (size_t)k * N);
        } else {
#pragma unroll
            for (int v = 0; v < VEC; ++v) d[v] = 0.0f;
        }
        acc.fma(r, a, d);
    }
    if (active) {
#pragma unroll
        for (int q = 0; q < CPG; ++q)
            if (r0 + q < M) store_vec<VEC>(C + (size_t)(r0 + q) * N + j, acc.row(q));
    }
}

template <int CPG, int VEC>
int launch_wave(const gcoo_matrix *h, const float *D, int N, float *C,
                const ConvStatus *guard, hipStream_t s)
{
    constexpr int U = 8;
    const int nGroupBlocks = (h->nGroup + kSpdmWaves - 1) / kSpdmWaves;
    const int tileCols = 64 * VEC;
    const int nColTiles = (N + tileCols - 1) / tileCols;
    const long long nblocks = (long long)nGroupBlocks * nColTiles;
    if (nblocks > INT32_MAX)
        return set_error(GCOO_ERR_OVERFLOW, "grid of %lld workgroups too large", nblocks);
    hipLaunchKernelGGL((spdm_f32_wave_kernel<CPG, VEC, U>), dim3((unsigned)nblocks),
                       dim3(kSpdmThreads), 0, s, h->vals, h->rows(), h->cols(),
                       h->groupIndex(), h->nGroup, h->M, D, N, C, nGroupBlocks,
                       guard, (long long)h->cap);
    GCOO_CHECK_LAUNCH("spdm_f32_wave_kernel");
    return GCOO_OK;
}

template <int CPG>
int launch_cpg(const gcoo_matrix *h, const float *D, int N, float *C,
               const ConvStatus *guard, hipStream_t s)
{
    const uintptr_t al = reinterpret_cast<uintptr_t>(D) | reinterpret_cast<uintptr_t>(C);
    if (N % 4 == 0 && (al & 15u) == 0) return launch_wave<CPG, 4>(h, D, N, C, guard, s);
    if (N % 2 == 0 && (al & 7u) == 0) return launch_wave<CPG, 2>(h, D, N, C, guard, s);
    return launch_wave<CPG, 1>(h, D, N, C, guard, s);
}

}  // namespace

int launch_spdm_f32(const gcoo_matrix *h, const float *D, int N, float *C,
                    const ConvStatus *guard, hipStream_t s)
{
    if (h->nGroup == 0 || N == 0) return GCOO_OK;
    switch (h->cpg) {
    case 1: return launch_cpg<1>(h, D, N, C, guard, s);
    case 2: return launch_cpg<2>(h, D, N, C, guard, s);
    case 4: return launch_cpg<4>(h, D, N, C, guard, s);
    default: return set_error(GCOO_ERR_INVALID, "cpg %d not supported", h->cpg);
    }
}

}  // namespace gcoo
