// gcoo_convert.hip -- dense S -> grouped COO on gfx950.
//
// Three kernels replace the reference's K1 / K2 / K3
// (src/gcoo_spdm_kernels.cu:4-28, :160-168, :30-90; launched from
// ShGCOOSpDM::convertToGroupCOOFormat, src/ShGCOOSpDM.cu:19-126):
//   count_kernel   one workgroup per group, float4 streaming reads of the
//                  group's cpg rows, wave64 shuffle + LDS reduction;
//   scan_kernel    one 1024-thread workgroup: exclusive scan of the counts
//                  (the reference runs this serially on one thread), 64-bit
//                  total and overflow flag into a device status word;
//   scatter_kernel one workgroup per group, k-chunks of 1024 columns, wave64
//                  prefix sums + LDS carry across the 4 waves, entries written
//                  in (k, r) order -- the reference's order, bit for bit.
// All three are HBM-streaming integer/byte work: no MFMA, coalesced 16-byte
// loads along rows of S.
#include "gcoo_internal.h"

#include <climits>

namespace gcoo {
namespace {

constexpr int kConvThreads = 256;
constexpr int kConvWaves = kConvThreads / 64;
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 8;

// `v != 0.0` exactly as the reference tests it (src/gcoo_spdm_kernels.cu:17):
// true for NaN and denormals, false for +0.0 and -0.0.
__device__ __forceinline__ int nz(float v) { return v != 0.0f ? 1 : 0; }

template <int CPG, bool VEC4>
__global__ __launch_bounds__(kConvThreads) void count_kernel(
    const float *__restrict__ S, int M, int K, int *__restrict__ nnzPerGroup)
{
    const int g = blockIdx.x;
    const int r0 = g * CPG;
    const int nrows = min(CPG, M - r0);
    int c = 0;
    if (VEC4) {
        const int K4 = K >> 2;
        const f32x4 *base = reinterpret_cast<const f32x4 *>(S + (size_t)r0 * K);
#pragma unroll 2
        for (int i = threadIdx.x; i < K4; i += kConvThreads) {
#pragma unroll
            for (int j = 0; j < CPG; ++j) {
                if (j < nrows) {
                    f32x4 v = __builtin_nontemporal_load(base + (size_t)j * K4 + i);
                    c += nz(v.x) + nz(v.y) + nz(v.z) + nz(v.w);
                }
            }
        }
    } else {
        const float *base = S + (size_t)r0 * K;
        for (int i = threadIdx.x; i < K; i += kConvThreads) {
#pragma unroll
            for (int j = 0; j < CPG; ++j)
                if (j < nrows) c += nz(base[(size_t)j * K + i]);
        }
    }
    __shared__ int part[kConvWaves];
    c = wave_sum(c);
    if (__lane_id() == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
#pragma unroll
        for (int w = 0; w < kConvWaves; ++w) t += part[w];
        nnzPerGroup[g] = t;
    }
}

// Exclusive scan of nnzPerGroup[0..nGroup) into groupIndex[0..nGroup];
// groupIndex[nGroup] = total, nnzPerGroup[nGroup] = 0 (the reference's memset
// of the shared allocation, src/ShGCOOSpDM.cu:37).
__global__ __launch_bounds__(kScanThreads) void scan_kernel(
    int *__restrict__ nnzPerGroup, int *__restrict__ groupIndex, int nGroup,
    ConvStatus *__restrict__ st)
{
    __shared__ long long wtot[kScanThreads / 64];
    const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    long long carry = 0;
    for (int base = 0; base < nGroup; base += kScanThreads * kScanItems) {
        int v[kScanItems];
        long long tsum = 0;
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            int idx = base + tid * kScanItems + i;
            v[i] = idx < nGroup ? nnzPerGroup[idx] : 0;
            tsum += v[i];
        }
        long long incl = tsum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            long long t = __shfl_up(incl, d, 64);
            if (lane >= d) incl += t;
        }
        if (lane == 63) wtot[wave] = incl;
        __syncthreads();
        long long before = 0, tile = 0;
#pragma unroll
        for (int w = 0; w < kScanThreads / 64; ++w) {
            long long x = wtot[w];
            before += (w < wave) ? x : 0;
            tile += x;
        }
        long long run = carry + before + incl - tsum;
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            int idx = base + tid * kScanItems + i;
            if (idx < nGroup) groupIndex[idx] = (int)run;
            run += v[i];
        }
        carry += tile;
        __syncthreads();
    }
    if (tid == 0) {
        groupIndex[nGroup] = (int)carry;
        nnzPerGroup[nGroup] = 0;
        st->total = carry;
        st->overflow = carry > (long long)INT_MAX ? 1 : 0;
    }
}

// One workgroup per group.  Chunk of kConvThreads * KPT columns; thread t owns
// columns [chunk + KPT*t, chunk + KPT*t + KPT) and emits its nonzeros in
// (k, r) order at an offset given by a workgroup-wide exclusive scan.
template <int CPG, bool VEC4>
__global__ __launch_bounds__(kConvThreads) void scatter_kernel(
    const float *__restrict__ S, int M, int K,
    const int *__restrict__ groupIndex, const ConvStatus *__restrict__ st,
    long long cap, float *__restrict__ vals, int *__restrict__ rows,
    int *__restrict__ cols)
{
    if (st->total > cap || st->overflow) return;
    constexpr int KPT = VEC4 ? 4 : 1;
    constexpr int CHUNK = kConvThreads * KPT;
    __shared__ int wtot[kConvWaves];
    const int g = blockIdx.x;
    const int r0 = g * CPG;
    const int nrows = min(CPG, M - r0);
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    int base = groupIndex[g];
    const float *grow = S + (size_t)r0 * K;
    for (int k0 = 0; k0 < K; k0 += CHUNK) {
        const int kb = k0 + threadIdx.x * KPT;
        float v[CPG][KPT];
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < CPG; ++j) {
            if (VEC4) {
                f32x4 x = {0.f, 0.f, 0.f, 0.f};
                if (j < nrows && kb < K)
                    x = __builtin_nontemporal_load(
                        reinterpret_cast<const f32x4 *>(grow + (size_t)j * K + kb));
                v[j][0] = x.x;
                if (KPT > 1) {
                    v[j][KPT > 1 ? 1 : 0] = x.y;
                    v[j][KPT > 2 ? 2 : 0] = x.z;
                    v[j][KPT > 3 ? 3 : 0] = x.w;
                }
            } else {
                v[j][0] = (j < nrows && kb < K) ? grow[(size_t)j * K + kb] : 0.f;
            }
#pragma unroll
            for (int q = 0; q < KPT; ++q) cnt += nz(v[j][q]);
        }
        int incl = wave_inclusive_sum(cnt);
        if (lane == 63) wtot[wave] = incl;
        __syncthreads();
        int before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kConvWaves; ++w) {
            int x = wtot[w];
            before += (w < wave) ? x : 0;
            total += x;
        }
        if (cnt) {
            int pos = base + before + incl - cnt;
#pragma unroll
            for (int q = 0; q < KPT; ++q) {
#pragma unroll
                for (int j = 0; j < CPG; ++j) {
                    if (nz(v[j][q])) {
                        vals[pos] = v[j][q];
                        rows[pos] = r0 + j;
                        cols[pos] = kb + q;
                        ++pos;
                    }
                }
            }
        }
        base += total;
        __syncthreads();
    }
}

template <int CPG>
int launch_count_t(gcoo_matrix *h, const float *S, bool vec4, hipStream_t s)
{
    dim3 grid(h->nGroup), block(kConvThreads);
    if (vec4)
        hipLaunchKernelGGL((count_kernel<CPG, true>), grid, block, 0, s, S, h->M,
                           h->K, h->nnzPerGroup());
    else
        hipLaunchKernelGGL((count_kernel<CPG, false>), grid, block, 0, s, S,
                           h->M, h->K, h->nnzPerGroup());
    GCOO_CHECK_LAUNCH("count_kernel");
    return GCOO_OK;
}

template <int CPG>
int launch_scatter_t(gcoo_matrix *h, const float *S, bool vec4, hipStream_t s)
{
    dim3 grid(h->nGroup), block(kConvThreads);
    if (vec4)
        hipLaunchKernelGGL((scatter_kernel<CPG, true>), grid, block, 0, s, S,
                           h->M, h->K, h->groupIndex(), h->d_status,
                           (long long)h->cap, h->vals, h->rows(), h->cols());
    else
        hipLaunchKernelGGL((scatter_kernel<CPG, false>), grid, block, 0, s, S,
                           h->M, h->K, h->groupIndex(), h->d_status,
                           (long long)h->cap, h->vals, h->rows(), h->cols());
    GCOO_CHECK_LAUNCH("scatter_kernel");
    return GCOO_OK;
}

bool use_vec4(const gcoo_matrix *h, const float *S)
{
    return (h->K % 4 == 0) && ((reinterpret_cast<uintptr_t>(S) & 15u) == 0);
}

}  // namespace

int launch_count_and_scan(gcoo_matrix *h, const float *S, hipStream_t s)
{
    const bool v4 = use_vec4(h, S);
    int rc = GCOO_OK;
    if (h->nGroup > 0) {
        switch (h->cpg) {
        case 1: rc = launch_count_t<1>(h, S, v4, s); break;
        case 2: rc = launch_count_t<2>(h, S, v4, s); break;
        case 4: rc = launch_count_t<4>(h, S, v4, s); break;
        default: return set_error(GCOO_ERR_INVALID, "cpg %d not supported", h->cpg);
        }
        if (rc) return rc;
    }
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(kScanThreads), 0, s,
                       h->nnzPerGroup(), h->groupIndex(), h->nGroup, h->d_status);
    GCOO_CHECK_LAUNCH("scan_kernel");
    return GCOO_OK;
}

int launch_scatter(gcoo_matrix *h, const float *S, hipStream_t s)
{
    if (h->nGroup == 0 || h->cap == 0) return GCOO_OK;
    const bool v4 = use_vec4(h, S);
    switch (h->cpg) {
    case 1: return launch_scatter_t<1>(h, S, v4, s);
    case 2: return launch_scatter_t<2>(h, S, v4, s);
    case 4: return launch_scatter_t<4>(h, S, v4, s);
    default: return set_error(GCOO_ERR_INVALID, "cpg %d not supported", h->cpg);
    }
}

}  // namespace gcoo
