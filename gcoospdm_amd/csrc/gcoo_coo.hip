// gcoo_coo.hip -- COO -> grouped COO on gfx950, for sparse matrices that are
// too large to exist densely (the reference converts only from a dense S and
// its int offsets overflow at M*K >= 2^31, src/gcoo_spdm_kernels.cu:16,57).
//
// The result is identical to the dense conversion of the equivalent dense
// matrix: same vals / rows / cols / groupIndex / nnzPerGroup, same row
// segments.  Steps (all deterministic although the first two use atomics,
// because only integer counts and slot reservations are atomic and the final
// order is fixed by unique sort keys):
//   1. coo_count_kernel    per-row counts of entries with val != 0.0 (integer
//                          atomics), range checks of rows / cols;
//   2. scan_kernel         row pointers, groupIndex, nnzPerGroup, total
//                          (shared with the dense path);
//   3. coo_place_kernel    each entry into its row's segment at an atomically
//                          reserved slot (unordered within the row);
//   4. sort_rows_*         every row segment sorted by k: a wave64 bitonic
//                          network in registers for rows of <= 64 entries, an
//                          LDS bitonic network for <= 16384, the same network
//                          in global memory beyond;
//   5. gcoo_from_rows      each entry's position in its group's (k, r) order
//                          = its rank in its row + binary searches in the
//                          group's other rows; duplicate (r, k) are rejected.
#include "gcoo_internal.h"

#include <climits>

namespace gcoo {
namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void coo_count_kernel(
    const int *__restrict__ rows, const int *__restrict__ cols, const float *__restrict__ vals,
    long long nnz, int M, int K, int *__restrict__ rowcnt, ConvStatus *__restrict__ st)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += stride) {
        const int r = rows[i], k = cols[i];
        if (r < 0 || r >= M || k < 0 || k >= K) {
            st->invalid = 1;
            continue;
        }
        if (vals[i] != 0.0f) atomicAdd(&rowcnt[r], 1);
    }
}

__global__ __launch_bounds__(kThreads) void coo_place_kernel(
    const int *__restrict__ rows, const int *__restrict__ cols, const float *__restrict__ vals,
    long long nnz, int M, int K, const int *__restrict__ rowptr, int *__restrict__ cursor,
    const ConvStatus *__restrict__ st, long long cap, int2 *__restrict__ seg)
{
    if (st->total > cap || st->overflow || st->invalid) return;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += stride) {
        const int r = rows[i], k = cols[i];
        const float v = vals[i];
        if (r < 0 || r >= M || k < 0 || k >= K || !(v != 0.0f)) continue;
        const int slot = atomicAdd(&cursor[r], 1);
        seg[rowptr[r] + slot] = make_int2(k, __float_as_int(v));
    }
}

// One wave per row with at most 64 entries: bitonic network over the lanes,
// key = k (unique within a row), padding lanes hold INT_MAX.
__global__ __launch_bounds__(kThreads) void sort_rows_wave_kernel(
    const int *__restrict__ rowptr, int M, const ConvStatus *__restrict__ st, long long cap,
    int2 *__restrict__ seg)
{
    if (st->total > cap || st->overflow || st->invalid) return;
    const int r = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (r >= M) return;
    const int p0 = rowptr[r], n = rowptr[r + 1] - p0;
    if (n <= 1 || n > 64) return;
    const int lane = __lane_id();
    int2 e = lane < n ? seg[p0 + lane] : make_int2(INT_MAX, 0);
    for (int size = 2; size <= 64; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const int partner = lane ^ stride;
            const int ox = __shfl(e.x, partner, 64);
            const int oy = __shfl(e.y, partner, 64);
            const bool up = ((lane & size) == 0);
            const bool lower = lane < partner;
            const bool take_min = (lower == up);
            const bool swap = take_min ? (ox < e.x) : (ox > e.x);
            if (swap) e = make_int2(ox, oy);
        }
    }
    if (lane < n) seg[p0 + lane] = e;
}

// Ascending bitonic sort of a[0..n) on key .x with virtual +inf padding up
// to the next power of two: every merge starts with the "flip" comparison
// (i with i ^ (size-1)), so all partial sequences stay ascending and pairs
// whose partner lies past n are left alone.
template <typename Access>
__device__ void bitonic_sort(Access a, int n)
{
    int np = 1;
    while (np < n) np <<= 1;
    for (int size = 2; size <= np; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < np; i += blockDim.x) {
                const int partner = (stride == (size >> 1)) ? (i ^ (size - 1)) : (i ^ stride);
                if (partner > i && partner < n) {
                    int2 x = a.get(i), y = a.get(partner);
                    if (y.x < x.x) {
                        a.set(i, y);
                        a.set(partner, x);
                    }
                }
            }
            __syncthreads();
        }
    }
}

struct LdsAccess {
    int2 *p;
    __device__ int2 get(int i) const { return p[i]; }
    __device__ void set(int i, int2 v) const { p[i] = v; }
};
struct GlobalAccess {
    int2 *p;
    __device__ int2 get(int i) const { return p[i]; }
    __device__ void set(int i, int2 v) const { p[i] = v; }
};

constexpr int kSortLds = 16384;  // entries sorted in LDS (128 KB)

// One workgroup per row with more than 64 entries.
__global__ __launch_bounds__(1024) void sort_rows_block_kernel(
    const int *__restrict__ rowptr, int M, const ConvStatus *__restrict__ st, long long cap,
    int2 *__restrict__ seg)
{
    if (st->total > cap || st->overflow || st->invalid) return;
    const int r = blockIdx.x;
    const int p0 = rowptr[r], n = rowptr[r + 1] - p0;
    if (n <= 64) return;
    if (n <= kSortLds) {
        __shared__ int2 sm[kSortLds];
        for (int i = threadIdx.x; i < n; i += blockDim.x) sm[i] = seg[p0 + i];
        __syncthreads();
        bitonic_sort(LdsAccess{sm}, n);
        for (int i = threadIdx.x; i < n; i += blockDim.x) seg[p0 + i] = sm[i];
    } else {
        bitonic_sort(GlobalAccess{seg + p0}, n);
    }
}

__device__ __forceinline__ int lower_bound_k(const int2 *__restrict__ s, int n, int k)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s[mid].x < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One wave per row: entry i of row r (sorted by k) goes to
//   groupIndex[g] + i + sum over the group's other rows r' of
//   #{k' <= k} for r' < r  and  #{k' < k} for r' > r,
// i.e. its place in the group's (k, r) order.
__global__ __launch_bounds__(kThreads) void gcoo_from_rows_kernel(
    const int *__restrict__ rowptr, const int *__restrict__ groupIndex, int M, int cpg,
    const int2 *__restrict__ seg, ConvStatus *__restrict__ st, long long cap,
    float *__restrict__ vals, int *__restrict__ rows, int *__restrict__ cols)
{
    if (st->total > cap || st->overflow || st->invalid) return;
    const int r = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (r >= M) return;
    const int g = r / cpg;
    const int rg0 = g * cpg;
    const int rg1 = min(rg0 + cpg, M);
    const int p0 = rowptr[r], n = rowptr[r + 1] - p0;
    for (int i = __lane_id(); i < n; i += 64) {
        const int2 e = seg[p0 + i];
        if (i > 0 && seg[p0 + i - 1].x == e.x) st->invalid = 2;  // duplicate (r, k)
        int pos = groupIndex[g] + i;
        for (int r2 = rg0; r2 < rg1; ++r2) {
            if (r2 == r) continue;
            const int q0 = rowptr[r2], qn = rowptr[r2 + 1] - q0;
            pos += lower_bound_k(seg + q0, qn, r2 < r ? e.x + 1 : e.x);
        }
        vals[pos] = __int_as_float(e.y);
        rows[pos] = r;
        cols[pos] = e.x;
    }
}

unsigned grid_for(long long n)
{
    long long b = (n + kThreads - 1) / kThreads;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return (unsigned)b;
}

}  // namespace

int convert_coo(gcoo_matrix *h, const int *rows, const int *cols, const float *vals,
                int64_t nnz, hipStream_t s)
{
    GCOO_HIP_TRY(hipMemsetAsync(h->rowcnt, 0, sizeof(int) * ((size_t)h->M + 1), s));
    GCOO_HIP_TRY(hipMemsetAsync(h->d_status, 0, sizeof(ConvStatus), s));
    if (nnz > 0) {
        hipLaunchKernelGGL(coo_count_kernel, dim3(grid_for(nnz)), dim3(kThreads), 0, s, rows,
                           cols, vals, (long long)nnz, h->M, h->K, h->rowcnt, h->d_status);
        GCOO_CHECK_LAUNCH("coo_count_kernel");
    }
    int rc = launch_row_scan(h, s, false);
    if (rc) return rc;
    GCOO_HIP_TRY(hipMemcpyAsync(h->h_status, h->d_status, sizeof(ConvStatus),
                                hipMemcpyDeviceToHost, s));
    GCOO_HIP_TRY(hipStreamSynchronize(s));
    if (h->h_status->invalid)
        return set_error(GCOO_ERR_INVALID, "COO entry outside the %d x %d matrix", h->M, h->K);
    if (h->h_status->overflow)
        return set_error(GCOO_ERR_OVERFLOW, "nnz %lld does not fit in int32",
                         (long long)h->h_status->total);
    const int64_t total = h->h_status->total;
    if ((rc = ensure_capacity(h, total + h->h_status->maxRow + kSegSlack))) return rc;
    if (total > 0) {
        GCOO_HIP_TRY(hipMemsetAsync(h->rowcnt, 0, sizeof(int) * ((size_t)h->M + 1), s));
        hipLaunchKernelGGL(coo_place_kernel, dim3(grid_for(nnz)), dim3(kThreads), 0, s, rows,
                           cols, vals, (long long)nnz, h->M, h->K, h->rowptr, h->rowcnt,
                           h->d_status, (long long)h->cap, h->seg);
        GCOO_CHECK_LAUNCH("coo_place_kernel");
        const unsigned rowBlocks = (unsigned)((h->M + kThreads / 64 - 1) / (kThreads / 64));
        hipLaunchKernelGGL(sort_rows_wave_kernel, dim3(rowBlocks), dim3(kThreads), 0, s,
                           h->rowptr, h->M, h->d_status, (long long)h->cap, h->seg);
        GCOO_CHECK_LAUNCH("sort_rows_wave_kernel");
        hipLaunchKernelGGL(sort_rows_block_kernel, dim3((unsigned)h->M), dim3(1024), 0, s,
                           h->rowptr, h->M, h->d_status, (long long)h->cap, h->seg);
        GCOO_CHECK_LAUNCH("sort_rows_block_kernel");
        hipLaunchKernelGGL(gcoo_from_rows_kernel, dim3(rowBlocks), dim3(kThreads), 0, s,
                           h->rowptr, h->groupIndex(), h->M, h->cpg, h->seg, h->d_status,
                           (long long)h->cap, h->vals, h->rows(), h->cols());
        GCOO_CHECK_LAUNCH("gcoo_from_rows_kernel");
    }
    GCOO_HIP_TRY(hipMemcpyAsync(h->h_status, h->d_status, sizeof(ConvStatus),
                                hipMemcpyDeviceToHost, s));
    GCOO_HIP_TRY(hipStreamSynchronize(s));
    if (h->h_status->invalid)
        return set_error(GCOO_ERR_INVALID, "duplicate (row, col) entries in the COO input");
    h->nnz = total;
    h->maxRow = h->h_status->maxRow;
    return GCOO_OK;
}

}  // namespace gcoo
