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
//                          atomics, one per run of equal rows in a wave),
//                          range checks of rows / cols;
//   2. scan_kernel         row pointers, groupIndex, nnzPerGroup, total
//                          (shared with the dense path), and the list of rows
//                          of more than 64 entries;
//   3. coo_place_kernel    each entry into its row's segment at an atomically
//                          reserved slot (one reservation per run of equal
//                          rows in a wave; unordered within the row);
//   4. sort_rows_*         every row segment sorted by k: a wave64 bitonic
//                          network in registers for rows of <= 64 entries;
//                          the rows the scan lists as longer, one workgroup
//                          each sized to the row (64, 256 or 1024 threads),
//                          in LDS up to 16384 entries and in global memory
//                          beyond;
//   5. gcoo_from_rows      one thread per entry: its position in its group's
//                          (k, r) order = its rank in its row + binary
//                          searches in the group's other rows; duplicate
//                          (r, k) are rejected.
#include "gcoo_internal.h"

#include <climits>

namespace gcoo {
namespace {

constexpr int kThreads = 256;
constexpr int kWaveSortMax = 64;  // rows up to this length are sorted by one wave

// Runs of equal rows among a wave's 64 consecutive entries (COO input sorted
// by row gives long runs): whether this lane starts a run, the run's length
// (valid in its first lane) and the lane that starts this lane's run.
__device__ __forceinline__ void row_run(int r, int lane, bool &head, int &len, int &headLane)
{
    const int rp = __shfl_up(r, 1, 64);
    head = lane == 0 || rp != r;
    const unsigned long long hm = __ballot(head);
    const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1;  // lanes <= this one
    const unsigned long long above = hm & ~upto;
    len = (above ? __ffsll((long long)above) - 1 : 64) - lane;
    headLane = 63 - __clzll(hm & upto);  // lane 0 always starts a run
}

// Grid-stride over whole waves: entry b + lane.  Entries outside the matrix
// set the invalid flag; entries whose value is 0.0 are not counted.  The
// first lane of a run of equal rows adds the run's length.
__global__ __launch_bounds__(kThreads) void coo_count_kernel(
    const int *__restrict__ rows, const int *__restrict__ cols, const float *__restrict__ vals,
    long long nnz, int M, int K, int *__restrict__ rowcnt, ConvStatus *__restrict__ st)
{
    const int lane = __lane_id();
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long b = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); b < nnz; b += stride) {
        const long long i = b + lane;
        int r = -1;
        if (i < nnz) {
            const int ri = rows[i], k = cols[i];
            if (ri < 0 || ri >= M || k < 0 || k >= K) st->invalid = 1;
            else if (vals[i] != 0.0f) r = ri;
        }
        bool head;
        int len, hl;
        row_run(r, lane, head, len, hl);
        if (head && r >= 0) atomicAdd(&rowcnt[r], len);
    }
}

// Each counted entry into its row's segment (and its row into `rowof` at
// the same place): the first lane of a run of equal rows reserves the run's
// slots with one atomic, the run's lanes take consecutive slots.
__global__ __launch_bounds__(kThreads) void coo_place_kernel(
    const int *__restrict__ rows, const int *__restrict__ cols, const float *__restrict__ vals,
    long long nnz, int M, int K, const int *__restrict__ rowptr, int *__restrict__ cursor,
    const ConvStatus *__restrict__ st, long long cap, int2 *__restrict__ seg,
    int *__restrict__ rowof)
{
    if (st->total > cap || st->overflow || st->invalid) return;
    const int lane = __lane_id();
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long b = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); b < nnz; b += stride) {
        const long long i = b + lane;
        int r = -1, k = 0;
        float v = 0.0f;
        if (i < nnz) {
            const int ri = rows[i];
            k = cols[i];
            v = vals[i];
            if (ri >= 0 && ri < M && k >= 0 && k < K && v != 0.0f) r = ri;
        }
        bool head;
        int len, hl;
        row_run(r, lane, head, len, hl);
        int slot = 0;
        if (head && r >= 0) slot = atomicAdd(&cursor[r], len);
        slot = __shfl(slot, hl, 64) + (lane - hl);
        if (r >= 0) {
            const int pos = rowptr[r] + slot;
            seg[pos] = make_int2(k, __float_as_int(v));
            rowof[pos] = r;
        }
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
    if (n <= 1 || n > kWaveSortMax) return;
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

// One workgroup of NT threads per listed row of minN < n <= maxN entries:
// its segment sorted in LDS when it has at most CAP entries, in global memory
// otherwise.  Three size classes (NT = 64, 256, 1024) give each thread about
// one compare-exchange per stage; a workgroup whose row is in another class
// exits at once.
template <int CAP, int NT>
__global__ __launch_bounds__(NT) void sort_rows_block_kernel(
    const int *__restrict__ rowptr, const int *__restrict__ longList, int minN, int maxN,
    const ConvStatus *__restrict__ st, long long cap, int2 *__restrict__ seg)
{
    if (st->total > cap || st->overflow || st->invalid) return;
    __shared__ int2 sm[CAP];
    const int r = longList[blockIdx.x];
    const int p0 = rowptr[r], n = rowptr[r + 1] - p0;
    if (n <= minN || n > maxN) return;
    if (n <= CAP) {
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

// One thread per entry e of the row segments: entry i of row r (sorted by
// k) goes to
//   groupIndex[g] + i + sum over the group's other rows r' of
//   #{k' <= k} for r' < r  and  #{k' < k} for r' > r,
// i.e. its place in the group's (k, r) order.
__global__ __launch_bounds__(kThreads) void gcoo_from_rows_kernel(
    const int *__restrict__ rowptr, const int *__restrict__ groupIndex,
    const int *__restrict__ rowof, int total, int M, int cpg, const int2 *__restrict__ seg,
    ConvStatus *__restrict__ st, long long cap, float *__restrict__ vals,
    int *__restrict__ rows, int *__restrict__ cols)
{
    if (st->total > cap || st->overflow || st->invalid) return;
    const int e = blockIdx.x * kThreads + threadIdx.x;
    if (e >= total) return;
    const int r = rowof[e];
    const int p0 = rowptr[r];
    const int i = e - p0;
    const int2 x = seg[e];
    if (i > 0 && seg[e - 1].x == x.x) st->invalid = 2;  // duplicate (r, k)
    const int g = r / cpg;
    const int rg0 = g * cpg;
    const int rg1 = min(rg0 + cpg, M);
    int pos = groupIndex[g] + i;
    for (int r2 = rg0; r2 < rg1; ++r2) {
        if (r2 == r) continue;
        const int q0 = rowptr[r2], qn = rowptr[r2 + 1] - q0;
        pos += lower_bound_k(seg + q0, qn, r2 < r ? x.x + 1 : x.x);
    }
    vals[pos] = __int_as_float(x.y);
    rows[pos] = r;
    cols[pos] = x.x;
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
    const size_t listBytes = sizeof(int) * ((size_t)h->M + 1);
    if (h->coo_tmp_bytes < listBytes) {
        if (h->coo_tmp) GCOO_HIP_TRY(hipFree(h->coo_tmp));
        h->coo_tmp = nullptr;
        h->coo_tmp_bytes = 0;
        if (hipMalloc(&h->coo_tmp, listBytes) != hipSuccess) {
            (void)hipGetLastError();
            h->coo_tmp = nullptr;
            return set_error(GCOO_ERR_ALLOC, "cannot allocate the long-row list");
        }
        h->coo_tmp_bytes = listBytes;
    }
    int *longList = static_cast<int *>(h->coo_tmp);
    int rc = launch_row_scan(h, s, false, longList, kWaveSortMax);
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
    if ((size_t)total > h->coo_rowof_elems) {
        if (h->coo_rowof) GCOO_HIP_TRY(hipFree(h->coo_rowof));
        h->coo_rowof = nullptr;
        h->coo_rowof_elems = 0;
        if (hipMalloc(&h->coo_rowof, sizeof(int) * (size_t)total) != hipSuccess) {
            (void)hipGetLastError();
            h->coo_rowof = nullptr;
            return set_error(GCOO_ERR_ALLOC, "cannot allocate the entry rows");
        }
        h->coo_rowof_elems = (size_t)total;
    }
    if (total > 0) {
        GCOO_HIP_TRY(hipMemsetAsync(h->rowcnt, 0, sizeof(int) * ((size_t)h->M + 1), s));
        hipLaunchKernelGGL(coo_place_kernel, dim3(grid_for(nnz)), dim3(kThreads), 0, s, rows,
                           cols, vals, (long long)nnz, h->M, h->K, h->rowptr, h->rowcnt,
                           h->d_status, (long long)h->cap, h->seg, h->coo_rowof);
        GCOO_CHECK_LAUNCH("coo_place_kernel");
        const unsigned rowBlocks = (unsigned)((h->M + kThreads / 64 - 1) / (kThreads / 64));
        hipLaunchKernelGGL(sort_rows_wave_kernel, dim3(rowBlocks), dim3(kThreads), 0, s,
                           h->rowptr, h->M, h->d_status, (long long)h->cap, h->seg);
        GCOO_CHECK_LAUNCH("sort_rows_wave_kernel");
        const int nLong = h->h_status->longRows;
        const int longest = h->h_status->maxRow;
        if (nLong > 0) {
            const dim3 grid((unsigned)nLong);
            hipLaunchKernelGGL((sort_rows_block_kernel<128, 64>), grid, dim3(64), 0, s,
                               h->rowptr, longList, kWaveSortMax, 128, h->d_status,
                               (long long)h->cap, h->seg);
            if (longest > 128)
                hipLaunchKernelGGL((sort_rows_block_kernel<512, 256>), grid, dim3(256), 0, s,
                                   h->rowptr, longList, 128, 512, h->d_status,
                                   (long long)h->cap, h->seg);
            if (longest > 512 && longest <= 4096)
                hipLaunchKernelGGL((sort_rows_block_kernel<4096, 1024>), grid, dim3(1024), 0, s,
                                   h->rowptr, longList, 512, INT_MAX, h->d_status,
                                   (long long)h->cap, h->seg);
            else if (longest > 4096)
                hipLaunchKernelGGL((sort_rows_block_kernel<kSortLds, 1024>), grid, dim3(1024), 0,
                                   s, h->rowptr, longList, 512, INT_MAX, h->d_status,
                                   (long long)h->cap, h->seg);
            GCOO_CHECK_LAUNCH("sort_rows_block_kernel");
        }
        hipLaunchKernelGGL(gcoo_from_rows_kernel,
                           dim3((unsigned)((total + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                           s, h->rowptr, h->groupIndex(), h->coo_rowof, (int)total, h->M, h->cpg,
                           h->seg, h->d_status, (long long)h->cap, h->vals, h->rows(),
                           h->cols());
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
