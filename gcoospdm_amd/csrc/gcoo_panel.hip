// gcoo_panel.hip -- panel lists for the LDS-staged multiply (variant 6).
//
// The LDS-staged multiply gives one workgroup a panel of kPanelRows rows and
// walks K in chunks of tk rows of D, staged in LDS and shared by all panel
// rows.  For each (panel, chunk) it needs that chunk's entries of the panel,
// row by row with k ascending inside a row, so that every row's fmaf chain
// stays the reference's (src/gcoo_spdm_kernels.cu:133-142: one chain per
// output in ascending k).  These lists are a permutation of the row segments
// (seg), built from them in three passes:
//   1. panel_rows_kernel<false>: per (row, chunk) entry counts, written at
//      poff[panel][chunk * kPanelRows + row in panel];
//   2. panel_tile_sum_kernel + panel_tile_scan_kernel: per panel, an
//      exclusive scan of those counts in (chunk, row) order, offset by the
//      panel's first entry -- poff then holds each (panel, chunk, row) list
//      start, and the element after the last chunk holds the panel's end;
//   3. panel_rows_kernel<true>: every entry moves to its list slot as
//      (((k - chunk start) * staged row bytes << 8) | row slot in its wave,
//      value bits): the LDS byte offset of its D row within the staged chunk
//      and the accumulator index, ready for the multiply's scalar unit.
// Passes 1 and 3 use one wave per row, 64 entries at a time; a row's entries
// of one chunk are a contiguous run of its k-sorted segment, so an entry's
// rank in its (row, chunk) list is its distance from the run's first entry.
#include "gcoo_internal.h"

namespace gcoo {
namespace {

template <bool SCATTER>
__global__ __launch_bounds__(256) void panel_rows_kernel(
    const int *__restrict__ rowptr, const int2 *__restrict__ seg, int M, int tk, int rowb,
    int nC, int *__restrict__ poff, uint2 *__restrict__ plist,
    const ConvStatus *__restrict__ guard, long long cap)
{
    if (guard != nullptr && guard->total > cap) return;
    const int wave = threadIdx.x >> 6;
    const int lane = __lane_id();
    const int row = blockIdx.x * 4 + wave;
    if (row >= M) return;
    const int p = row / kPanelRows;
    const int rl = row - p * kPanelRows;
    int *base = poff + (size_t)p * panel_stride(nC) + rl;
    const unsigned slot = (unsigned)(rl % kPanelRowsPerWave);
    const int b = rowptr[row];
    const int e = rowptr[row + 1];
    int carry = b;  // first entry of the run still open at the previous batch's end
    for (int i0 = b; i0 < e; i0 += 64) {
        const int i = i0 + lane;
        const bool act = i < e;
        const int2 kv = act ? seg[i] : make_int2(0, 0);
        const int c = kv.x / tk;
        const int cp = (act && i > b) ? seg[i - 1].x / tk : -1;
        const int cn = (act && i + 1 < e) ? seg[i + 1].x / tk : -1;
        const bool first = act && c != cp;
        const unsigned long long fm = __ballot(first);
        const unsigned long long mine = fm & ((2ull << lane) - 1);  // lanes <= this one
        const int s = mine ? i0 + 63 - __clzll(mine) : carry;
        if (fm) carry = i0 + 63 - __clzll(fm);
        if (!act) continue;
        if constexpr (!SCATTER) {
            if (c != cn) base[(size_t)c * kPanelRows] = i - s + 1;
        } else {
            const int pos = base[(size_t)c * kPanelRows] + (i - s);
            plist[pos] = make_uint2(((unsigned)((kv.x - c * tk) * rowb) << 8) | slot, (unsigned)kv.y);
        }
    }
}

// The scan runs over tiles of kScanTile counts, in two passes with one
// workgroup per (panel, tile): tile totals, then each tile's exclusive scan
// seeded by the panel's first entry and the totals of its earlier tiles.
constexpr int kScanTile = 4096;

__device__ __forceinline__ int4 load_counts(const int *a, long long i0, long long L)
{
    int4 v = make_int4(0, 0, 0, 0);
    if (i0 + 3 < L) {
        v = *reinterpret_cast<const int4 *>(a + i0);
    } else {
        if (i0 < L) v.x = a[i0];
        if (i0 + 1 < L) v.y = a[i0 + 1];
        if (i0 + 2 < L) v.z = a[i0 + 2];
    }
    return v;
}

__global__ __launch_bounds__(1024) void panel_tile_sum_kernel(
    const int *__restrict__ poff, int nC, int nT, int *__restrict__ tsum,
    const ConvStatus *__restrict__ guard, long long cap)
{
    if (guard != nullptr && guard->total > cap) return;
    __shared__ int wsum[16];
    const int p = blockIdx.x / nT;
    const int t = blockIdx.x - p * nT;
    const long long L = (long long)nC * kPanelRows + 1;
    const int *a = poff + (size_t)p * panel_stride(nC);
    const int4 v = load_counts(a, (long long)t * kScanTile + threadIdx.x * 4, L);
    const int w = wave_sum(v.x + v.y + v.z + v.w);
    if (__lane_id() == 0) wsum[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int i = 0; i < 16; ++i) tot += wsum[i];
        tsum[blockIdx.x] = tot;
    }
}

__global__ __launch_bounds__(1024) void panel_tile_scan_kernel(
    const int *__restrict__ rowptr, int *__restrict__ poff, int nC, int nT,
    const int *__restrict__ tsum, const ConvStatus *__restrict__ guard, long long cap)
{
    if (guard != nullptr && guard->total > cap) return;
    __shared__ int wsum[16];
    const int p = blockIdx.x / nT;
    const int t = blockIdx.x - p * nT;
    const long long L = (long long)nC * kPanelRows + 1;
    int *a = poff + (size_t)p * panel_stride(nC);
    int base = rowptr[p * kPanelRows];
    for (int i = 0; i < t; ++i) base += tsum[p * nT + i];
    const long long i0 = (long long)t * kScanTile + threadIdx.x * 4;
    const int4 v = load_counts(a, i0, L);
    const int mine = v.x + v.y + v.z + v.w;
    const int wave = threadIdx.x >> 6;
    const int lane = __lane_id();
    const int incl = wave_inclusive_sum(mine);
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int before = base;
    for (int i = 0; i < wave; ++i) before += wsum[i];
    int run = before + incl - mine;
    const int r[4] = {v.x, v.y, v.z, v.w};
    int out[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        out[k] = run;
        run += r[k];
    }
    if (i0 + 3 < L) {
        *reinterpret_cast<int4 *>(a + i0) = make_int4(out[0], out[1], out[2], out[3]);
    } else {
        for (int k = 0; k < 4; ++k)
            if (i0 + k < L) a[i0 + k] = out[k];
    }
}

}  // namespace

int ensure_panel_lists(const gcoo_matrix *h, const int2 *seg, int tk, int rowb,
                       const ConvStatus *guard, hipStream_t s)
{
    if (h->plist_src == seg && h->plist_tk == tk && h->plist_rowb == rowb) return GCOO_OK;
    if ((long long)tk * rowb >= (1ll << 24))
        return set_error(GCOO_ERR_UNSUPPORTED, "panel chunk of %d x %d bytes", tk, rowb);
    const int nP = (h->M + kPanelRows - 1) / kPanelRows;
    const int nC = (h->K + tk - 1) / tk;
    const size_t L = (size_t)nC * kPanelRows + 1;
    const size_t stride = panel_stride(nC);
    const int nT = (int)((L + kScanTile - 1) / kScanTile);
    // counts / offsets of every panel, then the tile totals of the scan
    const size_t elems = (size_t)nP * stride + (size_t)nP * nT;
    if (elems > h->poff_elems) {
        if (h->poff) GCOO_HIP_TRY(hipFree(h->poff));
        h->poff = nullptr;
        h->poff_elems = 0;
        if (hipMalloc(&h->poff, sizeof(int) * elems) != hipSuccess) {
            (void)hipGetLastError();
            h->poff = nullptr;
            return set_error(GCOO_ERR_ALLOC, "cannot allocate %zu panel offsets", elems);
        }
        h->poff_elems = elems;
    }
    if (!h->plist) {
        if (hipMalloc(&h->plist, sizeof(uint2) * (size_t)h->cap) != hipSuccess) {
            (void)hipGetLastError();
            h->plist = nullptr;
            return set_error(GCOO_ERR_ALLOC, "cannot allocate the panel lists");
        }
    }
    if (nP > 0) {
        GCOO_HIP_TRY(hipMemsetAsync(h->poff, 0, sizeof(int) * (size_t)nP * stride, s));
        const unsigned rowBlocks = (unsigned)((h->M + 3) / 4);
        hipLaunchKernelGGL(panel_rows_kernel<false>, dim3(rowBlocks), dim3(256), 0, s, h->rowptr,
                           seg, h->M, tk, rowb, nC, h->poff, h->plist, guard, (long long)h->cap);
        GCOO_CHECK_LAUNCH("panel_rows_kernel<count>");
        int *tsum = h->poff + (size_t)nP * stride;
        const unsigned tiles = (unsigned)((size_t)nP * nT);
        hipLaunchKernelGGL(panel_tile_sum_kernel, dim3(tiles), dim3(1024), 0, s, h->poff, nC, nT,
                           tsum, guard, (long long)h->cap);
        GCOO_CHECK_LAUNCH("panel_tile_sum_kernel");
        hipLaunchKernelGGL(panel_tile_scan_kernel, dim3(tiles), dim3(1024), 0, s, h->rowptr,
                           h->poff, nC, nT, tsum, guard, (long long)h->cap);
        GCOO_CHECK_LAUNCH("panel_tile_scan_kernel");
        hipLaunchKernelGGL(panel_rows_kernel<true>, dim3(rowBlocks), dim3(256), 0, s, h->rowptr,
                           seg, h->M, tk, rowb, nC, h->poff, h->plist, guard, (long long)h->cap);
        GCOO_CHECK_LAUNCH("panel_rows_kernel<scatter>");
    }
    h->plist_src = seg;
    h->plist_tk = tk;
    h->plist_rowb = rowb;
    return GCOO_OK;
}

}  // namespace gcoo
