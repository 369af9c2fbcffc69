// gcoo_panel.hip -- panel lists for the LDS-staged multiply (variant 6).
//
// The LDS-staged multiply gives one workgroup a panel of waves * rpw rows
// and walks K in chunks of tk rows of D, staged in LDS and shared by all
// panel rows; wave w of it owns the panel's rows [rpw w, rpw w + rpw).  For
// every
// (panel, wave, chunk) it needs that chunk's entries of the wave's rows, row
// by row with k ascending inside a row, so that every row's fmaf chain stays
// the reference's (src/gcoo_spdm_kernels.cu:133-142: one chain per output in
// ascending k).
//
// A wave's rpw rows are consecutive, so their entries are one contiguous
// range of the row segments, seg[rowptr[R0] .. rowptr[R0 + rpw]).  The wave's
// list is that same range reordered by (chunk, row, k); its chunk c list
// starts at poff[(panel * waves + wave) * (nC + 1) + c], and element nC holds
// its end (panel * waves + wave times rpw is the wave's first row).  No
// scan across waves or panels is needed.  One workgroup per (panel, wave):
//   1. per (chunk, row) entry counts: a row's entries of one chunk are a
//      contiguous run of its k-sorted segment, so the run's last entry
//      writes the run length (one wave per row, 64 entries at a time);
//   2. an exclusive scan of the counts in (chunk, row) order;
//   3. every entry moves to its list slot as
//      ((LDS byte offset of its D row from the first of the two chunk
//      buffers, odd chunks staged one chunk further on, / 16) << 16 |
//      row slot in its wave, value bits -- for BF16 segments the value widened
//      to FP32, BF16 bits << 16): the multiply forms the LDS address
//      with one multiply-add of the high half (hi16 * 16 + lane address) and
//      takes the accumulator index from the low byte.
// The counts sit in LDS when nC * rpw ints fit in 64 KB, else in a global
// scratch area after the offsets.
#include "gcoo_internal.h"

namespace gcoo {
namespace {

constexpr int kListThreads = 1024;  // 16 waves: one row of the 16 each
constexpr int kListLdsInts = 16384;  // 64 KB of counts in LDS

// For every active lane, the index of the first entry of the run (same
// chunk) that holds it; `carry` is the open run's first entry from the
// previous 64 entries.
__device__ __forceinline__ int run_start(bool first, int i0, int lane, int &carry)
{
    const unsigned long long fm = __ballot(first);
    const unsigned long long mine = fm & ((2ull << lane) - 1);  // lanes <= this one
    const int s = mine ? i0 + 63 - __clzll(mine) : carry;
    if (fm) carry = i0 + 63 - __clzll(fm);
    return s;
}

template <bool LDS, int NT>
__global__ __launch_bounds__(NT) void panel_wave_lists_kernel(
    const int *__restrict__ rowptr, const int2 *__restrict__ seg, int M, int tk, int rowb,
    int nC, int RPW, int widen, int *__restrict__ poff, int *__restrict__ gcnt,
    uint2 *__restrict__ plist, const ConvStatus *__restrict__ guard, long long cap,
    ConvStatus *__restrict__ hst)
{
    // the conversion's status word to the host (its host mirror's device
    // view), first of all, so that the host can check the conversion while
    // the multiply queued after this kernel runs
    if (hst && blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned *src = reinterpret_cast<const unsigned *>(guard);
        unsigned *dst = reinterpret_cast<unsigned *>(hst);
#pragma unroll
        for (int i = 0; i < (int)(sizeof(ConvStatus) / 4); ++i)
            __hip_atomic_store(dst + i, src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __threadfence_system();
    }
    if (guard_skip(guard, cap)) return;
    extern __shared__ int lcnt[];
    const int pw = blockIdx.x;  // panel * waves + wave
    const int tid = threadIdx.x, lane = __lane_id();
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int L = nC * RPW;
    int *cnt = LDS ? lcnt : gcnt + (size_t)pw * L;
    int *op = poff + (size_t)pw * (nC + 1);
    const long long r0 = (long long)pw * RPW;
    const int R0 = (int)(r0 < M ? r0 : M);
    const int R1 = (int)(r0 + RPW < M ? r0 + RPW : M);
    const int base = rowptr[R0];

    for (int i = tid; i < L; i += NT) cnt[i] = 0;
    __syncthreads();
    // 1. run lengths per (chunk, row)
    for (int rr = wave; rr < R1 - R0; rr += NT / 64) {
        const int b = rowptr[R0 + rr], e = rowptr[R0 + rr + 1];
        int carry = b;
        for (int i0 = b; i0 < e; i0 += 64) {
            const int i = i0 + lane;
            const bool act = i < e;
            const int c = act ? seg[i].x / tk : -1;
            const int cp = (act && i > b) ? seg[i - 1].x / tk : -1;
            const int cn = (act && i + 1 < e) ? seg[i + 1].x / tk : -1;
            const int s = run_start(act && c != cp, i0, lane, carry);
            if (act && c != cn) cnt[c * RPW + rr] = i - s + 1;
        }
    }
    __syncthreads();
    // 2. exclusive scan in (chunk, row) order: thread t takes a slice of
    // `per` counts
    {
        __shared__ int wsum[NT / 64];
        const int per = (L + NT - 1) / NT;
        const int a0 = tid * per;
        int sum = 0;
        for (int i = 0; i < per; ++i)
            if (a0 + i < L) sum += cnt[a0 + i];
        const int incl = wave_inclusive_sum(sum);
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int run = base + incl - sum;
        for (int w = 0; w < wave; ++w) run += wsum[w];
        __syncthreads();
        for (int i = 0; i < per; ++i) {
            if (a0 + i < L) {
                const int v = cnt[a0 + i];
                cnt[a0 + i] = run;
                run += v;
            }
        }
    }
    __syncthreads();
    for (int c = tid; c < nC; c += NT) op[c] = cnt[c * RPW];
    if (tid == 0) op[nC] = rowptr[R1];
    // 3. entries to their slots
    for (int rr = wave; rr < R1 - R0; rr += NT / 64) {
        const int b = rowptr[R0 + rr], e = rowptr[R0 + rr + 1];
        int carry = b;
        for (int i0 = b; i0 < e; i0 += 64) {
            const int i = i0 + lane;
            const bool act = i < e;
            const int2 kv = act ? seg[i] : make_int2(0, 0);
            const int c = act ? kv.x / tk : -1;
            const int cp = (act && i > b) ? seg[i - 1].x / tk : -1;
            const int s = run_start(act && c != cp, i0, lane, carry);
            if (act)
                plist[cnt[c * RPW + rr] + (i - s)] =
                    make_uint2(((unsigned)(((c & 1) * tk + kv.x - c * tk) * rowb) >> 4 << 16) |
                                   (unsigned)rr,
                               widen ? (unsigned)kv.y << 16 : (unsigned)kv.y);
        }
    }
}

}  // namespace

size_t panel_offset_elems(const gcoo_matrix *h, int tk, int rpw, int waves)
{
    const int rows = rpw * waves;
    const size_t nP = ((size_t)h->M + rows - 1) / rows;
    const int nC = (h->K + tk - 1) / tk;
    const size_t nW = nP * waves;
    const size_t L = (size_t)nC * rpw;
    const bool lds = L <= (size_t)kListLdsInts;
    // offsets of every wave, then (counts beyond LDS) the count scratch
    return nW * ((size_t)nC + 1) + (lds ? 0 : nW * L);
}

int build_panel_lists(const gcoo_matrix *h, const int2 *seg, int tk, int rowb, int rpw,
                      int waves, bool widen, int *poff, uint2 *plist, const ConvStatus *guard,
                      hipStream_t s, ConvStatus *hst, bool *exported)
{
    if (exported) *exported = false;
    // two chunk buffers of offsets / 16 in 16 bits
    if (2ll * tk * rowb > (1ll << 20) || rowb % 16 != 0)
        return set_error(GCOO_ERR_UNSUPPORTED, "panel chunk of %d x %d bytes", tk, rowb);
    if (rpw < 1 || rpw > 255 || waves < 1)
        return set_error(GCOO_ERR_UNSUPPORTED, "panel of %d waves x %d rows", waves, rpw);
    const int rows = rpw * waves;
    const size_t nP = ((size_t)h->M + rows - 1) / rows;
    const int nC = (h->K + tk - 1) / tk;
    const size_t nW = nP * waves;
    const size_t L = (size_t)nC * rpw;
    const bool lds = L <= (size_t)kListLdsInts;
    if (nW > 0) {
        int *gcnt = poff + nW * ((size_t)nC + 1);
        if (lds)
            hipLaunchKernelGGL((panel_wave_lists_kernel<true, kListThreads>), dim3((unsigned)nW),
                               dim3(kListThreads), sizeof(int) * L, s, h->rowptr, seg, h->M, tk,
                               rowb, nC, rpw, widen ? 1 : 0, poff, gcnt, plist, guard,
                               (long long)h->cap, guard ? hst : nullptr);
        else
            hipLaunchKernelGGL((panel_wave_lists_kernel<false, kListThreads>), dim3((unsigned)nW),
                               dim3(kListThreads), 0, s, h->rowptr, seg, h->M, tk, rowb, nC, rpw,
                               widen ? 1 : 0, poff, gcnt, plist, guard, (long long)h->cap,
                               guard ? hst : nullptr);
        GCOO_CHECK_LAUNCH("panel_wave_lists_kernel");
        if (exported) *exported = guard && hst;
    }
    return GCOO_OK;
}

int ensure_panel_lists(const gcoo_matrix *h, const int2 *seg, int tk, int rowb, int rpw,
                       int waves, bool widen, const ConvStatus *guard, hipStream_t s)
{
    if (h->plist_src == seg && h->plist_tk == tk && h->plist_rowb == rowb &&
        h->plist_rpw == rpw && h->plist_waves == waves)
        return GCOO_OK;
    const size_t elems = panel_offset_elems(h, tk, rpw, waves);
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
    // a conversion queued just before asked for its status word: the lists
    // kernel copies it to the host, and the event marks when it has
    ConvStatus *hst = (guard && guard == h->d_status) ? h->statusExport : nullptr;
    bool exported = false;
    ++h->plist_gen;
    int rc = build_panel_lists(h, seg, tk, rowb, rpw, waves, widen, h->poff, h->plist, guard, s,
                               hst, &exported);
    if (rc) return rc;
    if (exported) {
        GCOO_HIP_TRY(hipEventRecord(h->statusEvent, s));
        h->statusExport = nullptr;
        h->statusExported = true;
    }
    h->plist_src = seg;
    h->plist_tk = tk;
    h->plist_rowb = rowb;
    h->plist_rpw = rpw;
    h->plist_waves = waves;
    return GCOO_OK;
}

}  // namespace gcoo
