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

#include <climits>

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
    ConvStatus *__restrict__ hst, const int *__restrict__ rowmap, int waves)
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
    // Slot rr of this wave is row R0 + rr, or with a balanced order row
    // rowmap[pw * RPW + rr] (-1: none).  The panel's rows are the same
    // either way, so its waves' lists tile the same range of the segments;
    // a wave's list then starts after its panel's earlier waves' entries.
    auto rowof = [&](int rr) { return rowmap ? rowmap[(size_t)pw * RPW + rr] : R0 + rr; };
    const int nslots = rowmap ? RPW : R1 - R0;
    int base = rowptr[R0];
    int wend = rowptr[R1];
    if (rowmap) {
        __shared__ int sb[2];
        const int wv = pw % waves;
        const long long p0 = (long long)(pw - wv) * RPW;  // the panel's first row
        if (threadIdx.x < 64) {
            int before = 0, own = 0;
            for (int i = threadIdx.x; i < (wv + 1) * RPW; i += 64) {
                const int r = rowmap[p0 + i];
                const int len = r >= 0 ? rowptr[r + 1] - rowptr[r] : 0;
                if (i < wv * RPW) before += len;
                else own += len;
            }
            before = wave_sum(before);
            own = wave_sum(own);
            if (threadIdx.x == 0) {
                sb[0] = before;
                sb[1] = own;
            }
        }
        __syncthreads();
        base = rowptr[(int)(p0 < M ? p0 : M)] + sb[0];
        wend = base + sb[1];
    }

    for (int i = tid; i < L; i += NT) cnt[i] = 0;
    __syncthreads();
    // 1. run lengths per (chunk, row)
    for (int rr = wave; rr < nslots; rr += NT / 64) {
        const int rw = rowof(rr);
        if (rw < 0) continue;
        const int b = rowptr[rw], e = rowptr[rw + 1];
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
    if (tid == 0) op[nC] = wend;
    // 3. entries to their slots
    for (int rr = wave; rr < nslots; rr += NT / 64) {
        const int rw = rowof(rr);
        if (rw < 0) continue;
        const int b = rowptr[rw], e = rowptr[rw + 1];
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

// ---- balanced order of a panel's rows over its waves -------------------
// The multiply meets at a workgroup barrier every chunk, so a chunk takes as
// long as the wave with the most work in it.  With the panel's rows dealt
// to the waves in order, a wave's share of a chunk is a sum of 21 random
// rows' counts: at 1 % density the slowest of 16 waves carries about 1.45
// times the mean.  This kernel deals the rows instead so that every chunk's
// counts come out even: rows longest first, 16 at a time; each wave, the
// most loaded first, takes the row of the 16 that raises the sum over
// chunks of the per-chunk maximum least (the batched form of the greedy
// that picks one row at a time: the same result to 1 %, 21 steps instead of
// 336).  The waves that also stage D (kAssignStagers) start each chunk with
// kAssignHandicap entries' worth of load.  Measured at n = 8000, s = 0.99
// (16 x 21): 403.7 -> 389.3 us with the host's greedy order, 405.6 -> 397.0
// us with this kernel's.  Any order gives the same bits: each row's chain is
// still its entries in ascending k.  One workgroup per panel;
// rowmap[p * PR + w * rpw + q] = the row in slot q of wave w (-1: none).
// The picks are serial (one wave, 16 per step, each four lane exchanges),
// about 0.25 ms per build at that size: the order pays only for
// lists that serve many multiplies (batches; gcoo_set_panel_balance).
constexpr int kAssignThreads = 1024;
constexpr int kAssignMaxChunks = 128;  // longer K keeps the plain order
constexpr int kAssignMaxRows = 512;    // panel rows (waves * rpw) at most
constexpr int kAssignMaxWaves = 16;
constexpr int kAssignStagers = 5;      // the multiply's staging waves (0..4)
constexpr int kAssignHandicap = 4;

__global__ __launch_bounds__(kAssignThreads) void panel_assign_kernel(
    const int *__restrict__ rowptr, const int2 *__restrict__ seg, int M, int tk, int nC,
    int RPW, int waves, int *__restrict__ rowmap, const ConvStatus *__restrict__ guard,
    long long cap, ConvStatus *__restrict__ hst)
{
    constexpr int NT = kAssignThreads;
    // row strides one word past a multiple of the bank count, so that the
    // rows and waves a wave's lanes read at once fall in different banks
    constexpr int CW = kAssignMaxChunks / 4 + 1;  // count words per row (4 chunks each)
    constexpr int WS = kAssignMaxChunks + 1;      // words per wave of wc
    __shared__ unsigned cnt[kAssignMaxRows * CW];               // bytes: entries per (row, chunk)
    __shared__ int wc[kAssignMaxWaves * WS];                    // per (wave, chunk) load
    __shared__ int cur[kAssignMaxChunks];                       // per-chunk maximum over waves
    __shared__ int key[kAssignMaxRows], idx[kAssignMaxRows];    // rows by length
    __shared__ int part[NT / 256][256];                         // cost partial sums
    __shared__ int load[kAssignMaxWaves], pick[kAssignMaxWaves], worder[kAssignMaxWaves];
    __shared__ int rp[kAssignMaxRows + 1];                      // the panel's row offsets
    if (hst && blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned *src = reinterpret_cast<const unsigned *>(guard);
        unsigned *dst = reinterpret_cast<unsigned *>(hst);
#pragma unroll
        for (int i = 0; i < (int)(sizeof(ConvStatus) / 4); ++i)
            __hip_atomic_store(dst + i, src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __threadfence_system();
    }
    if (guard_skip(guard, cap)) return;
    const int tid = threadIdx.x, lane = __lane_id();
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int PR = waves * RPW;
    const long long r0l = (long long)blockIdx.x * PR;
    const int r0 = (int)r0l;
    const int nr = (int)(r0l + PR <= M ? PR : M - r0l);
    auto count = [&](int q, int c) -> int {
        return q < nr ? (int)((cnt[q * CW + (c >> 2)] >> ((c & 3) * 8)) & 255u) : 0;
    };

    // 1. entries per (row, chunk), and the rows by length
    for (int i = tid; i < PR * CW; i += NT) cnt[i] = 0;
    for (int i = tid; i <= nr; i += NT) rp[i] = rowptr[r0 + i];
    __syncthreads();
    for (int i = tid; i < kAssignMaxRows; i += NT) {
        key[i] = i < nr ? rp[i + 1] - rp[i] : (i < PR ? -1 : -2);
        idx[i] = i;
    }
    // the panel's entries are one range of the segments: every thread takes
    // every NT-th entry (independent loads) and finds its row in rp
    {
        const int e0 = rp[0], e1 = rp[nr];
#pragma unroll 4
        for (int i = e0 + tid; i < e1; i += NT) {
            const int k = seg[i].x;
            int lo = 0, hi = nr - 1;  // the last row whose start is <= i
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (rp[mid] <= i) lo = mid;
                else hi = mid - 1;
            }
            const int c = k / tk;
            atomicAdd(&cnt[lo * CW + (c >> 2)], 1u << ((c & 3) * 8));
        }
    }
    // bitonic sort of the 512 (key, index) pairs: key descending, index
    // ascending among equal keys
    for (int k = 2; k <= kAssignMaxRows; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            __syncthreads();
            if (tid < kAssignMaxRows) {
                const int l = tid ^ j;
                if (l > tid) {
                    const bool first = (key[tid] > key[l]) || (key[tid] == key[l] && idx[tid] < idx[l]);
                    // sections with (tid & k) == 0 end in order, the others reversed
                    if (((tid & k) == 0) != first) {
                        const int tk2 = key[tid], ti = idx[tid];
                        key[tid] = key[l];
                        idx[tid] = idx[l];
                        key[l] = tk2;
                        idx[l] = ti;
                    }
                }
            }
        }
    }
    for (int i = tid; i < waves * nC; i += NT)
        wc[(i / nC) * WS + i % nC] = i / nC < kAssignStagers ? kAssignHandicap : 0;
    if (tid < waves) load[tid] = (tid < kAssignStagers ? kAssignHandicap : 0) * nC;
    for (int c = tid; c < nC; c += NT) cur[c] = waves > 0 ? kAssignHandicap : 0;
    __syncthreads();

    // 2. RPW steps: the next `waves` rows by length, one to each wave
    const int pairs = waves * waves;
    for (int step = 0; step < RPW; ++step) {
        const int *brow = idx + step * waves;  // this step's rows (panel-relative)
        {
            const int pr = tid & 255, sl = tid >> 8;
            int sum = 0;
            if (pr < pairs) {
                const int w = pr / waves, q = brow[pr % waves];
                for (int c = sl; c < nC; c += NT / 256)
                    sum += max(cur[c], wc[w * WS + c] + count(q, c));
            }
            part[sl][pr] = sum;
        }
        __syncthreads();
        if (wv == 0) {
            // the waves by load, most loaded first (index ascending on ties)
            const int ld = lane < waves ? load[lane] : 0;
            int rank = 0;
            for (int o = 0; o < waves; ++o) {
                const int lo = __shfl(ld, o, 64);
                rank += (lo > ld) || (lo == ld && o < lane);
            }
            if (lane < waves) worder[rank] = lane;
            __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the LDS writes above are done
            __builtin_amdgcn_wave_barrier();
            // lane j holds the costs of batch row j for every wave.  A pick
            // is the untaken row of least cost, the lowest row among equal
            // costs: the least (cost << 5 | row) over lanes 0..15, found by
            // four exchanges (a cost is at most the panel's entries plus the
            // handicaps, below 2^24: 512 rows of at most 128 x 150 entries)
            int costs[kAssignMaxWaves];
#pragma unroll
            for (int w = 0; w < kAssignMaxWaves; ++w) {
                costs[w] = 0;
                if (w < waves && lane < waves)
#pragma unroll
                    for (int sl = 0; sl < NT / 256; ++sl) costs[w] += part[sl][w * waves + lane];
            }
            const int wo = lane < waves ? worder[lane] : 0;
            unsigned taken = 0;  // batch rows already picked (bit j)
            for (int t = 0; t < waves; ++t) {
                const int w = __builtin_amdgcn_readlane(wo, t);
                int mine = 0;
#pragma unroll
                for (int u = 0; u < kAssignMaxWaves; ++u)
                    if (u == w) mine = costs[u];
                const bool cand = lane < waves && !((taken >> lane) & 1u);
                int key = cand ? (mine << 5) | lane : INT_MAX;
#pragma unroll
                for (int d = 8; d >= 1; d >>= 1) key = min(key, __shfl_xor(key, d, 64));
                const int bj = __builtin_amdgcn_readfirstlane(key) & 31;
                taken |= 1u << bj;
                if (lane == 0) pick[w] = bj;
            }
        }
        __syncthreads();
        for (int i = tid; i < waves * nC; i += NT) {
            const int w = i / nC, c = i - w * nC;
            wc[w * WS + c] += count(brow[pick[w]], c);
        }
        if (tid < waves) {
            const int pos = step * waves + pick[tid];  // sorted position: key = its length
            const int q = idx[pos];
            load[tid] += q < nr ? key[pos] : 0;
            rowmap[r0l + (long long)tid * RPW + step] = q < nr ? r0 + q : -1;
        }
        __syncthreads();
        for (int c = tid; c < nC; c += NT) {
            int m = 0;
            for (int w = 0; w < waves; ++w) m = max(m, wc[w * WS + c]);
            cur[c] = m;
        }
        __syncthreads();
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

bool panel_balance_supported(const gcoo_matrix *h, int tk, int rpw, int waves)
{
    const int nC = (h->K + tk - 1) / tk;
    return nC <= kAssignMaxChunks && rpw * waves <= kAssignMaxRows && waves <= kAssignMaxWaves &&
           tk <= 255;
}

int build_panel_lists(const gcoo_matrix *h, const int2 *seg, int tk, int rowb, int rpw,
                      int waves, bool widen, int *poff, uint2 *plist, const ConvStatus *guard,
                      hipStream_t s, ConvStatus *hst, bool *exported, int *rowmap)
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
    if (rowmap && !panel_balance_supported(h, tk, rpw, waves))
        return set_error(GCOO_ERR_UNSUPPORTED, "balanced panel order for %d chunks of %d x %d rows",
                         nC, waves, rpw);
    if (nW > 0) {
        // the first kernel behind a conversion exports its status word
        ConvStatus *hx = guard ? hst : nullptr;
        if (rowmap) {
            hipLaunchKernelGGL(panel_assign_kernel, dim3((unsigned)nP), dim3(kAssignThreads), 0, s,
                               h->rowptr, seg, h->M, tk, nC, rpw, waves, rowmap, guard,
                               (long long)h->cap, hx);
            GCOO_CHECK_LAUNCH("panel_assign_kernel");
            hx = nullptr;
        }
        int *gcnt = poff + nW * ((size_t)nC + 1);
        if (lds)
            hipLaunchKernelGGL((panel_wave_lists_kernel<true, kListThreads>), dim3((unsigned)nW),
                               dim3(kListThreads), sizeof(int) * L, s, h->rowptr, seg, h->M, tk,
                               rowb, nC, rpw, widen ? 1 : 0, poff, gcnt, plist, guard,
                               (long long)h->cap, hx, rowmap, waves);
        else
            hipLaunchKernelGGL((panel_wave_lists_kernel<false, kListThreads>), dim3((unsigned)nW),
                               dim3(kListThreads), 0, s, h->rowptr, seg, h->M, tk, rowb, nC, rpw,
                               widen ? 1 : 0, poff, gcnt, plist, guard, (long long)h->cap, hx,
                               rowmap, waves);
        GCOO_CHECK_LAUNCH("panel_wave_lists_kernel");
        if (exported) *exported = guard && hst;
    }
    return GCOO_OK;
}

int ensure_panel_lists(const gcoo_matrix *h, const int2 *seg, int tk, int rowb, int rpw,
                       int waves, bool widen, const ConvStatus *guard, hipStream_t s,
                       bool balance)
{
    balance = balance && panel_balance_supported(h, tk, rpw, waves);
    if (h->plist_src == seg && h->plist_tk == tk && h->plist_rowb == rowb &&
        h->plist_rpw == rpw && h->plist_waves == waves && h->plist_balanced == balance)
        return GCOO_OK;
    if (balance) {
        const size_t rows = ((size_t)h->M + (size_t)rpw * waves - 1) / ((size_t)rpw * waves) *
                            ((size_t)rpw * waves);
        if (rows > h->prow_elems) {
            if (h->prow) GCOO_HIP_TRY(hipFree(h->prow));
            h->prow = nullptr;
            h->prow_elems = 0;
            if (hipMalloc(&h->prow, sizeof(int) * rows) != hipSuccess) {
                (void)hipGetLastError();
                h->prow = nullptr;
                return set_error(GCOO_ERR_ALLOC, "cannot allocate the panel row order");
            }
            h->prow_elems = rows;
        }
    }
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
                               hst, &exported, balance ? h->prow : nullptr);
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
    h->plist_balanced = balance;
    return GCOO_OK;
}

}  // namespace gcoo
