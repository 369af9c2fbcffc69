// gcoo_convert.hip -- dense S -> grouped COO on gfx950.
//
// Three kernels replace the reference's K1 / K2 / K3
// (src/gcoo_spdm_kernels.cu:4-28, :160-168, :30-90; launched from
// ShGCOOSpDM::convertToGroupCOOFormat, src/ShGCOOSpDM.cu:19-126):
//   count_kernel   one workgroup per group, 16-byte streaming reads of the
//                  group's cpg rows, per-row wave64 shuffle + LDS reductions;
//   scan_kernel    one 1024-thread workgroup: exclusive scan of the per-row
//                  counts into row pointers (the reference scans per group on
//                  a single thread), groupIndex / nnzPerGroup in the same
//                  pass; 64-bit total and overflow flag in a status word;
//   scatter_kernel one workgroup per group, k-chunks of 1024 columns, one
//                  64-bit wave64 prefix sum carrying the cpg per-row counts in
//                  16-bit fields; writes every entry twice:
//                    - in the reference GCOO order (k, r) -- bit for bit the
//                      reference's vals / rows / cols;
//                    - in row segments: (k, val) pairs of row r at
//                      rowptr[r] + rank, the order the multiply kernel walks
//                      (each group's entries split by row, k ascending).
// All three are HBM-streaming integer/byte work: no MFMA, coalesced 16-byte
// loads along rows of S.
//
// On the 16-byte path with cpg 4 (the reference's default) one kernel does
// all three in a single pass over S (convert_fused_kernel below: counts, a
// decoupled look-back scan across groups, and the scatter from LDS).  The
// three kernels remain for cpg 1 and 2, for K % 4 != 0 or an unaligned S,
// and when the look-back state cannot be allocated.
//
// On the 16-byte path (K % 4 == 0, S 16-byte aligned) the count pass also
// writes a nonzero mask of S -- per row and 256-column block, the four wave
// ballots of `v != 0` over the block's 16-byte pieces -- and the scatter takes
// its per-thread counts from the mask and loads only the pieces of S that
// hold an entry.  At 1% density about a quarter of S's 128-byte lines hold
// one, so the scatter reads a fraction of S instead of all of it a second
// time.  Same entries, same order: only where the values come from changes.
#include "gcoo_internal.h"

#include <algorithm>
#include <climits>

namespace gcoo {
namespace {

constexpr int kConvThreads = 256;
constexpr int kConvWaves = kConvThreads / 64;
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 16;  // rows per thread per pass

// `v != 0.0` exactly as the reference tests it (src/gcoo_spdm_kernels.cu:17):
// true for NaN and denormals, false for +0.0 and -0.0.
__device__ __forceinline__ int nz(float v) { return v != 0.0f ? 1 : 0; }

template <int CPG, bool VEC4>
__global__ __launch_bounds__(kConvThreads) void count_kernel(
    const float *__restrict__ S, int M, int K, int *__restrict__ rowcnt)
{
    const int g = blockIdx.x;
    const int r0 = g * CPG;
    const int nrows = min(CPG, M - r0);
    int c[CPG];
#pragma unroll
    for (int j = 0; j < CPG; ++j) c[j] = 0;
    if (VEC4) {
        const int K4 = K >> 2;
        const f32x4 *base = reinterpret_cast<const f32x4 *>(S + (size_t)r0 * K);
#pragma unroll 2
        for (int i = threadIdx.x; i < K4; i += kConvThreads) {
#pragma unroll
            for (int j = 0; j < CPG; ++j) {
                if (j < nrows) {
                    f32x4 v = __builtin_nontemporal_load(base + (size_t)j * K4 + i);
                    c[j] += nz(v.x) + nz(v.y) + nz(v.z) + nz(v.w);
                }
            }
        }
    } else {
        const float *base = S + (size_t)r0 * K;
        for (int i = threadIdx.x; i < K; i += kConvThreads) {
#pragma unroll
            for (int j = 0; j < CPG; ++j)
                if (j < nrows) c[j] += nz(base[(size_t)j * K + i]);
        }
    }
    __shared__ int part[kConvWaves][CPG];
#pragma unroll
    for (int j = 0; j < CPG; ++j) {
        int t = wave_sum(c[j]);
        if (__lane_id() == 0) part[threadIdx.x >> 6][j] = t;
    }
    __syncthreads();
    if (threadIdx.x < nrows) {
        int t = 0;
#pragma unroll
        for (int w = 0; w < kConvWaves; ++w) t += part[w][threadIdx.x];
        rowcnt[r0 + threadIdx.x] = t;
    }
}

// Count pass with the nonzero mask (16-byte path).  Wave w takes the
// 256-column blocks w, w + 4, ... of the group's rows; lane l loads the piece
// of columns [256 kb + 4 l, 256 kb + 4 l + 4) (zeros past K), the ballots of
// its four components are the block's mask words, and their popcounts the
// block's entries.
template <int CPG>
__global__ __launch_bounds__(kConvThreads) void count_mask_kernel(
    const float *__restrict__ S, int M, int K, int KB, int *__restrict__ rowcnt,
    unsigned long long *__restrict__ mask)
{
    const int g = blockIdx.x;
    const int r0 = g * CPG;
    const int nrows = min(CPG, M - r0);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = __lane_id();
    const int K4 = K >> 2;
    const f32x4 *base = reinterpret_cast<const f32x4 *>(S + (size_t)r0 * K);
    int c[CPG];
#pragma unroll
    for (int j = 0; j < CPG; ++j) c[j] = 0;
    for (int kb = wave; kb < KB; kb += kConvWaves) {
        const int i = kb * 64 + lane;
#pragma unroll
        for (int j = 0; j < CPG; ++j) {
            if (j < nrows) {
                f32x4 v = {0.f, 0.f, 0.f, 0.f};
                if (i < K4) v = __builtin_nontemporal_load(base + (size_t)j * K4 + i);
                const unsigned long long b0 = __ballot(nz(v.x)), b1 = __ballot(nz(v.y));
                const unsigned long long b2 = __ballot(nz(v.z)), b3 = __ballot(nz(v.w));
                c[j] += __popcll(b0) + __popcll(b1) + __popcll(b2) + __popcll(b3);
                const unsigned long long b = lane == 0 ? b0 : lane == 1 ? b1 : lane == 2 ? b2 : b3;
                if (lane < 4) mask[((size_t)(r0 + j) * KB + kb) * 4 + lane] = b;
            }
        }
    }
    __shared__ int part[kConvWaves][CPG];
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < CPG; ++j) part[wave][j] = c[j];
    }
    __syncthreads();
    if (threadIdx.x < nrows) {
        int t = 0;
#pragma unroll
        for (int w = 0; w < kConvWaves; ++w) t += part[w][threadIdx.x];
        rowcnt[r0 + threadIdx.x] = t;
    }
}

// rowptr[0..M] = exclusive scan of rowcnt[0..M) (rowptr[M] = total),
// groupIndex[g] = rowptr[min(g*cpg, M)] for g <= nGroup and
// nnzPerGroup[g] = groupIndex[g+1] - groupIndex[g], nnzPerGroup[nGroup] = 0
// (the reference's scan of the group counts, src/gcoo_spdm_kernels.cu:160-168,
// and the memset of the shared allocation, src/ShGCOOSpDM.cu:37).  With
// `longList` it also lists the rows of more than `longMin` entries in
// ascending order and reports their number in st->longRows.
// One workgroup: thread t takes kScanItems consecutive rows per pass (four
// 16-byte loads and stores), and a group's offsets come from the thread
// holding its first row (CPG divides kScanItems, so its end is there too).
template <int CPG>
__global__ __launch_bounds__(kScanThreads) void scan_kernel(
    const int *__restrict__ rowcnt, int *__restrict__ rowptr, int M, int nGroup,
    int *__restrict__ groupIndex, int *__restrict__ nnzPerGroup,
    ConvStatus *__restrict__ st, int clear_invalid, int *__restrict__ longList, int longMin)
{
    static_assert(kScanItems % 4 == 0 && kScanItems % CPG == 0, "rows per thread");
    __shared__ long long wtot[kScanThreads / 64];
    __shared__ int wlong[kScanThreads / 64];
    __shared__ int wmax[kScanThreads / 64];
    const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    long long carry = 0;
    int lcarry = 0, vmax = 0;
    for (int base = 0; base < M; base += kScanThreads * kScanItems) {
        const int r0 = base + tid * kScanItems;
        const bool whole = r0 + kScanItems <= M;
        int v[kScanItems];
        if (whole) {
            const int4 *p = reinterpret_cast<const int4 *>(rowcnt + r0);
#pragma unroll
            for (int q = 0; q < kScanItems / 4; ++q) {
                const int4 x = p[q];
                v[4 * q] = x.x;
                v[4 * q + 1] = x.y;
                v[4 * q + 2] = x.z;
                v[4 * q + 3] = x.w;
            }
        } else {
#pragma unroll
            for (int i = 0; i < kScanItems; ++i) v[i] = r0 + i < M ? rowcnt[r0 + i] : 0;
        }
        long long tsum = 0;
        int tlong = 0;
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            tsum += v[i];
            tlong += v[i] > longMin ? 1 : 0;
            vmax = max(vmax, v[i]);
        }
        long long incl = tsum;
        int lincl = tlong;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const long long t = __shfl_up(incl, d, 64);
            const int lt = __shfl_up(lincl, d, 64);
            if (lane >= d) {
                incl += t;
                lincl += lt;
            }
        }
        if (lane == 63) {
            wtot[wave] = incl;
            wlong[wave] = lincl;
        }
        __syncthreads();
        long long before = 0, tile = 0;
        int lbefore = 0, ltile = 0;
#pragma unroll
        for (int w = 0; w < kScanThreads / 64; ++w) {
            const long long x = wtot[w];
            const int lx = wlong[w];
            before += (w < wave) ? x : 0;
            lbefore += (w < wave) ? lx : 0;
            tile += x;
            ltile += lx;
        }
        long long run = carry + before + incl - tsum;
        int lrun = lcarry + lbefore + lincl - tlong;
        int out[kScanItems];
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            out[i] = (int)run;
            if (longList != nullptr && v[i] > longMin) longList[lrun++] = r0 + i;
            run += v[i];
        }
        const int endv = (int)run;  // rowptr at r0 + kScanItems (or at M)
        if (whole) {
            int4 *p = reinterpret_cast<int4 *>(rowptr + r0);
#pragma unroll
            for (int q = 0; q < kScanItems / 4; ++q)
                p[q] = make_int4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
        } else {
#pragma unroll
            for (int i = 0; i < kScanItems; ++i)
                if (r0 + i < M) rowptr[r0 + i] = out[i];
        }
#pragma unroll
        for (int i = 0; i < kScanItems; i += CPG) {
            if (r0 + i < M) {
                const int g = (r0 + i) / CPG;
                const int e = i + CPG < kScanItems ? out[i + CPG] : endv;
                groupIndex[g] = out[i];
                nnzPerGroup[g] = e - out[i];
            }
        }
        carry += tile;
        lcarry += ltile;
        __syncthreads();
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) vmax = max(vmax, __shfl_xor(vmax, d, 64));
    if (lane == 0) wmax[wave] = vmax;
    __syncthreads();
    if (tid == 0) {
        int m = 0;
        for (int w = 0; w < kScanThreads / 64; ++w) m = max(m, wmax[w]);
        rowptr[M] = (int)carry;
        groupIndex[nGroup] = (int)carry;
        nnzPerGroup[nGroup] = 0;
        st->total = carry;
        st->overflow = carry > (long long)INT_MAX ? 1 : 0;
        st->maxRow = m;
        st->stall = 0;
        st->longRows = longList != nullptr ? lcarry : 0;
        if (clear_invalid) st->invalid = 0;
    }
}

// One workgroup per group.  Chunk of kConvThreads * KPT columns; thread t owns
// columns [chunk + KPT*t, chunk + KPT*t + KPT).  Its per-row counts are packed
// in 16-bit fields of one 64-bit word (a chunk holds at most 1024 entries per
// row) so that one workgroup-wide exclusive scan yields both the group-order
// offset (sum of the fields) and every row segment's offset (each field).
template <int CPG, bool VEC4>
__global__ __launch_bounds__(kConvThreads) void scatter_kernel(
    const float *__restrict__ S, int M, int K, const int *__restrict__ groupIndex,
    const int *__restrict__ rowptr, const ConvStatus *__restrict__ st, long long cap,
    float *__restrict__ vals, int *__restrict__ rows, int *__restrict__ cols,
    int2 *__restrict__ seg)
{
    if (st->total > cap || st->overflow) return;
    constexpr int KPT = VEC4 ? 4 : 1;
    constexpr int CHUNK = kConvThreads * KPT;
    __shared__ unsigned long long wtot[kConvWaves];
    const int g = blockIdx.x;
    const int r0 = g * CPG;
    const int nrows = min(CPG, M - r0);
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    int gbase = groupIndex[g];
    int rbase[CPG];
#pragma unroll
    for (int j = 0; j < CPG; ++j) rbase[j] = j < nrows ? rowptr[r0 + j] : 0;
    const float *grow = S + (size_t)r0 * K;
    // this thread's elements of the group's rows in the chunk at k0 (zeros
    // past the matrix)
    auto load_chunk = [&](int k0, float (&v)[CPG][KPT]) {
        const int kb = k0 + threadIdx.x * KPT;
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
        }
    };
    float vn[CPG][KPT];
    load_chunk(0, vn);
    for (int k0 = 0; k0 < K; k0 += CHUNK) {
        const int kb = k0 + threadIdx.x * KPT;
        float v[CPG][KPT];
#pragma unroll
        for (int j = 0; j < CPG; ++j)
#pragma unroll
            for (int q = 0; q < KPT; ++q) v[j][q] = vn[j][q];
        // the next chunk's loads are in flight while this one is scanned and
        // written
        if (k0 + CHUNK < K) load_chunk(k0 + CHUNK, vn);
        unsigned long long packed = 0;
#pragma unroll
        for (int j = 0; j < CPG; ++j) {
            int cj = 0;
#pragma unroll
            for (int q = 0; q < KPT; ++q) cj += nz(v[j][q]);
            packed |= (unsigned long long)cj << (16 * j);
        }
        unsigned long long incl = packed;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            unsigned long long t = __shfl_up(incl, d, 64);
            if (lane >= d) incl += t;
        }
        if (lane == 63) wtot[wave] = incl;
        __syncthreads();
        unsigned long long before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kConvWaves; ++w) {
            unsigned long long x = wtot[w];
            before += (w < wave) ? x : 0ull;
            total += x;
        }
        const unsigned long long excl = before + incl - packed;
        if (packed) {
            int posr[CPG];
            int posg = gbase;
#pragma unroll
            for (int j = 0; j < CPG; ++j) {
                const int e = (int)((excl >> (16 * j)) & 0xFFFFu);
                posr[j] = rbase[j] + e;
                posg += e;
            }
#pragma unroll
            for (int q = 0; q < KPT; ++q) {
#pragma unroll
                for (int j = 0; j < CPG; ++j) {
                    if (nz(v[j][q])) {
                        vals[posg] = v[j][q];
                        rows[posg] = r0 + j;
                        cols[posg] = kb + q;
                        ++posg;
                        seg[posr[j]] = make_int2(kb + q, __float_as_int(v[j][q]));
                        ++posr[j];
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < CPG; ++j) {
            const int tj = (int)((total >> (16 * j)) & 0xFFFFu);
            rbase[j] += tj;
            gbase += tj;
        }
        __syncthreads();
    }
}

// The scatter on the 16-byte path with the count pass's mask: one workgroup
// per (group, chunk of 1024 columns), so that no workgroup walks a row.  The
// chunk's starting offsets are read off the mask: wave j counts the set bits
// of row j's mask words before the chunk (the entries of that row to the
// left), and the group-order base is the sum over the group's rows.  Within
// the chunk, wave w takes its w-th 256-column block; lane l's nibble of row j
// (bit q: column 256 kb + 4 l + q holds an entry) is bit l of the block's four
// mask words, S is read only where the nibble is not zero, and the counts, the
// prefix sum and the writes are those of scatter_kernel.
template <int CPG>
__global__ __launch_bounds__(kConvThreads) void scatter_mask_kernel(
    const float *__restrict__ S, int M, int K, int KB,
    const unsigned long long *__restrict__ mask, const int *__restrict__ groupIndex,
    const int *__restrict__ rowptr, const ConvStatus *__restrict__ st, long long cap,
    float *__restrict__ vals, int *__restrict__ rows, int *__restrict__ cols,
    int2 *__restrict__ seg)
{
    if (st->total > cap || st->overflow) return;
    __shared__ unsigned long long wtot[kConvWaves];
    __shared__ int left[CPG];
    const int g = blockIdx.x;
    const int cc = blockIdx.y;
    const int r0 = g * CPG;
    const int nrows = min(CPG, M - r0);
    const int lane = __lane_id();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int K4 = K >> 2;
    // entries of row `wave` left of the chunk: mask words [0, 16 cc)
    if (wave < nrows) {
        const unsigned long long *mr = mask + (size_t)(r0 + wave) * KB * 4;
        int n = 0;
        for (int w = lane; w < 16 * cc; w += 64) n += __popcll(mr[w]);
        n = wave_sum(n);
        if (lane == 0) left[wave] = n;
    }
    __syncthreads();
    int gbase = groupIndex[g];
    int rbase[CPG];
#pragma unroll
    for (int j = 0; j < CPG; ++j) {
        rbase[j] = 0;
        if (j < nrows) {
            rbase[j] = rowptr[r0 + j] + left[j];
            gbase += left[j];
        }
    }
    const f32x4 *grow = reinterpret_cast<const f32x4 *>(S + (size_t)r0 * K);
    const int kb = 4 * cc + wave;
    const int i = kb * 64 + lane;
    unsigned nib[CPG];
    float v[CPG][4];
    unsigned long long packed = 0;
#pragma unroll
    for (int j = 0; j < CPG; ++j) {
        nib[j] = 0;
        if (j < nrows && kb < KB) {
            const unsigned long long *mp = mask + ((size_t)(r0 + j) * KB + kb) * 4;
            nib[j] = (unsigned)((mp[0] >> lane) & 1ull) | (unsigned)((mp[1] >> lane) & 1ull) << 1 |
                     (unsigned)((mp[2] >> lane) & 1ull) << 2 | (unsigned)((mp[3] >> lane) & 1ull) << 3;
        }
        f32x4 x = {0.f, 0.f, 0.f, 0.f};
        if (nib[j]) x = __builtin_nontemporal_load(grow + (size_t)j * K4 + i);
        v[j][0] = x.x;
        v[j][1] = x.y;
        v[j][2] = x.z;
        v[j][3] = x.w;
        packed |= (unsigned long long)__popc(nib[j]) << (16 * j);
    }
    unsigned long long incl = packed;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        unsigned long long t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    if (lane == 63) wtot[wave] = incl;
    __syncthreads();
    unsigned long long before = 0;
#pragma unroll
    for (int w = 0; w < kConvWaves; ++w) before += (w < wave) ? wtot[w] : 0ull;
    if (!packed) return;
    const unsigned long long excl = before + incl - packed;
    int posr[CPG];
    int posg = gbase;
#pragma unroll
    for (int j = 0; j < CPG; ++j) {
        const int e = (int)((excl >> (16 * j)) & 0xFFFFu);
        posr[j] = rbase[j] + e;
        posg += e;
    }
    const int kcol = 4 * i;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int j = 0; j < CPG; ++j) {
            if ((nib[j] >> q) & 1u) {
                vals[posg] = v[j][q];
                rows[posg] = r0 + j;
                cols[posg] = kcol + q;
                ++posg;
                seg[posr[j]] = make_int2(kcol + q, __float_as_int(v[j][q]));
                ++posr[j];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Single-pass conversion (16-byte path, cpg 4): count, scan and scatter in one
// kernel that reads S once.  Workgroup g takes group g: each XCD dispatches
// its workgroups in index order, so every group a workgroup waits for is
// running or done.  (Device-scope atomics on one address serialize, so
// neither a ticket counter nor a finish counter is used.)
//   1. Wave j streams row j of the group in 256-column blocks, kFusedDepth
//      blocks of loads in flight.  Per block the four ballots of `v != 0` are the mask
//      words, kept in LDS with the row's running count at the block's start;
//      the row's entries (column, value bits) are compacted into an LDS list
//      in column order (rank = running count + set bits in lower lanes).
//   2. Decoupled look-back (one wave): the group publishes its entry count
//      and longest row, then reads up to 64 predecessors' flags at a time --
//      their own values, or values over everything up to them -- until it
//      reaches one of the latter, and publishes its own inclusive sum and
//      maximum; the last group's are the total and the longest row.  A flag
//      is one 64-bit word (state, epoch, value), two per group; the epoch
//      makes clearing between conversions unnecessary.  A wait that does not
//      end within a bounded number of polls raises the abort word (the
//      status word's stall field) that every waiting group sees; the host
//      then converts again with the count / scan / scatter kernels.
//   3. Thread t takes entries t, t + 256, ... of the group's LDS lists in row
//      order, so the row segments are a straight copy; an entry's place in
//      (k, r) order is its rank in its row plus, for every other row, that
//      row's entries left of it (at its column too for a lower row), read off
//      the block counts and mask words.  A group with a row longer than its
//      LDS list takes the block-parallel form: each wave takes blocks w,
//      w + 4, ...; a lane's nibble of each row is bit l of the block's mask
//      words, one wave-wide scan of the packed per-row counts places every
//      entry as in scatter_kernel, and values come from S.
// Rows, ranks and both orders are those of the count / scan / scatter kernels.
// ---------------------------------------------------------------------------
constexpr int kFusedThreads = 256;
constexpr int kFusedCpg = 4;
constexpr int kFusedLdsBytes = 40 * 1024;  // dynamic LDS per workgroup (4 per CU)
constexpr int kFusedDepth = 4;             // 256-column blocks of loads in flight per wave
constexpr unsigned long long kLbValue = (1ull << 40) - 1;
constexpr unsigned long long kLbTotal = 1ull << 62;   // the group's own total
constexpr unsigned long long kLbPrefix = 2ull << 62;  // total including all before
constexpr unsigned kLbEpochMask = (1u << 22) - 1;
constexpr int kLbSpinLimit = 1 << 20;

__device__ __forceinline__ unsigned long long lb_word(unsigned long long state, unsigned epoch,
                                                      long long v)
{
    return state | ((unsigned long long)epoch << 40) | ((unsigned long long)v & kLbValue);
}

// set bits of b in the lanes below this one
__device__ __forceinline__ int lanes_below(unsigned long long b)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
}

__device__ __forceinline__ long long wave_sum64(long long v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// flags: [nGroup][2] (entry count, longest row).  st->stall is the abort
// word: zero unless a group's look-back gave up (or the host set it); a
// group that gives up sets it, so that the others stop waiting too, and the
// host then clears it and converts again with the three-kernel path.
__global__ __launch_bounds__(kFusedThreads) void convert_fused_kernel(
    const float *__restrict__ S, int M, int K, int KB, int nGroup, int lcap, unsigned epoch,
    unsigned long long *__restrict__ flags, int *__restrict__ groupIndex, int *__restrict__ nnzPerGroup, int *__restrict__ rowptr,
    ConvStatus *__restrict__ st, long long cap, float *__restrict__ vals,
    int *__restrict__ rows, int *__restrict__ cols, int2 *__restrict__ seg)
{
    constexpr int CPG = kFusedCpg;
    extern __shared__ unsigned long long fused_lds[];
    unsigned long long *smask = fused_lds;                                   // [CPG][KB][4]
    int2 *slist = reinterpret_cast<int2 *>(smask + (size_t)CPG * KB * 4);  // [CPG][lcap]
    int *sbpre = reinterpret_cast<int *>(slist + (size_t)CPG * lcap);      // [CPG][KB]
    __shared__ int sCnt[CPG];
    __shared__ long long sBase;

    const int tid = threadIdx.x, lane = __lane_id();
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = blockIdx.x;
    unsigned *abort = reinterpret_cast<unsigned *>(&st->stall);
    const int r0 = g * CPG;
    const int nrows = min(CPG, M - r0);
    const int K4 = K >> 2;

    // 1. count, mask and compact row `wave`
    if (wave < nrows) {
        const f32x4 *rowp = reinterpret_cast<const f32x4 *>(S + (size_t)(r0 + wave) * K);
        unsigned long long *mrow = smask + (size_t)wave * KB * 4;
        int2 *lrow = slist + (size_t)wave * lcap;
        int *brow = sbpre + (size_t)wave * KB;
        int cnt = 0;
        // loads of blocks kb0 .. kb0 + kFusedDepth - 1 (the index clamped in
        // range, so that every load is issued; values past K are zeroed when
        // processed), the next set in flight while one is processed
        auto load = [&](int kb0, f32x4 (&v)[kFusedDepth]) {
#pragma unroll
            for (int u = 0; u < kFusedDepth; ++u) {
                const int i = min((kb0 + u) * 64 + lane, K4 - 1);
                v[u] = __builtin_nontemporal_load(rowp + i);
            }
        };
        auto process = [&](int kb0, f32x4 (&v)[kFusedDepth]) {
#pragma unroll
            for (int u = 0; u < kFusedDepth; ++u) {
                const int kb = kb0 + u;
                if (kb >= KB) break;
                if (kb * 64 + lane >= K4) v[u] = f32x4{0.f, 0.f, 0.f, 0.f};
                const unsigned long long b0 = __ballot(nz(v[u].x)), b1 = __ballot(nz(v[u].y));
                const unsigned long long b2 = __ballot(nz(v[u].z)), b3 = __ballot(nz(v[u].w));
                if (lane == 0) brow[kb] = cnt;
                if (lane < 4) mrow[kb * 4 + lane] = lane == 0 ? b0 : lane == 1 ? b1 : lane == 2 ? b2 : b3;
                int rank = cnt + lanes_below(b0) + lanes_below(b1) + lanes_below(b2) + lanes_below(b3);
                const int col = 4 * (kb * 64 + lane);
                const float e[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (nz(e[q])) {
                        if (rank < lcap) lrow[rank] = make_int2(col + q, __float_as_int(e[q]));
                        ++rank;
                    }
                }
                cnt += __popcll(b0) + __popcll(b1) + __popcll(b2) + __popcll(b3);
            }
        };
        // two sets of registers in turn: a set's next loads go out as soon as
        // it has been processed, so one set is always in flight
        constexpr int D = kFusedDepth;
        f32x4 va[D], vb[D];
        load(0, va);
        for (int kb0 = 0; kb0 < KB; kb0 += 2 * D) {
            if (kb0 + D < KB) load(kb0 + D, vb);
            process(kb0, va);
            if (kb0 + 2 * D < KB) load(kb0 + 2 * D, va);
            if (kb0 + D < KB) process(kb0 + D, vb);
        }
        if (lane == 0) sCnt[wave] = cnt;
    } else if (lane == 0) {
        sCnt[wave] = 0;
    }
    __syncthreads();
    long long agg = 0;
    int maxr = 0;
#pragma unroll
    for (int j = 0; j < CPG; ++j) {
        agg += sCnt[j];
        maxr = max(maxr, sCnt[j]);
    }

    // 2. look-back for the group's base (and the longest row so far)
    if (wave == 0) {
        long long excl = 0;
        int emax = 0;
        unsigned long long *fs = flags + 2 * (size_t)g;
        if (g == 0) {
            if (lane == 0) {
                __hip_atomic_store(&fs[0], lb_word(kLbPrefix, epoch, agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&fs[1], lb_word(kLbPrefix, epoch, maxr), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        } else {
            if (lane == 0) {
                __hip_atomic_store(&fs[0], lb_word(kLbTotal, epoch, agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&fs[1], lb_word(kLbTotal, epoch, maxr), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
            int p = g - 1;
            int spins = 0;
            for (;;) {
                const int q = p - lane;
                unsigned long long f0 = lb_word(kLbPrefix, epoch, 0);  // before group 0
                unsigned long long f1 = f0;
                if (q >= 0) {
                    f0 = __hip_atomic_load(&flags[2 * (size_t)q], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    f1 = __hip_atomic_load(&flags[2 * (size_t)q + 1], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                }
                const bool v0 = (unsigned)((f0 >> 40) & kLbEpochMask) == epoch && (f0 >> 62) != 0;
                const bool v1 = (unsigned)((f1 >> 40) & kLbEpochMask) == epoch && (f1 >> 62) != 0;
                const bool valid = v0 && v1;
                const bool pre0 = v0 && (f0 >> 62) == 2, pre1 = v1 && (f1 >> 62) == 2;
                // the nearest predecessor with any word already inclusive; it
                // must have both (the two stores of a group land apart)
                const unsigned long long anyPre = __ballot(pre0 || pre1);
                const unsigned long long bothPre = __ballot(pre0 && pre1);
                const unsigned long long bad = __ballot(!valid);
                const int fp = anyPre ? __ffsll((long long)anyPre) - 1 : 64;
                const unsigned long long need = fp >= 63 ? ~0ull : (2ull << fp) - 1;
                const bool ready = !(bad & need) && (fp == 64 || ((bothPre >> fp) & 1ull));
                if (!ready) {
                    bool give_up = ++spins > kLbSpinLimit;
                    if (!give_up && (spins & 63) == 1)
                        give_up = __hip_atomic_load(abort, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT) != 0;
                    if (__builtin_amdgcn_readfirstlane(give_up ? 1 : 0)) {
                        if (lane == 0) atomicExch(abort, 1u);
                        excl = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                const bool use = lane <= fp;
                excl += wave_sum64(use ? (long long)(f0 & kLbValue) : 0);
                int m = use ? (int)(f1 & kLbValue) : 0;
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) m = max(m, __shfl_xor(m, d, 64));
                emax = max(emax, m);
                if (fp < 64) break;
                p -= 64;
            }
            if (lane == 0) {
                __hip_atomic_store(&fs[0], lb_word(kLbPrefix, epoch, excl + agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&fs[1], lb_word(kLbPrefix, epoch, max(emax, maxr)),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (lane == 0) {
            sBase = excl;
            groupIndex[g] = (int)excl;
            nnzPerGroup[g] = (int)agg;
            long long r = excl;
            for (int j = 0; j < nrows; ++j) {
                rowptr[r0 + j] = (int)r;
                r += sCnt[j];
            }
            if (g == nGroup - 1) {
                const long long total = excl + agg;
                rowptr[M] = (int)total;
                groupIndex[nGroup] = (int)total;
                nnzPerGroup[nGroup] = 0;
                st->total = total;
                st->overflow = total > (long long)INT_MAX ? 1 : 0;
                st->maxRow = max(emax, maxr);
                st->invalid = 0;
            }
        }
    }
    __syncthreads();

    // 3. scatter (only when the group's entries fit the arrays)
    const long long gbase = sBase;
    bool anyBig = false;
#pragma unroll
    for (int j = 0; j < CPG; ++j) anyBig |= sCnt[j] > lcap;
    if (agg > 0 && gbase + agg <= cap && !anyBig) {
        // entry-parallel: thread t takes entries t, t + 256, ... of the
        // group's rows in row order (the LDS lists, back to back), so the
        // row segments are a straight copy; the entry's place in (k, r)
        // order is its rank in its row plus, for every other row, that row's
        // entries left of its column (at its column too for a lower row):
        // the row's count before the block plus the mask bits before it
        int rowend[CPG];
        int acc = 0;
#pragma unroll
        for (int j = 0; j < CPG; ++j) {
            acc += sCnt[j];
            rowend[j] = acc;
        }
        for (int e = tid; e < (int)agg; e += kFusedThreads) {
            int j = 0;
#pragma unroll
            for (int i = 0; i < CPG - 1; ++i) j += e >= rowend[i] ? 1 : 0;
            const int t = e - (j ? rowend[j - 1] : 0);
            const int2 en = slist[(size_t)j * lcap + t];
            const int c = en.x;
            const int kb = c >> 8, l = (c >> 2) & 63, q = c & 3;
            const unsigned long long below = l ? ~0ull >> (64 - l) : 0ull;
            int pos = t;
#pragma unroll
            for (int i = 0; i < CPG; ++i) {
                if (i == j || i >= nrows) continue;
                const unsigned long long *mp = smask + ((size_t)i * KB + kb) * 4;
                int n = sbpre[i * KB + kb];
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const unsigned long long m = mp[w];
                    n += __popcll(m & below);
                    if (w < q || (w == q && i < j)) n += (int)((m >> l) & 1ull);
                }
                pos += n;
            }
            vals[gbase + pos] = __int_as_float(en.y);
            rows[gbase + pos] = r0 + j;
            cols[gbase + pos] = c;
            seg[gbase + e] = en;
        }
    } else if (agg > 0 && gbase + agg <= cap) {
        // a row longer than its LDS list: block-parallel, values from S
        int rowoff[CPG];
        bool has[CPG], big[CPG];
        int off = 0;
#pragma unroll
        for (int j = 0; j < CPG; ++j) {
            rowoff[j] = off;
            off += sCnt[j];
            big[j] = sCnt[j] > lcap;
            has[j] = j < nrows;
        }
        const f32x4 *grow = reinterpret_cast<const f32x4 *>(S + (size_t)r0 * K);
        for (int kb = wave; kb < KB; kb += kFusedThreads / 64) {
            unsigned nib[CPG];
            int bpre[CPG];
            unsigned long long packed = 0;
            int gpre = 0;
#pragma unroll
            for (int j = 0; j < CPG; ++j) {
                nib[j] = 0;
                bpre[j] = 0;
                if (has[j]) {
                    const unsigned long long *mp = smask + ((size_t)j * KB + kb) * 4;
                    nib[j] = (unsigned)((mp[0] >> lane) & 1ull) | (unsigned)((mp[1] >> lane) & 1ull) << 1 |
                             (unsigned)((mp[2] >> lane) & 1ull) << 2 | (unsigned)((mp[3] >> lane) & 1ull) << 3;
                    bpre[j] = sbpre[j * KB + kb];
                }
                gpre += bpre[j];
                packed |= (unsigned long long)__popc(nib[j]) << (16 * j);
            }
            unsigned long long incl = packed;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const unsigned long long t = __shfl_up(incl, d, 64);
                if (lane >= d) incl += t;
            }
            if (!packed) continue;
            const unsigned long long ex = incl - packed;
            int rank[CPG];
            long long posg = gbase + gpre;
#pragma unroll
            for (int j = 0; j < CPG; ++j) {
                const int e = (int)((ex >> (16 * j)) & 0xFFFFu);
                rank[j] = bpre[j] + e;
                posg += e;
            }
            const int i = kb * 64 + lane;
            const int kcol = 4 * i;
            float sv[CPG][4];
#pragma unroll
            for (int j = 0; j < CPG; ++j) {
                if (big[j] && nib[j]) {
                    const f32x4 x = __builtin_nontemporal_load(grow + (size_t)j * K4 + i);
                    sv[j][0] = x.x;
                    sv[j][1] = x.y;
                    sv[j][2] = x.z;
                    sv[j][3] = x.w;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
#pragma unroll
                for (int j = 0; j < CPG; ++j) {
                    if ((nib[j] >> q) & 1u) {
                        const int bits = big[j] ? __float_as_int(sv[j][q])
                                                : slist[(size_t)j * lcap + rank[j]].y;
                        vals[posg] = __int_as_float(bits);
                        rows[posg] = r0 + j;
                        cols[posg] = kcol + q;
                        ++posg;
                        seg[gbase + rowoff[j] + rank[j]] = make_int2(kcol + q, bits);
                        ++rank[j];
                    }
                }
            }
        }
    }
}

// 256-column blocks of the mask.
int mask_blocks(const gcoo_matrix *h) { return (h->K + 255) / 256; }

// Make room for the mask of an M x K conversion (false if it cannot be had:
// the conversion then runs without it).
bool ensure_mask(gcoo_matrix *h)
{
    const size_t bytes = (size_t)h->M * (size_t)mask_blocks(h) * 4 * sizeof(unsigned long long);
    if (bytes <= h->mask_bytes) return true;
    if (h->mask) (void)hipFree(h->mask);
    h->mask = nullptr;
    h->mask_bytes = 0;
    if (hipMalloc(&h->mask, bytes) != hipSuccess) {
        (void)hipGetLastError();
        h->mask = nullptr;
        return false;
    }
    h->mask_bytes = bytes;
    return true;
}

template <int CPG>
int launch_count_t(gcoo_matrix *h, const float *S, bool vec4, hipStream_t s)
{
    dim3 grid(h->nGroup), block(kConvThreads);
    h->mask_valid = false;
    if (vec4 && ensure_mask(h)) {
        hipLaunchKernelGGL((count_mask_kernel<CPG>), grid, block, 0, s, S, h->M, h->K,
                           mask_blocks(h), h->rowcnt, h->mask);
        GCOO_CHECK_LAUNCH("count_mask_kernel");
        h->mask_valid = true;
        return GCOO_OK;
    }
    if (vec4)
        hipLaunchKernelGGL((count_kernel<CPG, true>), grid, block, 0, s, S, h->M,
                           h->K, h->rowcnt);
    else
        hipLaunchKernelGGL((count_kernel<CPG, false>), grid, block, 0, s, S,
                           h->M, h->K, h->rowcnt);
    GCOO_CHECK_LAUNCH("count_kernel");
    return GCOO_OK;
}

template <int CPG>
int launch_scatter_t(gcoo_matrix *h, const float *S, bool vec4, hipStream_t s)
{
    dim3 grid(h->nGroup), block(kConvThreads);
    const unsigned chunks = (unsigned)((h->K + kConvThreads * 4 - 1) / (kConvThreads * 4));
    if (vec4 && h->mask_valid && chunks <= 65535u) {  // chunks index grid y
        if (chunks == 0) return GCOO_OK;
        hipLaunchKernelGGL((scatter_mask_kernel<CPG>), dim3(h->nGroup, chunks), block, 0, s, S,
                           h->M, h->K,
                           mask_blocks(h), h->mask, h->groupIndex(), h->rowptr, h->d_status,
                           (long long)h->cap, h->vals, h->rows(), h->cols(), h->seg);
        GCOO_CHECK_LAUNCH("scatter_mask_kernel");
        return GCOO_OK;
    }
    if (vec4)
        hipLaunchKernelGGL((scatter_kernel<CPG, true>), grid, block, 0, s, S,
                           h->M, h->K, h->groupIndex(), h->rowptr, h->d_status,
                           (long long)h->cap, h->vals, h->rows(), h->cols(), h->seg);
    else
        hipLaunchKernelGGL((scatter_kernel<CPG, false>), grid, block, 0, s, S,
                           h->M, h->K, h->groupIndex(), h->rowptr, h->d_status,
                           (long long)h->cap, h->vals, h->rows(), h->cols(), h->seg);
    GCOO_CHECK_LAUNCH("scatter_kernel");
    return GCOO_OK;
}

bool use_vec4(const gcoo_matrix *h, const float *S)
{
    return (h->K % 4 == 0) && ((reinterpret_cast<uintptr_t>(S) & 15u) == 0);
}

}  // namespace

bool fused_conversion_ok(gcoo_matrix *h, const float *S)
{
    // the first loads of a row are issued before any test of K (their index
    // clamped to the row's last 16 bytes): the kernel needs K >= 4
    if (h->cpg != kFusedCpg || h->nGroup == 0 || h->K == 0 || !use_vec4(h, S)) return false;
    const int KB = mask_blocks(h);
    const long long fixed = 144ll * KB;  // masks + block counts
    if (fixed + 32ll * 256 > kFusedLdsBytes) return false;
    if (!h->lb_flags) {
        if (hipMalloc(&h->lb_flags, 2 * sizeof(unsigned long long) * (size_t)h->nGroup) != hipSuccess ||
            hipMemset(h->lb_flags, 0, 2 * sizeof(unsigned long long) * (size_t)h->nGroup) != hipSuccess) {
            (void)hipGetLastError();
            if (h->lb_flags) (void)hipFree(h->lb_flags);
            h->lb_flags = nullptr;
            return false;
        }
    }
    return true;
}

int launch_convert_fused(gcoo_matrix *h, const float *S, hipStream_t s)
{
    const int KB = mask_blocks(h);
    // LDS list length per row: what the budget allows, or -- once the handle
    // has converted before -- its last longest row plus an eighth, so that
    // more workgroups fit on a CU (a longer row only takes the slower form)
    int lcap = (int)((kFusedLdsBytes - 144ll * KB) / 32) & ~63;
    if (h->maxRow > 0) {
        const long long want = ((long long)h->maxRow + h->maxRow / 8 + 64 + 63) & ~63ll;
        lcap = (int)std::min<long long>(lcap, std::max<long long>(256, want));
    }
    const size_t lds = 144 * (size_t)KB + 32 * (size_t)lcap;
    h->lb_epoch = (h->lb_epoch + 1) & kLbEpochMask;
    if (h->lb_epoch == 0) h->lb_epoch = 1;
    h->mask_valid = false;
    hipLaunchKernelGGL(convert_fused_kernel, dim3(h->nGroup), dim3(kFusedThreads), lds, s, S, h->M,
                       h->K, KB, h->nGroup, lcap, h->lb_epoch, h->lb_flags, h->groupIndex(), h->nnzPerGroup(), h->rowptr, h->d_status,
                       (long long)h->cap, h->vals, h->rows(), h->cols(), h->seg);
    GCOO_CHECK_LAUNCH("convert_fused_kernel");
    return GCOO_OK;
}

int launch_count_and_scan(gcoo_matrix *h, const float *S, hipStream_t s)
{
    const bool v4 = use_vec4(h, S);
    int rc = GCOO_OK;
    h->mask_valid = false;
    if (h->nGroup > 0) {
        switch (h->cpg) {
        case 1: rc = launch_count_t<1>(h, S, v4, s); break;
        case 2: rc = launch_count_t<2>(h, S, v4, s); break;
        case 4: rc = launch_count_t<4>(h, S, v4, s); break;
        default: return set_error(GCOO_ERR_INVALID, "cpg %d not supported", h->cpg);
        }
        if (rc) return rc;
    }
    return launch_row_scan(h, s, true);
}

int launch_row_scan(gcoo_matrix *h, hipStream_t s, bool clear_invalid, int *longList,
                    int longMin)
{
    const int ci = clear_invalid ? 1 : 0;
    switch (h->cpg) {
    case 1:
        hipLaunchKernelGGL(scan_kernel<1>, dim3(1), dim3(kScanThreads), 0, s, h->rowcnt,
                           h->rowptr, h->M, h->nGroup, h->groupIndex(), h->nnzPerGroup(),
                           h->d_status, ci, longList, longMin);
        break;
    case 2:
        hipLaunchKernelGGL(scan_kernel<2>, dim3(1), dim3(kScanThreads), 0, s, h->rowcnt,
                           h->rowptr, h->M, h->nGroup, h->groupIndex(), h->nnzPerGroup(),
                           h->d_status, ci, longList, longMin);
        break;
    case 4:
        hipLaunchKernelGGL(scan_kernel<4>, dim3(1), dim3(kScanThreads), 0, s, h->rowcnt,
                           h->rowptr, h->M, h->nGroup, h->groupIndex(), h->nnzPerGroup(),
                           h->d_status, ci, longList, longMin);
        break;
    default: return set_error(GCOO_ERR_INVALID, "cpg %d not supported", h->cpg);
    }
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
