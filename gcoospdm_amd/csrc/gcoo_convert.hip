// gcoo_convert.hip -- dense S -> grouped COO on gfx950.
//
// Three kernels replace the reference's K1 / K2 / K3
// (src/gcoo_spdm_kernels.cu:4-28, :160-168, :30-90; launched from
// ShGCOOSpDM::convertToGroupCOOFormat, src/ShGCOOSpDM.cu:19-126):
//   count_kernel   one workgroup per group, 16-byte streaming reads of the
//                  group's cpg rows, per-row wave64 shuffle + LDS reductions;
//   scan_kernel    one 1024-thread workgroup: exclusive scan of the per-row
//                  counts into row pointers (the reference scans per group on
//                  a single thread), from which groupIndex / nnzPerGroup are
//                  read off; 64-bit total and overflow flag in a status word;
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
// On the 16-byte path (K % 4 == 0, S 16-byte aligned) the count pass also
// writes a nonzero mask of S -- per row and 256-column block, the four wave
// ballots of `v != 0` over the block's 16-byte pieces -- and the scatter takes
// its per-thread counts from the mask and loads only the pieces of S that
// hold an entry.  At 1% density about a quarter of S's 128-byte lines hold
// one, so the scatter reads a fraction of S instead of all of it a second
// time.  Same entries, same order: only where the values come from changes.
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

// rowptr[0..M] = exclusive scan of rowcnt[0..M) (rowptr[M] = total); then
// groupIndex[g] = rowptr[min(g*cpg, M)] for g <= nGroup and
// nnzPerGroup[g] = groupIndex[g+1] - groupIndex[g], nnzPerGroup[nGroup] = 0
// (the reference's scan of the group counts, src/gcoo_spdm_kernels.cu:160-168,
// and the memset of the shared allocation, src/ShGCOOSpDM.cu:37).
__global__ __launch_bounds__(kScanThreads) void scan_kernel(
    const int *__restrict__ rowcnt, int *__restrict__ rowptr, int M, int cpg, int nGroup,
    int *__restrict__ groupIndex, int *__restrict__ nnzPerGroup,
    ConvStatus *__restrict__ st, int clear_invalid)
{
    __shared__ long long wtot[kScanThreads / 64];
    __shared__ int wmax[kScanThreads / 64];
    const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    long long carry = 0;
    int vmax = 0;
    for (int base = 0; base < M; base += kScanThreads * kScanItems) {
        int v[kScanItems];
        long long tsum = 0;
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            int idx = base + tid * kScanItems + i;
            v[i] = idx < M ? rowcnt[idx] : 0;
            tsum += v[i];
            vmax = max(vmax, v[i]);
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
            if (idx < M) rowptr[idx] = (int)run;
            run += v[i];
        }
        carry += tile;
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
        st->total = carry;
        st->overflow = carry > (long long)INT_MAX ? 1 : 0;
        st->maxRow = m;
        if (clear_invalid) st->invalid = 0;
    }
    __syncthreads();
    for (int g = tid; g <= nGroup; g += kScanThreads) {
        const long long r = (long long)g * cpg;
        const int a = rowptr[r < M ? r : M];
        groupIndex[g] = a;
        if (g < nGroup) {
            const long long r1 = r + cpg;
            nnzPerGroup[g] = rowptr[r1 < M ? r1 : M] - a;
        } else {
            nnzPerGroup[g] = 0;
        }
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

int launch_row_scan(gcoo_matrix *h, hipStream_t s, bool clear_invalid)
{
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(kScanThreads), 0, s, h->rowcnt, h->rowptr,
                       h->M, h->cpg, h->nGroup, h->groupIndex(), h->nnzPerGroup(),
                       h->d_status, clear_invalid ? 1 : 0);
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
