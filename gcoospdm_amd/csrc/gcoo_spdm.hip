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
#include <type_traits>

#include "gcoo_internal.h"
#include "gcoo_panel_asm.h"

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
    if (guard_skip(guard, cap)) return;
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


// ---------------------------------------------------------------------------
// Storage types of D and of the entry values.  FP32: the reference's form.
// BF16: the upper 16 bits of an FP32 (RNE-rounded by gcoo_f32_to_bf16_async);
// widened exactly to FP32 before every FMA, so the arithmetic (and its order)
// is unchanged and only the bytes moved are halved.
// ---------------------------------------------------------------------------
struct F32 {
    static constexpr int bytes = 4;
    __device__ static float val(int bits) { return __int_as_float(bits); }
};
struct BF16 {
    static constexpr int bytes = 2;
    __device__ static float val(int bits) { return __int_as_float(bits << 16); }
};

constexpr int kBufferFlags = 0x00020000;  // raw 32-bit buffer, gfx9 dword3

__device__ __forceinline__ float lo_bf16(unsigned w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float hi_bf16(unsigned w) { return __uint_as_float(w & 0xffff0000u); }

// The bits of VEC consecutive elements of type E (at most 16 bytes), as
// loaded; 2-byte loads zero-extended.
template <typename E, int VEC>
struct RawBits {
    unsigned w[(VEC * E::bytes + 3) / 4];
};

// Those bits at byte offset voff + soff of the buffer.
template <typename E, int VEC>
__device__ __forceinline__ RawBits<E, VEC> buf_load_raw(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    constexpr int B = VEC * E::bytes;
    RawBits<E, VEC> x;
    if constexpr (B == 2) {
        x.w[0] = __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
    } else if constexpr (B == 4) {
        x.w[0] = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
    } else if constexpr (B == 8) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        x.w[0] = (unsigned)v[0];
        x.w[1] = (unsigned)v[1];
    } else {
        static_assert(B == 16, "16-byte loads at most");
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) x.w[i] = (unsigned)v[i];
    }
    return x;
}

// The loaded bits widened to FP32 (exact: a BF16 is the high half of an FP32).
template <typename E, int VEC>
__device__ __forceinline__ void widen(float (&d)[VEC], const RawBits<E, VEC> &x)
{
    if constexpr (E::bytes == 4) {
#pragma unroll
        for (int i = 0; i < VEC; ++i) d[i] = __uint_as_float(x.w[i]);
    } else if constexpr (VEC == 1) {
        d[0] = lo_bf16(x.w[0]);
    } else {
#pragma unroll
        for (int i = 0; i < VEC / 2; ++i) {
            d[2 * i] = lo_bf16(x.w[i]);
            d[2 * i + 1] = hi_bf16(x.w[i]);
        }
    }
}

// VEC consecutive elements of type E at byte offset voff + soff of the
// buffer, widened to FP32.
template <typename E, int VEC>
__device__ __forceinline__ void buf_load(float (&d)[VEC], __amdgpu_buffer_rsrc_t r, int voff,
                                         int soff)
{
    widen<E, VEC>(d, buf_load_raw<E, VEC>(r, voff, soff));
}

// Same, from LDS.
template <typename E, int VEC>
__device__ __forceinline__ void lds_load(float (&d)[VEC], const unsigned char *p)
{
    constexpr int B = VEC * E::bytes;
    if constexpr (B == 2) {
        d[0] = lo_bf16(*reinterpret_cast<const unsigned short *>(p));
    } else if constexpr (B == 4) {
        const unsigned x = *reinterpret_cast<const unsigned *>(p);
        if constexpr (VEC == 1) {
            d[0] = __uint_as_float(x);
        } else {
            d[0] = lo_bf16(x);
            d[VEC - 1] = hi_bf16(x);
        }
    } else if constexpr (B == 8) {
        typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
        const u32x2 x = *reinterpret_cast<const u32x2 *>(p);
        if constexpr (VEC == 2) {
            d[0] = __uint_as_float(x.x);
            d[1] = __uint_as_float(x.y);
        } else {
            d[0] = lo_bf16(x.x); d[1] = hi_bf16(x.x);
            d[2] = lo_bf16(x.y); d[3] = hi_bf16(x.y);
        }
    } else {
        typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
        const u32x4 x = *reinterpret_cast<const u32x4 *>(p);
        const unsigned w[4] = {x.x, x.y, x.z, x.w};
        if constexpr (VEC == 4) {
#pragma unroll
            for (int i = 0; i < 4; ++i) d[i] = __uint_as_float(w[i]);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                d[2 * i] = lo_bf16(w[i]);
                d[2 * i + 1] = hi_bf16(w[i]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Variant 2: row segments.  One wave per (rows, column tile): the rows are
// one row or a group's cpg rows (`nGroup` units of `cpg` rows here), each
// walked over its own k-sorted segment (rowptr / seg, written by the
// conversion scatter) with one accumulator set: no per-entry row dispatch.  D is read with buffer loads: the entry's row
// offset k*N*bytes is a scalar offset, the lane's column offset a vector
// offset, and lanes past N read zeros from the hardware range check.
//
// Long rows first.  A row's chain is serial, so on a skewed matrix the
// longest rows set the launch's end when they start late.  With longBlocks
// > 0 the first workgroups of each column tile take the rows listed in
// longList (longCtr[0] of them, one row per wave, listed on the device by
// length >= the threshold), and the row units skip those rows (skipLen).
// With xpt > 0 (8 / column tiles) XCD x serves tile x / xpt: its first
// workgroups are its share of the long rows (interleaved over the tile's
// xpt XCDs), the rest a contiguous range of the row units.
// ---------------------------------------------------------------------------
template <typename E, int VEC, int U>
__global__ __launch_bounds__(kSpdmThreads) void spdm_rowseg_kernel(
    const int *__restrict__ rowptr, const int2 *__restrict__ seg, int nGroup, int cpg,
    int M, const void *__restrict__ D, unsigned dBytes, int N, float *__restrict__ C,
    int nGroupBlocks, const ConvStatus *__restrict__ guard, long long cap, int skipLen,
    const int *__restrict__ longList, const unsigned *__restrict__ longCtr, int longBlocks,
    int xpt)
{
    if (guard_skip(guard, cap)) return;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = __lane_id();
    int ct, gb;  // column tile; row block, or long block when negative (-1 - index)
    if (xpt) {
        const int x = blockIdx.x % kNumXcd, i = blockIdx.x / kNumXcd;
        const int lbx = (longBlocks + xpt - 1) / xpt;  // long blocks per XCD
        const int rbx = (nGroupBlocks + xpt - 1) / xpt;
        ct = x / xpt;
        gb = i < lbx ? -1 - (i * xpt + x % xpt) : (x % xpt) * rbx + (i - lbx);
        if (gb < 0 ? -1 - gb >= longBlocks : gb >= nGroupBlocks) return;
    } else {
        const int L = xcd_remap(blockIdx.x, gridDim.x);
        const int per = longBlocks + nGroupBlocks;
        ct = L / per;
        gb = L - ct * per;
        gb = gb < longBlocks ? -1 - gb : gb - longBlocks;
    }
    int r0, r1;
    if (gb < 0) {
        const int u = (-1 - gb) * kSpdmWaves + wave;
        if (u >= (int)longCtr[0]) return;
        r0 = longList[u];
        r1 = r0 + 1;
        skipLen = INT_MAX;
    } else {
        const int g = gb * kSpdmWaves + wave;
        if (g >= nGroup) return;
        r0 = g * cpg;
        r1 = min(r0 + cpg, M);
    }
    const int j = ct * (64 * VEC) + lane * VEC;
    const bool active = j < N;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(D), (short)0, (int)dBytes, kBufferFlags);
    const int voff = active ? j * E::bytes : (int)dBytes;  // out of range -> zeros
    const int rowBytes = N * E::bytes;
    for (int r = r0; r < r1; ++r) {
        float acc[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) acc[v] = 0.0f;
        const int p0 = rowptr[r];
        const int p1 = rowptr[r + 1];
        if (p1 - p0 >= skipLen) continue;  // a long row: the long-row kernel owns it
        // B entries (the first n of them when guarded): all their D loads
        // are issued before the first FMA, then the chain runs in order.  The
        // loads keep their raw bits and are widened next to their FMAs: in a
        // guarded round, a widening beside its load would wait for that load
        // before the next one is issued.
        auto round = [&](auto bc, int p, int n) {
            constexpr int B = decltype(bc)::value;
            float a[B];
            RawBits<E, VEC> w[B];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                if (u < n) {
                    const int2 kv = seg[p + u];
                    a[u] = E::val(kv.y);
                    w[u] = buf_load_raw<E, VEC>(rsrc, voff, kv.x * rowBytes);
                }
            }
#pragma unroll
            for (int u = 0; u < B; ++u)
                if (u < n) {
                    float d[VEC];
                    widen<E, VEC>(d, w[u]);
#pragma unroll
                    for (int v = 0; v < VEC; ++v) fmac(acc[v], a[u], d[v]);
                }
        };
        int p = p0;
        for (; p + U <= p1; p += U) round(std::integral_constant<int, U>{}, p, U);
        if constexpr (U > 8)
            for (; p + 8 <= p1; p += 8) round(std::integral_constant<int, 8>{}, p, 8);
        if (p < p1) round(std::integral_constant<int, 8>{}, p, p1 - p);  // 1..7 entries
        if (active) {
            if constexpr (VEC <= 4) {
                store_vec<VEC>(C + (size_t)r * N + j, acc);
            } else {
                const float (&lo)[4] = *reinterpret_cast<const float(*)[4]>(&acc[0]);
                const float (&hi)[4] = *reinterpret_cast<const float(*)[4]>(&acc[4]);
                store_vec<4>(C + (size_t)r * N + j, lo);
                store_vec<4>(C + (size_t)r * N + j + 4, hi);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Variant 5: a group's rows in parallel.  One wave owns 4 consecutive rows
// (a cpg = 4 group) and a 16 * VEC-column tile: lane l serves row
// (l >> 4) and columns tile + VEC * (l & 15).  Every step the 4 row streams
// each take their row's next entry, so one buffer load moves 4 D-row pieces
// of 16*VEC elements -- 16 bytes per lane, the widest gather -- while the
// column tile (and so the D panel a tile sweeps) stays narrow enough for the
// XCD's L2.  Entries come from 4 scalar streams (one s_load_dwordx16 per 8
// entries per row); a row that has run out takes the entry (a = -0.0, k past
// the end): the load returns +0.0 and fmaf(-0.0, +0.0, acc) == acc exactly
// for every acc (+/-0, NaN, inf included), so no predication is needed and
// each row's chain is its own entries in ascending k.
// ---------------------------------------------------------------------------
template <typename E, int VEC>
__global__ __launch_bounds__(kSpdmThreads) void spdm_grouprows_kernel(
    const int *__restrict__ rowptr, const int2 *__restrict__ seg, int M, int nQuads,
    const void *__restrict__ D, unsigned dBytes, int N, float *__restrict__ C,
    const ConvStatus *__restrict__ guard, long long cap)
{
    constexpr int U = 8;                        // entries per scalar batch per row
    constexpr int TN = 16 * VEC;                // columns per tile
    // a row that runs out keeps reading (unused) entries up to the longest
    // sibling's length past its start: the segments must have nnz + maxRow +
    // kSegSlack entries of capacity (every conversion provides that)
    if (guard != nullptr && guard_skip(guard, cap, (long long)guard->maxRow + kSegSlack)) return;
    const int nQuadBlocks = (nQuads + kSpdmWaves - 1) / kSpdmWaves;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int ct = L / nQuadBlocks;
    const int qb = L - ct * nQuadBlocks;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int quad = qb * kSpdmWaves + wave;
    if (quad >= nQuads) return;
    const int lane = __lane_id();
    const int sub = lane >> 4;
    const int j = ct * TN + (lane & 15) * VEC;
    const bool active = j < N;
    const int r0 = quad * 4;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(D), (short)0, (int)dBytes, kBufferFlags);
    const unsigned rowBytes = (unsigned)N * E::bytes;
    const unsigned voffLane = active ? (unsigned)j * E::bytes : dBytes;
    // the 4 row streams (scalar)
    int p[4], pe[4];
    int steps = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = r0 + q;
        p[q] = r < M ? rowptr[r] : 0;
        pe[q] = r < M ? rowptr[r + 1] : 0;
        steps = max(steps, pe[q] - p[q]);
    }
    float acc[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[v] = 0.0f;
    for (int s = 0; s < steps; s += U) {
        // U entries of each row (reads past a row's end stay inside seg: the
        // allocation has at least 1024 entries of slack beyond nnz)
        int kq[4][U];
        float aq[4][U];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int2 e = seg[p[q] + s + u];
                const bool live = p[q] + s + u < pe[q];
                kq[q][u] = live ? e.x : -1;
                aq[q][u] = live ? E::val(e.y) : -0.0f;
            }
        }
        float d[U][VEC];
        float a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = sub == 0 ? kq[0][u] : sub == 1 ? kq[1][u] : sub == 2 ? kq[2][u] : kq[3][u];
            a[u] = sub == 0 ? aq[0][u] : sub == 1 ? aq[1][u] : sub == 2 ? aq[2][u] : aq[3][u];
            const unsigned off = k >= 0 ? voffLane + (unsigned)k * rowBytes : dBytes;
            buf_load<E, VEC>(d[u], rsrc, (int)off, 0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < VEC; ++v) acc[v] = __builtin_fmaf(a[u], d[u][v], acc[v]);
    }
    const int r = r0 + sub;
    if (active && r < M) {
        float *dst = C + (size_t)r * N + j;
        if constexpr (VEC <= 4) {
            store_vec<VEC>(dst, acc);
        } else {
            const float (&lo)[4] = *reinterpret_cast<const float(*)[4]>(&acc[0]);
            const float (&hi)[4] = *reinterpret_cast<const float(*)[4]>(&acc[4]);
            store_vec<4>(dst, lo);
            store_vec<4>(dst + 4, hi);
        }
    }
}

template <typename E, int VEC>
int launch_grouprows(const gcoo_matrix *h, const int2 *seg, const void *D, int N, float *C,
                     const ConvStatus *guard, hipStream_t s)
{
    const int nQuads = (h->M + 3) / 4;
    const int nQuadBlocks = (nQuads + kSpdmWaves - 1) / kSpdmWaves;
    const int nColTiles = (N + 16 * VEC - 1) / (16 * VEC);
    const long long nblocks = (long long)nQuadBlocks * nColTiles;
    if (nblocks > INT32_MAX)
        return set_error(GCOO_ERR_OVERFLOW, "grid of %lld workgroups too large", nblocks);
    const unsigned long long dBytes = (unsigned long long)h->K * N * E::bytes;
    hipLaunchKernelGGL((spdm_grouprows_kernel<E, VEC>), dim3((unsigned)nblocks), dim3(kSpdmThreads),
                       0, s, h->rowptr, seg, h->M, nQuads, D, (unsigned)dBytes, N, C, guard,
                       (long long)h->cap);
    GCOO_CHECK_LAUNCH("spdm_grouprows_kernel");
    return GCOO_OK;
}

// ---------------------------------------------------------------------------
// Variant 3: LDS-staged D panel.  A workgroup of WAVES waves owns a panel of
// WAVES*RPW consecutive rows and a TN = 64*VEC column tile.  It walks k in
// chunks of TK: every chunk's D rows D[k0 .. k0+TK)[tile] are staged once in
// LDS (double-buffered: the next chunk's global loads are in flight while the
// current one is consumed) and read from there by every row of the panel, so
// each staged value serves panel_rows * density FMAs instead of one.  Each
// wave keeps RPW rows' accumulators in VGPRs and walks every row's k-sorted
// segment across the chunks in order: per row the FMA chain is still k
// ascending, bit for bit the reference's.  The next entry of every row is
// held in SGPRs so a chunk with no entry for a row costs one compare.
// ---------------------------------------------------------------------------
template <typename E, int VEC, int RPW, int WAVES, int TK>
__global__ __launch_bounds__(WAVES * 64) void spdm_lds_kernel(
    const int *__restrict__ rowptr, const int2 *__restrict__ seg, int M, int K,
    const void *__restrict__ D, unsigned dBytes, int N, float *__restrict__ C,
    int nPanels, const ConvStatus *__restrict__ guard, long long cap)
{
    constexpr int TN = 64 * VEC;                 // columns per tile
    constexpr int ROWB = TN * E::bytes;          // bytes of one staged D row
    constexpr int NT = WAVES * 64;
    constexpr int CHUNK_P = TK * ROWB / 16;      // 16-byte pieces per chunk
    constexpr int PER_T = CHUNK_P / NT;          // pieces per thread per chunk
    constexpr int PIECES_ROW = ROWB / 16;
    static_assert(CHUNK_P % NT == 0, "chunk must split evenly over threads");
    __shared__ f32x4 sD[2][CHUNK_P];

    if (guard_skip(guard, cap)) return;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int ct = L / nPanels;
    const int panel = L - ct * nPanels;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = __lane_id();
    const int col0 = ct * TN;
    const int j = col0 + lane * VEC;
    const bool active = j < N;

    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(D), (short)0, (int)dBytes, kBufferFlags);
    constexpr int EPP = 16 / E::bytes;           // elements per 16-byte piece
    int s_voff[PER_T];
    int s_row[PER_T];
#pragma unroll
    for (int i = 0; i < PER_T; ++i) {
        const int f = threadIdx.x + i * NT;
        s_row[i] = f / PIECES_ROW;
        const int c = col0 + (f % PIECES_ROW) * EPP;
        s_voff[i] = c < N ? c * E::bytes : (int)dBytes;  // past N -> zeros
    }
    const int rowBytes = N * E::bytes;

    const int rbase = panel * (WAVES * RPW) + wave * RPW;
    int p[RPW], pe[RPW], kn[RPW];
    float vn[RPW];
    float acc[RPW][VEC];
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int r = rbase + q;
        p[q] = r < M ? rowptr[r] : 0;
        pe[q] = r < M ? rowptr[r + 1] : 0;
        if (p[q] < pe[q]) {
            const int2 e = seg[p[q]];
            kn[q] = e.x;
            vn[q] = E::val(e.y);
        } else {
            kn[q] = INT_MAX;
            vn[q] = 0.0f;
        }
#pragma unroll
        for (int v = 0; v < VEC; ++v) acc[q][v] = 0.0f;
    }

    f32x4 stage[PER_T];
    auto issue = [&](int k0) {
#pragma unroll
        for (int i = 0; i < PER_T; ++i) {
            const int k = k0 + s_row[i];
            stage[i] = __builtin_bit_cast(
                f32x4, __builtin_amdgcn_raw_buffer_load_b128(
                           rsrc, k < K ? s_voff[i] : (int)dBytes, k < K ? k * rowBytes : 0, 0));
        }
    };
    auto commit = [&](int buf) {
#pragma unroll
        for (int i = 0; i < PER_T; ++i) sD[buf][threadIdx.x + i * NT] = stage[i];
    };

    issue(0);
    commit(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += TK) {
        const int kend = k0 + TK;
        const bool more = kend < K;
        if (more) issue(kend);
        const unsigned char *sb =
            reinterpret_cast<const unsigned char *>(sD[buf]) + lane * VEC * E::bytes;
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            while (kn[q] < kend) {
                float d[VEC];
                lds_load<E, VEC>(d, sb + (kn[q] - k0) * ROWB);
#pragma unroll
                for (int v = 0; v < VEC; ++v) fmac(acc[q][v], vn[q], d[v]);
                ++p[q];
                if (p[q] < pe[q]) {
                    const int2 e = seg[p[q]];
                    kn[q] = e.x;
                    vn[q] = E::val(e.y);
                } else {
                    kn[q] = INT_MAX;
                }
            }
        }
        if (more) commit(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    if (active) {
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            if (rbase + q < M) {
                float *dst = C + (size_t)(rbase + q) * N + j;
                if constexpr (VEC <= 4) {
                    store_vec<VEC>(dst, acc[q]);
                } else {
                    const float (&lo)[4] = *reinterpret_cast<const float(*)[4]>(&acc[q][0]);
                    const float (&hi)[4] = *reinterpret_cast<const float(*)[4]>(&acc[q][4]);
                    store_vec<4>(dst, lo);
                    store_vec<4>(dst + 4, hi);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Variant 6: panel-staged.  One workgroup of WAVES waves owns a panel of
// WAVES * RPW rows and a 256-column tile of C (4 columns per lane); wave w
// keeps the accumulators of panel rows [RPW w, RPW w + RPW) in VGPRs (FP32:
// 16 waves of 16 rows, 12 of 28 or 8 of 42, so 256 or 336 rows share each
// staged chunk; BF16: 16 of 16).  K is
// walked in chunks of TK rows of D: a chunk's tile piece is copied to LDS once,
// by direct buffer-to-LDS loads issued one chunk ahead into the other of two
// buffers, and serves every entry of the panel in that chunk -- on average
// 256 * density entries per staged D row instead of one L2 gather per entry.
//   * Waves 0..4 stage the chunks (15 wave loads each); the last wave stages
//     nothing and, in one workgroup of each column tile per chunk
//     (panel == chunk % panels), touches the tile piece of the chunk
//     kPanelPrefetch chunks ahead so that it is in L2 when the stagers load
//     it.  The panels of a tile run in step, so without this every chunk
//     would be a simultaneous L2 miss of all of them.  The touch wave never
//     waits for those loads inside the loop: vector memory completions are
//     counted in issue order, and a wait would stall on the miss it is there
//     to absorb.  The hardware issues from the oldest (lowest) waves first, so
//     the staging sits with them and the younger waves, which finish their
//     entries last, carry no loads of their own.
//   * Each wave takes its rows' entries of the chunk from the panel list
//     (built by gcoo_panel.hip from the row segments, ordered by row then k)
//     through scalar loads.  A chunk is one generated asm block
//     (gcoo_panel_asm.h): the first 16 entries' scalar loads go out before
//     the chunk barrier (their latency passes while the wave waits there),
//     then the barrier and the staging of the next chunk, then entries 0..15
//     in groups of 4 with the next group's LDS reads in flight, then batches
//     of 8, and the last 1..7 entries in one guarded round.  Before each batch
//     a wave sets its issue priority from the entries it has left in the
//     chunk, so that the wave the barrier waits for is served first.
//   * The step's row is wave-uniform but varies per entry.  The 4 * RPW
//     accumulators sit in v0.. (component v of row slot q in v[RPW v + q])
//     and each step is 4 v_fma_f32 in VGPR-index mode, source 2 and
//     destination relative to M0 = q: acc[q] = fmaf(x, d, acc[q]) with no
//     moves and no branches.  A list entry's first word is (LDS offset / 16)
//     << 16 | row slot: one v_mad_u32_u16 of its high half gives the LDS
//     address, and the index instructions take the slot from its low byte.
//   * BF16 D is staged as it is (150-row chunks of 512-byte rows, two rows
//     per wave load), read 8 bytes per lane and widened exactly to FP32
//     before the steps, the lane's first column by a second, 16-bit read into
//     the high half of a register with a zero low half; the list values are
//     stored widened.
// Every row's chain is its entries in ascending k, as in all variants.
// ---------------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kPanelPrefetch = 6;  // chunks between the L2 touch and the staging
constexpr double kPanelMinDensity = 0.0025;  // automatic choice of variant 6 (FP32) ...
constexpr double kPanelMinDensityBf16 = 0.005;  // ... and for BF16
constexpr long long kPanelMinWorkgroups = 256;  // ... with at least one workgroup per CU

typedef float f32x8 __attribute__((ext_vector_type(8)));

// One workgroup's tile: panel `panel` of the row panels, column tile ct.
// The workgroups that stage the same D tile at the same time on one XCD
// take turns at the L2 touch: this one touches chunk c when c % tmod ==
// tidx (a single launch: all panels of a tile, tmod = nPanels, tidx = panel).
template <typename E, int TK, int WAVES, int RPW>
__device__ __forceinline__ void panel_tile(const int *__restrict__ poff,
                                           const uint2 *__restrict__ plist, int M, int K,
                                           int nC, const void *__restrict__ D, unsigned dBytes,
                                           int N, float *__restrict__ C, int ct, int panel,
                                           int tmod, int tidx, const int *__restrict__ rowmap)
{
    constexpr int VEC = 4;
    constexpr int PR = WAVES * RPW;           // panel rows
    static_assert(E::bytes == 4 ? (RPW == 16 || RPW == 21 || RPW == 28 || RPW == 42)
                                : (RPW == 16 || RPW == 21 || RPW == 28),
                  "rows per wave of the generated asm blocks");
    constexpr int TN = 64 * VEC;              // columns per tile
    constexpr int LB = VEC * E::bytes;        // bytes of a D row per lane
    constexpr int ROWB = 64 * LB;             // bytes of a staged D row
    constexpr int CHUNKB = TK * ROWB;         // bytes of a staged chunk
    constexpr int NLOAD = CHUNKB / 1024;      // 1 KB wave loads per chunk
    constexpr int SW = 5;                     // staging waves
    constexpr int TW = WAVES - 1;             // L2 touch wave
    static_assert(TW >= SW, "the touch wave stages nothing");
    static_assert(CHUNKB % 1024 == 0 && NLOAD % SW == 0, "chunk loads must split over the stagers");
    constexpr int LPW = NLOAD / SW;           // wave loads per staging wave per chunk
    constexpr int PPR = ROWB / 16;            // 16-byte pieces per staged D row
    constexpr int RPL = 64 / PPR;             // D rows per wave load
    constexpr int LINES = ROWB / 128;         // 128-byte lines per staged D row
    constexpr int RPP = 64 / LINES;           // D rows per prefetch load
    constexpr int NPF = (TK + RPP - 1) / RPP; // prefetch loads per chunk
    constexpr int DUMMY = 2 * CHUNKB / 16;    // 256-byte target of the prefetch loads
    static_assert(2 * CHUNKB + 256 <= 160 * 1024, "LDS");
    __shared__ f32x4 sD[DUMMY + 16];

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = __lane_id();
    const int col0 = ct * TN;
    const int rowBytes = N * E::bytes;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(D), (short)0, (int)dBytes, kBufferFlags);

    // Staging: this lane's 16-byte piece of every wave load, D row (first
    // row of the load + prow), columns from pcol; pieces past N read zeros.
    // Rows past K are never read by an entry: a load of whole rows past K is
    // skipped, and a row past K inside a load reads zeros.  With one D row
    // per load (FP32) the test is wave-uniform, so a load costs no vector
    // instructions: the chunk's loads are issued right after the barrier,
    // ahead of the entries.
    const int prow = lane / PPR;
    const int pcol = col0 + (lane % PPR) * (16 / E::bytes);
    const int pvoff = pcol < N ? pcol * E::bytes + prow * rowBytes : (int)dBytes;
    // The wave's loads are consecutive: the LDS destination and the D row
    // advance by a constant per load (base + immediate, running sum), so no
    // per-load constants are kept in scalar registers across the loop.
    typedef __attribute__((address_space(3))) unsigned char lds_u8;
    lds_u8 *const ldsBase = (lds_u8 *)sD + wave * LPW * 1024;
    auto stage = [&](int c, int buf) {
        if (wave < SW) {
            const int k0 = c * TK + wave * LPW * RPL;  // the wave's first D row
            int soff = k0 * rowBytes;
            lds_u8 *dst = ldsBase + (buf ? CHUNKB : 0);
#pragma unroll
            for (int i = 0; i < LPW; ++i) {
                const int k = k0 + i * RPL;
                if (k >= K) break;
                const int voff = (RPL == 1 || k + prow < K) ? pvoff : (int)dBytes;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (__attribute__((address_space(3))) void *)(dst + i * 1024),
                                                         16, voff, soff, 0, 0);
                // an opaque add: the compiler would otherwise keep one
                // precomputed row offset per load in a scalar register
                asm("s_add_u32 %0, %0, %1" : "+s"(soff) : "s"(RPL * rowBytes));
            }
        }
    };
    // L2 touch of chunk c by the last wave: 4 bytes of every 128-byte line of the
    // tile piece (lane: row lane / LINES, line lane % LINES).
    const int frow = lane / LINES;
    const int fbyte = col0 * E::bytes + (lane % LINES) * 128;
    auto touch = [&](int c) {
#pragma unroll
        for (int i = 0; i < NPF; ++i) {
            const int r = i * RPP + frow;
            const bool ok = r < TK && c * TK + r < K && fbyte < rowBytes;
            const int voff = ok ? fbyte + frow * rowBytes : (int)dBytes;
            // soffset must be wave-uniform (a per-lane choice here would
            // make the compiler loop over its distinct values)
            const int soff = c * TK + i * RPP < K ? (c * TK + i * RPP) * rowBytes : 0;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                rsrc, (__attribute__((address_space(3))) void *)&sD[DUMMY], 4, voff, soff, 0, 0);
        }
    };

    // 4 * RPW accumulators pinned to v0.. in f32x8 blocks, f32x4 blocks for
    // an odd RPW (component c of row slot q in register RPW c + q)
    constexpr int AW = RPW % 2 == 0 ? 8 : 4;
    typedef float accblk __attribute__((ext_vector_type(AW)));
    accblk acc[4 * RPW / AW];
#pragma unroll
    for (int i = 0; i < 4 * RPW / AW; ++i) acc[i] = 0.0f;
    // this wave's chunk list starts (nC + 1 of them, the last its end)
    const int *op = poff + ((size_t)panel * WAVES + wave) * ((size_t)nC + 1);
    const unsigned char *ldsLane = reinterpret_cast<const unsigned char *>(sD) + lane * LB;

    // This wave's list bounds of chunks [cw, cw + 64), lane i holding chunk
    // cw + i, read with v_readlane: a scalar load of the next chunk's bounds
    // would be waited for before the entry loop, a full round trip per chunk.
    // The window is reloaded (and waited for) once every 64 chunks.
    int cw = 0;
    int vpb = 0, vpe = 0;
    auto load_window = [&]() {
        vpb = 0;
        vpe = 0;
        if (cw + lane < nC) {
            vpb = op[cw + lane];
            vpe = op[cw + lane + 1];
        }
        asm volatile("" : "+v"(vpb), "+v"(vpe));  // the wait for the loads lands here
    };
    load_window();
    stage(0, 0);
    {
        // The chunk loop in asm (gcoo_panel_asm.h), one block per
        // segment of consecutive chunks of a 64-chunk window.  Per chunk: the
        // barrier, the scalar loads of the next chunk's first 16 entries
        // (landing while this chunk runs), the staging of the next chunk,
        // and the entry loop.  The L2 touch wave ends a segment before each
        // chunk at which it touches, and touches between segments.
        static_assert(LPW == 15 && RPL == (E::bytes == 4 ? 1 : 2) && CHUNKB == 75 * 1024,
                      "the asm blocks stage 15 loads of one (FP32) or two (BF16) D rows "
                      "into 75 KB buffers");
        typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
        const u32x4 rs = {(unsigned)reinterpret_cast<uintptr_t>(D),
                          (unsigned)(reinterpret_cast<uintptr_t>(D) >> 32) & 0xffffu,
                          (unsigned)dBytes, (unsigned)kBufferFlags};
        const unsigned vb0 =
            (unsigned)(size_t)((const __attribute__((address_space(3))) unsigned char *)ldsLane);
        const unsigned sdst = __builtin_amdgcn_readfirstlane((unsigned)(size_t)ldsBase);
        // staging: row offsets below kb are loaded, and when the wave's last
        // row of a chunk is below K too (offset <= lim) without tests; the
        // other waves have kb = 0 and stage nothing.  A load advances RPL
        // rows (uload bytes); BF16 lanes of a load's second row take voffz
        // when that row is past K.
        const unsigned urow = (unsigned)rowBytes;
        const unsigned uload = (unsigned)RPL * urow;
        const unsigned kb = __builtin_amdgcn_readfirstlane(wave < SW ? dBytes : 0u);
        const unsigned lim = __builtin_amdgcn_readfirstlane(
            (wave < SW && dBytes >= (unsigned)LPW * uload) ? dBytes - (unsigned)LPW * uload : 0u);
        const unsigned tkrowb = (unsigned)TK * urow;
        unsigned sso = __builtin_amdgcn_readfirstlane((unsigned)(TK + wave * LPW * RPL) * urow);
        const int pvoffz = (RPL == 1 || prow == 0) ? pvoff : (int)dBytes;
        int c = 0;
        for (;;) {
            // the window as list byte offsets from its first chunk's entries
            const int pb0 = __builtin_amdgcn_readlane(vpb, 0);
            const uint2 *pbase = plist + pb0;
            const bool inw = cw + lane < nC;
            const unsigned vpbo = inw ? (unsigned)(vpb - pb0) * 8u : 0u;
            const unsigned vpn = inw ? (unsigned)(vpe - vpb) : 0u;
            const int wend = nC < cw + 64 ? nC : cw + 64;
            while (c < wend) {
                int cend = wend;
                if (wave == TW) {
                    if (c % tmod == tidx && c + kPanelPrefetch < nC) touch(c + kPanelPrefetch);
                    int t = (tidx - (c + 1)) % tmod;
                    t = c + 1 + (t < 0 ? t + tmod : t);  // next chunk that is touched at
                    if (t < cend && t + kPanelPrefetch < nC) cend = t;
                }
                unsigned sli = __builtin_amdgcn_readfirstlane((unsigned)(c - cw));
                const unsigned liend = __builtin_amdgcn_readfirstlane((unsigned)(cend - cw));
                const unsigned par = __builtin_amdgcn_readfirstlane((unsigned)c & 1u);
#define GCOO_PANEL_SEGMENT(T, N)                                                           \
    asm volatile(GCOO_PANEL_SEGMENT_ASM_##T##R##N                                          \
                 : GCOO_PANEL_ACC_OPERANDS_R##N(acc), [sli] "+s"(sli), [sso] "+s"(sso)     \
                 : [pbase] "s"(pbase), [vpbo] "v"(vpbo), [vpn] "v"(vpn), [liend] "s"(liend), \
                   [par] "s"(par), [vb0] "v"(vb0), [sdst] "s"(sdst),                      \
                   [lim] "s"(lim), [kb] "s"(kb), [rowb] "s"(uload), [row1] "s"(urow),     \
                   [tkrowb] "s"(tkrowb), [voff] "v"(pvoff), [voffz] "v"(pvoffz),          \
                   [rsrc] "s"(rs)                                                          \
                 : GCOO_PANEL_SEGMENT_CLOBBERS_##T##R##N)
                if constexpr (E::bytes == 4) {
                    if constexpr (RPW == 16) GCOO_PANEL_SEGMENT(, 16);
                    else if constexpr (RPW == 21) GCOO_PANEL_SEGMENT(, 21);
                    else if constexpr (RPW == 28) GCOO_PANEL_SEGMENT(, 28);
                    else GCOO_PANEL_SEGMENT(, 42);
                } else {
                    if constexpr (RPW == 16) GCOO_PANEL_SEGMENT(BF16_, 16);
                    else if constexpr (RPW == 21) GCOO_PANEL_SEGMENT(BF16_, 21);
                    else GCOO_PANEL_SEGMENT(BF16_, 28);
                }
#undef GCOO_PANEL_SEGMENT
                c = cend;
            }
            if (c >= nC) break;
            cw = c;
            load_window();
        }
    }
    // no LDS write may land after the workgroup's LDS is released
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");

    const int j = col0 + lane * VEC;
    if (j < N) {
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            // slot q's row: in order, or the lists' balanced order (-1: none)
            const int r = rowmap ? rowmap[panel * PR + wave * RPW + q] : panel * PR + wave * RPW + q;
            float out[VEC];
#pragma unroll
            for (int c = 0; c < VEC; ++c) out[c] = acc[(RPW * c + q) / AW][(RPW * c + q) % AW];
            if (r >= 0 && r < M) store_vec<VEC>(C + (size_t)r * N + j, out);
        }
    }
}

template <typename E, int TK, int WAVES, int RPW>
__global__ __launch_bounds__(WAVES * 64) void spdm_panel_kernel(
    const int *__restrict__ poff, const uint2 *__restrict__ plist, int M, int K, int nC,
    const void *__restrict__ D, unsigned dBytes, int N, float *__restrict__ C, int nPanels,
    const ConvStatus *__restrict__ guard, long long cap, const int *__restrict__ rowmap)
{
    if (guard_skip(guard, cap)) return;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int ct = L / nPanels;
    const int panel = L - ct * nPanels;
    panel_tile<E, TK, WAVES, RPW>(poff, plist, M, K, nC, D, dBytes, N, C, ct, panel, nPanels,
                                  panel, rowmap);
}

// The panel multiply of many jobs in one launch (the panel jobs of a batch):
// workgroup b takes tile units[b] = (job, the job's tile index as in the
// single-job launch: column tile * panels + panel, L2-touch turn tidx of
// tmod), or nothing for job -1.  Workgroups are dispatched to the XCDs
// round-robin (b % 8); the host deals the units over the XCDs by estimated
// work and keeps a job's units on one XCD next to each other, so that they
// stage the same D tile together through that XCD's L2.
template <typename E, int TK, int WAVES, int RPW>
__global__ __launch_bounds__(WAVES * 64) void spdm_panel_batch_kernel(
    const PanelUnit *__restrict__ units, const PanelJob *__restrict__ jobs)
{
    const PanelUnit u = units[blockIdx.x];
    if (u.job < 0) return;
    const PanelJob &j = jobs[u.job];
    const int ct = u.tile / j.nPanels;
    panel_tile<E, TK, WAVES, RPW>(j.poff, j.plist, j.M, j.K, j.nC, j.D, j.dBytes, j.N, j.C, ct,
                                  u.tile - ct * j.nPanels, u.tmod, u.tidx, j.rowmap);
}

// ---------------------------------------------------------------------------
// Long rows.  A row with thousands of entries walked by one wave is a serial
// chain of dependent load -> FMA steps: latency-bound, and on skewed matrices
// it is the tail of the whole launch.  Long rows are instead listed (order
// irrelevant: rows are independent) and pulled by a persistent kernel in
// which one workgroup owns one (long row, TN-column tile) at a time: for T
// entries at a time, all NT threads load the T D-row pieces into LDS at once
// (T * TN elements in flight per workgroup, the entries of the chunk after
// next already staged), then thread t runs column t's fmaf chain over the T
// entries from LDS -- still row r's entries in ascending k, so the bits are
// the reference's.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void longrow_classify_kernel(
    const int *__restrict__ rowptr, int M, int minLen, int *__restrict__ list,
    unsigned *__restrict__ ctr, const ConvStatus *__restrict__ guard, long long cap)
{
    // behind a conversion that overflowed its capacity the segments are
    // stale: list nothing (the host re-runs the multiply)
    if (guard_skip(guard, cap)) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < M && rowptr[r + 1] - rowptr[r] >= minLen) list[atomicAdd(&ctr[0], 1u)] = r;
}

__global__ __launch_bounds__(64) void longrow_reset_kernel(unsigned *ctr)
{
    if (threadIdx.x < 4) ctr[threadIdx.x] = 0;
}

template <typename E, int W, int T>
__global__ __launch_bounds__(W * 64) void spdm_longrow_kernel(
    const int *__restrict__ rowptr, const int2 *__restrict__ seg, const int *__restrict__ list,
    unsigned *__restrict__ ctr, int colTiles, const void *__restrict__ D, unsigned dBytes, int N,
    float *__restrict__ C)
{
    constexpr int NT = W * 64;              // threads = columns per tile
    constexpr int ROWB = NT * E::bytes;     // bytes of one staged D row piece
    constexpr int PIECES_ROW = ROWB / 16;
    constexpr int CHUNK_P = T * PIECES_ROW;
    constexpr int PER_T = CHUNK_P / NT;
    static_assert(CHUNK_P % NT == 0 && T <= NT, "chunk geometry");
    __shared__ f32x4 sD[2][CHUNK_P];
    __shared__ int2 sE[2][T];
    __shared__ int sItem;

    const int tid = threadIdx.x;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(D), (short)0, (int)dBytes, kBufferFlags);
    const unsigned rowBytes = (unsigned)N * E::bytes;
    constexpr int EPP = 16 / E::bytes;
    for (;;) {
        if (tid == 0) sItem = (int)atomicAdd(&ctr[1], 1u);
        __syncthreads();
        const int item = sItem;
        __syncthreads();
        if (item >= (int)ctr[0] * colTiles) return;
        const int row = list[item / colTiles];
        const int ct = item % colTiles;
        const int col0 = ct * NT;
        const int p0 = rowptr[row], p1 = rowptr[row + 1];
        const int nChunks = (p1 - p0 + T - 1) / T;
        // per-thread staging geometry
        unsigned s_voff[PER_T];
        int s_ent[PER_T];
#pragma unroll
        for (int i = 0; i < PER_T; ++i) {
            const int f = tid + i * NT;
            s_ent[i] = f / PIECES_ROW;
            const int c = col0 + (f % PIECES_ROW) * EPP;
            s_voff[i] = c < N ? (unsigned)c * E::bytes : dBytes;
        }
        auto load_entries = [&](int chunk, int buf) {
            const int p = p0 + chunk * T + tid;
            if (tid < T) sE[buf][tid] = (chunk < nChunks && p < p1) ? seg[p] : make_int2(-1, 0);
        };
        f32x4 stage[PER_T];
        auto issue = [&](int buf) {  // D pieces of the chunk whose entries are in sE[buf]
#pragma unroll
            for (int i = 0; i < PER_T; ++i) {
                const int k = sE[buf][s_ent[i]].x;
                const unsigned off = k >= 0 ? s_voff[i] + (unsigned)k * rowBytes : dBytes;
                stage[i] = __builtin_bit_cast(
                    f32x4, __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)off, 0, 0));
            }
        };
        auto commit = [&](int buf) {
#pragma unroll
            for (int i = 0; i < PER_T; ++i) sD[buf][tid + i * NT] = stage[i];
        };
        load_entries(0, 0);
        load_entries(1, 1);
        __syncthreads();
        issue(0);
        commit(0);
        __syncthreads();
        float acc = 0.0f;
        const int j = col0 + tid;
        const unsigned char *colp = nullptr;
        for (int c = 0; c < nChunks; ++c) {
            const int buf = c & 1;
            const bool more = c + 1 < nChunks;
            int2 eNext = make_int2(-1, 0);
            if (more) {
                issue(buf ^ 1);                       // D of chunk c+1
                const int p = p0 + (c + 2) * T + tid;  // entries of chunk c+2
                if (tid < T && c + 2 < nChunks && p < p1) eNext = seg[p];
            }
            const int n = min(T, p1 - (p0 + c * T));
            colp = reinterpret_cast<const unsigned char *>(sD[buf]) + tid * E::bytes;
            for (int e = 0; e < n; ++e) {
                const float a = E::val(sE[buf][e].y);
                float d;
                if constexpr (E::bytes == 4)
                    d = *reinterpret_cast<const float *>(colp + e * ROWB);
                else
                    d = lo_bf16(*reinterpret_cast<const unsigned short *>(colp + e * ROWB));
                acc = __builtin_fmaf(a, d, acc);
            }
            __syncthreads();  // everyone is done with sD[buf] / sE[buf]
            if (more) {
                commit(buf ^ 1);
                if (tid < T) sE[buf][tid] = eNext;
            }
            __syncthreads();
        }
        if (j < N) C[(size_t)row * N + j] = acc;
    }
}

template <typename E, int VEC, int RPW, int WAVES, int TK>
int launch_lds(const gcoo_matrix *h, const int2 *seg, const void *D, int N, float *C,
               const ConvStatus *guard, hipStream_t s)
{
    const int panelRows = WAVES * RPW;
    const int nPanels = (h->M + panelRows - 1) / panelRows;
    const int nColTiles = (N + 64 * VEC - 1) / (64 * VEC);
    const long long nblocks = (long long)nPanels * nColTiles;
    if (nblocks > INT32_MAX)
        return set_error(GCOO_ERR_OVERFLOW, "grid of %lld workgroups too large", nblocks);
    const unsigned long long dBytes = (unsigned long long)h->K * N * E::bytes;
    hipLaunchKernelGGL((spdm_lds_kernel<E, VEC, RPW, WAVES, TK>), dim3((unsigned)nblocks),
                       dim3(WAVES * 64), 0, s, h->rowptr, seg, h->M, h->K, D,
                       (unsigned)dBytes, N, C, nPanels, guard, (long long)h->cap);
    GCOO_CHECK_LAUNCH("spdm_lds_kernel");
    return GCOO_OK;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

template <typename E, int TK, int WAVES, int RPW>
int launch_panel(const gcoo_matrix *h, const int2 *seg, const void *D, int N, float *C,
                 const ConvStatus *guard, hipStream_t s, const ExtScratch *xs)
{
    constexpr int ROWB = 256 * E::bytes;  // bytes of a staged D row (256 columns)
    constexpr int PR = WAVES * RPW;       // panel rows
    const int *poff = h->poff;
    const uint2 *plist = h->plist;
    const int *rowmap = nullptr;
    if (xs) {
        // lists rebuilt into the caller's scratch: offsets, then the entries
        // (plus the slack the kernel's batched entry loads may read past a
        // list's end)
        const size_t offB = align256(sizeof(int) * panel_offset_elems(h, TK, RPW, WAVES));
        const size_t listB = sizeof(uint2) * ((size_t)h->nnz + kSegSlack);
        if (offB + listB > xs->bytes)
            return set_error(GCOO_ERR_BUDGET,
                             "panel lists of %lld entries need %zu bytes of scratch, the "
                             "no-allocation entry point has %zu (variant 2 needs none)",
                             (long long)h->nnz, offB + listB, xs->bytes);
        int *po = reinterpret_cast<int *>(xs->base);
        uint2 *pl = reinterpret_cast<uint2 *>(xs->base + offB);
        int rc = build_panel_lists(h, seg, TK, ROWB, RPW, WAVES, E::bytes == 2, po, pl, guard, s);
        if (rc) return rc;
        poff = po;
        plist = pl;
    } else {
        // the balanced row order (gcoo_panel.hip) costs a kernel per list
        // build: only when asked for (mode 1: where the lists outlive the
        // call, not right behind a conversion in the same call; 2: always)
        const bool balance = h->panel_balance == 2 || (h->panel_balance == 1 && !guard);
        int rc = ensure_panel_lists(h, seg, TK, ROWB, RPW, WAVES, E::bytes == 2, guard, s, balance);
        if (rc) return rc;
        poff = h->poff;
        plist = h->plist;
        rowmap = panel_rowmap(h);
    }
    const unsigned long long dBytes = (unsigned long long)h->K * N * E::bytes;
    const int nPanels = (h->M + PR - 1) / PR;
    const int nC = (h->K + TK - 1) / TK;
    const int tiles = (N + 255) / 256;
    const long long nblocks = (long long)nPanels * tiles;
    hipLaunchKernelGGL((spdm_panel_kernel<E, TK, WAVES, RPW>), dim3((unsigned)nblocks),
                       dim3(WAVES * 64), 0, s,
                       poff, plist, h->M, h->K, nC, D, (unsigned)dBytes, N, C, nPanels,
                       guard, (long long)h->cap, rowmap);
    GCOO_CHECK_LAUNCH("spdm_panel_kernel");
    return GCOO_OK;
}

// Threshold of the long-rows-first order of variant 2 (entries; 0: off).
// Configured as 2^waves (waves 2..16, 1: off), else automatic: on when the
// last conversion's longest row holds at least 16 times the mean and 512
// entries, with the power of two at or above 16 times the mean (at least
// 256).  On the 65536^2 Zipf matrix (mean 16, longest 1334 rows, 476 rows
// of 512 or more holding 37 % of the entries) the threshold 512 took 146
// against 164 us (FP32, 1 column per lane, rounds of 16).
int long_first_threshold(const gcoo_matrix *h)
{
    const int w = h->spdm_cfg.waves;
    if (w == 1) return 0;
    if (w >= 2) return 1 << w;
    if (h->M <= 0 || h->nnz <= 0) return 0;
    const double mean = (double)h->nnz / h->M;
    if (h->maxRow < 512 || h->maxRow < 16.0 * mean) return 0;
    int t = 256;
    while (t < 16.0 * mean && t < (1 << 30)) t <<= 1;
    return t;
}

template <typename E, int VEC, int U>
int launch_rowseg_u(const gcoo_matrix *h, const int2 *seg, const void *D, int N, float *C,
                    const ConvStatus *guard, hipStream_t s, int skipLen, const ExtScratch *xs)
{
    // A wave walks `rpw` consecutive rows: one row, or the group's cpg rows
    // (the segments are per row, so any split gives the same bits).  One row
    // per wave shortens each wave's serial chain when rows are long enough
    // to keep its loads busy: at n = 8000, 370 -> 347 us at s = 0.997 and
    // 591 -> 564 us (BF16) at s = 0.99, 23 -> 18 us at 800^2, s = 0.9; with
    // 8 entries per row (s = 0.999) the group per wave is faster (150
    // against 161 us).  Automatic: one row once rows average 12 entries
    // (the handle's last conversion); config chunk 1 forces one row, another
    // nonzero chunk the group.
    const gcoo_spdm_config &cfg = h->spdm_cfg;
    const bool longRows = h->M > 0 && h->nnz >= 12ll * h->M;
    const int rpw = (cfg.chunk == 1 || (cfg.chunk == 0 && longRows)) ? 1 : h->cpg;
    const int nUnits = (h->M + rpw - 1) / rpw;
    const int nGroupBlocks = (nUnits + kSpdmWaves - 1) / kSpdmWaves;
    const int tileCols = 64 * VEC;
    const int nColTiles = (N + tileCols - 1) / tileCols;
    // long rows first: listed on the device (rows with >= lf entries), in
    // the handle's long-row scratch or, for the no-allocation entry point,
    // in xs (too little of it: the natural order, same bits)
    const int lf = skipLen == INT_MAX ? long_first_threshold(h) : 0;
    int longBlocks = 0;
    int *longList = nullptr;
    unsigned *longCtr = nullptr;
    if (lf > 0) {
        const size_t listB = align256(sizeof(int) * ((size_t)h->M + 1));
        if (xs) {
            if (listB + 256 <= xs->bytes) {
                longList = reinterpret_cast<int *>(xs->base);
                longCtr = reinterpret_cast<unsigned *>(xs->base + listB);
            }
        } else {
            if (!h->longList) {
                if (hipMalloc(&h->longList, sizeof(int) * ((size_t)h->M + 1)) != hipSuccess ||
                    hipMalloc(&h->longCtr, 256) != hipSuccess) {
                    (void)hipGetLastError();
                    if (h->longList) (void)hipFree(h->longList);
                    h->longList = nullptr;
                    return set_error(GCOO_ERR_ALLOC, "long-row scratch");
                }
            }
            longList = h->longList;
            longCtr = h->longCtr;
        }
    }
    if (longList) {
        const long long maxLong = std::min<long long>(h->M, h->nnz / lf);
        longBlocks = (int)((maxLong + kSpdmWaves - 1) / kSpdmWaves);
        hipLaunchKernelGGL(longrow_reset_kernel, dim3(1), dim3(64), 0, s, longCtr);
        GCOO_CHECK_LAUNCH("longrow_reset_kernel");
        hipLaunchKernelGGL(longrow_classify_kernel, dim3((h->M + 255) / 256), dim3(256), 0, s,
                           h->rowptr, h->M, lf, longList, longCtr, guard, (long long)h->cap);
        GCOO_CHECK_LAUNCH("longrow_classify_kernel");
        skipLen = lf;
    }
    const int xpt = (longBlocks > 0 && nColTiles <= kNumXcd && kNumXcd % nColTiles == 0)
                        ? kNumXcd / nColTiles : 0;
    const long long nblocks =
        xpt ? (long long)kNumXcd * ((longBlocks + xpt - 1) / xpt + (nGroupBlocks + xpt - 1) / xpt)
            : ((long long)longBlocks + nGroupBlocks) * nColTiles;
    if (nblocks > INT32_MAX)
        return set_error(GCOO_ERR_OVERFLOW, "grid of %lld workgroups too large", nblocks);
    const unsigned long long dBytes = (unsigned long long)h->K * N * E::bytes;
    hipLaunchKernelGGL((spdm_rowseg_kernel<E, VEC, U>), dim3((unsigned)nblocks),
                       dim3(kSpdmThreads), 0, s, h->rowptr, seg, nUnits, rpw, h->M, D,
                       (unsigned)dBytes, N, C, nGroupBlocks, guard, (long long)h->cap, skipLen,
                       (const int *)longList, (const unsigned *)longCtr, longBlocks, xpt);
    GCOO_CHECK_LAUNCH("spdm_rowseg_kernel");
    return GCOO_OK;
}

// unroll (entries in flight per wave): 8, 16 or 32, then rounds of 8, then
// one guarded round of the last 1..7 entries.  16 by default: on the 65536^2
// Zipf matrix (rows up to 4096 entries) 161 against 203 us (FP32, vec 1),
// within +-5 us of 8 at n = 8000 for s = 0.99 .. 0.999.
template <typename E, int VEC>
int launch_rowseg(const gcoo_matrix *h, const int2 *seg, const void *D, int N, float *C,
                  const ConvStatus *guard, hipStream_t s, int skipLen = INT_MAX,
                  const ExtScratch *xs = nullptr)
{
    const int u = h->spdm_cfg.unroll ? h->spdm_cfg.unroll : 16;
    if (u <= 8) return launch_rowseg_u<E, VEC, 8>(h, seg, D, N, C, guard, s, skipLen, xs);
    if (u <= 16 || VEC * E::bytes > 8)
        return launch_rowseg_u<E, VEC, 16>(h, seg, D, N, C, guard, s, skipLen, xs);
    return launch_rowseg_u<E, VEC, 32>(h, seg, D, N, C, guard, s, skipLen, xs);
}

// Rows with at least minLen entries through the long-row kernel (listed on
// the device, no host synchronisation), the rest through the row-segment
// kernel.  Scratch (the row list and its counters) lives in the handle.
template <typename E, int VEC>
int launch_split(const gcoo_matrix *h, const int2 *seg, const void *D, int N, float *C,
                 const ConvStatus *guard, hipStream_t s, int minLen, const ExtScratch *xs)
{
    int *longList = h->longList;
    unsigned *longCtr = h->longCtr;
    if (xs) {
        const size_t listB = align256(sizeof(int) * ((size_t)h->M + 1));
        if (listB + 256 > xs->bytes)
            return set_error(GCOO_ERR_BUDGET,
                             "the long-row list of %d rows needs %zu bytes of scratch, the "
                             "no-allocation entry point has %zu", h->M, listB + 256, xs->bytes);
        longList = reinterpret_cast<int *>(xs->base);
        longCtr = reinterpret_cast<unsigned *>(xs->base + listB);
    } else if (!h->longList) {
        if (hipMalloc(&h->longList, sizeof(int) * ((size_t)h->M + 1)) != hipSuccess ||
            hipMalloc(&h->longCtr, 256) != hipSuccess) {
            (void)hipGetLastError();
            if (h->longList) (void)hipFree(h->longList);
            h->longList = nullptr;
            return set_error(GCOO_ERR_ALLOC, "long-row scratch");
        }
        longList = h->longList;
        longCtr = h->longCtr;
    }
    constexpr int W = 4, T = 32;
    const int colTiles = (N + W * 64 - 1) / (W * 64);
    const unsigned long long dBytes = (unsigned long long)h->K * N * E::bytes;
    hipLaunchKernelGGL(longrow_reset_kernel, dim3(1), dim3(64), 0, s, longCtr);
    GCOO_CHECK_LAUNCH("longrow_reset_kernel");
    hipLaunchKernelGGL(longrow_classify_kernel, dim3((h->M + 255) / 256), dim3(256), 0, s,
                       h->rowptr, h->M, minLen, longList, longCtr, guard,
                       (long long)h->cap);
    GCOO_CHECK_LAUNCH("longrow_classify_kernel");
    int dev = 0, cus = 256, perCu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, spdm_longrow_kernel<E, W, T>,
                                                     W * 64, 0) != hipSuccess || perCu < 1) {
        (void)hipGetLastError();
        perCu = 1;
    }
    hipLaunchKernelGGL((spdm_longrow_kernel<E, W, T>), dim3(cus * perCu), dim3(W * 64), 0, s,
                       h->rowptr, seg, longList, longCtr, colTiles, D, (unsigned)dBytes, N,
                       C);
    GCOO_CHECK_LAUNCH("spdm_longrow_kernel");
    return launch_rowseg<E, VEC>(h, seg, D, N, C, guard, s, minLen);
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
int launch_group_order(const gcoo_matrix *h, int vec, const float *D, int N, float *C,
                       const ConvStatus *guard, hipStream_t s)
{
    if (vec >= 4) return launch_wave<CPG, 4>(h, D, N, C, guard, s);
    if (vec == 2) return launch_wave<CPG, 2>(h, D, N, C, guard, s);
    return launch_wave<CPG, 1>(h, D, N, C, guard, s);
}

// Widest vector of `elemBytes`-byte elements (at most 16 bytes per lane) that
// N and the alignment of D and C allow.
int max_vec(const void *D, int N, const float *C, int elemBytes)
{
    const uintptr_t aD = reinterpret_cast<uintptr_t>(D);
    const uintptr_t aC = reinterpret_cast<uintptr_t>(C);
    for (int v = 16 / elemBytes; v > 1; v >>= 1) {
        const unsigned dAl = (unsigned)(v * elemBytes) - 1;
        const unsigned cAl = (unsigned)(v * 4 > 16 ? 16 : v * 4) - 1;
        if (N % v == 0 && (aD & dAl) == 0 && (aC & cAl) == 0) return v;
    }
    return 1;
}

// Shape (waves x rows per wave) of the panel kernel's workgroup: the
// configured one -- `waves` 16, 12 or 8 (8 for FP32 only) with `chunk` rows
// per wave (16 waves: 16 or 21; 12: 28; 8: 42; 0: the wave count's first) --
// or automatically 16 x 16 or 16 x 21, whichever the estimate below puts
// lower.  A workgroup holds a CU's LDS, so the grid runs in rounds of one
// workgroup per CU; a workgroup's time is about its panel's entries plus its
// staging of the K x 256 D tile, which does not depend on the panel's rows.
// Taller panels share each staged chunk among more rows (336 against 256
// rows: a quarter less staging); 16 x 21 keeps 4 waves per SIMD (128 VGPRs),
// where 12 x 28 has 3 and hides less of the entry loop's latencies (n =
// 8000, FP32: 414 / 456 / 472 us for 16 x 21 / 16 x 16 / 12 x 28 at 1 %
// density, 346 / 408 / 382 at 0.7 %, 319 / 386 / 340 at 0.5 %; BF16 16 x 21
// / 12 x 28: 460 / 492 us at 1 %).
struct PanelShape {
    int waves, rows;
};
template <typename E>
PanelShape panel_shape(const gcoo_matrix *h, int N)
{
    const gcoo_spdm_config &cfg = h->spdm_cfg;
    const int r = cfg.chunk;
    if (cfg.waves == 16 || (cfg.waves == 0 && (r == 16 || r == 21)))
        return {16, r == 21 ? 21 : 16};
    if (cfg.waves == 12 || (cfg.waves == 0 && r == 28)) return {12, 28};
    if ((cfg.waves == 8 || (cfg.waves == 0 && r == 42)) && E::bytes == 4) return {8, 42};
    // rounds of workgroups x (entries of the panel, in rows at 1 % density,
    // plus the staging, counted as 256 such rows)
    const double cells = (double)h->M * (double)h->K;
    const double d100 = cells > 0 ? 100.0 * (double)h->nnz / cells : 0.0;
    const long long tiles = (N + 255) / 256;
    auto cost = [&](int pr) {
        const long long wgs = (long long)((h->M + pr - 1) / pr) * tiles;
        return (double)((wgs + 255) / 256) * (pr * d100 + 256.0);
    };
    return cost(336) <= cost(256) ? PanelShape{16, 21} : PanelShape{16, 16};
}

// The variant (2..6) and columns per lane that a launch of h on these
// operands runs: the handle's configuration, or the automatic choice, after
// the fall-backs each variant's operand requirements impose.
template <typename E>
void resolve_variant(const gcoo_matrix *h, const void *D, int N, const float *C, int *variantOut,
                     int *vecOut)
{
    const gcoo_spdm_config &cfg = h->spdm_cfg;
    const int vmax = max_vec(D, N, C, E::bytes);
    const int vdef = E::bytes == 4 ? 2 : 4;
    int vec = cfg.vec ? cfg.vec : vdef;
    if (vec > vmax) vec = vmax;
    int variant = cfg.variant;
    if (variant == 0) {
        // Panel staging costs a fixed K x N read of D per panel of rows, row
        // segments cost one D-row gather per entry: the panel kernel wins from
        // about 0.25 % density on (n = 8000, FP32, 336-row panels: 311
        // against 351 us at s = 0.997, 316 against 248 us at 0.998).  BF16
        // gathers move half the bytes and the BF16 panel kernel widens every
        // D value (4 more vector instructions per entry), so for BF16 it only
        // wins from about 0.5 % (16 x 21 panels: 460 against 565 us at 1 %,
        // 365 against 400 us at 0.7 %, 287 against 292 us at 0.5 %, 212
        // against 194 us at 0.3 %).  The density is the handle's last completed conversion (a
        // first optimistic launch does not know the new count yet); every
        // variant gives the same bits.  It also needs a workgroup per CU at
        // least (256 x 256 tiles of C).
        const double cells = (double)h->M * (double)h->K;
        const double minDensity = E::bytes == 4 ? kPanelMinDensity : kPanelMinDensityBf16;
        const bool dense = cells > 0 && (double)h->nnz >= minDensity * cells;
        const PanelShape ps = panel_shape<E>(h, N);
        const int pr = ps.waves * ps.rows;
        const long long wgs = (long long)((h->M + pr - 1) / pr) * ((N + 255) / 256);
        variant = (dense && wgs >= kPanelMinWorkgroups) ? 6 : 2;
    }
    if (variant == 1) variant = 2;
    // variants 3, 4 and 6 stage 16-byte pieces of D rows: they need whole
    // pieces per row (N * bytes % 16 == 0) and a 16-byte aligned D; variant
    // 6 also stores 4 columns per lane (vec 4 allowed)
    const bool pieces =
        (N * E::bytes) % 16 == 0 && (reinterpret_cast<uintptr_t>(D) & 15u) == 0;
    if ((variant == 3 || variant == 4) && !pieces) variant = 2;
    if (variant == 6 && !(pieces && vmax >= 4)) variant = 2;
    if (variant == 6) vec = 4;  // 256 columns per workgroup, 4 per lane
    if (variant == 5) vec = vmax;
    *variantOut = variant;
    *vecOut = vec;
}

template <typename E>
int launch_typed(const gcoo_matrix *h, const int2 *seg, const void *D, int N, float *C,
                 const ConvStatus *guard, hipStream_t s, const ExtScratch *xs = nullptr)
{
    const gcoo_spdm_config &cfg = h->spdm_cfg;
    const unsigned long long dBytes = (unsigned long long)h->K * N * E::bytes;
    if (dBytes >= 0x7fffffffull)
        return set_error(GCOO_ERR_UNSUPPORTED,
                         "dense operand of %llu bytes: the buffer-addressed kernels need < 2 GiB",
                         dBytes);
    int variant = 2, vec = 1;
    resolve_variant<E>(h, D, N, C, &variant, &vec);
    if (variant == 3) {
        const int waves = cfg.waves ? cfg.waves : 16;
        if constexpr (E::bytes == 4) {
            if (vec >= 4) {
                if (waves <= 8) return launch_lds<E, 4, 16, 8, 64>(h, seg, D, N, C, guard, s);
                return launch_lds<E, 4, 16, 16, 64>(h, seg, D, N, C, guard, s);
            }
            if (vec == 2) return launch_lds<E, 2, 16, 16, 128>(h, seg, D, N, C, guard, s);
            return launch_lds<E, 1, 32, 16, 256>(h, seg, D, N, C, guard, s);
        } else {
            if (vec >= 8) return launch_lds<E, 8, 8, 16, 64>(h, seg, D, N, C, guard, s);
            if (vec == 4) return launch_lds<E, 4, 16, 16, 128>(h, seg, D, N, C, guard, s);
            return launch_lds<E, 2, 16, 16, 256>(h, seg, D, N, C, guard, s);
        }
    }
    if (variant == 6) {
        constexpr int TK = E::bytes == 4 ? 75 : 150;
        const PanelShape ps = panel_shape<E>(h, N);
        if (ps.rows == 21) return launch_panel<E, TK, 16, 21>(h, seg, D, N, C, guard, s, xs);
        if (ps.rows == 28) return launch_panel<E, TK, 12, 28>(h, seg, D, N, C, guard, s, xs);
        if constexpr (E::bytes == 4)
            if (ps.rows == 42) return launch_panel<E, TK, 8, 42>(h, seg, D, N, C, guard, s, xs);
        return launch_panel<E, TK, 16, 16>(h, seg, D, N, C, guard, s, xs);
    }
    if (variant == 5) {
        // widest 16-byte-per-lane vector the operands allow
        if constexpr (E::bytes == 4) {
            if (vec >= 4) return launch_grouprows<E, 4>(h, seg, D, N, C, guard, s);
            if (vec == 2) return launch_grouprows<E, 2>(h, seg, D, N, C, guard, s);
            return launch_grouprows<E, 1>(h, seg, D, N, C, guard, s);
        } else {
            if (vec >= 8) return launch_grouprows<E, 8>(h, seg, D, N, C, guard, s);
            if (vec == 4) return launch_grouprows<E, 4>(h, seg, D, N, C, guard, s);
            if (vec == 2) return launch_grouprows<E, 2>(h, seg, D, N, C, guard, s);
            return launch_grouprows<E, 1>(h, seg, D, N, C, guard, s);
        }
    }
    // Long rows through the long-row kernel (variant 4, chosen explicitly or
    // by the autotuner's table; measured slower than variant 2 on the Zipf
    // workload so far, so never chosen automatically).
    if (variant == 4) {
        const int minLen = cfg.chunk > 0 ? cfg.chunk : 256;
        if constexpr (E::bytes == 4) {
            if (vec >= 4) return launch_split<E, 4>(h, seg, D, N, C, guard, s, minLen, xs);
            if (vec == 2) return launch_split<E, 2>(h, seg, D, N, C, guard, s, minLen, xs);
            return launch_split<E, 1>(h, seg, D, N, C, guard, s, minLen, xs);
        } else {
            if (vec >= 8) return launch_split<E, 8>(h, seg, D, N, C, guard, s, minLen, xs);
            if (vec == 4) return launch_split<E, 4>(h, seg, D, N, C, guard, s, minLen, xs);
            if (vec == 2) return launch_split<E, 2>(h, seg, D, N, C, guard, s, minLen, xs);
            return launch_split<E, 1>(h, seg, D, N, C, guard, s, minLen, xs);
        }
    }
    if constexpr (E::bytes == 4) {
        if (vec >= 4) return launch_rowseg<E, 4>(h, seg, D, N, C, guard, s, INT_MAX, xs);
        if (vec == 2) return launch_rowseg<E, 2>(h, seg, D, N, C, guard, s, INT_MAX, xs);
        return launch_rowseg<E, 1>(h, seg, D, N, C, guard, s, INT_MAX, xs);
    } else {
        if (vec >= 8) return launch_rowseg<E, 8>(h, seg, D, N, C, guard, s, INT_MAX, xs);
        if (vec == 4) return launch_rowseg<E, 4>(h, seg, D, N, C, guard, s, INT_MAX, xs);
        if (vec == 2) return launch_rowseg<E, 2>(h, seg, D, N, C, guard, s, INT_MAX, xs);
        return launch_rowseg<E, 1>(h, seg, D, N, C, guard, s, INT_MAX, xs);
    }
}

}  // namespace

bool panel_operands_ok(const void *D, int N, const float *C, int elemBytes)
{
    return (N * elemBytes) % 16 == 0 && (reinterpret_cast<uintptr_t>(D) & 15u) == 0 &&
           max_vec(D, N, C, elemBytes) >= 4;
}

int launch_panel_batch(bool bf16, const PanelUnit *units, int nUnits, const PanelJob *jobs,
                       hipStream_t s)
{
    static_assert(kBatchPanelWaves == 16 && kBatchPanelRpw == 21, "batch panel shape");
    if (nUnits <= 0) return GCOO_OK;
    if (bf16)
        hipLaunchKernelGGL((spdm_panel_batch_kernel<BF16, 150, 16, 21>), dim3((unsigned)nUnits),
                           dim3(16 * 64), 0, s, units, jobs);
    else
        hipLaunchKernelGGL((spdm_panel_batch_kernel<F32, 75, 16, 21>), dim3((unsigned)nUnits),
                           dim3(16 * 64), 0, s, units, jobs);
    GCOO_CHECK_LAUNCH("spdm_panel_batch_kernel");
    return GCOO_OK;
}

int launch_spdm_f32_cfg(const gcoo_matrix *h, const float *D, int N, float *C,
                        const ConvStatus *guard, hipStream_t s, const ExtScratch *xs);

int launch_spdm_f32_ext(const gcoo_matrix *h, const float *D, int N, float *C,
                        const ExtScratch *xs, hipStream_t s)
{
    return launch_spdm_f32_cfg(h, D, N, C, nullptr, s, xs);
}

int launch_spdm_f32(const gcoo_matrix *h, const float *D, int N, float *C,
                    const ConvStatus *guard, hipStream_t s)
{
    return launch_spdm_f32_cfg(h, D, N, C, guard, s, nullptr);
}

int launch_spdm_f32_cfg(const gcoo_matrix *h, const float *D, int N, float *C,
                        const ConvStatus *guard, hipStream_t s, const ExtScratch *xs)
{
    if (h->nGroup == 0 || N == 0) return GCOO_OK;
    const gcoo_spdm_config &cfg = h->spdm_cfg;
    const unsigned long long dBytes = (unsigned long long)h->K * N * 4;
    // Variant 1 (group order, 64-bit addressing) is also the path for dense
    // operands of 2 GiB or more.
    if (cfg.variant == 1 || dBytes >= 0x7fffffffull) {
        int vec = cfg.vec ? cfg.vec : 4;
        const int vmax = max_vec(D, N, C, 4);
        if (vec > vmax) vec = vmax;
        switch (h->cpg) {
        case 1: return launch_group_order<1>(h, vec, D, N, C, guard, s);
        case 2: return launch_group_order<2>(h, vec, D, N, C, guard, s);
        case 4: return launch_group_order<4>(h, vec, D, N, C, guard, s);
        default: return set_error(GCOO_ERR_INVALID, "cpg %d not supported", h->cpg);
        }
    }
    return launch_typed<F32>(h, h->seg, D, N, C, guard, s, xs);
}

int spdm_choice(const gcoo_matrix *h, int dtype, const void *D, int N, const float *C,
                int *variant, int *vec, int *waves, int *rows)
{
    *waves = 0;
    *rows = 0;
    if (dtype == 0) {
        const unsigned long long dBytes = (unsigned long long)h->K * N * 4;
        if (h->spdm_cfg.variant == 1 || dBytes >= 0x7fffffffull) {
            const int vmax = max_vec(D, N, C, 4);
            const int v = h->spdm_cfg.vec ? h->spdm_cfg.vec : 4;
            *variant = 1;
            *vec = v > vmax ? vmax : v;
            return GCOO_OK;
        }
        resolve_variant<F32>(h, D, N, C, variant, vec);
        if (*variant == 6) {
            const PanelShape ps = panel_shape<F32>(h, N);
            *waves = ps.waves;
            *rows = ps.rows;
        }
    } else {
        resolve_variant<BF16>(h, D, N, C, variant, vec);
        if (*variant == 6) {
            const PanelShape ps = panel_shape<BF16>(h, N);
            *waves = ps.waves;
            *rows = ps.rows;
        }
    }
    return GCOO_OK;
}

int launch_spdm_bf16(const gcoo_matrix *h, const uint16_t *D, int N, float *C, hipStream_t s)
{
    if (h->nGroup == 0 || N == 0) return GCOO_OK;
    return launch_typed<BF16>(h, h->seg16, D, N, C, nullptr, s);
}

}  // namespace gcoo
