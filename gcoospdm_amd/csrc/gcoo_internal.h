// gcoo_internal.h -- shared declarations of the HIP translation units of
// libgcoospdm.so.  Not part of the public ABI (that is include/gcoo_spdm.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/gcoo_spdm.h"

namespace gcoo {

// Native clang vectors (HIP's float4 wrapper is not accepted by the
// nontemporal builtins).
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Thread-local last-error message and status helpers.
int set_error(int code, const char *fmt, ...);
void clear_error();

#define GCOO_HIP_TRY(expr)                                                    \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess)                                                 \
            return ::gcoo::set_error(GCOO_ERR_HIP, "%s failed: %s (%s:%d)",   \
                                     #expr, hipGetErrorString(e_), __FILE__,  \
                                     __LINE__);                               \
    } while (0)

#define GCOO_CHECK_LAUNCH(name)                                               \
    do {                                                                      \
        hipError_t e_ = hipGetLastError();                                    \
        if (e_ != hipSuccess)                                                 \
            return ::gcoo::set_error(GCOO_ERR_HIP, "launch of %s failed: %s", \
                                     name, hipGetErrorString(e_));            \
    } while (0)

// Device-side status word of a conversion: total nnz (64-bit, written by the
// scan), an overflow flag (total exceeds int32) and an invalid-input flag
// (COO entries out of range or duplicated).
struct ConvStatus {
    long long total;
    int overflow;
    int invalid;
    int maxRow;  // most entries in one row
    // Abort word of the single-pass conversion: set by a group whose bounded
    // look-back wait gave up (the offsets of that conversion are then not
    // consistent), read by the other groups' waits, cleared only by the host
    // (which then converts again with the count / scan / scatter kernels).
    int stall;
    int longRows;  // rows listed by the row scan as longer than its threshold
};

// True when a kernel queued behind a conversion must leave the entry arrays
// alone: the entries (plus `extra`) outgrew the capacity, or the single-pass
// conversion stalled.  A null guard never skips.
__device__ __forceinline__ bool guard_skip(const ConvStatus *g, long long cap,
                                           long long extra = 0)
{
    return g != nullptr && (g->total + extra > cap || g->stall != 0);
}

}  // namespace gcoo

// The handle.  Array layout mirrors the reference (src/ShGCOOSpDM.cu:33-35,
// 94-97): rows and cols share one allocation, groupIndex and nnzPerGroup
// share another.  Capacity is kept across conversions so a repeated
// conversion of same-sized inputs allocates nothing.
struct gcoo_matrix {
    int M = 0, K = 0, cpg = 4, nGroup = 0;
    int64_t nnz = 0;
    bool converted = false;
    int device = 0;
    // Bumped by every conversion (and so by every reallocation of the entry
    // arrays): a batch records it per job and refuses to run once it moved.
    uint64_t generation = 0;

    // recorded behind the status word's copy to the host, so that a
    // conversion can be checked while the multiply queued behind it runs
    hipEvent_t statusEvent = nullptr;
    float *vals = nullptr;     // [cap]
    int *rowcol = nullptr;     // [2 * cap]: rows at 0, cols at cap
    int *idx = nullptr;        // [2 * (nGroup + 1)]: groupIndex, nnzPerGroup
    int64_t cap = 0;           // entry capacity of vals / rowcol / seg

    // Row segments: the same entries grouped by row, k ascending within a
    // row, as (k, bits of val) pairs; row r's segment is
    // seg[rowptr[r] .. rowptr[r+1]).  Group g's segments are contiguous and
    // span exactly [groupIndex[g], groupIndex[g+1]).
    int2 *seg = nullptr;       // [cap]
    // BF16 storage form of the same segments (value bits in the low 16 bits
    // of .y); allocated on first use, rebuilt after every conversion.
    int2 *seg16 = nullptr;     // [cap]
    bool seg16_valid = false;
    int *rowptr = nullptr;     // [M + 1]
    int *rowcnt = nullptr;     // [M] per-row counts (conversion scratch)
    int maxRow = 0;            // most entries in one row (last conversion)

    // Nonzero mask of the last dense conversion (16-byte path only), written
    // by the count pass and read by the scatter so that it loads only the
    // pieces of S that hold an entry.  Per row and 256-column block, four
    // 64-bit words: bit l of word c is column block * 256 + 4 * l + c.
    unsigned long long *mask = nullptr;  // [M * maskBlocks * 4]
    size_t mask_bytes = 0;
    bool mask_valid = false;   // written for the S of the current conversion

    // long-row path scratch, allocated on first use by the multiply
    mutable int *longList = nullptr;       // [M + 1] rows of the long-row kernel
    mutable unsigned *longCtr = nullptr;   // [0] listed rows, [1] next work item

    // Panel lists of the LDS-staged multiply (variant 6), built on first use
    // after a conversion: the entries of each (row panel, k chunk) pair,
    // ordered by (row, k), as ((LDS offset of the D row << 8) | row slot,
    // value bits), and their start offsets (see gcoo_panel.hip).
    mutable uint2 *plist = nullptr;        // [cap]
    mutable int *poff = nullptr;           // [poff_elems]
    mutable size_t poff_elems = 0;
    mutable const int2 *plist_src = nullptr;  // segments the lists hold (null: stale)
    mutable int plist_tk = 0;                 // chunk length they were built for
    mutable int plist_rowb = 0;               // bytes of a staged D row they encode
    mutable int plist_rpw = 0;                // rows per wave they were built for
    mutable int plist_waves = 0;              // waves per panel they were built for
    // Bumped whenever the lists are rebuilt or forgotten: a batch that runs
    // panel jobs from these lists records it and refuses to run once it moved.
    mutable uint64_t plist_gen = 0;
    // Balanced row order of the lists' panels (see gcoo_panel.hip): when
    // plist_balanced, slot q of wave w of panel p is row prow[p * PR + w *
    // rpw + q] (-1: none) instead of p * PR + w * rpw + q.
    mutable int *prow = nullptr;
    mutable size_t prow_elems = 0;
    mutable bool plist_balanced = false;
    int panel_balance = 0;  // gcoo_set_panel_balance: balanced lists for this handle's multiplies
    // Set by a conversion with a panel multiply behind it: the panel-list
    // kernel copies the status word to this device view of h_status (so no
    // separate copy sits on the multiply's path) and the lists' launch
    // records statusEvent; `statusExported` tells the conversion it happened.
    mutable gcoo::ConvStatus *statusExport = nullptr;
    mutable bool statusExported = false;
    gcoo::ConvStatus *h_status_dev = nullptr;  // device view of h_status

    // single-pass conversion: per-group look-back flags (two (state, epoch,
    // value) words per group, never cleared: each conversion has its own
    // epoch); its abort word is d_status->stall
    unsigned long long *lb_flags = nullptr;  // [2 * nGroup]
    unsigned lb_epoch = 0;

    // multiply-kernel configuration (gcoo_set_spdm_config)
    gcoo_spdm_config spdm_cfg = {0, 0, 0, 0, 0};

    gcoo::ConvStatus *d_status = nullptr;  // device
    gcoo::ConvStatus *h_status = nullptr;  // pinned host mirror

    // scratch for the COO path, grown on demand: the rows the scan lists as
    // long, and the row of every entry of the row segments
    void *coo_tmp = nullptr;
    size_t coo_tmp_bytes = 0;
    int *coo_rowof = nullptr;
    size_t coo_rowof_elems = 0;

    int *rows() const { return rowcol; }
    int *cols() const { return rowcol + cap; }
    int *groupIndex() const { return idx; }
    int *nnzPerGroup() const { return idx + (nGroup + 1); }
};

namespace gcoo {

// ---- conversion (gcoo_convert.hip) ----
// Enqueue count + scan into h->idx and h->d_status.
int launch_count_and_scan(gcoo_matrix *h, const float *S, hipStream_t s);
// Count, scan and scatter in one kernel that reads S once (16-byte path,
// cpg 4); fused_conversion_ok says whether it applies (and allocates its
// look-back state on first use).  Groups whose entries do not fit the
// capacity skip their writes; the status word carries the total as after
// launch_count_and_scan.
bool fused_conversion_ok(gcoo_matrix *h, const float *S);
int launch_convert_fused(gcoo_matrix *h, const float *S, hipStream_t s);
// Enqueue the row-pointer scan alone (rowcnt -> rowptr, groupIndex, ...);
// with longList (M ints) also the rows of more than longMin entries.
int launch_row_scan(gcoo_matrix *h, hipStream_t s, bool clear_invalid, int *longList = nullptr,
                    int longMin = 0);
// Enqueue the scatter.  Writes nothing when d_status reports overflow.
int launch_scatter(gcoo_matrix *h, const float *S, hipStream_t s);
// Grow entry capacity to at least need (discarding contents).
int ensure_capacity(gcoo_matrix *h, int64_t need);
// COO path (gcoo_coo.hip).
int convert_coo(gcoo_matrix *h, const int *rows, const int *cols,
                const float *vals, int64_t nnz, hipStream_t s);

// ---- SpDM (gcoo_spdm.hip) ----
// guard: when non-null, the kernel exits if guard->total > h->cap.
int launch_spdm_f32(const gcoo_matrix *h, const float *D, int N, float *C,
                    const ConvStatus *guard, hipStream_t s);
// The multiply variant (1..6), columns per lane and (variant 6) waves per
// workgroup a launch of h on these operands runs (dtype 0 FP32, 1 BF16).
int spdm_choice(const gcoo_matrix *h, int dtype, const void *D, int N, const float *C,
                int *variant, int *vec, int *waves, int *rows);
// BF16 storage form: h->seg16 must hold the BF16 entry values.
int launch_spdm_bf16(const gcoo_matrix *h, const uint16_t *D, int N, float *C,
                     hipStream_t s);

// Enqueue (on s) the BF16 copy of h's segments unless it is current; the
// copy is usable by later work on s (or by anyone after s is synchronised).
int ensure_seg16(gcoo_matrix *h, hipStream_t s);

// ---- panel lists of the LDS-staged multiply (gcoo_panel.hip) ----
// A panel is waves * rpw consecutive rows; wave w of the multiply owns panel
// rows [w * rpw, (w + 1) * rpw).
// Enqueue (on s) the panel lists of h's entries from seg (h->seg or
// h->seg16) for k chunks of tk rows staged at rowb bytes per D row, unless
// they are current (with widen, the list values are BF16 segment values
// widened to FP32).  Wave w of panel p has nC + 1 list offsets at
// poff[(p * waves + w) * (nC + 1)]: its chunk c entries are
// plist[off[c] .. off[c + 1]), ordered by (row, k).  Kernels exit early when
// guard_skip(guard, h->cap) holds.
int ensure_panel_lists(const gcoo_matrix *h, const int2 *seg, int tk, int rowb, int rpw,
                       int waves, bool widen, const ConvStatus *guard, hipStream_t s,
                       bool balance = false);
// The row order of the lists' panels for the multiply's C stores: h->prow
// when the current lists are balanced, else null (rows in order).
inline const int *panel_rowmap(const gcoo_matrix *h) { return h->plist_balanced ? h->prow : nullptr; }
// True when the balanced order is available for these lists (at most 128
// chunks, 512 panel rows, 16 waves).
bool panel_balance_supported(const gcoo_matrix *h, int tk, int rpw, int waves);
// Forget the panel lists (after a conversion or a capacity change).
inline void invalidate_panel_lists(const gcoo_matrix *h)
{
    h->plist_src = nullptr;
    ++h->plist_gen;
}
// Ints of the offset (and count) area the panel lists of h need for k
// chunks of tk rows and panels of waves x rpw rows; the lists themselves
// need h->nnz uint2 entries.
size_t panel_offset_elems(const gcoo_matrix *h, int tk, int rpw, int waves);
// Enqueue the build of the panel lists into caller-given buffers (poff of
// panel_offset_elems ints, plist of h->nnz entries); touches no handle state.
// With rowmap (waves * rpw ints per panel), the panels' rows are dealt to
// the waves in the balanced order, written to rowmap.
int build_panel_lists(const gcoo_matrix *h, const int2 *seg, int tk, int rowb, int rpw,
                      int waves, bool widen, int *poff, uint2 *plist, const ConvStatus *guard,
                      hipStream_t s, ConvStatus *hst = nullptr, bool *exported = nullptr,
                      int *rowmap = nullptr);

// Device scratch handed to the multiply by a caller that cannot allocate:
// the panel lists and the long-row list are carved out of it instead of
// the handle's own buffers.  base is 256-byte aligned.
struct ExtScratch {
    char *base;
    size_t bytes;
};
// FP32 multiply with every scratch buffer taken from xs (rebuilt on every
// call); returns GCOO_ERR_BUDGET, before enqueuing anything, when the
// chosen kernel needs more than xs->bytes.
int launch_spdm_f32_ext(const gcoo_matrix *h, const float *D, int N, float *C,
                        const ExtScratch *xs, hipStream_t s);

// ---- the panel multiply of many jobs in one launch (gcoo_spdm.hip) ----
// A batch's panel jobs: one descriptor each (16-byte aligned), the same
// arguments as a single panel launch.
struct PanelJob {
    const int *poff;
    const uint2 *plist;
    const void *D;
    float *C;
    const int *rowmap;  // balanced row order of the lists, or null
    int M, K, N, nC, nPanels;
    unsigned dBytes;
};
// The one workgroup shape batches use: 16 waves x 21 rows (336-row panels),
// chunks of 75 FP32 or 150 BF16 D rows of 256 columns.
constexpr int kBatchPanelWaves = 16, kBatchPanelRpw = 21;
inline int batch_panel_tk(bool bf16) { return bf16 ? 150 : 75; }
inline int batch_panel_rowb(bool bf16) { return bf16 ? 512 : 1024; }
// True when D, N and C allow the panel kernel (16-byte pieces of D rows, 4
// columns per lane).
bool panel_operands_ok(const void *D, int N, const float *C, int elemBytes);
// Workgroup b of a launch: job (-1: none), tile = column tile * nPanels +
// panel, and its turn tidx of tmod at the L2 touch among the units of the
// same job and column tile on its XCD.
struct PanelUnit {
    int job, tile, tmod, tidx;
};
// Enqueue one launch of nUnits workgroups over the jobs (FP32 or BF16 D).
int launch_panel_batch(bool bf16, const PanelUnit *units, int nUnits, const PanelJob *jobs,
                       hipStream_t s);

// ---- batches (gcoo_batch.hip) ----
// gcoo_batch_commit without the final stream synchronisation.
int batch_commit_async(gcoo_batch_t b, gcoo_stream s);

// ---- generators and element conversions (gcoo_gen.hip) ----
int launch_fill(float *dst, int64_t n, uint64_t seed, uint64_t stream_id,
                float scale, uint64_t keep_threshold, bool sparse,
                hipStream_t s);
// y = RNE bf16 of x (NaN kept quiet).
int launch_f32_to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t s);
// seg16[i] = (seg[i].x, bf16 bits of seg[i].y) for i < n.
int launch_seg_to_bf16(const int2 *seg, int2 *seg16, int64_t n, hipStream_t s);

// Entries of segment capacity beyond nnz + maxRow that every conversion
// provides, so kernels may read a few entries past a row's end.
constexpr int kSegSlack = 64;

// Number of XCDs a launch is spread over (round-robin dispatch of
// workgroups).  gfx950 (MI355X) has 8.
constexpr int kNumXcd = 8;

// Workgroup index remapped so that consecutive logical indices run on the
// same XCD: hardware places workgroup b on XCD b % kNumXcd.  Returns a
// bijection of [0, nblocks).
__device__ __forceinline__ int xcd_remap(int b, int nblocks)
{
    int q = nblocks / kNumXcd, r = nblocks % kNumXcd;
    int x = b % kNumXcd, i = b / kNumXcd;
    return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Wave64 inclusive prefix sum of an int.
__device__ __forceinline__ int wave_inclusive_sum(int v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

}  // namespace gcoo
