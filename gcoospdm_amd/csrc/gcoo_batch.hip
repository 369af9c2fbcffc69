// gcoo_batch.hip -- many independent SpDM jobs in one persistent launch.
//
// The reference runs one multiply per launch with a fixed tile
// (src/ShGCOOSpDM.cu:153-167, src/constants.h:4-5).  Here a batch of jobs
// C_i = S_i * D_i (each S_i a converted handle, D_i FP32 or BF16, any
// M_i / K_i / N_i) is cut into work items -- `rows` consecutive rows x one
// column tile of 64*vec columns (or two or four row streams of 32 / 16 lanes
// when N_i is narrow) -- and a persistent kernel sized to the machine pulls
// items from a device counter until none are left.  Every output element is
// the same fmaf chain (row r's entries, k ascending) as in the plain kernels,
// so a job's C is bit-identical whether it runs alone or in any batch.
//
// Jobs dense enough for the panel-staged multiply (gcoo_spdm.hip, variant 6)
// run it instead, all of them in one launch per element type ahead of the
// persistent kernel: a cost model (below) picks them at commit, their panel
// lists are built then, and a table gives each workgroup of the launch one
// (job, panel, column tile) unit, whole jobs per XCD with balanced work.
//
// The run is one counter-reset kernel, the panel launches and the persistent
// kernel; it is graph-capturable (no memset, no host synchronisation, no
// allocation).
#include "gcoo_internal.h"

#include <type_traits>

#include <algorithm>
#include <new>
#include <vector>

namespace gcoo {
namespace {

constexpr int kBatchWaves = 4;
constexpr int kBatchThreads = 64 * kBatchWaves;

// Device descriptor of one job (16-byte aligned).
struct JobDesc {
    const int *rowptr;
    const int2 *seg;
    const void *D;
    float *C;
    int M, K, N;
    int bytes;        // 4 (FP32) or 2 (BF16)
    int vec;          // columns per lane
    int lanes;        // lanes per row stream: 64, 32 or 16
    int rows;         // rows per work item (multiple of 64 / lanes)
    int colTiles;     // ceil(N / (lanes * vec))
    unsigned dBytes;  // K * N * bytes (< 2^31)
    int pad;
};

__device__ __forceinline__ float lo16(unsigned w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float hi16(unsigned w) { return __uint_as_float(w & 0xffff0000u); }

// The bits of VEC elements of BYTES bytes each (at most 16 bytes), as
// loaded; 2-byte loads zero-extended.
template <int BYTES, int VEC>
struct ElemBits {
    unsigned w[(BYTES * VEC + 3) / 4];
};

// Those bits at byte offset voff (per lane) + soff (wave-uniform) of the
// buffer; out-of-range lanes read zeros.
template <int BYTES, int VEC>
__device__ __forceinline__ ElemBits<BYTES, VEC> load_bits(__amdgpu_buffer_rsrc_t r, int voff,
                                                          int soff)
{
    constexpr int B = BYTES * VEC;
    static_assert(B == 2 || B == 4 || B == 8 || B == 16, "at most 16 bytes per lane");
    ElemBits<BYTES, VEC> x;
    if constexpr (B == 16) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) x.w[i] = (unsigned)v[i];
    } else if constexpr (B == 8) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        x.w[0] = (unsigned)v[0];
        x.w[1] = (unsigned)v[1];
    } else if constexpr (B == 4) {
        x.w[0] = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
    } else {
        x.w[0] = __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
    }
    return x;
}

// The loaded bits widened to FP32 (exact: a BF16 is the high half of an FP32).
template <int BYTES, int VEC>
__device__ __forceinline__ void widen_elems(float (&d)[VEC], const ElemBits<BYTES, VEC> &x)
{
    if constexpr (BYTES == 4) {
#pragma unroll
        for (int i = 0; i < VEC; ++i) d[i] = __uint_as_float(x.w[i]);
    } else if constexpr (VEC == 1) {
        d[0] = lo16(x.w[0]);
    } else {
#pragma unroll
        for (int i = 0; i < VEC / 2; ++i) {
            d[2 * i] = lo16(x.w[i]);
            d[2 * i + 1] = hi16(x.w[i]);
        }
    }
}

// VEC widened elements of BYTES bytes each at byte offset voff (per lane) +
// soff (wave-uniform) of the buffer; out-of-range lanes read zeros.  The
// width is a template parameter, so the entry loops carry no branches on it.
template <int BYTES, int VEC>
__device__ __forceinline__ void load_elems(float (&d)[VEC], __amdgpu_buffer_rsrc_t r, int voff,
                                           int soff)
{
    widen_elems<BYTES, VEC>(d, load_bits<BYTES, VEC>(r, voff, soff));
}

// An entry's value word as FP32 (BF16 segments hold the bits in the low half).
template <int BYTES>
__device__ __forceinline__ float entry_val(int w)
{
    return BYTES == 2 ? __int_as_float(w << 16) : __int_as_float(w);
}

__device__ __forceinline__ void fmac_v(float &acc, float a, float d)
{
    asm("v_fmac_f32 %0, %1, %2" : "+v"(acc) : "v"(a), "v"(d));
}

// the same step with a wave-uniform multiplier held in a scalar register
__device__ __forceinline__ void fmac_s(float &acc, float a, float d)
{
    asm("v_fmac_f32 %0, %1, %2" : "+v"(acc) : "s"(a), "v"(d));
}

// One row stream of a work item: lane (sub, lc) walks row r's segment and
// accumulates columns j .. j+vec-1, U entries' loads in flight at a time.
template <int BYTES, int VEC>
__device__ __forceinline__ void run_row(const JobDesc &jd, __amdgpu_buffer_rsrc_t rsrc, int r,
                                        int j, bool active)
{
    constexpr int U = 4;
    float acc[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[v] = 0.0f;
    const bool valid = r < jd.M;
    int p = valid ? jd.rowptr[r] : 0;
    const int pend = valid ? jd.rowptr[r + 1] : 0;
    // byte offsets are unsigned: a lane past N starts at dBytes, so every one
    // of its loads is out of range and returns zeros
    const unsigned rowBytes = (unsigned)jd.N * BYTES;
    const unsigned voffc = active ? (unsigned)j * BYTES : jd.dBytes;
    if constexpr (BYTES == 4) {
        // FP32: rounds of U entries, each loaded and then its D loads issued
        // (a software-pipelined form measured slower for FP32: more vector
        // registers, fewer waves)
        while (__builtin_amdgcn_ballot_w64(p < pend) != 0) {
            int2 e[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                e[u] = (p + u < pend) ? jd.seg[p + u] : make_int2(0, 0);
            float d[U][VEC];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const unsigned voff =
                    (p + u < pend) ? voffc + (unsigned)e[u].x * rowBytes : jd.dBytes;
                load_elems<BYTES, VEC>(d[u], rsrc, (int)voff, 0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (p + u < pend) {
                    const float a = entry_val<BYTES>(e[u].y);
#pragma unroll
                    for (int v = 0; v < VEC; ++v) fmac_v(acc[v], a, d[u][v]);
                }
            }
            p += U;
        }
    } else {
        // BF16: each round issues the next round's entry loads and its own
        // D loads before its FMAs, so it waits for one load latency, not for
        // an entry load and then the D load it addresses.  The loads are
        // unconditional, so the waits count exactly: an entry past the row's
        // end reads entry pend - 1 (entry 0 when pend is 0: the segments'
        // capacity is never zero), its D load goes out of range and returns
        // +0.0, and its multiplier is -0.0: fmaf(-0.0, +0.0, acc) == acc for
        // every acc, so the chain is the row's own entries in ascending k.
        // An entry is read as one 64-bit word (k, value bits) through a
        // global pointer: a vector-memory load only.
        typedef __attribute__((address_space(1))) const long long gword;
        gword *sg = (gword *)jd.seg;  // a C-style cast may change the address space
        const int last = pend > 0 ? pend - 1 : 0;
        auto entry = [&](int q) {
            const long long w = sg[min(q, last)];
            return make_int2((int)w, (int)(w >> 32));
        };
        int2 e[U];
#pragma unroll
        for (int u = 0; u < U; ++u) e[u] = entry(p + u);
        while (__builtin_amdgcn_ballot_w64(p < pend) != 0) {
            // the next round's entries first: loads complete in issue order,
            // so the wait for a D load of this round also covers them
            int2 en[U];
#pragma unroll
            for (int u = 0; u < U; ++u) en[u] = entry(p + U + u);
            float d[U][VEC];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const unsigned voff =
                    (p + u < pend) ? voffc + (unsigned)e[u].x * rowBytes : jd.dBytes;
                load_elems<BYTES, VEC>(d[u], rsrc, (int)voff, 0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float a = (p + u < pend) ? entry_val<BYTES>(e[u].y) : -0.0f;
#pragma unroll
                for (int v = 0; v < VEC; ++v) fmac_v(acc[v], a, d[u][v]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) e[u] = en[u];
            p += U;
        }
    }
    if (valid && active) {
        float *dst = jd.C + (size_t)r * jd.N + j;
#pragma unroll
        for (int v = 0; v < VEC; ++v) __builtin_nontemporal_store(acc[v], dst + v);
    }
}

// One row per wave (64-lane streams): the row is wave-uniform, so its entries
// are read through the constant address space (scalar loads; the segments
// are not written during the batch) and an entry's D row offset is a scalar
// buffer offset.  Rounds of R entries (16, or 8 at 8 columns per lane, so
// the D values take at most 64 registers) load all their entries, then issue
// all their D loads, then run the FMAs; then rounds of 8, then one guarded
// round of the last 1..7 -- the same chain order, so the same bits.
template <int BYTES, int VEC>
__device__ __forceinline__ void run_row_wave(const JobDesc &jd, __amdgpu_buffer_rsrc_t rsrc,
                                             int r, int j, bool active)
{
    typedef __attribute__((address_space(4))) const int cint;
    float acc[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[v] = 0.0f;
    const bool valid = r < jd.M;
    int p = __builtin_amdgcn_readfirstlane(valid ? jd.rowptr[r] : 0);
    const int pend = __builtin_amdgcn_readfirstlane(valid ? jd.rowptr[r + 1] : 0);
    const int rowBytes = jd.N * BYTES;
    const int voffc = active ? j * BYTES : (int)jd.dBytes;
    cint *sg = (cint *)(jd.seg);  // (k, value bits) pairs; a C-style cast may change the address space
    // B entries from p: unguarded (every one exists), so the entry loads merge
    auto round = [&](auto bc, int q) {
        constexpr int B = decltype(bc)::value;
        int ek[B], ev[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            ek[u] = sg[2 * (q + u)];
            ev[u] = sg[2 * (q + u) + 1];
        }
        float d[B][VEC];
#pragma unroll
        for (int u = 0; u < B; ++u) load_elems<BYTES, VEC>(d[u], rsrc, voffc, ek[u] * rowBytes);
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const float a = entry_val<BYTES>(ev[u]);
#pragma unroll
            for (int v = 0; v < VEC; ++v) fmac_s(acc[v], a, d[u][v]);
        }
    };
    constexpr int R = VEC >= 8 ? 8 : 16;
    for (; p + R <= pend; p += R) round(std::integral_constant<int, R>{}, p);
    if constexpr (R > 8)
        for (; p + 8 <= pend; p += 8) round(std::integral_constant<int, 8>{}, p);
    if (p < pend) {
        // the last 1..7 entries; the loads keep their raw bits and are
        // widened next to their FMAs (a widening beside its guarded load
        // would wait for that load before the next one is issued)
        const int n = pend - p;
        float a[7];
        ElemBits<BYTES, VEC> w[7];
#pragma unroll
        for (int u = 0; u < 7; ++u) {
            if (u < n) {
                a[u] = entry_val<BYTES>(sg[2 * (p + u) + 1]);
                w[u] = load_bits<BYTES, VEC>(rsrc, voffc, sg[2 * (p + u)] * rowBytes);
            }
        }
#pragma unroll
        for (int u = 0; u < 7; ++u)
            if (u < n) {
                float d[VEC];
                widen_elems<BYTES, VEC>(d, w[u]);
#pragma unroll
                for (int v = 0; v < VEC; ++v) fmac_s(acc[v], a[u], d[v]);
            }
    }
    if (valid && active) {
        float *dst = jd.C + (size_t)r * jd.N + j;
#pragma unroll
        for (int v = 0; v < VEC; ++v) __builtin_nontemporal_store(acc[v], dst + v);
    }
}

// The rows of one work item, in streams of jd.lanes lanes.
template <int BYTES, int VEC>
__device__ __forceinline__ void run_item(const JobDesc &jd, __amdgpu_buffer_rsrc_t rsrc, int r0,
                                         int sub, int j, bool active, int streams)
{
    for (int q = 0; q < jd.rows; q += streams) {
        const int r = r0 + q + sub;
        if (streams == 1) run_row_wave<BYTES, VEC>(jd, rsrc, r, j, active);
        else run_row<BYTES, VEC>(jd, rsrc, r, j, active);
    }
}

// One item queue per XCD: queue x holds the items [start[x], start[x + 1]),
// a contiguous run of whole jobs (mostly) chosen so that the queues carry
// about equal work, and its counter sits on a cache line of its own.
// Default ratios of the panel-job choice (see plan_panels): on the 512-job
// workload FP32 was fastest at 1.0 (1.63 ms; 0.5: 1.98, 2.0: 1.71, none:
// 2.04) and BF16 at 1.0 among the ratios (1.61 ms; 0.5: 1.75, 2.0: 1.69,
// 8.0: 1.81), though with no panel jobs at all it takes 1.60 ms.
constexpr double kPanelRatioF32 = 1.0;
constexpr double kPanelRatioBf16 = 1.0;

struct BatchQueues {
    int start[kNumXcd + 1];
};
constexpr int kCounterStride = 32;  // unsigned ints between two queue counters
// Resident persistent workgroups per CU at most: the FP32 kernel fits 7, but
// on the 512-job workload (every job persistent) 4 / 5 / 6 / 7 / 8 per CU
// took 2.07 / 2.00 / 2.00 / 2.02 / 2.02 ms; the BF16 kernel fits 5 (1.59 ms;
// 6 and 8: 1.61, 1.63 ms).
constexpr int kBatchMaxPerCu = 6;

__global__ __launch_bounds__(64) void batch_reset_kernel(unsigned *counter)
{
    if (threadIdx.x < kNumXcd) counter[threadIdx.x * kCounterStride] = 0;
}

// Element types a persistent kernel serves (a bit set): the kernel holds the
// row loops of those types only, since its register count, and so the waves
// per SIMD, is the largest any of its loops needs.
constexpr int kTypeF32 = 1, kTypeBf16 = 2;

// One dispatch per item on the element size and width.
template <int TYPES>
__device__ __forceinline__ void dispatch_item(const JobDesc &jd, __amdgpu_buffer_rsrc_t rsrc,
                                              int r0, int sub, int j, bool active, int streams)
{
    if constexpr ((TYPES & kTypeF32) != 0) {
        if (TYPES == kTypeF32 || jd.bytes == 4) {
            switch (jd.vec) {
            case 1: run_item<4, 1>(jd, rsrc, r0, sub, j, active, streams); break;
            case 2: run_item<4, 2>(jd, rsrc, r0, sub, j, active, streams); break;
            default: run_item<4, 4>(jd, rsrc, r0, sub, j, active, streams); break;
            }
            return;
        }
    }
    if constexpr ((TYPES & kTypeBf16) != 0) {
        switch (jd.vec) {
        case 1: run_item<2, 1>(jd, rsrc, r0, sub, j, active, streams); break;
        case 2: run_item<2, 2>(jd, rsrc, r0, sub, j, active, streams); break;
        case 4: run_item<2, 4>(jd, rsrc, r0, sub, j, active, streams); break;
        default: run_item<2, 8>(jd, rsrc, r0, sub, j, active, streams); break;
        }
    }
}

// Workgroup b runs on XCD b % kNumXcd (the hardware dispatches workgroups
// round-robin over the XCDs) and takes items from that XCD's queue first, so
// a job's D and C tiles are fetched into one XCD's L2 instead of all eight;
// when its queue is empty it takes items from the others in turn.  TYPES:
// the element types of the batch's items (kTypeF32 | kTypeBf16).
template <int TYPES>
__global__ __launch_bounds__(kBatchThreads) void batch_persistent_kernel(
    const JobDesc *__restrict__ jobs, const int *__restrict__ itemStart, int nJobs,
    BatchQueues queues, unsigned *__restrict__ counter)
{
    const int lane = __lane_id();
    const int home = blockIdx.x % kNumXcd;
    // the job of the last item (its item range cached in scalar registers)
    int lo = 0, jbeg = 0, jend = 0;
    for (int qi = 0; qi < kNumXcd; ++qi) {
        const int q = (home + qi) % kNumXcd;
        const int qbeg = queues.start[q], qend = queues.start[q + 1];
        if (qbeg >= qend) continue;
        for (;;) {
        int it = 0;
        if (lane == 0) it = (int)atomicAdd(&counter[q * kCounterStride], 1u);
        it = qbeg + __builtin_amdgcn_readfirstlane(it);
        if (it >= qend) break;
        if (it < jbeg || it >= jend) {
            // job = last index with itemStart[job] <= it (scalar binary
            // search, only when the item leaves the cached job)
            lo = 0;
            int hi = nJobs - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (itemStart[mid] <= it) lo = mid;
                else hi = mid - 1;
            }
            jbeg = itemStart[lo];
            jend = itemStart[lo + 1];
        }
        const JobDesc jd = jobs[lo];
        const int local = it - jbeg;
        const int rb = local / jd.colTiles;
        const int ct = local - rb * jd.colTiles;
        const int sub = lane / jd.lanes;
        const int lc = lane - sub * jd.lanes;
        const int streams = 64 / jd.lanes;
        const int j = (ct * jd.lanes + lc) * jd.vec;
        const bool active = j < jd.N;
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<void *>(jd.D), (short)0, (int)jd.dBytes, 0x00020000);
        const int r0 = rb * jd.rows;
        dispatch_item<TYPES>(jd, rsrc, r0, sub, j, active, streams);
        }
    }
}

}  // namespace
}  // namespace gcoo

using namespace gcoo;

struct gcoo_batch {
    std::vector<JobDesc> jobs;
    std::vector<int> itemStart;  // nJobs + 1 (panel jobs have none once committed)
    std::vector<int> jobItems;   // work items of each job on the persistent kernel
    std::vector<gcoo_handle> handles;
    // each handle's conversion generation when the job table was committed:
    // a run refuses to start once one moved (the table would point at
    // re-converted or freed entry arrays)
    std::vector<uint64_t> gens;
    // resident workgroups per CU of each persistent kernel (by its element
    // types, 1..3), queried once at creation
    int cus = 256, perCu[4] = {4, 4, 4, 4};
    int types = kTypeF32;          // element types of the items, set at commit
    BatchQueues queues{};          // per-XCD item ranges, set at commit
    JobDesc *d_jobs = nullptr;
    int *d_itemStart = nullptr;
    unsigned *d_counter = nullptr;
    JobDesc *h_jobs = nullptr;     // pinned staging of the descriptor upload
    int *h_itemStart = nullptr;
    size_t capJobs = 0;
    bool committed = false;
    int gridBlocks = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;

    // panel jobs (chosen at commit): 1 = by the cost model, 0 = none, 2 =
    // every job whose operands allow it
    int panelMode = 1;
    double panelRatio[2] = {kPanelRatioF32, kPanelRatioBf16};  // FP32, BF16 jobs
    std::vector<int> panelOf;          // per job: index into pjobs, or -1
    std::vector<uint64_t> plistGens;   // per panel job: its handle's list generation
    std::vector<PanelJob> pjobs;
    std::vector<PanelUnit> units;      // FP32 launch's units, then BF16's
    int nUnitsF32 = 0, nUnitsBf16 = 0;
    PanelJob *d_pjobs = nullptr, *h_pjobs = nullptr;
    PanelUnit *d_units = nullptr, *h_units = nullptr;
    size_t capPJobs = 0, capUnits = 0;
};

namespace {

int pick_lanes(int N, int vec)
{
    const int cols = (N + vec - 1) / vec;
    if (cols <= 16) return 16;
    if (cols <= 32) return 32;
    return 64;
}

void destroy_graph(gcoo_batch *b)
{
    if (b->graphExec) (void)hipGraphExecDestroy(b->graphExec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graphExec = nullptr;
    b->graph = nullptr;
}

// GCOO_OK when every job's handle still holds the conversion the job table
// was committed with.
int check_generations(const gcoo_batch *b)
{
    for (size_t i = 0; i < b->handles.size(); ++i) {
        const gcoo_handle h = b->handles[i];
        if (i >= b->gens.size() || !h->converted || h->generation != b->gens[i])
            return set_error(GCOO_ERR_STATE,
                             "job %zu: its matrix was converted again after the batch was "
                             "committed; commit (and capture) the batch again", i);
        const int pj = i < b->panelOf.size() ? b->panelOf[i] : -1;
        if (pj >= 0 && (h->plist_gen != b->plistGens[pj] || h->plist != b->pjobs[pj].plist ||
                        h->poff != b->pjobs[pj].poff || panel_rowmap(h) != b->pjobs[pj].rowmap))
            return set_error(GCOO_ERR_STATE,
                             "job %zu: its matrix's panel lists were rebuilt (another multiply "
                             "of it used another panel shape) after the batch was committed; "
                             "commit (and capture) the batch again", i);
    }
    return GCOO_OK;
}

bool has_work(const gcoo_batch *b)
{
    return !b->jobs.empty() && (b->itemStart.back() > 0 || b->nUnitsF32 + b->nUnitsBf16 > 0);
}

// The panel launches, then the persistent kernel, on s.  (Run beside each
// other on two streams, the persistent kernel's workgroups held their CUs
// until its queues drained and the panel units, which need a whole CU each,
// waited: the panel launch stretched from 0.8 to 1.8 ms and the batch took
// 1.85 ms either way.)
int enqueue_run(gcoo_batch *b, hipStream_t s)
{
    const bool items = b->itemStart.back() > 0;
    if (int rc = launch_panel_batch(false, b->d_units, b->nUnitsF32, b->d_pjobs, s)) return rc;
    if (int rc = launch_panel_batch(true, b->d_units + b->nUnitsF32, b->nUnitsBf16, b->d_pjobs, s))
        return rc;
    if (items) {
        hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, b->d_counter);
        GCOO_CHECK_LAUNCH("batch_reset_kernel");
        const int nJobs = (int)b->jobs.size();
        if (b->types == kTypeF32)
            hipLaunchKernelGGL(batch_persistent_kernel<kTypeF32>, dim3(b->gridBlocks),
                               dim3(kBatchThreads), 0, s, b->d_jobs, b->d_itemStart, nJobs,
                               b->queues, b->d_counter);
        else if (b->types == kTypeBf16)
            hipLaunchKernelGGL(batch_persistent_kernel<kTypeBf16>, dim3(b->gridBlocks),
                               dim3(kBatchThreads), 0, s, b->d_jobs, b->d_itemStart, nJobs,
                               b->queues, b->d_counter);
        else
            hipLaunchKernelGGL(batch_persistent_kernel<kTypeF32 | kTypeBf16>, dim3(b->gridBlocks),
                               dim3(kBatchThreads), 0, s, b->d_jobs, b->d_itemStart, nJobs,
                               b->queues, b->d_counter);
        GCOO_CHECK_LAUNCH("batch_persistent_kernel");
    }
    return GCOO_OK;
}

// Cost model of the choice, in GPU-seconds.  Panel kernel: one workgroup
// per CU; a unit takes about a ns per entry of its panel, b ns per D row
// staged and c us of fixed costs (C tile store, first chunk), fitted on
// single-job launches of the 512-job workload (FP32: 3.1 ns, 6.85 ns,
// 10.3 us; BF16: 4.2 ns, 2.84 ns, 12.4 us), times 1.3 for running beside
// the other jobs, over the CUs.  Persistent kernel: bound by the D gathers
// into the CUs (one element per FMA), about 8.4 TFLOP/s over the whole GPU
// on that workload, FP32 and BF16 alike.  A job is a panel job when its
// panel estimate is below `panelRatio` times its persistent-kernel estimate.
constexpr double kPanelBatchFactor = 1.3;
constexpr double kGatherFlops = 8.4e12;

double panel_unit_seconds(const gcoo_matrix *h, bool bf16)
{
    const double panelRows = std::min(h->M, kBatchPanelWaves * kBatchPanelRpw);
    const double entries = h->M > 0 ? (double)h->nnz * panelRows / h->M : 0.0;
    return bf16 ? 4.2e-9 * entries + 2.84e-9 * h->K + 12.4e-6
                : 3.1e-9 * entries + 6.85e-9 * h->K + 10.3e-6;
}

long long panel_units(const gcoo_matrix *h, int N)
{
    const int pr = kBatchPanelWaves * kBatchPanelRpw;
    return (long long)((h->M + pr - 1) / pr) * ((N + 255) / 256);
}

double gather_seconds(const JobDesc &jd, const gcoo_matrix *h)
{
    return 2.0 * (double)h->nnz * jd.N / kGatherFlops;
}

// Choose the panel jobs, build their lists (the handles keep them) and lay
// out the units of the two launches.
int plan_panels(gcoo_batch *b, hipStream_t s)
{
    const size_t n = b->jobs.size();
    b->panelOf.assign(n, -1);
    b->plistGens.clear();
    b->pjobs.clear();
    b->units.clear();
    b->nUnitsF32 = b->nUnitsBf16 = 0;
    if (b->panelMode == 0) return GCOO_OK;
    // the segments whose lists a handle holds for this batch: a second job of
    // the same matrix in the other element type would rebuild them, so it
    // takes the gather kernel
    std::vector<std::pair<const gcoo_matrix *, const int2 *>> claimed;
    std::vector<double> unitSec;
    std::vector<bool> pjBf16;
    for (size_t i = 0; i < n; ++i) {
        const JobDesc &jd = b->jobs[i];
        const gcoo_matrix *h = b->handles[i];
        if (jd.N <= 0 || h->M <= 0 || h->K <= 0 || h->nnz <= 0) continue;
        if (!panel_operands_ok(jd.D, jd.N, jd.C, jd.bytes)) continue;
        const bool bf16 = jd.bytes == 2;
        const long long units = panel_units(h, jd.N);
        const double us = panel_unit_seconds(h, bf16);
        if (b->panelMode == 1 &&
            !(kPanelBatchFactor * (double)units * us / b->cus <
              b->panelRatio[bf16 ? 1 : 0] * gather_seconds(jd, h)))
            continue;
        bool clash = false;
        for (const auto &c : claimed)
            if (c.first == h && c.second != jd.seg) clash = true;
        if (clash) continue;
        // the lists serve every run of the batch: balanced row order
        int rc = ensure_panel_lists(h, jd.seg, batch_panel_tk(bf16), batch_panel_rowb(bf16),
                                    kBatchPanelRpw, kBatchPanelWaves, bf16, nullptr, s, true);
        if (rc) return rc;
        claimed.emplace_back(h, jd.seg);
        PanelJob pj{};
        pj.poff = h->poff;
        pj.plist = h->plist;
        pj.rowmap = panel_rowmap(h);
        pj.D = jd.D;
        pj.C = jd.C;
        pj.M = h->M;
        pj.K = h->K;
        pj.N = jd.N;
        pj.nC = (h->K + batch_panel_tk(bf16) - 1) / batch_panel_tk(bf16);
        pj.nPanels = (h->M + kBatchPanelWaves * kBatchPanelRpw - 1) /
                     (kBatchPanelWaves * kBatchPanelRpw);
        pj.dBytes = jd.dBytes;
        b->panelOf[i] = (int)b->pjobs.size();
        b->pjobs.push_back(pj);
        b->plistGens.push_back(h->plist_gen);
        unitSec.push_back(us);
        pjBf16.push_back(bf16);
    }
    // per element type: the jobs with the longest units first, each column
    // tile's units in groups of at most one XCD's CUs (32), each group to the
    // XCD with the least estimated work so far.  A group's units run together
    // on one XCD, stage the same D tile through its L2 (taking turns at the
    // L2 touch) and so fetch it once; a job larger than one XCD spreads over
    // several, and no XCD idles while another runs a long job.  Workgroup b
    // runs on XCD b % 8, so XCD x's k-th unit goes to workgroup 8 k + x; XCDs
    // with fewer units get empty ones.
    constexpr int kGroupUnits = 32;
    for (int bf = 0; bf < 2; ++bf) {
        std::vector<int> js;
        for (size_t p = 0; p < b->pjobs.size(); ++p)
            if (pjBf16[p] == (bf == 1)) js.push_back((int)p);
        if (js.empty()) continue;
        std::stable_sort(js.begin(), js.end(), [&](int x, int y) { return unitSec[x] > unitSec[y]; });
        std::vector<std::vector<PanelUnit>> lists(kNumXcd);
        double load[kNumXcd] = {};
        for (int p : js) {
            const int tiles = (b->pjobs[p].N + 255) / 256;
            const int np = b->pjobs[p].nPanels;
            for (int ct = 0; ct < tiles; ++ct) {
                for (int p0 = 0; p0 < np; p0 += kGroupUnits) {
                    const int cnt = std::min(kGroupUnits, np - p0);
                    int x = 0;
                    for (int y = 1; y < kNumXcd; ++y)
                        if (load[y] < load[x]) x = y;
                    for (int k = 0; k < cnt; ++k)
                        lists[x].push_back(PanelUnit{p, ct * np + p0 + k, cnt, k});
                    load[x] += cnt * unitSec[p];
                }
            }
        }
        size_t most = 0;
        for (int x = 0; x < kNumXcd; ++x) most = std::max(most, lists[x].size());
        if (most * kNumXcd > (size_t)INT32_MAX)
            return set_error(GCOO_ERR_OVERFLOW, "too many panel workgroups");
        const size_t first = b->units.size();
        b->units.resize(first + most * kNumXcd, PanelUnit{-1, 0, 1, 0});
        for (int x = 0; x < kNumXcd; ++x)
            for (size_t k = 0; k < lists[x].size(); ++k) b->units[first + k * kNumXcd + x] = lists[x][k];
        (bf ? b->nUnitsBf16 : b->nUnitsF32) = (int)(most * kNumXcd);
    }
    return GCOO_OK;
}

}  // namespace

extern "C" {

int gcoo_batch_create(gcoo_batch_t *out)
{
    if (!out) return set_error(GCOO_ERR_INVALID, "null batch out");
    auto *b = new (std::nothrow) gcoo_batch;
    if (!b) return set_error(GCOO_ERR_ALLOC, "batch");
    if (hipMalloc(&b->d_counter, sizeof(unsigned) * kCounterStride * kNumXcd) != hipSuccess) {
        (void)hipGetLastError();
        delete b;
        return set_error(GCOO_ERR_ALLOC, "batch counter");
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        b->cus < 1) {
        (void)hipGetLastError();
        b->cus = 256;
    }
    const void *kernels[4] = {nullptr, (const void *)batch_persistent_kernel<kTypeF32>,
                              (const void *)batch_persistent_kernel<kTypeBf16>,
                              (const void *)batch_persistent_kernel<kTypeF32 | kTypeBf16>};
    for (int t = 1; t <= 3; ++t) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->perCu[t], kernels[t], kBatchThreads,
                                                         0) != hipSuccess ||
            b->perCu[t] < 1) {
            (void)hipGetLastError();
            b->perCu[t] = 4;
        }
    }
    b->itemStart.push_back(0);
    *out = b;
    return GCOO_OK;
}

int gcoo_batch_destroy(gcoo_batch_t b)
{
    if (!b) return GCOO_OK;
    destroy_graph(b);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->d_itemStart) (void)hipFree(b->d_itemStart);
    if (b->d_counter) (void)hipFree(b->d_counter);
    if (b->h_jobs) (void)hipHostFree(b->h_jobs);
    if (b->h_itemStart) (void)hipHostFree(b->h_itemStart);
    if (b->d_pjobs) (void)hipFree(b->d_pjobs);
    if (b->h_pjobs) (void)hipHostFree(b->h_pjobs);
    if (b->d_units) (void)hipFree(b->d_units);
    if (b->h_units) (void)hipHostFree(b->h_units);
    delete b;
    return GCOO_OK;
}

int gcoo_batch_clear(gcoo_batch_t b)
{
    if (!b) return set_error(GCOO_ERR_INVALID, "null batch");
    destroy_graph(b);
    b->jobs.clear();
    b->handles.clear();
    b->gens.clear();
    b->itemStart.assign(1, 0);
    b->jobItems.clear();
    b->panelOf.clear();
    b->plistGens.clear();
    b->pjobs.clear();
    b->units.clear();
    b->nUnitsF32 = b->nUnitsBf16 = 0;
    b->committed = false;
    return GCOO_OK;
}

int gcoo_batch_add(gcoo_batch_t b, gcoo_handle h, const void *D, int dtype, int N, float *C,
                   int rows, int vec, int *job_index)
{
    if (!b || !h) return set_error(GCOO_ERR_INVALID, "null batch or handle");
    if (!h->converted) return set_error(GCOO_ERR_STATE, "matrix not converted");
    if (dtype != 0 && dtype != 1) return set_error(GCOO_ERR_INVALID, "dtype %d (0 f32, 1 bf16)", dtype);
    if (N < 0 || (N > 0 && h->M > 0 && (!C || (!D && h->K > 0))))
        return set_error(GCOO_ERR_INVALID, "bad operands");
    const int bytes = dtype == 0 ? 4 : 2;
    const unsigned long long dBytes = (unsigned long long)h->K * N * bytes;
    if (dBytes >= 0x7fffffffull) return set_error(GCOO_ERR_UNSUPPORTED, "dense operand >= 2 GiB");
    if (rows < 0 || rows > 64) return set_error(GCOO_ERR_INVALID, "rows %d (0..64)", rows);
    const int vmaxElems = 16 / bytes;
    if (vec == 0) vec = bytes == 4 ? (N >= 256 ? 4 : N >= 128 ? 2 : 1) : (N >= 256 ? 4 : 2);
    if (vec != 1 && vec != 2 && vec != 4 && vec != 8) return set_error(GCOO_ERR_INVALID, "vec %d", vec);
    if (vec > vmaxElems) vec = vmaxElems;
    // D is read vec elements at a time: N and D's alignment must allow it
    const uintptr_t al = reinterpret_cast<uintptr_t>(D);
    while (vec > 1 && (N % vec != 0 || (al & (uintptr_t)(vec * bytes - 1)) != 0)) vec >>= 1;
    JobDesc jd{};
    jd.rowptr = h->rowptr;  // entry arrays bound at commit
    jd.seg = nullptr;
    jd.D = D;
    jd.C = C;
    jd.M = h->M;
    jd.K = h->K;
    jd.N = N;
    jd.bytes = bytes;
    jd.vec = vec;
    jd.lanes = pick_lanes(N, vec);
    const int streams = 64 / jd.lanes;
    if (rows == 0) rows = 4;
    rows = (rows + streams - 1) / streams * streams;
    jd.rows = rows;
    jd.colTiles = N > 0 ? (N + jd.lanes * vec - 1) / (jd.lanes * vec) : 0;
    jd.dBytes = (unsigned)dBytes;
    const long long items = (N > 0 && h->M > 0) ? (long long)((h->M + rows - 1) / rows) * jd.colTiles : 0;
    const long long start = b->itemStart.back();
    if (start + items > INT32_MAX) return set_error(GCOO_ERR_OVERFLOW, "too many work items");
    b->jobs.push_back(jd);
    b->handles.push_back(h);
    b->itemStart.push_back((int)(start + items));
    b->jobItems.push_back((int)items);
    b->committed = false;
    destroy_graph(b);
    if (job_index) *job_index = (int)b->jobs.size() - 1;
    return GCOO_OK;
}

int gcoo_batch_commit(gcoo_batch_t b, gcoo_stream stream)
{
    int rc = gcoo::batch_commit_async(b, stream);
    if (rc) return rc;
    GCOO_HIP_TRY(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
    return GCOO_OK;
}

}  // extern "C"

// Upload the job table through pinned staging on `stream` without waiting:
// the staging belongs to the batch, so the caller must not commit this batch
// again before the upload has executed (the serving loop commits a slot's
// batch only after that slot's previous run completed).
int gcoo::batch_commit_async(gcoo_batch_t b, gcoo_stream stream)
{
    if (!b) return set_error(GCOO_ERR_INVALID, "null batch");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // bind every job to its handle's current entry arrays (BF16 jobs to the
    // BF16 segments, derived here when stale) and record the generation
    b->gens.resize(b->jobs.size());
    for (size_t i = 0; i < b->jobs.size(); ++i) {
        gcoo_handle h = b->handles[i];
        if (!h->converted) return set_error(GCOO_ERR_STATE, "job %zu: matrix not converted", i);
        if (b->jobs[i].bytes == 2) {
            int rc = ensure_seg16(h, s);
            if (rc) return rc;
            b->jobs[i].seg = h->seg16;
        } else {
            b->jobs[i].seg = h->seg;
        }
        b->jobs[i].rowptr = h->rowptr;
        b->gens[i] = h->generation;
    }
    const size_t n = b->jobs.size();
    // the panel jobs leave the persistent kernel's items
    if (int rc = plan_panels(b, s)) return rc;
    b->itemStart.assign(1, 0);
    for (size_t i = 0; i < n; ++i)
        b->itemStart.push_back(b->itemStart.back() + (b->panelOf[i] >= 0 ? 0 : b->jobItems[i]));
    const size_t np = b->pjobs.size(), nu = b->units.size();
    if (np > b->capPJobs || nu > b->capUnits) {
        if (b->d_pjobs) GCOO_HIP_TRY(hipFree(b->d_pjobs));
        if (b->h_pjobs) GCOO_HIP_TRY(hipHostFree(b->h_pjobs));
        if (b->d_units) GCOO_HIP_TRY(hipFree(b->d_units));
        if (b->h_units) GCOO_HIP_TRY(hipHostFree(b->h_units));
        b->d_pjobs = b->h_pjobs = nullptr;
        b->d_units = b->h_units = nullptr;
        b->capPJobs = b->capUnits = 0;
        const size_t cj = std::max<size_t>(np, 16), cu = std::max<size_t>(nu, 256);
        if (hipMalloc(&b->d_pjobs, sizeof(PanelJob) * cj) != hipSuccess ||
            hipHostMalloc(&b->h_pjobs, sizeof(PanelJob) * cj, hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&b->d_units, sizeof(PanelUnit) * cu) != hipSuccess ||
            hipHostMalloc(&b->h_units, sizeof(PanelUnit) * cu, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(GCOO_ERR_ALLOC, "batch panel tables");
        }
        b->capPJobs = cj;
        b->capUnits = cu;
    }
    if (np > 0) {
        std::copy(b->pjobs.begin(), b->pjobs.end(), b->h_pjobs);
        std::copy(b->units.begin(), b->units.end(), b->h_units);
        GCOO_HIP_TRY(hipMemcpyAsync(b->d_pjobs, b->h_pjobs, sizeof(PanelJob) * np,
                                    hipMemcpyHostToDevice, s));
        GCOO_HIP_TRY(hipMemcpyAsync(b->d_units, b->h_units, sizeof(PanelUnit) * nu,
                                    hipMemcpyHostToDevice, s));
    }
    if (n > b->capJobs) {
        if (b->d_jobs) GCOO_HIP_TRY(hipFree(b->d_jobs));
        if (b->d_itemStart) GCOO_HIP_TRY(hipFree(b->d_itemStart));
        if (b->h_jobs) GCOO_HIP_TRY(hipHostFree(b->h_jobs));
        if (b->h_itemStart) GCOO_HIP_TRY(hipHostFree(b->h_itemStart));
        b->d_jobs = nullptr;
        b->d_itemStart = nullptr;
        b->h_jobs = nullptr;
        b->h_itemStart = nullptr;
        const size_t cap = std::max<size_t>(n, 64);
        if (hipMalloc(&b->d_jobs, sizeof(JobDesc) * cap) != hipSuccess ||
            hipMalloc(&b->d_itemStart, sizeof(int) * (cap + 1)) != hipSuccess ||
            hipHostMalloc(&b->h_jobs, sizeof(JobDesc) * cap, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc(&b->h_itemStart, sizeof(int) * (cap + 1), hipHostMallocDefault) !=
                hipSuccess) {
            (void)hipGetLastError();
            b->capJobs = 0;
            return set_error(GCOO_ERR_ALLOC, "batch descriptors");
        }
        b->capJobs = cap;
    }
    if (n > 0) {
        std::copy(b->jobs.begin(), b->jobs.end(), b->h_jobs);
        std::copy(b->itemStart.begin(), b->itemStart.end(), b->h_itemStart);
        GCOO_HIP_TRY(hipMemcpyAsync(b->d_jobs, b->h_jobs, sizeof(JobDesc) * n,
                                    hipMemcpyHostToDevice, s));
        GCOO_HIP_TRY(hipMemcpyAsync(b->d_itemStart, b->h_itemStart, sizeof(int) * (n + 1),
                                    hipMemcpyHostToDevice, s));
    }
    const long long items = b->itemStart.back();
    // the element types among the jobs with items: the kernel launched
    b->types = 0;
    for (size_t i = 0; i < b->jobs.size(); ++i)
        if (b->itemStart[i + 1] > b->itemStart[i])
            b->types |= b->jobs[i].bytes == 4 ? kTypeF32 : kTypeBf16;
    if (b->types == 0) b->types = kTypeF32;
    // persistent: as many workgroups as can be resident at once, at most
    // kBatchMaxPerCu per CU, a multiple of the XCD count so that every XCD
    // has the same number
    long long blocks = (long long)b->cus * std::min(b->perCu[b->types], kBatchMaxPerCu);
    const long long needed = (items + kBatchWaves - 1) / kBatchWaves;
    if (blocks > needed) blocks = needed;
    blocks = (blocks + kNumXcd - 1) / kNumXcd * kNumXcd;
    b->gridBlocks = (int)blocks;
    // split the items into kNumXcd contiguous ranges of about equal work: a
    // job's items share its estimated work (one FMA per entry and column, one
    // store per output element) equally
    {
        double total = 0;
        std::vector<double> w(b->jobs.size());
        for (size_t i = 0; i < b->jobs.size(); ++i) {
            const JobDesc &jd = b->jobs[i];
            w[i] = (double)b->handles[i]->nnz * jd.N + (double)jd.M * jd.N;
            total += w[i];
        }
        b->queues.start[0] = 0;
        double acc = 0;
        int x = 1;
        for (size_t i = 0; i < b->jobs.size() && x < kNumXcd; ++i) {
            const int i0 = b->itemStart[i], i1 = b->itemStart[i + 1];
            const double per = i1 > i0 ? w[i] / (i1 - i0) : 0.0;
            // place the boundaries that fall inside this job's items
            while (x < kNumXcd && acc + w[i] >= total * x / kNumXcd) {
                const double need = total * x / kNumXcd - acc;
                int cut = per > 0 ? i0 + (int)(need / per + 0.5) : i0;
                cut = std::min(std::max(cut, i0), i1);
                b->queues.start[x] = std::max(cut, b->queues.start[x - 1]);
                ++x;
            }
            acc += w[i];
        }
        for (; x <= kNumXcd; ++x) b->queues.start[x] = (int)items;
        b->queues.start[kNumXcd] = (int)items;
    }
    b->committed = true;
    destroy_graph(b);
    return GCOO_OK;
}

extern "C" {

int gcoo_batch_run_async(gcoo_batch_t b, gcoo_stream stream)
{
    if (!b) return set_error(GCOO_ERR_INVALID, "null batch");
    if (!b->committed) return set_error(GCOO_ERR_STATE, "batch not committed");
    if (int rc = check_generations(b)) return rc;
    if (!has_work(b)) return GCOO_OK;
    return enqueue_run(b, reinterpret_cast<hipStream_t>(stream));
}

int gcoo_batch_graph_capture(gcoo_batch_t b, gcoo_stream stream)
{
    if (!b) return set_error(GCOO_ERR_INVALID, "null batch");
    if (!b->committed) return set_error(GCOO_ERR_STATE, "batch not committed");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!s) return set_error(GCOO_ERR_INVALID, "graph capture needs a non-null stream");
    if (int rc = check_generations(b)) return rc;
    destroy_graph(b);
    GCOO_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = GCOO_OK;
    if (has_work(b)) rc = enqueue_run(b, s);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return set_error(GCOO_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    e = hipGraphInstantiate(&b->graphExec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        destroy_graph(b);
        return set_error(GCOO_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    return GCOO_OK;
}

int gcoo_batch_graph_launch(gcoo_batch_t b, gcoo_stream stream)
{
    if (!b) return set_error(GCOO_ERR_INVALID, "null batch");
    if (!b->graphExec) return set_error(GCOO_ERR_STATE, "no captured graph");
    if (int rc = check_generations(b)) return rc;
    GCOO_HIP_TRY(hipGraphLaunch(b->graphExec, reinterpret_cast<hipStream_t>(stream)));
    return GCOO_OK;
}

int gcoo_batch_info(gcoo_batch_t b, int *jobs, int64_t *items, int *grid_blocks)
{
    if (!b) return set_error(GCOO_ERR_INVALID, "null batch");
    if (jobs) *jobs = (int)b->jobs.size();
    if (items) *items = b->itemStart.back();
    if (grid_blocks) *grid_blocks = b->gridBlocks;
    return GCOO_OK;
}

int gcoo_batch_set_panel_mode(gcoo_batch_t b, int mode, double ratio_f32, double ratio_bf16)
{
    if (!b) return set_error(GCOO_ERR_INVALID, "null batch");
    if (mode < 0 || mode > 2) return set_error(GCOO_ERR_INVALID, "panel mode %d (0, 1 or 2)", mode);
    if (!(ratio_f32 >= 0.0 && ratio_f32 <= 1e6 && ratio_bf16 >= 0.0 && ratio_bf16 <= 1e6))
        return set_error(GCOO_ERR_INVALID, "panel ratios %g, %g", ratio_f32, ratio_bf16);
    b->panelMode = mode;
    b->panelRatio[0] = ratio_f32 > 0.0 ? ratio_f32 : kPanelRatioF32;
    b->panelRatio[1] = ratio_bf16 > 0.0 ? ratio_bf16 : kPanelRatioBf16;
    b->committed = false;
    destroy_graph(b);
    return GCOO_OK;
}

int gcoo_batch_panel_info(gcoo_batch_t b, int *panel_jobs, int *workgroups)
{
    if (!b) return set_error(GCOO_ERR_INVALID, "null batch");
    if (panel_jobs) *panel_jobs = b->committed ? (int)b->pjobs.size() : 0;
    if (workgroups) *workgroups = b->committed ? b->nUnitsF32 + b->nUnitsBf16 : 0;
    return GCOO_OK;
}

int gcoo_batch_job_is_panel(gcoo_batch_t b, int job, int *panel)
{
    if (!b || job < 0 || job >= (int)b->jobs.size() || !panel)
        return set_error(GCOO_ERR_INVALID, "bad batch, job index or output");
    *panel = b->committed && job < (int)b->panelOf.size() && b->panelOf[job] >= 0;
    return GCOO_OK;
}

int gcoo_batch_job_tile(gcoo_batch_t b, int job, int *rows, int *vec, int *lanes)
{
    if (!b || job < 0 || job >= (int)b->jobs.size())
        return set_error(GCOO_ERR_INVALID, "bad batch or job index");
    if (rows) *rows = b->jobs[job].rows;
    if (vec) *vec = b->jobs[job].vec;
    if (lanes) *lanes = b->jobs[job].lanes;
    return GCOO_OK;
}

}  // extern "C"
