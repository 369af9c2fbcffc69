// gcoo_api.hip -- the extern "C" boundary of libgcoospdm.so
// (declared in include/gcoo_spdm.h).
#include "gcoo_internal.h"

#include <climits>
#include <cstdarg>
#include <cstdio>
#include <new>

namespace gcoo {

namespace {
thread_local char g_err[512] = "";
}

const char *last_error_message() { return g_err; }

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

void clear_error() { g_err[0] = '\0'; }

int ensure_capacity(gcoo_matrix *h, int64_t need)
{
    if (need <= h->cap) return GCOO_OK;
    // Slack of 1/16 so that re-converting a same-sized matrix of similar
    // density does not reallocate.
    int64_t cap = need + need / 16 + 1024;
    if (cap > (int64_t)INT32_MAX) cap = INT32_MAX;
    if (h->vals) GCOO_HIP_TRY(hipFree(h->vals));
    if (h->rowcol) GCOO_HIP_TRY(hipFree(h->rowcol));
    if (h->seg) GCOO_HIP_TRY(hipFree(h->seg));
    if (h->seg16) GCOO_HIP_TRY(hipFree(h->seg16));
    if (h->plist) GCOO_HIP_TRY(hipFree(h->plist));
    h->plist = nullptr;
    invalidate_panel_lists(h);
    h->vals = nullptr;
    h->rowcol = nullptr;
    h->seg = nullptr;
    h->seg16 = nullptr;
    h->seg16_valid = false;
    h->cap = 0;
    if (hipMalloc(&h->vals, sizeof(float) * (size_t)cap) != hipSuccess ||
        hipMalloc(&h->rowcol, sizeof(int) * 2 * (size_t)cap) != hipSuccess ||
        hipMalloc(&h->seg, sizeof(int2) * (size_t)cap) != hipSuccess) {
        (void)hipGetLastError();
        if (h->vals) (void)hipFree(h->vals);
        if (h->rowcol) (void)hipFree(h->rowcol);
        h->vals = nullptr;
        h->rowcol = nullptr;
        h->seg = nullptr;
        return set_error(GCOO_ERR_ALLOC, "cannot allocate GCOO arrays for %lld entries",
                         (long long)cap);
    }
    h->cap = cap;
    return GCOO_OK;
}

int ensure_seg16(gcoo_matrix *h, hipStream_t s)
{
    if (!h->seg16_valid) {
        // allocated even without entries (a converted handle's capacity is
        // never zero): the batch's BF16 row streams read entry 0 of an
        // empty row's array, unused, so that their loads need no guard
        if (!h->seg16 && h->cap > 0) {
            if (hipMalloc(&h->seg16, sizeof(int2) * (size_t)h->cap) != hipSuccess) {
                (void)hipGetLastError();
                h->seg16 = nullptr;
                return set_error(GCOO_ERR_ALLOC, "cannot allocate the BF16 segments");
            }
        }
        if (h->nnz > 0) {
            int rc = launch_seg_to_bf16(h->seg, h->seg16, h->nnz, s);
            if (rc) return rc;
        }
    }
    h->seg16_valid = true;
    return GCOO_OK;
}

}  // namespace gcoo

using namespace gcoo;

namespace {

inline hipStream_t as_stream(gcoo_stream s) { return reinterpret_cast<hipStream_t>(s); }

// Internal result of status_check: the single-pass conversion stalled.
constexpr int kStalled = -1;

// The conversion status word's readback (the replacement of the reference's
// nnz readback, src/ShGCOOSpDM.cu:83-90) in two halves: status_copy queues
// the small D2H copy and an event behind it; status_check waits for that
// event only -- work queued after the copy (the multiply) keeps running --
// and returns kStalled after a stalled single-pass conversion, having queued
// the clearing of its abort word.
int status_copy(gcoo_handle h, hipStream_t s)
{
    GCOO_HIP_TRY(hipMemcpyAsync(h->h_status, h->d_status, sizeof(ConvStatus),
                                hipMemcpyDeviceToHost, s));
    GCOO_HIP_TRY(hipEventRecord(h->statusEvent, s));
    return GCOO_OK;
}

// The status copy and the multiply behind a conversion.  When the multiply
// is the panel kernel, the panel-list kernel (rebuilt after every
// conversion, first on the stream) copies the status word to the host and
// its launch records the event, so no separate copy sits between the
// conversion and the multiply; otherwise the copy goes first.
int status_and_multiply(gcoo_handle h, const float *D, int N, float *C, hipStream_t s)
{
    int variant = 0, vec = 0, waves = 0, rows = 0;
    if (h->h_status_dev &&
        spdm_choice(h, 0, D, N, C, &variant, &vec, &waves, &rows) == GCOO_OK && variant == 6) {
        h->statusExport = h->h_status_dev;
        h->statusExported = false;
        const int rc = launch_spdm_f32(h, D, N, C, h->d_status, s);
        h->statusExport = nullptr;
        if (rc) return rc;
        // no lists kernel ran (lists still valid, or no rows): copy after
        return h->statusExported ? GCOO_OK : status_copy(h, s);
    }
    int rc = status_copy(h, s);
    if (!rc) rc = launch_spdm_f32(h, D, N, C, h->d_status, s);
    return rc;
}

int status_check(gcoo_handle h, hipStream_t s)
{
    GCOO_HIP_TRY(hipEventSynchronize(h->statusEvent));
    if (h->h_status->stall) {
        GCOO_HIP_TRY(hipMemsetAsync(&h->d_status->stall, 0, sizeof(int), s));
        return kStalled;
    }
    if (h->h_status->overflow)
        return set_error(GCOO_ERR_OVERFLOW, "nnz %lld does not fit in int32",
                         (long long)h->h_status->total);
    return GCOO_OK;
}

// Enqueue count + scan + scatter with the current capacity; when the scan's
// total exceeds it, grow and scatter again.  On return the conversion is
// complete and h->nnz is set.  When D is given, the multiply is enqueued
// right behind the status copy (guarded by the status word) and the host
// waits for the copy only, so a conversion that fits the existing capacity
// returns while its multiply runs (as the reference returns with its kernel
// launched); the caller synchronises the stream before reading C.  When the
// single-pass kernel's look-back gives up (st->stall), the kernels queued
// behind it skip their work and the conversion runs again on the count /
// scan / scatter kernels, whose launches always complete (the reference's
// three launches, src/ShGCOOSpDM.cu:44-111).
int convert_impl(gcoo_handle h, const float *S, const float *D, int N, float *C,
                 hipStream_t s, bool allow_fused)
{
    h->converted = false;
    h->seg16_valid = false;
    ++h->generation;
    invalidate_panel_lists(h);
    // one pass over S where the single-pass kernel applies, else count + scan
    // then scatter
    const bool fused = allow_fused && fused_conversion_ok(h, S);
    int rc = fused ? launch_convert_fused(h, S, s) : launch_count_and_scan(h, S, s);
    if (rc) return rc;
    if (h->cap > 0 && !fused && (rc = launch_scatter(h, S, s))) return rc;
    rc = (h->cap > 0 && D) ? status_and_multiply(h, D, N, C, s) : status_copy(h, s);
    if (rc) return rc;
    rc = status_check(h, s);
    if (rc == kStalled) return convert_impl(h, S, D, N, C, s, false);
    if (rc) return rc;
    const int64_t total = h->h_status->total;
    const int64_t need = total + h->h_status->maxRow + kSegSlack;
    if (need > h->cap || h->cap == 0) {
        if ((rc = ensure_capacity(h, need))) return rc;
        h->nnz = total;  // the multiply's automatic variant choice reads it
        h->maxRow = h->h_status->maxRow;
        if ((rc = fused ? launch_convert_fused(h, S, s) : launch_scatter(h, S, s))) return rc;
        rc = D ? status_and_multiply(h, D, N, C, s) : status_copy(h, s);
        if (rc) return rc;
        rc = status_check(h, s);
        if (rc == kStalled) return convert_impl(h, S, D, N, C, s, false);
        if (rc) return rc;
    }
    h->nnz = total;
    h->maxRow = h->h_status->maxRow;
    h->converted = true;
    return GCOO_OK;
}

int convert_and_maybe_multiply(gcoo_handle h, const float *S, const float *D, int N, float *C,
                               hipStream_t s)
{
    return convert_impl(h, S, D, N, C, s, true);
}

int check_spdm_args(gcoo_handle h, const float *D, int K, int N, float *C)
{
    if (!h) return set_error(GCOO_ERR_INVALID, "null handle");
    if (K != h->K)
        return set_error(GCOO_ERR_INVALID, "dense operand has %d rows, sparse has %d columns",
                         K, h->K);
    if (N < 0) return set_error(GCOO_ERR_INVALID, "N = %d", N);
    if (N > 0 && h->M > 0 && (!D && h->K > 0)) return set_error(GCOO_ERR_INVALID, "null D");
    if (N > 0 && h->M > 0 && !C) return set_error(GCOO_ERR_INVALID, "null C");
    if ((long long)h->M * N > (long long)INT32_MAX * 64)
        return set_error(GCOO_ERR_OVERFLOW, "output too large");
    return GCOO_OK;
}

}  // namespace

struct gcoo_timer {
    hipEvent_t a = nullptr, b = nullptr;
};

extern "C" {

const char *gcoo_last_error(void) { return last_error_message(); }

const char *gcoo_version(void) { return "gcoospdm-mi355x 0.1 (gfx950)"; }

int gcoo_device_count(int *count)
{
    if (!count) return set_error(GCOO_ERR_INVALID, "null count");
    GCOO_HIP_TRY(hipGetDeviceCount(count));
    return GCOO_OK;
}

int gcoo_set_device(int device)
{
    GCOO_HIP_TRY(hipSetDevice(device));
    return GCOO_OK;
}

int gcoo_dev_malloc(void **ptr, size_t bytes)
{
    if (!ptr) return set_error(GCOO_ERR_INVALID, "null ptr");
    *ptr = nullptr;
    if (bytes == 0) return GCOO_OK;
    if (hipMalloc(ptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return set_error(GCOO_ERR_ALLOC, "hipMalloc of %zu bytes failed", bytes);
    }
    return GCOO_OK;
}

int gcoo_dev_free(void *ptr)
{
    if (ptr) GCOO_HIP_TRY(hipFree(ptr));
    return GCOO_OK;
}

int gcoo_memcpy_h2d(void *dst, const void *src, size_t bytes, gcoo_stream s)
{
    if (bytes == 0) return GCOO_OK;
    if (!dst || !src) return set_error(GCOO_ERR_INVALID, "null pointer");
    GCOO_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(s)));
    GCOO_HIP_TRY(hipStreamSynchronize(as_stream(s)));
    return GCOO_OK;
}

int gcoo_memcpy_d2h(void *dst, const void *src, size_t bytes, gcoo_stream s)
{
    if (bytes == 0) return GCOO_OK;
    if (!dst || !src) return set_error(GCOO_ERR_INVALID, "null pointer");
    GCOO_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(s)));
    GCOO_HIP_TRY(hipStreamSynchronize(as_stream(s)));
    return GCOO_OK;
}

int gcoo_memset_async(void *dst, int value, size_t bytes, gcoo_stream s)
{
    if (bytes == 0) return GCOO_OK;
    if (!dst) return set_error(GCOO_ERR_INVALID, "null pointer");
    GCOO_HIP_TRY(hipMemsetAsync(dst, value, bytes, as_stream(s)));
    return GCOO_OK;
}

int gcoo_stream_create(gcoo_stream *s)
{
    if (!s) return set_error(GCOO_ERR_INVALID, "null stream out");
    hipStream_t st;
    GCOO_HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *s = st;
    return GCOO_OK;
}

int gcoo_stream_destroy(gcoo_stream s)
{
    if (s) GCOO_HIP_TRY(hipStreamDestroy(as_stream(s)));
    return GCOO_OK;
}

int gcoo_stream_sync(gcoo_stream s)
{
    GCOO_HIP_TRY(hipStreamSynchronize(as_stream(s)));
    return GCOO_OK;
}

int gcoo_device_sync(void)
{
    GCOO_HIP_TRY(hipDeviceSynchronize());
    return GCOO_OK;
}

int gcoo_timer_create(gcoo_timer_t *t)
{
    if (!t) return set_error(GCOO_ERR_INVALID, "null timer out");
    auto *x = new (std::nothrow) gcoo_timer;
    if (!x) return set_error(GCOO_ERR_ALLOC, "timer");
    if (hipEventCreate(&x->a) != hipSuccess || hipEventCreate(&x->b) != hipSuccess) {
        delete x;
        return set_error(GCOO_ERR_HIP, "hipEventCreate failed");
    }
    *t = x;
    return GCOO_OK;
}

int gcoo_timer_destroy(gcoo_timer_t t)
{
    if (!t) return GCOO_OK;
    (void)hipEventDestroy(t->a);
    (void)hipEventDestroy(t->b);
    delete t;
    return GCOO_OK;
}

int gcoo_timer_start(gcoo_timer_t t, gcoo_stream s)
{
    if (!t) return set_error(GCOO_ERR_INVALID, "null timer");
    GCOO_HIP_TRY(hipEventRecord(t->a, as_stream(s)));
    return GCOO_OK;
}

int gcoo_timer_stop(gcoo_timer_t t, gcoo_stream s)
{
    if (!t) return set_error(GCOO_ERR_INVALID, "null timer");
    GCOO_HIP_TRY(hipEventRecord(t->b, as_stream(s)));
    return GCOO_OK;
}

int gcoo_timer_elapsed_ms(gcoo_timer_t t, float *ms)
{
    if (!t || !ms) return set_error(GCOO_ERR_INVALID, "null argument");
    GCOO_HIP_TRY(hipEventSynchronize(t->b));
    GCOO_HIP_TRY(hipEventElapsedTime(ms, t->a, t->b));
    return GCOO_OK;
}

int gcoo_fill_uniform_async(float *dst, int64_t n, uint64_t seed, uint64_t stream_id,
                            float scale, gcoo_stream s)
{
    if (n < 0 || (n > 0 && !dst)) return set_error(GCOO_ERR_INVALID, "bad fill args");
    return launch_fill(dst, n, seed, stream_id, scale, 0, false, as_stream(s));
}

int gcoo_fill_sparse_async(float *dst, int64_t n, uint64_t seed, uint64_t stream_id,
                           float scale, uint64_t keep_threshold, gcoo_stream s)
{
    if (n < 0 || (n > 0 && !dst)) return set_error(GCOO_ERR_INVALID, "bad fill args");
    return launch_fill(dst, n, seed, stream_id, scale, keep_threshold, true, as_stream(s));
}

int gcoo_create(int M, int K, int cpg, gcoo_handle *out)
{
    if (!out) return set_error(GCOO_ERR_INVALID, "null handle out");
    *out = nullptr;
    if (M < 0 || K < 0) return set_error(GCOO_ERR_INVALID, "negative shape %d x %d", M, K);
    if (cpg != 1 && cpg != 2 && cpg != 4)
        return set_error(GCOO_ERR_INVALID, "cpg must be 1, 2 or 4 (got %d)", cpg);
    auto *h = new (std::nothrow) gcoo_matrix;
    if (!h) return set_error(GCOO_ERR_ALLOC, "handle");
    h->M = M;
    h->K = K;
    h->cpg = cpg;
    h->nGroup = (M + cpg - 1) / cpg;
    (void)hipGetDevice(&h->device);
    if (hipMalloc(&h->idx, sizeof(int) * 2 * ((size_t)h->nGroup + 1)) != hipSuccess ||
        hipMalloc(&h->rowptr, sizeof(int) * ((size_t)M + 1)) != hipSuccess ||
        hipMalloc(&h->rowcnt, sizeof(int) * ((size_t)M + 1)) != hipSuccess ||
        hipMalloc(&h->d_status, sizeof(ConvStatus)) != hipSuccess ||
        hipHostMalloc(&h->h_status, sizeof(ConvStatus), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&h->statusEvent, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        gcoo_destroy(h);
        return set_error(GCOO_ERR_ALLOC, "cannot allocate GCOO index arrays");
    }
    if (hipHostGetDevicePointer(reinterpret_cast<void **>(&h->h_status_dev), h->h_status, 0) !=
        hipSuccess) {
        (void)hipGetLastError();
        h->h_status_dev = nullptr;  // the status is then always copied
    }
    if (hipMemset(h->d_status, 0, sizeof(ConvStatus)) != hipSuccess) {
        (void)hipGetLastError();
        gcoo_destroy(h);
        return set_error(GCOO_ERR_HIP, "cannot clear the status word");
    }
    *h->h_status = ConvStatus{};
    *out = h;
    return GCOO_OK;
}

int gcoo_destroy(gcoo_handle h)
{
    if (!h) return GCOO_OK;
    if (h->statusEvent) (void)hipEventDestroy(h->statusEvent);
    if (h->vals) (void)hipFree(h->vals);
    if (h->rowcol) (void)hipFree(h->rowcol);
    if (h->idx) (void)hipFree(h->idx);
    if (h->longList) (void)hipFree(h->longList);
    if (h->longCtr) (void)hipFree(h->longCtr);
    if (h->seg) (void)hipFree(h->seg);
    if (h->seg16) (void)hipFree(h->seg16);
    if (h->plist) (void)hipFree(h->plist);
    if (h->poff) (void)hipFree(h->poff);
    if (h->prow) (void)hipFree(h->prow);
    if (h->rowptr) (void)hipFree(h->rowptr);
    if (h->rowcnt) (void)hipFree(h->rowcnt);
    if (h->mask) (void)hipFree(h->mask);
    if (h->lb_flags) (void)hipFree(h->lb_flags);
    if (h->d_status) (void)hipFree(h->d_status);
    if (h->h_status) (void)hipHostFree(h->h_status);
    if (h->coo_tmp) (void)hipFree(h->coo_tmp);
    if (h->coo_rowof) (void)hipFree(h->coo_rowof);
    delete h;
    return GCOO_OK;
}

int gcoo_convert_dense(gcoo_handle h, const float *S, gcoo_stream s)
{
    if (!h) return set_error(GCOO_ERR_INVALID, "null handle");
    if (!S && (long long)h->M * h->K > 0) return set_error(GCOO_ERR_INVALID, "null S");
    return convert_and_maybe_multiply(h, S, nullptr, 0, nullptr, as_stream(s));
}

int gcoo_convert_coo(gcoo_handle h, const int *rows, const int *cols, const float *vals,
                     int64_t nnz, gcoo_stream s)
{
    if (!h) return set_error(GCOO_ERR_INVALID, "null handle");
    if (nnz < 0 || (nnz > 0 && (!rows || !cols || !vals)))
        return set_error(GCOO_ERR_INVALID, "bad COO arguments");
    if (nnz > INT32_MAX) return set_error(GCOO_ERR_OVERFLOW, "nnz %lld > int32", (long long)nnz);
    h->converted = false;
    h->seg16_valid = false;
    ++h->generation;
    invalidate_panel_lists(h);
    int rc = convert_coo(h, rows, cols, vals, nnz, as_stream(s));
    if (rc) return rc;
    h->converted = true;
    return GCOO_OK;
}

int gcoo_get_info(gcoo_handle h, int *M, int *K, int *cpg, int *nGroup, int64_t *nnz)
{
    if (!h) return set_error(GCOO_ERR_INVALID, "null handle");
    if (M) *M = h->M;
    if (K) *K = h->K;
    if (cpg) *cpg = h->cpg;
    if (nGroup) *nGroup = h->nGroup;
    if (nnz) *nnz = h->nnz;
    return GCOO_OK;
}

int gcoo_get_arrays(gcoo_handle h, const float **vals, const int **rows, const int **cols,
                    const int **groupIndex, const int **nnzPerGroup)
{
    if (!h) return set_error(GCOO_ERR_INVALID, "null handle");
    if (!h->converted) return set_error(GCOO_ERR_STATE, "matrix not converted");
    if (vals) *vals = h->vals;
    if (rows) *rows = h->rows();
    if (cols) *cols = h->cols();
    if (groupIndex) *groupIndex = h->groupIndex();
    if (nnzPerGroup) *nnzPerGroup = h->nnzPerGroup();
    return GCOO_OK;
}

int gcoo_spdm_f32_async(gcoo_handle h, const float *D, int K, int N, float *C, gcoo_stream s)
{
    int rc = check_spdm_args(h, D, K, N, C);
    if (rc) return rc;
    if (!h->converted) return set_error(GCOO_ERR_STATE, "matrix not converted");
    return launch_spdm_f32(h, D, N, C, nullptr, as_stream(s));
}

int gcoo_spdm_f32(gcoo_handle h, const float *D, int K, int N, float *C, gcoo_stream s)
{
    int rc = gcoo_spdm_f32_async(h, D, K, N, C, s);
    if (rc) return rc;
    GCOO_HIP_TRY(hipStreamSynchronize(as_stream(s)));
    return GCOO_OK;
}

int gcoo_f32_to_bf16_async(const float *x, uint16_t *y, int64_t n, gcoo_stream s)
{
    if (n < 0 || (n > 0 && (!x || !y))) return set_error(GCOO_ERR_INVALID, "bad arguments");
    return launch_f32_to_bf16(x, y, n, as_stream(s));
}

int gcoo_spdm_bf16_async(gcoo_handle h, const uint16_t *D, int K, int N, float *C,
                         gcoo_stream s)
{
    int rc = check_spdm_args(h, reinterpret_cast<const float *>(D), K, N, C);
    if (rc) return rc;
    if (!h->converted) return set_error(GCOO_ERR_STATE, "matrix not converted");
    if ((rc = ensure_seg16(h, as_stream(s)))) return rc;
    return launch_spdm_bf16(h, D, N, C, as_stream(s));
}

int gcoo_spdm_bf16(gcoo_handle h, const uint16_t *D, int K, int N, float *C, gcoo_stream s)
{
    int rc = gcoo_spdm_bf16_async(h, D, K, N, C, s);
    if (rc) return rc;
    GCOO_HIP_TRY(hipStreamSynchronize(as_stream(s)));
    return GCOO_OK;
}

int gcoo_set_spdm_config(gcoo_handle h, const gcoo_spdm_config *cfg)
{
    if (!h) return set_error(GCOO_ERR_INVALID, "null handle");
    if (!cfg) {
        h->spdm_cfg = gcoo_spdm_config{0, 0, 0, 0, 0};
        return GCOO_OK;
    }
    if (cfg->variant < 0 || cfg->variant > 6)
        return set_error(GCOO_ERR_INVALID, "variant %d", cfg->variant);
    if (cfg->vec != 0 && cfg->vec != 1 && cfg->vec != 2 && cfg->vec != 4 && cfg->vec != 8)
        return set_error(GCOO_ERR_INVALID, "vec %d (0, 1, 2, 4 or 8)", cfg->vec);
    if (cfg->waves < 0 || cfg->waves > 16)
        return set_error(GCOO_ERR_INVALID, "waves %d (0..16)", cfg->waves);
    if (cfg->chunk < 0) return set_error(GCOO_ERR_INVALID, "chunk %d", cfg->chunk);
    if (cfg->unroll != 0 && cfg->unroll != 8 && cfg->unroll != 16 && cfg->unroll != 32)
        return set_error(GCOO_ERR_INVALID, "unroll %d (0, 8, 16 or 32)", cfg->unroll);
    h->spdm_cfg = *cfg;
    return GCOO_OK;
}

int gcoo_set_panel_balance(gcoo_handle h, int mode)
{
    if (!h) return set_error(GCOO_ERR_INVALID, "null handle");
    if (mode < 0 || mode > 2) return set_error(GCOO_ERR_INVALID, "panel balance %d (0, 1 or 2)", mode);
    h->panel_balance = mode;
    return GCOO_OK;
}

int gcoo_get_spdm_config(gcoo_handle h, gcoo_spdm_config *cfg)
{
    if (!h || !cfg) return set_error(GCOO_ERR_INVALID, "null argument");
    *cfg = h->spdm_cfg;
    return GCOO_OK;
}

int gcoo_get_spdm_choice(gcoo_handle h, int dtype, const void *D, int N, const float *C,
                         int *variant, int *vec, int *waves, int *rows)
{
    if (!h || !variant || !vec || !waves || !rows)
        return set_error(GCOO_ERR_INVALID, "null argument");
    if (dtype != 0 && dtype != 1) return set_error(GCOO_ERR_INVALID, "dtype %d", dtype);
    if (N < 0) return set_error(GCOO_ERR_INVALID, "N = %d", N);
    return spdm_choice(h, dtype, D, N, C, variant, vec, waves, rows);
}

int gcoo_test_set_conversion_abort(gcoo_handle h, gcoo_stream s)
{
    if (!h) return set_error(GCOO_ERR_INVALID, "null handle");
    GCOO_HIP_TRY(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(&h->d_status->stall), 1, 1,
                                   as_stream(s)));
    GCOO_HIP_TRY(hipStreamSynchronize(as_stream(s)));
    return GCOO_OK;
}

int gcoo_multiply_by(gcoo_handle h, const float *S, const float *D, int K, int N, float *C,
                     gcoo_stream s)
{
    int rc = check_spdm_args(h, D, K, N, C);
    if (rc) return rc;
    if (!S && (long long)h->M * h->K > 0) return set_error(GCOO_ERR_INVALID, "null S");
    return convert_and_maybe_multiply(h, S, D, N, C, as_stream(s));
}

}  // extern "C"
