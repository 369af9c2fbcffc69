/*
 * gcoo_spdm.h -- C ABI of the MI355X GCOOSpDM library (libgcoospdm.so).
 *
 * The library implements the reference's hot path -- dense -> grouped-COO
 * (GCOO) conversion and the grouped-COO sparse x dense multiply -- as HIP
 * kernels for gfx950.  Every entry point takes plain pointers and sizes; all
 * matrix pointers are DEVICE pointers owned by the caller unless stated.
 * Storage is row-major FP32: S is the sparse operand (M x K), D the dense
 * operand (K x N), C = S * D (M x N); this is the product the reference
 * computes (reference src/main.cu:131 with src/utils.cpp:30-46).
 *
 * Errors: every function returns GCOO_OK (0) or a GCOO_ERR_* code and never
 * exits the process (the reference prints and exits, src/ShGCOOSpDM.cu via
 * checkCudaErrors); gcoo_last_error() gives the message of the calling
 * thread's last failure.
 *
 * Streams: gcoo_stream is a hipStream_t passed as void* (NULL = the null
 * stream).  Entry points named *_async only enqueue work; the others return
 * after the work is complete, like the reference's methods, which end in
 * cudaDeviceSynchronize.
 */
#ifndef GCOO_SPDM_H
#define GCOO_SPDM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gcoo_matrix *gcoo_handle;
typedef void *gcoo_stream;

enum gcoo_status {
    GCOO_OK = 0,
    GCOO_ERR_INVALID = 1,     /* bad argument (size, pointer, cpg, ...)     */
    GCOO_ERR_HIP = 2,         /* a HIP runtime call failed                  */
    GCOO_ERR_ALLOC = 3,       /* device or host allocation failed           */
    GCOO_ERR_OVERFLOW = 4,    /* nnz or an index does not fit in int32      */
    GCOO_ERR_UNSUPPORTED = 5, /* algorithm / option not available           */
    GCOO_ERR_BUDGET = 6,      /* request exceeds the serving HBM budget     */
    GCOO_ERR_STATE = 7        /* handle not in the state the call needs     */
};

/* Message of the calling thread's last error ("" if none). */
const char *gcoo_last_error(void);
/* Library version string. */
const char *gcoo_version(void);

/* ---- Device memory, copies, synchronisation and timing --------------------
 * For hosts that do not link the HIP runtime themselves.  They replace the
 * driver's direct CUDA runtime calls: cudaMalloc / cudaMemcpy / cudaFree
 * (reference src/main.cu:33-38,85,91-93) and the cudaEvent timing of the
 * repeat loop (src/main.cu:63-81). */
int gcoo_device_count(int *count);
int gcoo_set_device(int device);
int gcoo_dev_malloc(void **ptr, size_t bytes);
int gcoo_dev_free(void *ptr);
int gcoo_memcpy_h2d(void *dst, const void *src, size_t bytes, gcoo_stream s);
int gcoo_memcpy_d2h(void *dst, const void *src, size_t bytes, gcoo_stream s);
int gcoo_memset_async(void *dst, int value, size_t bytes, gcoo_stream s);
int gcoo_stream_create(gcoo_stream *s);
int gcoo_stream_destroy(gcoo_stream s);
int gcoo_stream_sync(gcoo_stream s);
int gcoo_device_sync(void);
typedef struct gcoo_timer *gcoo_timer_t;
int gcoo_timer_create(gcoo_timer_t *t);
int gcoo_timer_destroy(gcoo_timer_t t);
int gcoo_timer_start(gcoo_timer_t t, gcoo_stream s);
int gcoo_timer_stop(gcoo_timer_t t, gcoo_stream s);
/* Waits for the stop event and returns start->stop in milliseconds. */
int gcoo_timer_elapsed_ms(gcoo_timer_t t, float *ms);

/* ---- Synthetic inputs ------------------------------------------------------
 * Device-side generators for the benchmark workloads.  They play the role of
 * the driver's randomInit (reference src/utils.cpp:5-19) with a portable,
 * counter-based generator instead of glibc rand():
 *   h(i)  = mix64(key + (i + 1) * 0x9E3779B97F4A7C15),
 *   key   = mix64(seed * 0x9E3779B97F4A7C15 + stream_id),
 *   mix64 = the SplitMix64 finaliser;
 *   value = (2 * (h >> 40) - 2^24 + 1) * 2^-24 * scale   (odd numerator, so
 *           never 0, uniform in (-scale, scale));
 *   kept  = (h & (2^40 - 1)) >= keep_threshold           (Bernoulli mask).
 * keep_threshold = floor(sparsity * 2^40); an element is zero when not kept. */
int gcoo_fill_uniform_async(float *dst, int64_t n, uint64_t seed,
                            uint64_t stream_id, float scale, gcoo_stream s);
int gcoo_fill_sparse_async(float *dst, int64_t n, uint64_t seed,
                           uint64_t stream_id, float scale,
                           uint64_t keep_threshold, gcoo_stream s);

/* ---- The GCOO matrix -------------------------------------------------------
 * A handle owns the grouped-COO arrays of one sparse matrix S (M x K) with
 * cpg consecutive rows per group (nGroup = ceil(M / cpg)), in the reference
 * layout (reference src/ShGCOOSpDM.h:25-32, src/ShGCOOSpDM.cu:33-35,94-97):
 *   vals        f32[nnz]   S[r][k] bit copies,
 *   rows, cols  i32[nnz]   absolute row r and column k,
 *   groupIndex  i32[nGroup+1], exclusive scan of nnzPerGroup (last = nnz),
 *   nnzPerGroup i32[nGroup+1], last entry 0,
 * entries of group g stored at [groupIndex[g], groupIndex[g+1]) sorted by
 * (k, r); S[r][k] is a nonzero iff S[r][k] != 0.0 (so -0.0 is dropped and NaN
 * kept).  cpg must be 1, 2 or 4 (the reference's COLUMN_PER_GROUP range,
 * src/constants.h:4 and src/gcoo_spdm_kernels.cu:141). */

/* Empty handle for an M x K sparse matrix.  Replaces the ShGCOOSpDM
 * constructor (reference src/ShGCOOSpDM.cu:13-17). */
int gcoo_create(int M, int K, int cpg, gcoo_handle *out);
int gcoo_destroy(gcoo_handle h);

/* Dense -> GCOO: per-group count, wave64/LDS exclusive scan, scatter.
 * Replaces ShGCOOSpDM::convertToGroupCOOFormat and build()
 * (reference src/ShGCOOSpDM.cu:19-126,182-193; kernels
 * src/gcoo_spdm_kernels.cu:4-90,160-168).  S is a device pointer to the
 * dense row-major M x K matrix.  Synchronous. */
int gcoo_convert_dense(gcoo_handle h, const float *S, gcoo_stream s);

/* COO -> GCOO for matrices too large to hold densely (the reference's int
 * offsets overflow at M*K >= 2^31, src/gcoo_spdm_kernels.cu:16,57).  rows,
 * cols, vals are device arrays of nnz entries in any order, duplicates not
 * allowed; entries with value == 0.0 are dropped as in the dense path.
 * The result is identical to gcoo_convert_dense of the equivalent dense S. */
int gcoo_convert_coo(gcoo_handle h, const int *rows, const int *cols,
                     const float *vals, int64_t nnz, gcoo_stream s);

/* Shape and arrays of a converted handle (device pointers, owned by h). */
int gcoo_get_info(gcoo_handle h, int *M, int *K, int *cpg, int *nGroup,
                  int64_t *nnz);
int gcoo_get_arrays(gcoo_handle h, const float **vals, const int **rows,
                    const int **cols, const int **groupIndex,
                    const int **nnzPerGroup);

/* C = S * D with the converted S.  D is K x N, C is M x N, row-major FP32,
 * device pointers.  Every C[r][j] is one FP32 fmaf chain from +0.0f over
 * row r's nonzeros in ascending k: bit-identical to the reference kernel
 * sparse_dense_groupcoo_mat_mul_kernel (src/gcoo_spdm_kernels.cu:92-158).
 * Replaces ShGCOOSpDM::runKernelTest (src/ShGCOOSpDM.cu:195-214). */
int gcoo_spdm_f32_async(gcoo_handle h, const float *D, int K, int N, float *C,
                        gcoo_stream s);
int gcoo_spdm_f32(gcoo_handle h, const float *D, int K, int N, float *C,
                  gcoo_stream s);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): gcoo_spdm_f32_async with the same kernels and the same
 * bits, but calling no hipMalloc / hipFree.  The scratch its kernels need --
 * the panel lists of the LDS-staged kernel, rebuilt on every call, and the
 * long-row list and counter -- lives in one 16 MiB device array per GPU,
 * held by the library until the process ends.  A multiply whose scratch
 * would exceed 16 MiB (panel lists take 8 bytes per entry, so about 2 M
 * entries) returns GCOO_ERR_BUDGET and enqueues nothing.  Variant 2 needs
 * only the long-row list when its long rows go first (4 bytes per row);
 * without room for it, it keeps the natural row order (the same bits).  All calls on one GPU share that array: calls that overlap on a
 * GPU (other streams, other threads) race, so the caller serialises them. */
int gcoo_spdm_f32_noalloc_async(gcoo_handle h, const float *D, int K, int N, float *C,
                                gcoo_stream s);

/* BF16 storage form (no counterpart in the reference, which is FP32 only):
 * S's entry values and D stored as BF16 -- the upper 16 bits of an FP32,
 * rounded to nearest-even by gcoo_f32_to_bf16_async -- widened exactly to
 * FP32 and accumulated with the same per-output fmaf chain (k ascending) in
 * FP32; C is FP32.  Moves half the bytes of D.  The BF16 values of S are
 * derived from the converted FP32 entries on first use after a conversion.
 * D is K x N BF16 (uint16 bits), device pointer. */
int gcoo_f32_to_bf16_async(const float *x, uint16_t *y, int64_t n, gcoo_stream s);
int gcoo_spdm_bf16_async(gcoo_handle h, const uint16_t *D, int K, int N, float *C,
                         gcoo_stream s);
int gcoo_spdm_bf16(gcoo_handle h, const uint16_t *D, int K, int N, float *C,
                   gcoo_stream s);

/* Multiply-kernel tuning of one handle (the reference fixes these at compile
 * time: COLUMN_PER_GROUP / BS in src/constants.h:4-5, grid-searched by
 * src/autotune.py:61-86).  Zero fields mean "choose automatically".
 *   variant 1  group order: walks the group's (k, r) list with one wave per
 *              group and per-row accumulators;
 *   variant 2  row segments: walks each row's k-sorted segment with one
 *              accumulator set per row (no per-entry row dispatch);
 *   variant 3  LDS-staged: a workgroup stages k-chunks of a D column panel
 *              in LDS, shared by all rows of a row panel.
 *   variant 4  long rows split off: rows with >= chunk entries (default
 *              256) go to a persistent kernel in which a workgroup stages the
 *              D rows of 32 entries at a time in LDS and walks the row's
 *              chain from there; the others take variant 2.
 *   variant 5  group rows: one wave walks the 4 rows of a row group at once,
 *              16 lanes per row, each lane loading 16 bytes of D per entry.
 *   variant 6  panel-staged: a workgroup owns a panel of rows (waves x
 *              rows per wave: 16 x 16, 16 x 21, 12 x 28 or, FP32 only,
 *              8 x 42) x 256 columns; k-chunks of the D tile are staged in
 *              LDS once and serve all the panel's rows (needs N * element
 *              bytes % 16 == 0, a 16-byte aligned D and C; otherwise
 *              variant 2 runs).
 *   vec        columns per lane (1, 2, 4; 8 for BF16), at most 16 bytes of
 *              D per lane; a wave covers 64*vec columns.
 *   waves      waves per workgroup (1..16) of variant 3; for variant 6 16,
 *              12 or (FP32) 8 (0: automatic, 16 x 16 or 16 x 21 by the grid's
 *              rounds of one workgroup per CU); for variant 2 the long-rows-
 *              first order: rows of 2^waves entries or more (waves 2..16)
 *              are listed on the device and walked by the first workgroups
 *              of each column tile, so that the longest serial chains start
 *              first (1: natural row order; 0: automatic, on when the
 *              longest row holds 512 entries and 16 times the mean).
 *   chunk      k-chunk length of variant 3; long-row threshold of variant 4;
 *              rows per wave of variant 2 (1: one row, another nonzero value:
 *              the group's cpg rows, 0: one row once rows average 12 entries);
 *              rows per wave of variant 6 (16 or 21 with 16 waves, 28 with
 *              12, 42 with 8; 0: 16, 28, 42 by the waves).
 *   unroll     entries whose D loads are in flight per wave in variant 2
 *              (8, 16 or 32; 0 means 16).
 * All variants give bit-identical C. */
typedef struct gcoo_spdm_config {
    int variant;
    int vec;
    int waves;
    int chunk;
    int unroll;
} gcoo_spdm_config;
int gcoo_set_spdm_config(gcoo_handle h, const gcoo_spdm_config *cfg);
int gcoo_get_spdm_config(gcoo_handle h, gcoo_spdm_config *cfg);
/* Row order of the panel-staged multiply (variant 6).  Its workgroup meets
 * at a barrier every k chunk, so a chunk lasts as long as its busiest wave;
 * the lists can deal each panel's rows to the waves so that every chunk's
 * entry counts come out even (the multiply at n = 8000, s = 0.99: 405.6 ->
 * 397.0 us), at the price of an extra kernel per list build (about 0.25 ms
 * there), so it pays only for lists that serve many multiplies.
 * mode 0 (default): rows in order; 1: balanced when the lists are built for
 * a multiply of an already converted matrix (gcoo_spdm_*), whose lists
 * outlive the call, in order inside gcoo_multiply_by; 2: always balanced.
 * Batches always balance their panel jobs' lists (built once at commit).
 * The bits of C do not depend on it.  No counterpart in the reference. */
int gcoo_set_panel_balance(gcoo_handle h, int mode);
/* The variant (1..6), columns per lane and, for variant 6, waves per
 * workgroup and rows per wave (else 0 and 0) that a multiply of h
 * with these operands (dtype 0: gcoo_spdm_f32*, 1: gcoo_spdm_bf16*) runs: the
 * configuration, or the automatic choice from the last conversion's density
 * and the shape, after the fall-backs of each variant's operand
 * requirements.  No counterpart in the reference (one fixed kernel); it
 * lets measurement tools name the kernel they time.  D and C are only
 * inspected for alignment. */
int gcoo_get_spdm_choice(gcoo_handle h, int dtype, const void *D, int N, const float *C,
                         int *variant, int *vec, int *waves, int *rows);

/* Convert S then multiply: ShGCOOSpDM::mutiplyBy (src/ShGCOOSpDM.cu:131-180),
 * without the reference's per-call allocations once capacity is reached.
 * Returns once the conversion has been checked (its status read back, as
 * the reference reads nnz back); the multiply may still be running on s, as
 * the reference's kernel is when mutiplyBy returns: synchronise s before
 * reading C, and issue later calls on the handle on the same stream. */
int gcoo_multiply_by(gcoo_handle h, const float *S, const float *D, int K,
                     int N, float *C, gcoo_stream s);

/* Testing hook, no counterpart in the reference: raises the abort word of
 * the single-pass dense conversion, as a group whose look-back wait gave up
 * does, so that the next gcoo_convert_dense / gcoo_multiply_by takes the
 * fallback (count, scan and scatter kernels, the reference's three launches
 * src/ShGCOOSpDM.cu:44-111) and must still return the same arrays.  Every
 * kernel still exits normally. */
int gcoo_test_set_conversion_abort(gcoo_handle h, gcoo_stream s);

/* ---- The "csr" algorithm -------------------------------------------------
 * The reference's row-wise CSR SpDM ablation, factory key "csr" (reference
 * src/ShCsrSpGemm.h, src/ShCsrSpGemm.cu:9-146, src/SpGemmBuilder.h:12-23).
 * A CSR matrix holds the standard three arrays of S (M x K): rowPtr[M + 1],
 * colInd[nnz], vals[nnz], each row's entries in ascending column order (what
 * the reference's cusparseSnnz + cusparseSdense2csr produce,
 * src/ShCsrSpGemm.cu:67-110; the same entries as the GCOO conversion:
 * S[r][k] != 0.0).  The multiply follows the reference kernel
 * (src/ShCsrSpGemm.cu:9-43): one workgroup per (row, 256-column tile), the
 * row's entries staged in LDS, one output column per thread; each C element
 * is the fmaf chain over the row's entries in ascending k, the same bits as
 * gcoo_spdm_f32.  No transposes of D and C (the reference makes them with
 * cublasSgeam because its kernel reads column-major operands). */
typedef struct gcoo_csr *gcoo_csr_t;
int gcoo_csr_create(int M, int K, gcoo_csr_t *out);
int gcoo_csr_destroy(gcoo_csr_t c);
/* Dense S (device, row-major M x K) -> CSR.  Synchronous. */
int gcoo_csr_convert_dense(gcoo_csr_t c, const float *S, gcoo_stream s);
/* Device arrays of the last conversion (owned by c). */
int gcoo_csr_get_arrays(gcoo_csr_t c, const int **rowPtr, const int **colInd,
                        const float **vals, int64_t *nnz);
/* C = S * D with D K x N and C M x N (device, row-major). */
int gcoo_csr_spmm_f32(gcoo_csr_t c, const float *D, int N, float *C, gcoo_stream s);
int gcoo_csr_spmm_f32_async(gcoo_csr_t c, const float *D, int N, float *C, gcoo_stream s);

/* ---- Batches: many independent multiplies in one persistent launch -------
 * No counterpart in the reference (one multiply per launch,
 * src/ShGCOOSpDM.cu:153-167).  A batch holds jobs C_i = S_i * D_i, each S_i
 * a converted handle, D_i FP32 (dtype 0) or BF16 bits (dtype 1, the handle's
 * BF16 values are derived at commit), any shapes.  Each job is cut into work
 * items of `rows` rows x one column tile of `vec` columns per lane (0 = the
 * built-in choice; tuned values come from gcoospdm_amd/autotune.py).  A run
 * is the panel jobs' launches (below), a counter-reset kernel and one
 * persistent kernel whose waves pull work items until none are left; it
 * allocates nothing and synchronises nothing,
 * so it can be captured into a HIP graph and replayed.  Every C element is
 * the same fmaf chain as in gcoo_spdm_f32 / gcoo_spdm_bf16: results are
 * bit-identical to running each job alone.  The batch keeps the pointers it
 * was given; handles and operands must outlive its runs.  Commit binds each
 * job to its matrix's current conversion: once a job's matrix is converted
 * again (which may reallocate its arrays), run, capture and graph launch
 * return GCOO_ERR_STATE, launching nothing, until the batch is committed
 * (and captured) again. */
typedef struct gcoo_batch *gcoo_batch_t;
int gcoo_batch_create(gcoo_batch_t *b);
int gcoo_batch_destroy(gcoo_batch_t b);
int gcoo_batch_clear(gcoo_batch_t b);
int gcoo_batch_add(gcoo_batch_t b, gcoo_handle h, const void *D, int dtype, int N, float *C,
                   int rows, int vec, int *job_index);
/* Upload the job table (synchronous); required after the last add. */
int gcoo_batch_commit(gcoo_batch_t b, gcoo_stream s);
int gcoo_batch_run_async(gcoo_batch_t b, gcoo_stream s);
/* Capture one run into a HIP graph on stream s (non-null), then replay it. */
int gcoo_batch_graph_capture(gcoo_batch_t b, gcoo_stream s);
int gcoo_batch_graph_launch(gcoo_batch_t b, gcoo_stream s);
/* items: work items of the persistent kernel (panel jobs have none). */
int gcoo_batch_info(gcoo_batch_t b, int *jobs, int64_t *items, int *grid_blocks);
int gcoo_batch_job_tile(gcoo_batch_t b, int job, int *rows, int *vec, int *lanes);
/* Panel jobs.  Jobs dense enough for the panel-staged multiply (the LDS
 * kernel gcoo_spdm_f32 / gcoo_spdm_bf16 use from about 0.25 % / 0.5 %
 * density; 336-row panels x 256-column tiles, so N % 4 == 0 and 16-byte
 * aligned D and C) run it instead of the persistent kernel: one launch per
 * element type ahead of it, its units dealt over the XCDs by estimated work.
 * Commit chooses them (mode 1, the default: where a cost model puts the panel
 * kernel's time below `ratio_f32` / `ratio_bf16` times the persistent
 * kernel's, 0 meaning the built-in 1.0 / 1.0; 0: none; 2: every job whose
 * operands allow it) and builds their panel lists, which the handle keeps; a
 * later multiply of the same matrix with another panel shape rebuilds them,
 * after which run,
 * capture and graph launch return GCOO_ERR_STATE until the batch is
 * committed again.  A matrix in two jobs of different element types runs
 * the second on the persistent kernel.  Same bits either way. */
int gcoo_batch_set_panel_mode(gcoo_batch_t b, int mode, double ratio_f32,
                              double ratio_bf16);
/* After commit: panel jobs and the workgroups of their launches. */
int gcoo_batch_panel_info(gcoo_batch_t b, int *panel_jobs, int *workgroups);
int gcoo_batch_job_is_panel(gcoo_batch_t b, int job, int *panel);

/* ---- Serving: continuous batching of requests on resident operands -------
 * No counterpart in the reference.  Operands (a converted handle and its D,
 * FP32 or BF16) are registered once and stay resident.  Requests
 * "C = S_op * D_op" are submitted from any thread; a scheduler thread inside
 * the server admits them in arrival order while their outputs fit the HBM
 * budget (one device arena of hbm_budget_bytes, first fit; a request larger
 * than the arena is rejected with GCOO_ERR_BUDGET), and whenever one of
 * `slots` streams is idle launches all admitted requests (at most
 * max_batch_jobs) as one persistent batch on it, so requests arriving while
 * other batches run are batched on another stream and overlap on the device.
 * With max_batch_jobs == 1 the server is a plain FIFO instead: each request
 * is one launch of the plain multiply (gcoo_spdm_f32_async /
 * gcoo_spdm_bf16_async) on a single stream, in arrival order, with up to 64
 * launches in flight (`slots` is not used).  A thread per stream waits on
 * its launches' completion events, so completion is event-driven.
 * Latency is arrival (submit) to completion.  Results are bit-identical to
 * the same multiply run alone; they stay in the arena until released.
 * Failures: when a launch cannot be enqueued or a launched batch reports an
 * error, its requests and every request not yet launched fail with that
 * error code (gcoo_server_wait and gcoo_server_result return it, release
 * frees them), and later submits are refused with it.  A registered handle
 * must not be converted again while the server may run its requests. */
typedef struct gcoo_server *gcoo_server_t;
typedef struct gcoo_server_config_s {
    size_t hbm_budget_bytes; /* output arena size: the admission budget */
    int slots;               /* concurrent batches / streams, 1..8 (2 serves best
                                on one MI355X: with more streams than the
                                process's hardware queues, completions stall) */
    int max_batch_jobs;      /* requests per launched batch */
} gcoo_server_config;
typedef struct gcoo_server_stats_s {
    uint64_t completed;
    uint64_t batches;
    double p50_ms, p99_ms, mean_ms, max_ms;
    double wall_s;           /* first arrival to last completion */
    size_t peak_bytes;       /* most arena bytes in use at once */
    size_t budget_bytes;
} gcoo_server_stats_t;
int gcoo_server_create(const gcoo_server_config *cfg, gcoo_server_t *s);
int gcoo_server_destroy(gcoo_server_t s);
/* rows / vec: batch tile of this operand's requests (0 = automatic). */
int gcoo_server_register(gcoo_server_t s, gcoo_handle h, const void *D, int dtype, int N,
                         int rows, int vec, int *op_id);
int gcoo_server_submit(gcoo_server_t s, int op_id, uint64_t *req_id);
int gcoo_server_wait(gcoo_server_t s, uint64_t req_id, double *latency_ms);
/* Device pointer to the completed request's C (M x N FP32, row-major). */
int gcoo_server_result(gcoo_server_t s, uint64_t req_id, const float **C, int *M, int *N);
int gcoo_server_release(gcoo_server_t s, uint64_t req_id);
int gcoo_server_stats(gcoo_server_t s, gcoo_server_stats_t *st);
int gcoo_server_reset_stats(gcoo_server_t s);
/* Testing hook, no counterpart in the reference: the server's next launch
 * fails with `code` (a GCOO_ERR_* value) before anything is enqueued, as a
 * failing launch call would, so that tests can check the failure path
 * without making the GPU fail. */
int gcoo_test_server_inject_failure(gcoo_server_t s, int code);

#ifdef __cplusplus
}
#endif
#endif
