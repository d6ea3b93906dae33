/*
 * trident_hip.h — C-ABI of the MI355X (gfx950) local SpGEMM path.
 *
 * This is the drop-in boundary for the single-GPU share of Trident's
 * SpGEMM: the local tile multiply, the bitmask tile compression and the
 * stage-wise partial-C accumulation.  Plain C: device pointers, sizes and
 * status codes, no C++ or HIP types in the signatures (the stream is passed
 * as void*).  The C++ shim include/KokkosWrap.hpp rebuilds the reference's
 * KokkosWrap::LocalMatrix interface on top of it.
 *
 * Matrices are int32 CSR/CSC with fp64 values (the reference instantiates
 * <int32_t,float>, src/hns_spgemm.cu:1116-1118; fp64 is the precision this
 * path is built for).  Outputs produced by the library are stored in the
 * reference's contiguous tile layout (include/tile_holder.cuh:74-89):
 *     buf = [ double val[nnz] | int32 idx[nnz] | int32 ptr[ptrdim+1] ]
 * with contig = 1, buf/buf_size set, and are owned by the context until
 * released with tr_csx_free.
 *
 * Semantics (the contract this path defines where the reference delegates
 * to Kokkos-Kernels 4.7.01 / cuSPARSE):
 *  - output columns are sorted ascending in every row;
 *  - structural zeros are kept (an entry exists iff a product lands on it);
 *  - C = A*B accumulates, for every output, the products in the order of
 *    A's stored entries: c_ij = fma(a_ik, b_kj, c_ij) from +0.0, so results
 *    are bit-identical run to run, alone or batched;
 *  - inputs of spgemm/spadd need sorted columns within rows (validated; an
 *    unsorted operand is rejected with TR_ERR_INVALID and tr_csx_sort_rows
 *    sorts one in place).
 *
 * Every call is ordered on the context's stream.  Calls that return sizes
 * (spgemm, spadd, compress, symbolic) synchronise that stream once to read
 * them, like the reference's own host round trips.  A context must not be
 * used by two host threads at once (the reference keeps one stream per
 * thread, include/common.h:13).  Errors are returned, never exit().
 */
#ifndef TRIDENT_HIP_H
#define TRIDENT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TR_OK            0
#define TR_ERR_INVALID   1   /* bad arguments or malformed/unsorted CSR   */
#define TR_ERR_OOM       2   /* device allocation failed                  */
#define TR_ERR_OVERFLOW  3   /* result nnz exceeds INT32_MAX (int32 tile) */
#define TR_ERR_HIP       4   /* a HIP runtime call failed                 */

/* majordim: mirrors mmio::MajorDim (ROWS = CSR, COLS = CSC). */
#define TR_ROWS 0
#define TR_COLS 1

/* Mirrors mmio::CSX (fields as used at include/KokkosWrap.hpp:95-104,
 * include/tile_holder.cuh:215-226, include/utils.cuh:455-466).  All
 * pointers are device pointers.  ptr_vec == NULL marks an uninitialised
 * matrix (LocalMatrix::initialized == false, KokkosWrap.hpp:49,208). */
typedef struct tr_csx {
    int32_t  majordim;  /* TR_ROWS or TR_COLS                                  */
    int32_t  nrows;
    int32_t  ncols;
    int32_t  contig;    /* 1 if val|idx|ptr live in one buffer `buf`            */
    int64_t  nnz;
    int32_t *ptr_vec;   /* (majordim == ROWS ? nrows : ncols) + 1 entries       */
    int32_t *idx_vec;   /* nnz                                                  */
    double  *val;       /* nnz                                                  */
    void    *buf;       /* base of the contiguous buffer (NULL if not contig)   */
    size_t   buf_size;  /* bytes of buf                                         */
} tr_csx_t;

typedef struct tr_ctx *tr_ctx_t;

/* ---- context ------------------------------------------------------------ */

/* Create a context on `device`.  hip_stream: an existing hipStream_t to order
 * work on, or NULL to create a private default-priority stream. */
int tr_ctx_create(tr_ctx_t *ctx, int device, void *hip_stream);
int tr_ctx_destroy(tr_ctx_t ctx);
/* Message of the last failing call on this context ("" if none). */
const char *tr_last_error(tr_ctx_t ctx);
/* Block until all work queued on the context's stream has finished. */
int tr_ctx_synchronize(tr_ctx_t ctx);
/* The hipStream_t work is ordered on. */
void *tr_ctx_stream(tr_ctx_t ctx);

/* ---- local multiply and accumulation (KokkosWrap::LocalMatrix) ---------- */

/* C = A*B.  Replaces LocalMatrix::spgemm (include/KokkosWrap.hpp:367-391).
 * A, B: TR_ROWS.  C_out receives a context-owned contiguous tile.
 * TR_ERR_OVERFLOW if nnz(C) > INT32_MAX (use tr_spgemm_panels). */
int tr_spgemm(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, tr_csx_t *C_out);

/* C = A*B split into row panels of at most max_panel_nnz entries each
 * (<= INT32_MAX), for products too large for one int32 tile.  Panel p holds
 * rows [row_begin[p], row_begin[p+1]) of C.  panels/row_begin have room for
 * max_panels (+1) entries; *npanels receives the count.  No counterpart in
 * the reference, whose int32 size_type cannot hold such a C
 * (KokkosWrap.hpp:46); it is how a C bigger than a tile is returned as tiles. */
int tr_spgemm_panels(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, int64_t max_panel_nnz,
                     tr_csx_t *panels, int32_t *row_begin, int32_t max_panels, int32_t *npanels);

/* C += A*B; if C is uninitialised (ptr_vec == NULL) C := A*B.
 * Replaces LocalMatrix::sp_mma (include/KokkosWrap.hpp:315-364) and
 * sp_mma_hybrid (include/cusparse_helpers.cuh:708-739: an empty product
 * leaves C untouched).  A context-owned C is released and replaced by the
 * new accumulator; a caller-owned C is left intact and replaced. */
int tr_sp_mma(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, tr_csx_t *C_inout);

/* C = A + C.  Replaces LocalMatrix::spadd (include/KokkosWrap.hpp:394-439)
 * and cusparse_spgeam (include/cusparse_helpers.cuh:799-905, alpha=beta=1).
 * A.nnz == 0: no-op; C uninitialised: C := copy of A. */
int tr_spadd(tr_ctx_t ctx, const tr_csx_t *A, tr_csx_t *C_inout);

/* out = A + B (both CSR with sorted unique columns, same shape) into
 * caller-owned arrays: out->ptr_vec (nrows + 1 entries), out->idx_vec and
 * out->val (room for `capacity` entries); on success out->nnz and the
 * shape fields are set.  *nnz_needed receives nnz(A + B); if it exceeds
 * `capacity` nothing is written and TR_ERR_OVERFLOW is returned, so the
 * caller can grow its buffers and call again.  The output must not alias
 * an operand.  Replaces cusparse_spgeam (include/cusparse_helpers.cuh:
 * 799-905), which grows C_accum's CsxBuffers between its nnz and numeric
 * calls (:866-867); this is the double-buffer add of sp_mma_hybrid
 * (:708-739). */
int tr_spadd_into(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, tr_csx_t *out, int64_t capacity,
                  int64_t *nnz_needed);

/* tr_spadd_into that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory while they run
 * (tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and errors as tr_spadd_into; its temporaries live in a fixed
 * device array of TR_NOALLOC_SCRATCH_BYTES per GPU that the library holds
 * until the process ends, and operands needing more (about
 * 28 * nrows + (nnz(A) + nnz(B)) / 256 + max(nnz(A), nnz(B)) / 8 bytes) are
 * rejected with TR_ERR_INVALID.  Calls on one GPU share that array: calls
 * that overlap on one GPU (several streams or threads at once) race and
 * must be serialised by the caller.  Like tr_spadd_into it synchronises the
 * context's stream to read nnz(A + B). */
#define TR_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int tr_spadd_into_noalloc(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, tr_csx_t *out, int64_t capacity,
                          int64_t *nnz_needed);

/* Exact nnz(C) and an upper bound of the device bytes tr_spgemm takes for
 * C = A*B from a context pool with nothing cached (every scratch array and
 * work list at the size the product requests, the output tile, and the
 * product arena of a batch): the admission bound used by the serving loop.
 * Runs the product's symbolic steps.  No counterpart in the reference (its
 * buffers grow on demand, include/utils.cuh:671-716). */
int tr_symbolic(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, int64_t *nnz_C,
                size_t *bytes_needed);

/* Number of multiplications sum_k nnz(A[:,k]) * nnz(B[k,:]) (device pass). */
int tr_spgemm_flops(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, int64_t *flops);

/* ---- multiply tuning --------------------------------------------------------
 * No counterpart in the reference: Kokkos-Kernels picks its spgemm algorithm
 * internally (the handle of include/KokkosWrap.hpp:371-375 is created with
 * its default).  The cutovers of this multiply are parameters, chosen per
 * workload by trident_amd/autotune.py and loaded from its JSON output:
 *  - light_max_flops: rows with more multiplications than this take the
 *    heavy (column-windowed) path, [64, 1024];
 *  - window_entries: heavy window packing quantum H (a window holds at most
 *    H + 127 entries), [64, 3 * heavy_table / 4 - 127];
 *  - item_windows: heavy windows per work item, [1, 64] (rows of more than
 *    2048 entries use a quarter of it, at least 1);
 *  - heavy_table: slots of the heavy numeric table, 2048 (256-thread
 *    workgroups) or 4096 (512-thread workgroups);
 *  - dense_width: heavy windows spanning at most this many columns use a
 *    dense accumulator (slot = column offset: no hashing or probing, and
 *    the slots in order are the sorted output), [0, heavy_table]; 0 hashes
 *    every window.
 * Results do not depend on the tuning (bit-identical), only the time does.
 * tr_tuning_default gives the built-in values; tr_set_default_tuning sets
 * what contexts created afterwards start with (server workers included;
 * NULL restores the built-in values); tr_ctx_set_tuning changes one context
 * (NULL: the process default).  A batch keeps the tuning of its creation. */
typedef struct tr_tuning {
    int32_t light_max_flops;
    int32_t window_entries;
    int32_t item_windows;
    int32_t heavy_table;
    int32_t dense_width;
} tr_tuning_t;
int tr_tuning_default(tr_tuning_t *t);
int tr_set_default_tuning(const tr_tuning_t *t);
int tr_ctx_set_tuning(tr_ctx_t ctx, const tr_tuning_t *t);
int tr_ctx_get_tuning(tr_ctx_t ctx, tr_tuning_t *t);

/* ---- batched stage executor ------------------------------------------------ */

/* One SUMMA / Trident stage of tile products on this GPU, run as a single
 * hipGraph launch: for every task t, C_out_t = C_in_t + compress(A_t) *
 * compress(B_t).  This is the per-stage work each rank does through
 * sp_mma_hybrid (src/hns_spgemm.cu:953-969, src/sparse_summa.cu:97-99), all
 * tile products of the stage at once.
 *  - A[t], B[t]: CSR tiles (borrowed, sorted), A[t].ncols == B[t].nrows;
 *  - C_in: NULL or ntasks accumulators (borrowed; ptr_vec == NULL entries
 *    are empty);
 *  - flags: TR_BATCH_COMPRESS applies bitmask compression with
 *    mask = non-empty columns of A[t] AND non-empty rows of B[t] (the filter
 *    of SpaComm, include/spacomm.cuh:165-352, :962-1091); TR_BATCH_NO_GRAPH
 *    skips the capture (one-shot batches, e.g. the serving loop): then
 *    tr_batch_run enqueues the same fixed-size launch sequence eagerly,
 *    kernel by kernel, which also lets per-kernel timing see it.
 * tr_batch_create plans the stage (one eager pass that sizes every buffer
 * and already produces the results) and captures it; tr_batch_run replays
 * the graph on the context stream with no host round trip.  Between create
 * and run the operands must keep their addresses, shapes and sparsity
 * patterns (ptr_vec and idx_vec contents): only values may change, since a
 * replay reuses every size of the plan.  Each replay checks this on the
 * device before it writes into planned storage; on a mismatch it writes
 * nothing out of bounds and tr_batch_check (which synchronises the stream)
 * returns TR_ERR_INVALID for that run.  Results are batch-owned
 * (tr_batch_result, callable from any thread) and stay valid until the next
 * run or tr_batch_destroy. */
#define TR_BATCH_COMPRESS 1
#define TR_BATCH_NO_GRAPH 2
typedef struct tr_batch *tr_batch_t;
int tr_batch_create(tr_ctx_t ctx, int32_t ntasks, const tr_csx_t *A, const tr_csx_t *B, const tr_csx_t *C_in,
                    int32_t flags, tr_batch_t *batch);
int tr_batch_run(tr_batch_t batch);
int tr_batch_result(tr_batch_t batch, int32_t task, tr_csx_t *C_out);
/* Task t's result as consecutive row panels of at most max_panel_nnz
 * (<= INT32_MAX) entries each: the way to read a result that is larger than
 * an int32 tile (possible for a batch without accumulators, whose results
 * are its products; tr_batch_result returns TR_ERR_OVERFLOW for such a
 * task).  Panel p holds rows [row_begin[p], row_begin[p+1]) of the result;
 * the panels are views into the batch's storage with int32 row pointers of
 * their own, valid until tr_batch_destroy.  Uses the batch's context: call
 * it from the thread that owns the context.  No counterpart in the
 * reference, whose int32 size_type cannot hold such a tile
 * (include/KokkosWrap.hpp:46); it is tr_spgemm_panels for a batch. */
int tr_batch_result_panels(tr_batch_t batch, int32_t task, int64_t max_panel_nnz, tr_csx_t *panels,
                           int32_t *row_begin, int32_t max_panels, int32_t *npanels);
int tr_batch_check(tr_batch_t batch);
int tr_batch_stats(tr_batch_t batch, int64_t *flops, int64_t *nnz_products, int64_t *nnz_out);
int tr_batch_destroy(tr_batch_t batch);

/* ---- serving loop ------------------------------------------------------------ */

/* Continuous-batching request loop on one GPU.  Operands are registered once
 * and stay resident (borrowed device CSR; a Galerkin operator P also keeps
 * an on-device P^T, built on first use).  Requests:
 *   TR_REQ_SPGEMM   C = op_a * op_b
 *   TR_REQ_GALERKIN C = P^T (A P) with A = op_a, P = op_b (two product
 *                   steps; AP stays resident between them).
 * `workers` threads, each with its own stream, admit waiting steps in
 * arrival order while their symbolic memory bound (tr_symbolic) fits
 * `mem_budget`, and run them as one batch of up to `max_batch` steps;
 * steps that arrive while a batch runs join the next one.  One worker is
 * the efficient choice on one GPU: batches of several workers overlap on
 * the device, but concurrent products compete for its caches (measured
 * 3.4x lower throughput with two workers on MI355X).  Each worker reserves
 * mem_budget / workers of device memory when the server is created and
 * places its batches' blocks in it, so a running server makes no device
 * allocation.  Results are bit-identical to the same product run alone.
 * No counterpart in the reference, which has no serving layer. */
#define TR_REQ_SPGEMM 0
#define TR_REQ_GALERKIN 1
#define TR_REQ_DISCARD 0x100  /* or-ed into the kind: drop the result when done */
typedef struct tr_server *tr_server_t;
int tr_server_create(int device, size_t mem_budget, int32_t workers, int32_t max_batch, tr_server_t *server);
int tr_server_register(tr_server_t server, const tr_csx_t *M, int32_t *op_id);
int tr_server_submit(tr_server_t server, int32_t kind_flags, int32_t op_a, int32_t op_b, int64_t *req_id);
/* Block until the request finished; C (optional) receives a view of the
 * result, valid until tr_server_release. */
int tr_server_wait(tr_server_t server, int64_t req_id, tr_csx_t *C, double *latency_ms);
int tr_server_release(tr_server_t server, int64_t req_id);
/* Completed requests, batches run, latency percentiles since the last
 * reset, and throughput (2 * multiplications / wall time). */
int tr_server_stats(tr_server_t server, int64_t *completed, int64_t *batches, double *p50_ms, double *p99_ms,
                    double *gflops);
int tr_server_reset_stats(tr_server_t server);
int tr_server_destroy(tr_server_t server);

/* ---- bitmask tile compression (SpaComm) --------------------------------- */

/* mask[w] bit b set iff segment 64*w+b of ptr is non-empty; ceil(n/64)
 * uint64 words.  Replaces SpaComm::gen_bitmask (include/spacomm.cuh:29-117)
 * with 64-bit words (MASK_SIZE=64); on little-endian hosts the words are the
 * same bytes as the reference's MASK_SIZE=8 uint8 mask. */
int tr_bitmask_gen(tr_ctx_t ctx, const int32_t *ptr, int32_t n, uint64_t *mask);

/* out[w] = a[w] & b[w].  Replaces SpaComm::intersect_bitmasks
 * (include/spacomm.cuh:126-148). */
int tr_bitmask_and(tr_ctx_t ctx, const uint64_t *a, const uint64_t *b, int32_t words, uint64_t *out);

/* Keep the entries of major segment i (row for TR_ROWS, column for TR_COLS)
 * iff bit i of mask is set; dimensions and majordim unchanged, dropped
 * segments become empty.  `out` receives a context-owned contiguous tile.
 * Replaces SpaCommHandler::Compress (include/spacomm.cuh:962-1091). */
int tr_compress(tr_ctx_t ctx, const tr_csx_t *M, const uint64_t *mask, tr_csx_t *out);

/* ---- conversions --------------------------------------------------------- */

/* Transpose: out = CSR of M^T (equivalently M re-expressed in the other
 * major dimension), context-owned, rows sorted.  Replaces
 * LocalMatrix::csc_to_csr / cusparse_csc_to_csr (KokkosWrap.hpp:441-531,
 * cusparse_helpers.cuh:607-705) when applied to a CSC tile. */
int tr_csx_transpose(tr_ctx_t ctx, const tr_csx_t *M, tr_csx_t *out);

/* CSC tile -> CSR tile of the same matrix (majordim TR_COLS in, TR_ROWS out). */
int tr_csc_to_csr(tr_ctx_t ctx, const tr_csx_t *M_csc, tr_csx_t *out);

/* CSC tile -> CSR tile of the same matrix into caller-owned arrays
 * (ptr: nrows + 1, idx/val: nnz entries), rows sorted.  Replaces
 * cusparse_csc_to_csr with conversion buffers (include/cusparse_helpers.cuh:
 * 607-705) and LocalMatrix::csc_to_csr(handle, csx, buffs)
 * (include/KokkosWrap.hpp:441-531). */
int tr_csc_to_csr_into(tr_ctx_t ctx, const tr_csx_t *M_csc, int32_t *ptr, int32_t *idx, double *val);

/* Sort the columns of every row in place (values follow). */
int tr_csx_sort_rows(tr_ctx_t ctx, tr_csx_t *M);

/* Device validation: *status bit0 ptr malformed, bit1 idx out of range,
 * bit2 rows unsorted or duplicate columns. */
int tr_csx_validate(tr_ctx_t ctx, const tr_csx_t *M, int32_t *status);

/* ---- memory ---------------------------------------------------------------- */

/* Release a context-owned matrix (LocalMatrix::freeBuffers,
 * KokkosWrap.hpp:294-302); caller-owned storage is not touched.  The struct
 * is reset to uninitialised. */
int tr_csx_free(tr_ctx_t ctx, tr_csx_t *M);

/* Build a context-owned contiguous device tile from host arrays
 * (CSX_create_contig_device, include/utils.cuh:451-466). */
int tr_csx_upload(tr_ctx_t ctx, int32_t majordim, int32_t nrows, int32_t ncols, int64_t nnz,
                  const int32_t *h_ptr, const int32_t *h_idx, const double *h_val, tr_csx_t *out);

/* Copy a device matrix to host arrays (ptrdim+1, nnz, nnz). */
int tr_csx_download(tr_ctx_t ctx, const tr_csx_t *M, int32_t *h_ptr, int32_t *h_idx, double *h_val);

/* Plain device memory through the context's caching pool. */
int tr_malloc(tr_ctx_t ctx, void **ptr, size_t bytes);
int tr_free(tr_ctx_t ctx, void *ptr);
int tr_memcpy_h2d(tr_ctx_t ctx, void *dst, const void *src, size_t bytes);
int tr_memcpy_d2h(tr_ctx_t ctx, void *dst, const void *src, size_t bytes);
int tr_memcpy_d2d(tr_ctx_t ctx, void *dst, const void *src, size_t bytes);
/* Fill `bytes` bytes at dst with `value`, ordered on the context's stream
 * (the cudaMemset of the reference's buffers, include/utils.cuh:569). */
int tr_memset(tr_ctx_t ctx, void *dst, int value, size_t bytes);

/* Context-free device allocation on `device` (thread-safe, outside any
 * context's pool): storage that outlives or is shared between the contexts
 * of several host threads, such as the CsxBuffers of the partial C that the
 * accumulation thread grows (include/utils.cuh:527-784, whose ensure()
 * reallocates with cudaMalloc, :637-667). */
int tr_device_malloc(int device, void **ptr, size_t bytes);
int tr_device_free(void *ptr);

/* Cross-stream ordering between contexts: tr_event_record marks the work
 * queued so far on ctx's stream; after tr_ctx_wait_event, work queued on
 * ctx (another context) waits for the marked work.  This is how a product
 * made on the main thread's stream is handed to the accumulation thread's
 * stream (AccumThreadHandle, include/hns_spgemm.cuh:32-130). */
typedef struct tr_event *tr_event_t;
int tr_event_create(tr_event_t *ev);
int tr_event_record(tr_ctx_t ctx, tr_event_t ev);
int tr_ctx_wait_event(tr_ctx_t ctx, tr_event_t ev);
int tr_event_synchronize(tr_event_t ev);
int tr_event_destroy(tr_event_t ev);

/* Return cached, unused pool blocks to the device. */
int tr_pool_trim(tr_ctx_t ctx);
/* Bytes currently handed out / cached by the pool. */
int tr_pool_stats(tr_ctx_t ctx, size_t *in_use, size_t *cached);
/* Test hook: make allocations beyond `bytes` in use fail with TR_ERR_OOM
 * (0 = no limit). */
int tr_pool_set_limit(tr_ctx_t ctx, size_t bytes);

/* ---- instrumentation ------------------------------------------------------- */

/* When enabled, every kernel launch on the context is bracketed by HIP
 * events on its stream and its duration is accumulated per kernel name. */
int tr_ctx_set_timing(tr_ctx_t ctx, int enable);
/* Clear accumulated kernel timings. */
int tr_ctx_reset_timing(tr_ctx_t ctx);
/* Number of distinct kernel names timed so far. */
int tr_ctx_timing_count(tr_ctx_t ctx);
/* Entry i: name (copied into `name`, up to name_len bytes), launches and
 * total milliseconds.  Synchronises the stream. */
int tr_ctx_timing_get(tr_ctx_t ctx, int i, char *name, int name_len, int64_t *launches, double *total_ms);

/* dst = src with a plain streaming copy kernel of this library (16 bytes per
 * lane, grid-stride): the reference workload for calibrating hardware
 * traffic counters (FETCH_SIZE / WRITE_SIZE) against bytes known to move.
 * Pointers 16-byte aligned, bytes a multiple of 16.  No counterpart in the
 * reference. */
int tr_calibration_copy(tr_ctx_t ctx, void *dst, const void *src, size_t bytes);

/* Library build string (architecture and tuning constants). */
const char *tr_version(void);

#ifdef __cplusplus
}
#endif
#endif
