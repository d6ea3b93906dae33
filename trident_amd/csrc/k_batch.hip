// Batched stage executor: one SUMMA / Trident stage's tile products on one
// GPU as a single hipGraph launch.
//
// For every task t of the stage (a tile product that the reference runs as
// one sp_mma_hybrid call per stage per rank, src/hns_spgemm.cu:953-969,
// src/sparse_summa.cu:97-99):
//   1. bitmask compression (optional): mask_t = non-empty columns of A_t AND
//      non-empty rows of B_t (the filter SpaComm derives, include/spacomm.cuh
//      :165-352 and :962-1091); A_t keeps the entries of masked columns, B_t
//      the masked rows;
//   2. the product P_t = A_t * B_t through the multiply pipeline, all tasks
//      in one task table, so every kernel launch covers all tile products
//      (uneven products fill the GPU together instead of one after another);
//   3. the accumulation C_t = C_in_t + P_t through the merge-path add, again
//      one launch per phase over all tasks.
// The plan step runs the whole stage once eagerly (host reads size every
// list and output), then captures the same launch sequence, with every size
// fixed, into a hipGraph; tr_batch_run replays it with no host round trip.
// Inputs are borrowed and must not change shape between plan and run.
#include "dev_common.h"
#include "tr_internal.h"
#include "spgemm_types.h"

#include <algorithm>
#include <cstring>
#include <memory>

namespace tr {
namespace {

using namespace dev;

__global__ void k_fill_u64(uint64_t *__restrict__ p, int64_t n, uint64_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_col_presence(const int32_t *__restrict__ idx, int64_t nnz, uint64_t *__restrict__ mask) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = idx[p];
        atomicOr((unsigned long long *)&mask[c >> 6], 1ull << (c & 63));
    }
}

// mask[w] &= (bit b set iff row 64 w + b of ptr is non-empty)
__global__ void k_and_row_presence(const int32_t *__restrict__ ptr, int32_t n, uint64_t *__restrict__ mask) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool nonempty = false;
    if (i < n) nonempty = ptr[i] != ptr[i + 1];
    const uint64_t word = ballot(nonempty);
    if (lane_id() == 0 && i < n) mask[i >> 6] &= word;
}

// CSR column filter: keep the entries whose column bit is set.  A thread per
// row for rows up to 32 entries, the whole wave for longer rows.
template <bool kWrite>
__global__ __launch_bounds__(256) void k_colfilter(const int32_t *__restrict__ ap, const int32_t *__restrict__ ai,
                                                  const double *__restrict__ av, int32_t m,
                                                  const uint64_t *__restrict__ mask, int32_t *__restrict__ kept,
                                                  const int64_t *__restrict__ np64, int32_t *__restrict__ optr,
                                                  int32_t *__restrict__ oi, double *__restrict__ ov, int64_t cap) {
    const int l = lane_id();
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t a0 = 0, a1 = 0;
    if (r < m) {
        a0 = ap[r];
        a1 = ap[r + 1];
    }
    int64_t base = 0;
    if (kWrite && r < m) {
        base = np64[r];
        optr[r] = (int32_t)base;
        if (r == m - 1) optr[m] = (int32_t)np64[m];
    }
    int32_t c = 0;
    if (a1 - a0 <= 32) {
        for (int32_t p = a0; p < a1; ++p) {
            const int32_t col = ai[p];
            if ((mask[col >> 6] >> (col & 63)) & 1ull) {
                if (kWrite && base + c < cap) {
                    oi[base + c] = col;
                    ov[base + c] = av[p];
                }
                ++c;
            }
        }
    }
    uint64_t longm = ballot(a1 - a0 > 32);
    while (longm) {
        const int src = __ffsll((unsigned long long)longm) - 1;
        longm &= longm - 1;
        const int32_t s0 = __shfl(a0, src, kWave), s1 = __shfl(a1, src, kWave);
        const int64_t sb = __shfl(base, src, kWave);
        int32_t run = 0;
        for (int32_t pb = s0; pb < s1; pb += kWave) {
            const int32_t p = pb + l;
            bool keep = false;
            int32_t col = 0;
            if (p < s1) {
                col = ai[p];
                keep = (mask[col >> 6] >> (col & 63)) & 1ull;
            }
            const uint64_t mk = ballot(keep);
            if (kWrite && keep) {
                const int64_t q = sb + run + rank_below(mk);
                if (q < cap) {
                    oi[q] = col;
                    ov[q] = av[p];
                }
            }
            run += __popcll(mk);
        }
        if (l == src) c = run;
    }
    if (!kWrite && r < m) kept[r] = c;
}

// ---- replay guard ----
// A replay reuses every size of the plan: work-list capacities, the product
// and output arenas and each task's place in them.  It is correct only for
// operands with the planned sparsity pattern (values may change).  The
// guards compare what a replay recomputes with the plan before each phase
// that writes into planned storage; on a mismatch they record it in the
// guard word (tr_batch_check reports it) and zero the device-side counts the
// later kernels loop over, so nothing is written outside the batch's
// buffers.  Bits: 8 operand nnz changed, 1 products differ from the plan,
// 2/4 (the add) chunk or output offsets differ.
struct GuardOperand {
    const int32_t *ptr;
    int32_t n;      // ptr has n + 1 entries
    int32_t pad;
    int64_t nnz;    // planned ptr[n]
};

__global__ void k_guard_reset(int32_t *__restrict__ guard) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *guard = 0;
}

__global__ void k_guard_operands(const GuardOperand *__restrict__ ops, int nops, int32_t *__restrict__ guard) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nops) {
        const GuardOperand o = ops[i];
        if (o.ptr[0] != 0 || (int64_t)o.ptr[o.n] != o.nnz) atomicOr(guard, 8);
    }
}

// after the analysis: a flagged replay runs no symbolic or numeric rows
__global__ void k_guard_counts(Counters *__restrict__ cnt, const int32_t *__restrict__ guard) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && *guard) {
        for (int c = 0; c < kNumClasses; ++c) cnt->cls[c] = 0;
        cnt->items = 0;
    }
}

// after the symbolic phase: every task's products start where the plan put
// them (expect[t], expect[nt] = total), the heavy work lists fit and the
// cursor total is the planned one; otherwise no numeric row runs
__global__ void k_guard_products(const TaskDev *__restrict__ tasks, int nt, const int64_t *__restrict__ rowptr,
                                 int32_t rows, const int64_t *__restrict__ expect, Counters *__restrict__ cnt,
                                 const int64_t *__restrict__ cbase, int32_t item_cap, int64_t cursor_cap,
                                 int32_t *__restrict__ guard) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int bad = cnt->overflow & 1;
    for (int t = 0; t <= nt && !bad; ++t)
        if (rowptr[t < nt ? tasks[t].row_base : rows] != expect[t]) bad = 1;
    if (cbase && cbase[item_cap] > cursor_cap) bad = 1;
    if (bad) atomicOr(guard, 1);
    if (*guard) {
        for (int c = 0; c < kNumClasses; ++c) cnt->cls[c] = 0;
        cnt->items = 0;
    }
}

__global__ void k_zero_i32(int32_t *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

}  // namespace
}  // namespace tr

using namespace tr;

// Operands of one task after compression (device arrays, batch-owned when
// compressed, borrowed otherwise).
struct BatchOperand {
    const int32_t *ptr = nullptr, *idx = nullptr;
    const double *val = nullptr;
    int32_t *optr = nullptr, *oidx = nullptr;
    double *oval = nullptr;
};

struct tr_batch {
    tr_ctx *ctx = nullptr;
    int nt = 0;
    int32_t mode = 0;
    std::vector<tr_csx_t> A, B, Cin;
    // compression
    std::vector<uint64_t *> masks;
    std::vector<BatchOperand> Ac, Bc;
    std::vector<int32_t *> kept_tmp;
    std::vector<int64_t *> np64;
    std::vector<int64_t *> scan_tmp;
    std::vector<void *> owned;  // pool blocks owned by the batch
    int32_t *zero_ptr = nullptr;
    // pipelines
    std::unique_ptr<Workspace> ws_mul, ws_add;
    MulPlan mp;
    AddPlan ap;
    Panel *d_panel = nullptr;
    int64_t ncur = 0;
    // product and output arenas
    int32_t *p_idx = nullptr;
    double *p_val = nullptr;
    int64_t p_total = 0;
    int32_t *o_idx = nullptr;
    double *o_val = nullptr;
    int32_t *o_ptr = nullptr;
    int64_t o_total = 0;
    std::vector<int64_t> o_base;     // per task: first entry in the output arena
    std::vector<int64_t> optr_off;   // per task: first row pointer in o_ptr
    std::vector<int64_t> o_nnz;
    int64_t flops = 0;
    // replay guard: flag word, planned operand ends, planned per-task
    // product and output offsets (nt + 1 each)
    int32_t *guard = nullptr;
    GuardOperand *guard_ops = nullptr;
    int guard_nops = 0;
    int64_t *expect_p = nullptr, *expect_o = nullptr;
    bool no_accum = false;  // no task has an accumulator: results are the products
    bool single_run = false;  // workspace in the context's scratch slots: no replay
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool ran = false;

    void *own(size_t bytes) {
        void *p = ctx->pool.alloc(bytes);
        owned.push_back(p);
        return p;
    }
    ~tr_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        ws_mul.reset();
        ws_add.reset();
        for (void *p : owned) {
            try {
                ctx->pool.release(p);
            } catch (...) {
            }
        }
    }
};

namespace {

// enqueue compression for every task
void batch_compress(tr_batch *b) {
    tr_ctx *ctx = b->ctx;
    for (int t = 0; t < b->nt; ++t) {
        const tr_csx_t &A = b->A[t], &B = b->B[t];
        const int32_t k = A.ncols;
        const int64_t words = ((int64_t)k + 63) / 64;
        uint64_t *mask = b->masks[t];
        TR_LAUNCH(ctx, "mask_fill", k_fill_u64, div_up((uint64_t)std::max<int64_t>(words, 1), 256), 256, 0, mask,
                  words, 0ull);
        if (A.nnz)
            TR_LAUNCH(ctx, "mask_cols", k_col_presence, div_up((uint64_t)std::min<int64_t>(A.nnz, 1 << 22), 256), 256,
                      0, A.idx_vec, A.nnz, mask);
        TR_LAUNCH(ctx, "mask_rows", k_and_row_presence, div_up((uint64_t)std::max(k, 1), 256), 256, 0, B.ptr_vec, k,
                  mask);
        // A: column filter
        BatchOperand &ac = b->Ac[t];
        TR_LAUNCH(ctx, "colfilter_count", k_colfilter<false>, div_up((uint64_t)std::max(A.nrows, 1), 256), 256, 0,
                  A.ptr_vec, A.idx_vec, A.val, A.nrows, mask, b->kept_tmp[t], (const int64_t *)nullptr,
                  (int32_t *)nullptr, (int32_t *)nullptr, (double *)nullptr, (int64_t)0);
        launch_exclusive_scan_i32_to_i64(ctx, b->kept_tmp[t], b->np64[t], A.nrows, b->scan_tmp[t]);
        TR_LAUNCH(ctx, "colfilter_write", k_colfilter<true>, div_up((uint64_t)std::max(A.nrows, 1), 256), 256, 0,
                  A.ptr_vec, A.idx_vec, A.val, A.nrows, mask, (int32_t *)nullptr, b->np64[t], ac.optr, ac.oidx,
                  ac.oval, A.nnz);
        if (A.nrows == 0) TR_LAUNCH(ctx, "zero_ptr", k_zero_i32, 1, 64, 0, ac.optr, (int64_t)1);
        // B: row segments
        BatchOperand &bc = b->Bc[t];
        if (B.nrows == 0) TR_LAUNCH(ctx, "zero_ptr", k_zero_i32, 1, 64, 0, bc.optr, (int64_t)1);
        launch_compress_fixed(ctx, B.nrows, B.ptr_vec, B.idx_vec, B.val, B.nnz, mask, b->kept_tmp[t], b->np64[t],
                              b->scan_tmp[t], bc.optr, bc.oidx, bc.oval);
    }
}

// enqueue multiply + add; capture == true: sizes from the eager run
void batch_body(tr_batch *b, bool capture) {
    tr_ctx *ctx = b->ctx;
    const bool guarded = capture && b->guard;  // replays check the plan
    if (guarded) {
        TR_LAUNCH(ctx, "batch_guard", k_guard_reset, 1, 64, 0, b->guard);
        TR_LAUNCH(ctx, "batch_guard", k_guard_operands, div_up((uint64_t)b->guard_nops, 64), 64, 0, b->guard_ops,
                  b->guard_nops, b->guard);
    }
    if (b->mode) batch_compress(b);
    mul_analyze(ctx, b->mp);
    if (guarded) TR_LAUNCH(ctx, "batch_guard", k_guard_counts, 1, 64, 0, b->mp.cnt, b->guard);
    if (!capture) mul_read_counts(ctx, b->mp);
    mul_symbolic(ctx, *b->ws_mul, b->mp, true, capture);
    const int64_t ncur = mul_cursor_prep(ctx, *b->ws_mul, b->mp, capture);
    if (guarded)
        TR_LAUNCH(ctx, "batch_guard", k_guard_products, 1, 64, 0, b->mp.d_tasks, b->nt, b->mp.rowptr,
                  b->mp.rows_total, b->expect_p, b->mp.cnt, b->mp.h.items > 0 ? b->mp.cbase : nullptr,
                  b->mp.item_cap, b->mp.cursor_cap, b->guard);
    if (!capture) {
        b->p_total = read_scalar<int64_t>(ctx, b->mp.rowptr + b->mp.rows_total);
        mul_read_counts(ctx, b->mp);
        if (b->mp.h.overflow & 1) fail(TR_ERR_HIP, "batch: heavy work list capacity exceeded");
        b->mp.cursor_cap = ncur;
        if (b->mp.h.items > 0)
            b->mp.cursors = (int4 *)b->ws_mul->get(S_CURSORS, sizeof(int4) * (size_t)std::max<int64_t>(ncur, 1));
        b->p_idx = (int32_t *)b->own(sizeof(int32_t) * (size_t)std::max<int64_t>(b->p_total, 1));
        b->p_val = (double *)b->own(sizeof(double) * (size_t)std::max<int64_t>(b->p_total, 1));
        Panel hp{};
        hp.nnz_begin = 0;
        hp.row_begin = 0;
        hp.idx = b->p_idx;
        hp.val = b->p_val;
        b->d_panel = (Panel *)b->own(sizeof(Panel));
        TR_HIP(hipMemcpyAsync(b->d_panel, &hp, sizeof(Panel), hipMemcpyHostToDevice, ctx->stream));
        // a product added into a partial-C tile must fit an int32 tile; a
        // product that is itself the result may be larger (it is then read
        // as row panels, tr_batch_result_panels)
        std::vector<int64_t> rp(b->nt + 1);
        for (int t = 0; t <= b->nt; ++t) {
            const int32_t g = t < b->nt ? b->mp.h_tasks[t].row_base : b->mp.rows_total;
            rp[t] = read_scalar<int64_t>(ctx, b->mp.rowptr + g);
        }
        for (int t = 0; t < b->nt; ++t)
            if ((!b->no_accum || b->single_run) && rp[t + 1] - rp[t] > INT32_MAX)
                fail(TR_ERR_OVERFLOW, "batch: a tile product exceeds INT32_MAX entries");
    }
    mul_numeric(ctx, b->mp, b->d_panel, 1);
    if (!capture) {
        mul_check(ctx, b->mp);
        b->optr_off.assign(b->nt, 0);
        int64_t optr_total = 0;
        for (int t = 0; t < b->nt; ++t) {
            b->optr_off[t] = optr_total;
            optr_total += b->A[t].nrows + 1;
        }
        b->o_ptr = (int32_t *)b->own(sizeof(int32_t) * (size_t)std::max<int64_t>(optr_total, 1));
    }
    if (b->no_accum) {
        // no accumulators: the products are the results (per-task int32 row
        // pointers rebased from the 64-bit product row pointer)
        for (int t = 0; t < b->nt; ++t) {
            const int32_t rb = b->mp.h_tasks[t].row_base;
            launch_rowptr_i64_to_i32(ctx, b->mp.rowptr, rb, rb + b->A[t].nrows, b->o_ptr + b->optr_off[t]);
        }
        if (!capture) {
            b->o_idx = b->p_idx;
            b->o_val = b->p_val;
            b->o_total = b->p_total;
        }
        return;
    }
    if (!capture) {
        // add tasks: X = products (64-bit global rows), Y = C_in
        std::vector<AddTaskDev> at(b->nt);
        int64_t ny = 0;
        for (int t = 0; t < b->nt; ++t) {
            AddTaskDev &a = at[t];
            a = AddTaskDev{};
            a.xp64 = b->mp.rowptr;
            a.xi = b->p_idx;
            a.xv = b->p_val;
            const tr_csx_t &C = b->Cin[t];
            if (C.ptr_vec) {
                a.yp = C.ptr_vec;
                a.yi = C.idx_vec;
                a.yv = C.val;
                ny += C.nnz;
            } else {
                a.yp = b->zero_ptr;
                a.yi = nullptr;
                a.yv = nullptr;
            }
            a.optr = b->o_ptr + b->optr_off[t];
            a.m = b->A[t].nrows;
        }
        add_setup(ctx, *b->ws_add, b->ap, at, b->p_total, ny);
    }
    add_count(ctx, b->ap, capture);
    if (!capture) {
        b->o_total = add_total(ctx, b->ap);
        b->o_idx = (int32_t *)b->own(sizeof(int32_t) * (size_t)std::max<int64_t>(b->o_total, 1));
        b->o_val = (double *)b->own(sizeof(double) * (size_t)std::max<int64_t>(b->o_total, 1));
    }
    add_write(ctx, b->ap, b->o_idx, b->o_val);
}

}  // namespace

template <typename F>
static int bguard(tr_ctx *ctx, F &&body) {
    try {
        body();
        if (ctx) ctx->last_error.clear();
        return TR_OK;
    } catch (const Error &e) {
        if (ctx) ctx->last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->last_error = "host allocation failed";
        return TR_ERR_OOM;
    } catch (...) {
        if (ctx) ctx->last_error = "unexpected internal error";
        return TR_ERR_HIP;
    }
}

extern "C" {

int tr_batch_create(tr_ctx_t ctx, int32_t ntasks, const tr_csx_t *A, const tr_csx_t *B, const tr_csx_t *C_in,
                    int32_t flags, tr_batch_t *out) {
    return tr::batch_create(ctx, ntasks, A, B, C_in, flags, true, out);
}

}  // extern "C"

// validate == false: the caller vouches for the operands (the server's
// registered and produced operands), which skips one device check and host
// read per task
int tr::batch_create(tr_ctx_t ctx, int32_t ntasks, const tr_csx_t *A, const tr_csx_t *B, const tr_csx_t *C_in,
                     int32_t flags, bool validate, tr_batch_t *out) {
    if (!ctx || !out) return TR_ERR_INVALID;
    *out = nullptr;
    tr_batch *b = nullptr;
    const int rc = bguard(ctx, [&] {
        TR_REQUIRE(ntasks >= 1 && A && B, "tr_batch_create: need at least one task");
        TR_REQUIRE((flags & ~(TR_BATCH_COMPRESS | TR_BATCH_NO_GRAPH | kBatchSingleRun)) == 0,
                   "tr_batch_create: unknown flags");
        TR_REQUIRE(!(flags & kBatchSingleRun) || (flags & TR_BATCH_NO_GRAPH), "batch: single-run batches have no graph");
        const int32_t compress_mode = (flags & TR_BATCH_COMPRESS) ? 1 : 0;
        b = new tr_batch();
        b->ctx = ctx;
        b->nt = ntasks;
        b->mode = compress_mode;
        b->A.assign(A, A + ntasks);
        b->B.assign(B, B + ntasks);
        b->Cin.resize(ntasks);
        int32_t max_m = 0;
        for (int t = 0; t < ntasks; ++t) {
            b->Cin[t] = C_in ? C_in[t] : tr_csx_t{};
            if (validate) validate_operands(ctx, &b->A[t], &b->B[t]);
            else check_operand_shapes(&b->A[t], &b->B[t]);
            const tr_csx_t &C = b->Cin[t];
            if (C.ptr_vec) {
                check_csx(&C, "C_in");
                TR_REQUIRE(C.majordim == TR_ROWS && C.nrows == b->A[t].nrows && C.ncols == b->B[t].ncols,
                           "tr_batch_create: accumulator shape mismatch");
                int32_t *st = (int32_t *)scratch(ctx, S_TMP_CSX, 64);
                TR_HIP(hipMemsetAsync(st, 0, sizeof(int32_t), ctx->stream));
                launch_validate(ctx, &C, st);
                int32_t hs = 0;
                read_array(ctx, &hs, st, sizeof(hs));
                TR_REQUIRE(hs == 0, "tr_batch_create: accumulator is not a sorted CSR");
            }
            max_m = std::max(max_m, b->A[t].nrows);
        }
        b->no_accum = true;
        for (int t = 0; t < ntasks; ++t)
            if (b->Cin[t].ptr_vec) b->no_accum = false;
        b->zero_ptr = (int32_t *)b->own(sizeof(int32_t) * ((size_t)max_m + 1));
        TR_HIP(hipMemsetAsync(b->zero_ptr, 0, sizeof(int32_t) * ((size_t)max_m + 1), ctx->stream));
        // compression buffers (capacity = input nnz)
        std::vector<TaskDev> tasks(ntasks);
        b->masks.assign(ntasks, nullptr);
        b->Ac.resize(ntasks);
        b->Bc.resize(ntasks);
        b->kept_tmp.assign(ntasks, nullptr);
        b->np64.assign(ntasks, nullptr);
        b->scan_tmp.assign(ntasks, nullptr);
        for (int t = 0; t < ntasks; ++t) {
            const tr_csx_t &a = b->A[t], &bb = b->B[t];
            TaskDev &T = tasks[t];
            T = TaskDev{};
            T.m = a.nrows;
            T.n = bb.ncols;
            if (compress_mode) {
                const int32_t k = a.ncols;
                b->masks[t] = (uint64_t *)b->own(sizeof(uint64_t) * (size_t)(((int64_t)k + 63) / 64 + 1));
                const int32_t rmax = std::max(a.nrows, bb.nrows);
                b->kept_tmp[t] = (int32_t *)b->own(sizeof(int32_t) * ((size_t)rmax + 1));
                b->np64[t] = (int64_t *)b->own(sizeof(int64_t) * ((size_t)rmax + 1));
                b->scan_tmp[t] = (int64_t *)b->own(scan_tmp_bytes(rmax));
                BatchOperand &ac = b->Ac[t], &bc = b->Bc[t];
                ac.optr = (int32_t *)b->own(sizeof(int32_t) * ((size_t)a.nrows + 1));
                ac.oidx = (int32_t *)b->own(sizeof(int32_t) * (size_t)std::max<int64_t>(a.nnz, 1));
                ac.oval = (double *)b->own(sizeof(double) * (size_t)std::max<int64_t>(a.nnz, 1));
                bc.optr = (int32_t *)b->own(sizeof(int32_t) * ((size_t)bb.nrows + 1));
                bc.oidx = (int32_t *)b->own(sizeof(int32_t) * (size_t)std::max<int64_t>(bb.nnz, 1));
                bc.oval = (double *)b->own(sizeof(double) * (size_t)std::max<int64_t>(bb.nnz, 1));
                T.ap = ac.optr;
                T.ai = ac.oidx;
                T.av = ac.oval;
                T.bp = bc.optr;
                T.bi = bc.oidx;
                T.bv = bc.oval;
            } else {
                T.ap = a.ptr_vec;
                T.ai = a.idx_vec;
                T.av = a.val;
                T.bp = bb.ptr_vec;
                T.bi = bb.idx_vec;
                T.bv = bb.val;
            }
        }
        // single-run batches without accumulators (the server's) take their
        // multiply workspace from the context's scratch slots instead of
        // allocating one per batch
        b->single_run = (flags & kBatchSingleRun) && b->no_accum;
        b->ws_mul.reset(new Workspace(ctx, !b->single_run));
        b->ws_add.reset(new Workspace(ctx, true));
        mul_setup(ctx, *b->ws_mul, b->mp, tasks);
        // eager run: sizes every list and output
        const bool timing = ctx->timing;
        ctx->timing = false;
        batch_body(b, false);
        b->flops = (int64_t)b->mp.h.flops;
        b->o_base.assign(ntasks, 0);
        b->o_nnz.assign(ntasks, 0);
        for (int t = 0; t < ntasks; ++t) {
            if (b->no_accum) {
                b->o_base[t] = read_scalar<int64_t>(ctx, b->mp.rowptr + b->mp.h_tasks[t].row_base);
                continue;
            }
            const int32_t g = b->ap.h[t].row_base;
            const int64_t cs = read_scalar<int64_t>(ctx, b->ap.cstart + g);
            b->o_base[t] = read_scalar<int64_t>(ctx, b->ap.coff + cs);
        }
        for (int t = 0; t < ntasks; ++t)
            b->o_nnz[t] = (t + 1 < ntasks ? b->o_base[t + 1] : b->o_total) - b->o_base[t];
        for (int t = 0; t < ntasks; ++t)
            if (!b->no_accum && b->o_nnz[t] > INT32_MAX)
                fail(TR_ERR_OVERFLOW, "batch: an accumulated tile exceeds INT32_MAX entries");
        b->ran = true;  // the eager pass already produced the results
        if (!b->single_run) {
            // what a replay must reproduce: operand ends and task offsets
            std::vector<GuardOperand> go;
            auto add_op = [&](const tr_csx_t &m) {
                if (m.ptr_vec) go.push_back(GuardOperand{m.ptr_vec, m.nrows, 0, m.nnz});
            };
            for (int t = 0; t < ntasks; ++t) {
                add_op(b->A[t]);
                add_op(b->B[t]);
                add_op(b->Cin[t]);
            }
            std::vector<int64_t> ep(ntasks + 1), eo(ntasks + 1);
            for (int t = 0; t <= ntasks; ++t) {
                const int32_t g = t < ntasks ? b->mp.h_tasks[t].row_base : b->mp.rows_total;
                ep[t] = read_scalar<int64_t>(ctx, b->mp.rowptr + g);
                eo[t] = t < ntasks ? b->o_base[t] : b->o_total;
            }
            b->guard_nops = (int)go.size();
            b->guard = (int32_t *)b->own(sizeof(int32_t));
            b->guard_ops = (GuardOperand *)b->own(sizeof(GuardOperand) * go.size());
            b->expect_p = (int64_t *)b->own(sizeof(int64_t) * ep.size());
            b->expect_o = (int64_t *)b->own(sizeof(int64_t) * eo.size());
            TR_HIP(hipMemcpyAsync(b->guard_ops, go.data(), sizeof(GuardOperand) * go.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
            TR_HIP(hipMemcpyAsync(b->expect_p, ep.data(), sizeof(int64_t) * ep.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
            TR_HIP(hipMemcpyAsync(b->expect_o, eo.data(), sizeof(int64_t) * eo.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
            TR_LAUNCH(ctx, "batch_guard", k_guard_reset, 1, 64, 0, b->guard);
            b->ap.guard = b->guard;
            b->ap.expect = b->expect_o;
            TR_HIP(hipStreamSynchronize(ctx->stream));  // host staging of the copies above ends here
        }
        if (flags & TR_BATCH_NO_GRAPH) {
            ctx->timing = timing;
            return;
        }
        TR_HIP(hipStreamSynchronize(ctx->stream));
        // capture the same sequence with fixed sizes
        TR_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        try {
            batch_body(b, true);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(ctx->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            ctx->timing = timing;
            throw;
        }
        TR_HIP(hipStreamEndCapture(ctx->stream, &b->graph));
        TR_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        ctx->timing = timing;
    });
    if (rc != TR_OK) {
        delete b;
        return rc;
    }
    *out = b;
    return TR_OK;
}

extern "C" {

int tr_batch_run(tr_batch_t b) {
    if (!b) return TR_ERR_INVALID;
    tr_ctx *ctx = b->ctx;
    return bguard(ctx, [&] {
        TR_REQUIRE(!b->single_run, "tr_batch_run: this batch was created for a single run");
        if (b->exec) {
            ::tr::TimerScope ts(ctx, "batch_graph");
            TR_HIP(hipGraphLaunch(b->exec, ctx->stream));
        } else {
            batch_body(b, true);  // the sequence the graph would hold, launched eagerly
        }
        b->ran = true;
    });
}

// A view of task t's result.  Reads only fields fixed when the batch was
// created and never touches the context (its last-error string belongs to
// the thread that owns the context), so any thread may call it while the
// owner keeps using the context.
int tr_batch_result(tr_batch_t b, int32_t t, tr_csx_t *C_out) {
    if (!b || !C_out || t < 0 || t >= b->nt) return TR_ERR_INVALID;
    if (b->o_nnz[t] > INT32_MAX) return TR_ERR_OVERFLOW;  // read it with tr_batch_result_panels
    tr_csx_t c{};
    c.majordim = TR_ROWS;
    c.nrows = b->A[t].nrows;
    c.ncols = b->B[t].ncols;
    c.nnz = b->o_nnz[t];
    c.ptr_vec = b->o_ptr + b->optr_off[t];
    c.idx_vec = b->o_idx + b->o_base[t];
    c.val = b->o_val + b->o_base[t];
    c.contig = 0;
    *C_out = c;
    return TR_OK;
}

}  // extern "C"

// Device bytes the batch holds between runs (its own blocks: operands after
// compression, product and output arenas, row pointers), without the
// context scratch slots a single-run batch borrows while it runs.
size_t tr::batch_retained_bytes(tr_batch_t b) {
    size_t t = 0;
    for (void *p : b->owned) t += b->ctx->pool.size_of(p);
    if (b->ws_mul && b->ws_mul->owned)
        for (size_t i = 0; i < b->ws_mul->p.size(); ++i) t += b->ws_mul->p[i] ? b->ws_mul->sz[i] : 0;
    if (b->ws_add && b->ws_add->owned)
        for (size_t i = 0; i < b->ws_add->p.size(); ++i) t += b->ws_add->p[i] ? b->ws_add->sz[i] : 0;
    return t;
}

extern "C" {

int tr_batch_result_panels(tr_batch_t b, int32_t t, int64_t max_panel_nnz, tr_csx_t *panels, int32_t *row_begin,
                           int32_t max_panels, int32_t *npanels) {
    if (!b) return TR_ERR_INVALID;
    tr_ctx *ctx = b->ctx;
    return bguard(ctx, [&] {
        TR_REQUIRE(t >= 0 && t < b->nt && panels && row_begin && npanels && max_panels >= 1,
                   "tr_batch_result_panels: bad arguments");
        TR_REQUIRE(max_panel_nnz >= 1 && max_panel_nnz <= INT32_MAX,
                   "tr_batch_result_panels: max_panel_nnz must be in [1, INT32_MAX]");
        const int32_t m = b->A[t].nrows;
        // 64-bit row pointers of the task's result: the product's (no
        // accumulators) or, for an accumulated tile (always an int32 tile),
        // its own row pointers widened
        std::vector<int64_t> rp((size_t)m + 1);
        if (b->no_accum) {
            read_array(ctx, rp.data(), b->mp.rowptr + b->mp.h_tasks[t].row_base, sizeof(int64_t) * ((size_t)m + 1));
        } else {
            std::vector<int32_t> r32((size_t)m + 1);
            read_array(ctx, r32.data(), b->o_ptr + b->optr_off[t], sizeof(int32_t) * ((size_t)m + 1));
            for (size_t i = 0; i <= (size_t)m; ++i) rp[i] = b->o_base[t] + r32[i];
        }
        std::vector<int32_t> rb{0};
        int64_t start = rp[0];
        for (int32_t r = 0; r < m; ++r) {
            if (rp[r + 1] - rp[r] > max_panel_nnz) fail(TR_ERR_OVERFLOW, "tr_batch_result_panels: a row exceeds the panel capacity");
            if (rp[r + 1] - start > max_panel_nnz) {
                rb.push_back(r);
                start = rp[r];
            }
        }
        rb.push_back(m);
        const int32_t np = (int32_t)rb.size() - 1;
        *npanels = np;
        if (np > max_panels) fail(TR_ERR_OVERFLOW, "tr_batch_result_panels: needs " + std::to_string(np) + " panels");
        for (int32_t p = 0; p < np; ++p) {
            const int32_t r0 = rb[p], r1 = rb[p + 1];
            // int32 row pointers of the panel, rebased to its first entry
            // (batch-owned: valid until tr_batch_destroy)
            int32_t *ptr = (int32_t *)b->own(sizeof(int32_t) * ((size_t)(r1 - r0) + 1));
            if (b->no_accum) {
                const int32_t g0 = b->mp.h_tasks[t].row_base;
                launch_rowptr_i64_to_i32(ctx, b->mp.rowptr, g0 + r0, g0 + r1, ptr);
            } else {
                std::vector<int32_t> h((size_t)(r1 - r0) + 1);
                for (int32_t r = r0; r <= r1; ++r) h[r - r0] = (int32_t)(rp[r] - rp[r0]);
                TR_HIP(hipMemcpyAsync(ptr, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice, ctx->stream));
                TR_HIP(hipStreamSynchronize(ctx->stream));
            }
            tr_csx_t c{};
            c.majordim = TR_ROWS;
            c.nrows = r1 - r0;
            c.ncols = b->B[t].ncols;
            c.nnz = rp[r1] - rp[r0];
            c.ptr_vec = ptr;
            c.idx_vec = b->o_idx + rp[r0];
            c.val = b->o_val + rp[r0];
            panels[p] = c;
            row_begin[p] = r0;
        }
        row_begin[np] = m;
    });
}

int tr_batch_check(tr_batch_t b) {
    if (!b) return TR_ERR_INVALID;
    tr_ctx *ctx = b->ctx;
    return bguard(ctx, [&] {
        if (!b->guard) return;  // single-run batches are never replayed
        const int32_t g = read_scalar<int32_t>(ctx, b->guard);
        if (g)
            fail(TR_ERR_INVALID, "tr_batch_check: the last run found operands whose sparsity pattern differs from "
                                 "the planned one (guard bits " + std::to_string(g) + "); its results are not valid");
    });
}

int tr_batch_stats(tr_batch_t b, int64_t *flops, int64_t *nnz_products, int64_t *nnz_out) {
    if (!b) return TR_ERR_INVALID;
    if (flops) *flops = b->flops;
    if (nnz_products) *nnz_products = b->p_total;
    if (nnz_out) *nnz_out = b->o_total;
    return TR_OK;
}

int tr_batch_destroy(tr_batch_t b) {
    if (!b) return TR_ERR_INVALID;
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    return TR_OK;
}

}  // extern "C"
