// Sorted sparse add C = X + Y on gfx950 (partial-C accumulation).
//
// Replaces the spadd of LocalMatrix::spadd / sp_mma
// (include/KokkosWrap.hpp:341-352, :394-439: KokkosSparse spadd with
// alpha = beta = 1) and cusparse_spgeam (include/cusparse_helpers.cuh:
// 799-905: csrgeam2 bufferSize / Nnz / compute with a host sync after each).
//
// Merge-path formulation: every row's union is cut into chunks of kChunk
// merged positions; a wave owns a chunk, splits its start by binary search
// in global memory (ties take X's entry first, so an equal column appears as
// an adjacent (X, Y) pair that collapses into one output entry x + y) and
// walks it in LDS-staged sub-steps.  Pass 1 counts the kept positions per
// chunk, one scan over all chunks gives both the row pointers and every
// chunk's output offset, pass 2 writes.  Long rows are spread over many
// waves.
//
// Like the multiply, the add runs over a table of tasks (a plain call is a
// table of one; the batched stage executor adds all tile products into
// their partial-C tiles in one pass).  X rows are addressed either by an
// int32 pointer per task or by a 64-bit pointer over the global row space
// (the product arena of a batch).
#include "dev_common.h"
#include "tr_internal.h"
#include "spgemm_types.h"

namespace tr {
namespace {

using namespace dev;

constexpr int kChunk = 4096;  // merged positions per wave work unit

__device__ __forceinline__ int add_task_of(const AddTaskDev *__restrict__ tasks, int nt, int32_t g) {
    int lo = 0, hi = nt - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tasks[mid].row_base <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct RowSpan {
    int64_t xs, ys;
    int32_t lx, ly;
};

__device__ __forceinline__ RowSpan row_span(const AddTaskDev &T, int32_t g) {
    const int32_t r = g - T.row_base;
    RowSpan s;
    if (T.xp64) {
        const TR_GLOBAL int64_t *xp64 = as_global(T.xp64);
        s.xs = xp64[g];
        s.lx = (int32_t)(xp64[g + 1] - s.xs);
    } else {
        const TR_GLOBAL int32_t *xp = as_global(T.xp);
        s.xs = xp[r];
        s.lx = xp[r + 1] - xp[r];
    }
    const TR_GLOBAL int32_t *yp = as_global(T.yp);
    s.ys = yp[r];
    s.ly = yp[r + 1] - yp[r];
    return s;
}

__global__ void k_add_chunks(const AddTaskDev *__restrict__ tasks, int nt, int32_t rows, int32_t *__restrict__ nch) {
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= rows) return;
    const RowSpan s = row_span(tasks[add_task_of(tasks, nt, g)], g);
    const int64_t len = (int64_t)s.lx + (int64_t)s.ly;
    nch[g] = (int32_t)((len + kChunk - 1) / kChunk);
}

__global__ void k_add_chunk_rows(const int64_t *__restrict__ cstart, int32_t rows, int32_t *__restrict__ crow,
                                 int64_t cap) {
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= rows) return;
    const int64_t c1 = min(cstart[g + 1], cap);  // cap: the planned chunk capacity (replays)
    for (int64_t c = cstart[g]; c < c1; ++c) crow[c] = g;
}

// Replay guards of a planned add (a batch's graph): the chunk count must fit
// the planned chunk capacity and every task's output must start where the
// plan put it.  On a mismatch the guard word records it and the device-side
// chunk count (cstart[rows]) is zeroed, so the add passes write nothing.
__global__ void k_add_guard_chunks(int64_t *__restrict__ cstart, int32_t rows, int64_t cap,
                                   int32_t *__restrict__ guard) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (cstart[rows] > cap) atomicOr(guard, 2);
    if (*guard) cstart[rows] = 0;
}

__global__ void k_add_guard_output(const AddTaskDev *__restrict__ tasks, int nt, int64_t *__restrict__ cstart,
                                   const int64_t *__restrict__ coff, int32_t rows, int64_t nchunks,
                                   const int64_t *__restrict__ expect, int32_t *__restrict__ guard) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= nt && *guard == 0) {
        // expect[t]: planned first output entry of task t (expect[nt]: total)
        const int64_t got = t < nt ? coff[cstart[tasks[t].row_base]] : coff[nchunks];
        if (got != expect[t]) atomicOr(guard, 4);
    }
}

__global__ void k_add_guard_apply(int64_t *__restrict__ cstart, int32_t rows, const int32_t *__restrict__ guard) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && *guard) cstart[rows] = 0;
}

// merge-path split of diagonal t: number of X entries among the first t
// merged positions (ties take X first)
__device__ __forceinline__ int32_t merge_split(const TR_GLOBAL int32_t *__restrict__ xi, int32_t lx,
                                               const TR_GLOBAL int32_t *__restrict__ yi, int32_t ly, int32_t t) {
    int32_t lo = t - ly > 0 ? t - ly : 0;
    int32_t hi = t < lx ? t : lx;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (xi[mid] <= yi[t - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

constexpr int32_t kGuardLo = INT32_MIN;   // "no X entry before the tile"
constexpr int32_t kGuardHi = INT32_MAX;   // "no Y entry after the row"

// Setup of one chunk, loaded ahead of its use (values in lane 0)
struct ChunkAhead {
    int64_t xs, ys, cs, out;
    int32_t lx, ly, g, task;
};

// Lane 0's vector loads of chunk c's row spans, chunk start and output
// offset; the task (found through the scalar cache) and its row pointer
// arrays are given
template <bool kWrite>
__device__ __forceinline__ ChunkAhead chunk_ahead(const AddTaskDev &T, int task, const int64_t *__restrict__ cstart,
                                                  const int64_t *__restrict__ coff, int64_t c, int32_t g) {
    ChunkAhead a;
    a.g = g;
    a.task = task;
    const RowSpan s = row_span(T, g);
    a.xs = s.xs;
    a.ys = s.ys;
    a.lx = s.lx;
    a.ly = s.ly;
    a.cs = as_global(cstart)[g];
    a.out = kWrite ? as_global(coff)[c] : 0;
    return a;
}

// One wave per chunk of kChunk merged positions of a row.  The chunk start
// is split once in global memory (merge path, ties to X); the wave then walks
// the chunk in sub-steps of kSub positions.  Each sub-step stages the next
// kSub entries of X and Y (plus the X entry before them and the Y entry after
// them, which decide folding at the edges) in LDS with coalesced loads, every
// lane splits its own kPerLane positions inside the staged tile and merges
// them, and the wave advances by the X entries it consumed (a wave sum), so
// no further global search is needed.  A Y entry equal to the preceding X
// entry is folded into it (output x + y at the X position).  The chunk count
// is read on the device (cstart[rows]); grid-stride so the launch can be
// replayed from a graph.
template <bool kWrite>
__global__ __launch_bounds__(256, kWrite ? 6 : 8) void k_add_pass(const AddTaskDev *__restrict__ tasks, int nt, int32_t rows,
                                                  const int64_t *__restrict__ cstart,
                                                  const int32_t *__restrict__ crow, int32_t *__restrict__ ccount,
                                                  const int64_t *__restrict__ coff, int32_t *__restrict__ oi,
                                                  double *__restrict__ ov) {
    constexpr int kSub = kWrite ? 256 : 512;          // merged positions per LDS-staged sub-step
    constexpr int kPerLane = kSub / kWave;            // merged positions per lane per sub-step
    constexpr int kStage = (kSub + 1) / kWave + 1;    // tile entries (kSub + 1) per lane, rounded up
    constexpr int kVW = kWrite ? 4 : 1;
    constexpr int kVN = kWrite ? kSub + 2 : 1;
    __shared__ int32_t s_xi[4][kSub + 2], s_yi[4][kSub + 2];
    __shared__ double s_xv[kVW][kVN], s_yv[kVW][kVN];
    const int l = lane_id(), w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);  // wave-uniform
    int32_t *XI = s_xi[w], *YI = s_yi[w];
    double *XV = s_xv[kWrite ? w : 0], *YV = s_yv[kWrite ? w : 0];
    const int64_t nchunks = cstart[rows];
    const int64_t stride = (int64_t)gridDim.x * 4;
    int64_t c = (int64_t)blockIdx.x * 4 + w;
    // A chunk's setup (row -> task -> row spans, chunk start, output offset)
    // is a chain of dependent loads.  It runs one chunk ahead, the row two
    // ahead, in lane 0 with vector loads (they count on vmcnt; scalar loads
    // would share the LDS counter and be waited for by the merge's first
    // LDS wait).  The task search reads the small task table through the
    // scalar cache.
    ChunkAhead nx = {};
    int32_t g_after = 0;  // lane 0: row of the chunk after the next one
    if (c < nchunks) {
        const int32_t g = crow[c];
        const int ti = add_task_of(tasks, nt, g);
        const AddTaskDev &T = tasks[ti];
        if (l == 0) nx = chunk_ahead<kWrite>(T, ti, cstart, coff, c, g);
    }
    if (l == 0 && c + stride < nchunks) g_after = as_global(crow)[c + stride];
    for (; c < nchunks; c += stride) {
        const ChunkAhead cur = nx;
        const int32_t g = __builtin_amdgcn_readfirstlane(cur.g);
        const int ti = __builtin_amdgcn_readfirstlane(cur.task);
        const AddTaskDev &T = tasks[ti];
        RowSpan s;
        s.xs = ((int64_t)__builtin_amdgcn_readfirstlane((int32_t)(cur.xs >> 32)) << 32) |
               (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)cur.xs);
        s.ys = ((int64_t)__builtin_amdgcn_readfirstlane((int32_t)(cur.ys >> 32)) << 32) |
               (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)cur.ys);
        s.lx = __builtin_amdgcn_readfirstlane(cur.lx);
        s.ly = __builtin_amdgcn_readfirstlane(cur.ly);
        const int64_t cs = ((int64_t)__builtin_amdgcn_readfirstlane((int32_t)(cur.cs >> 32)) << 32) |
                           (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)cur.cs);
        int64_t out = kWrite ? (((int64_t)__builtin_amdgcn_readfirstlane((int32_t)(cur.out >> 32)) << 32) |
                                (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)cur.out))
                             : 0;
        // the next chunk's setup and the row of the one after it
        const int64_t cn = c + stride;
        if (cn < nchunks) {
            const int32_t gn = __builtin_amdgcn_readfirstlane(g_after);
            const int tn = add_task_of(tasks, nt, gn);
            const AddTaskDev &TN = tasks[tn];
            if (l == 0) nx = chunk_ahead<kWrite>(TN, tn, cstart, coff, cn, gn);
            if (l == 0 && cn + stride < nchunks) g_after = as_global(crow)[cn + stride];
        }
        const TR_GLOBAL int32_t *__restrict__ xi = as_global(T.xi) + s.xs;
        const TR_GLOBAL double *__restrict__ xv = as_global(T.xv) + s.xs;
        const TR_GLOBAL int32_t *__restrict__ yi = as_global(T.yi) + s.ys;
        const TR_GLOBAL double *__restrict__ yv = as_global(T.yv) + s.ys;
        const int32_t t0 = (int32_t)(c - cs) * kChunk;
        const int32_t t1 = min(t0 + kChunk, s.lx + s.ly);
        int32_t sp = 0;
        if (l == 0 && t0 > 0) sp = merge_split(xi, s.lx, yi, s.ly, t0);
        int32_t i0 = __shfl(sp, 0, kWave);
        int32_t j0 = t0 - i0;
        int32_t kept_total = 0;
        for (int32_t tb = t0; tb < t1; tb += kSub) {
            const int32_t n = min(kSub, t1 - tb);
            const int32_t nxa = min(kSub, s.lx - i0), nya = min(kSub, s.ly - j0);
            // XI[u] = X[i0 - 1 + u] for u in [0, nxa]; YI[u] = Y[j0 + u] for u in [0, nya];
            // all loads of the tile are issued before the first LDS store
            int32_t rxi[kStage], ryi[kStage];
            double rxv[kStage], ryv[kStage];
#pragma unroll
            for (int k = 0; k < kStage; ++k) {
                const int u = l + k * kWave;
                const int32_t ii = i0 - 1 + u, jj = j0 + u;
                rxi[k] = kGuardLo;
                ryi[k] = kGuardHi;
                rxv[k] = ryv[k] = 0.0;
                if (u <= nxa && ii >= 0) {
                    rxi[k] = xi[ii];
                    if (kWrite) rxv[k] = xv[ii];
                }
                if (u <= nya && jj < s.ly) {
                    ryi[k] = yi[jj];
                    if (kWrite) ryv[k] = yv[jj];
                }
            }
#pragma unroll
            for (int k = 0; k < kStage; ++k) {
                const int u = l + k * kWave;
                if (u <= nxa) {
                    XI[u] = rxi[k];
                    if (kWrite) XV[u] = rxv[k];
                }
                if (u <= nya) {
                    YI[u] = ryi[k];
                    if (kWrite) YV[u] = ryv[k];
                }
            }
            wave_sync();
            // this lane's first position and its split inside the staged tile
            const int32_t d = l * kPerLane;
            int32_t i = 0, j = 0;
            if (d < n) {
                int32_t lo = d - nya > 0 ? d - nya : 0, hi = d < nxa ? d : nxa;
                while (lo < hi) {
                    const int32_t mid = (lo + hi) >> 1;
                    if (XI[1 + mid] <= YI[d - 1 - mid]) lo = mid + 1;
                    else hi = mid;
                }
                i = lo;
                j = d - lo;
            }
            int32_t keys[kPerLane];
            double vals[kPerLane];
            uint32_t keep = 0;
            int32_t xcnt = 0;
#pragma unroll
            for (int u = 0; u < kPerLane; ++u) {
                keys[u] = 0;
                vals[u] = 0.0;
                if (d + u >= n) continue;
                const bool fromX = (i < nxa) && (j >= nya || XI[1 + i] <= YI[j]);
                if (fromX) {
                    keys[u] = XI[1 + i];
                    if (kWrite) {
                        vals[u] = XV[1 + i];
                        if (YI[j] == keys[u]) vals[u] = vals[u] + YV[j];
                    }
                    keep |= 1u << u;
                    ++i;
                    ++xcnt;
                } else {
                    keys[u] = YI[j];
                    if (XI[i] != keys[u]) {  // else folded into the preceding X entry
                        if (kWrite) vals[u] = YV[j];
                        keep |= 1u << u;
                    }
                    ++j;
                }
            }
            const int nkeep = __popc(keep);
            const int32_t incl = wave_incl_scan((int32_t)nkeep);
            const int32_t kept = __shfl(incl, kWave - 1, kWave);
            if (kWrite) {
                // kept entries staged at their wave-local offsets, then written
                // with consecutive lanes on consecutive entries
                wave_sync();
                int32_t q = incl - nkeep;
#pragma unroll
                for (int u = 0; u < kPerLane; ++u)
                    if (keep & (1u << u)) {
                        XI[q] = keys[u];
                        XV[q] = vals[u];
                        ++q;
                    }
                wave_sync();
                for (int32_t e = l; e < kept; e += kWave) {
                    __builtin_nontemporal_store(XI[e], &oi[out + e]);
                    __builtin_nontemporal_store(XV[e], &ov[out + e]);
                }
            }
            out += kept;
            kept_total += kept;
            const int32_t xs = wave_sum(xcnt);
            i0 += xs;
            j0 += n - xs;
            wave_sync();  // XI/YI are rewritten by the next sub-step
        }
        if (!kWrite && l == 0) ccount[c] = kept_total;
    }
}

// int32 row pointers of every task's output, rebased to the task
__global__ void k_add_rowptr(const AddTaskDev *__restrict__ tasks, int nt, int32_t rows,
                             const int64_t *__restrict__ cstart, const int64_t *__restrict__ coff) {
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < rows) {
        const AddTaskDev &T = tasks[add_task_of(tasks, nt, g)];
        const int64_t base = coff[cstart[T.row_base]];
        const int32_t r = g - T.row_base;
        TR_GLOBAL int32_t *optr = as_global(T.optr);
        optr[r] = (int32_t)(coff[cstart[g]] - base);
        if (r == T.m - 1) optr[T.m] = (int32_t)(coff[cstart[g + 1]] - base);
    }
    if (g < nt && tasks[g].m == 0) as_global(tasks[g].optr)[0] = 0;
}

// the task table of an allocation-free add, written by a kernel from its
// by-value argument (no host staging buffer)
__global__ void k_put_add_task(AddTaskDev *__restrict__ dst, AddTaskDev t) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = t;
}

constexpr size_t kAlign = 256;
inline size_t align_up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

// Byte sizes of the pieces of a one-task add over `rows` rows, in the order
// they are carved: task, nch, cstart, crow, ccount, coff, scan temp.
struct AddFixedLayout {
    size_t piece[7];
    int64_t chunk_cap;
    size_t total() const {
        size_t t = 0;
        for (size_t p : piece) t += align_up(p);
        return t;
    }
};

AddFixedLayout add_fixed_layout(int32_t rows, int64_t nnz_x, int64_t nnz_y) {
    AddFixedLayout L;
    const size_t R = (size_t)rows + 1;
    L.chunk_cap = (int64_t)rows + (nnz_x + nnz_y) / kChunk + 1;
    L.piece[0] = sizeof(AddTaskDev);
    L.piece[1] = sizeof(int32_t) * R;
    L.piece[2] = sizeof(int64_t) * R;
    L.piece[3] = sizeof(int32_t) * (size_t)L.chunk_cap;
    L.piece[4] = sizeof(int32_t) * ((size_t)L.chunk_cap + 1);
    L.piece[5] = sizeof(int64_t) * ((size_t)L.chunk_cap + 1);
    L.piece[6] = scan_tmp_bytes(std::max<int64_t>(L.chunk_cap, rows));
    return L;
}

}  // namespace

size_t spadd_fixed_bytes(int32_t rows, int64_t nnz_x, int64_t nnz_y) {
    return add_fixed_layout(rows, nnz_x, nnz_y).total();
}

void add_setup(tr_ctx *ctx, Workspace &ws, AddPlan &ap, const std::vector<AddTaskDev> &tasks, int64_t nnz_x,
               int64_t nnz_y) {
    ap.nt = (int)tasks.size();
    ap.h = tasks;
    int64_t rows = 0;
    for (auto &t : ap.h) {
        t.row_base = (int32_t)rows;
        rows += t.m;
    }
    TR_REQUIRE(rows < INT32_MAX, "spadd: total rows exceed INT32_MAX");
    ap.rows_total = (int32_t)rows;
    ap.chunk_cap = rows + (nnz_x + nnz_y) / kChunk + 1;
    const size_t R = (size_t)rows + 1;
    ap.d = (AddTaskDev *)ws.get(S_PANELS, sizeof(AddTaskDev) * (size_t)std::max(ap.nt, 1));
    ap.nch = (int32_t *)ws.get(S_MISC, sizeof(int32_t) * R);
    ap.cstart = (int64_t *)ws.get(S_ROWPTR64, sizeof(int64_t) * R);
    ap.crow = (int32_t *)ws.get(S_BINS, sizeof(int32_t) * (size_t)ap.chunk_cap);
    ap.ccount = (int32_t *)ws.get(S_MISC2, sizeof(int32_t) * ((size_t)ap.chunk_cap + 1));
    ap.coff = (int64_t *)ws.get(S_FLOPS, sizeof(int64_t) * ((size_t)ap.chunk_cap + 1));
    if (ws.owned)
        ap.scan_tmp = (int64_t *)ws.get(S_SCAN_TMP, scan_tmp_bytes(std::max<int64_t>(ap.chunk_cap, rows)));
    TR_HIP(hipMemcpyAsync(ap.d, ap.h.data(), sizeof(AddTaskDev) * (size_t)ap.nt, hipMemcpyHostToDevice,
                          ctx->stream));
}

void add_count(tr_ctx *ctx, AddPlan &ap, bool capture) {
    const int32_t rows = ap.rows_total;
    TR_LAUNCH(ctx, "add_chunks", k_add_chunks, div_up((uint64_t)rows, 256), 256, 0, ap.d, ap.nt, rows, ap.nch);
    launch_exclusive_scan_i32_to_i64(ctx, ap.nch, ap.cstart, rows, ap.scan_tmp);
    TR_LAUNCH(ctx, "add_chunk_rows", k_add_chunk_rows, div_up((uint64_t)rows, 256), 256, 0, ap.cstart, rows, ap.crow,
              ap.chunk_cap);
    if (capture && ap.guard)
        TR_LAUNCH(ctx, "add_guard", k_add_guard_chunks, 1, 64, 0, ap.cstart, rows, ap.chunk_cap, ap.guard);
    if (!capture) {
        ap.nchunks = read_scalar<int64_t>(ctx, ap.cstart + rows);
        TR_REQUIRE(ap.nchunks <= ap.chunk_cap, "spadd: chunk bound violated");
    }
    TR_LAUNCH(ctx, "add_count", k_add_pass<false>, div_up((uint64_t)std::max<int64_t>(ap.nchunks, 1), 4), 256, 0,
              ap.d, ap.nt, rows, ap.cstart, ap.crow, ap.ccount, (const int64_t *)nullptr, (int32_t *)nullptr,
              (double *)nullptr);
    launch_exclusive_scan_i32_to_i64(ctx, ap.ccount, ap.coff, ap.nchunks, ap.scan_tmp);
    if (capture && ap.guard && ap.expect) {
        TR_LAUNCH(ctx, "add_guard", k_add_guard_output, div_up((uint64_t)ap.nt + 1, 64), 64, 0, ap.d, ap.nt,
                  ap.cstart, ap.coff, rows, ap.nchunks, ap.expect, ap.guard);
        TR_LAUNCH(ctx, "add_guard", k_add_guard_apply, 1, 64, 0, ap.cstart, rows, ap.guard);
    }
}

int64_t add_total(tr_ctx *ctx, AddPlan &ap) { return read_scalar<int64_t>(ctx, ap.coff + ap.nchunks); }

void add_write(tr_ctx *ctx, AddPlan &ap, int32_t *oi, double *ov) {
    const int32_t rows = ap.rows_total;
    TR_LAUNCH(ctx, "add_rowptr", k_add_rowptr, div_up((uint64_t)std::max<int32_t>(rows, ap.nt), 256), 256, 0, ap.d,
              ap.nt, rows, ap.cstart, ap.coff);
    TR_LAUNCH(ctx, "add_write", k_add_pass<true>, div_up((uint64_t)std::max<int64_t>(ap.nchunks, 1), 4), 256, 0,
              ap.d, ap.nt, rows, ap.cstart, ap.crow, (int32_t *)nullptr, ap.coff, oi, ov);
}

// optr/oi/ov = A + C (both CSR, sorted) into caller arrays with room for
// `capacity` entries; returns nnz(A + C) and writes nothing if it exceeds
// the capacity.
int64_t spadd_into(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *C, int32_t *optr, int32_t *oi, double *ov,
                   int64_t capacity) {
    Workspace ws(ctx, false);
    AddPlan ap;
    AddTaskDev t{};
    t.xp = A->ptr_vec;
    t.xi = A->idx_vec;
    t.xv = A->val;
    t.yp = C->ptr_vec;
    t.yi = C->idx_vec;
    t.yv = C->val;
    t.m = A->nrows;
    t.optr = optr;
    add_setup(ctx, ws, ap, {t}, A->nnz, C->nnz);
    add_count(ctx, ap, false);
    const int64_t nnz = add_total(ctx, ap);
    if (nnz > capacity) return nnz;
    add_write(ctx, ap, oi, ov);
    return nnz;
}

int64_t spadd_into_fixed(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *C, int32_t *optr, int32_t *oi, double *ov,
                         int64_t capacity, char *base, size_t bytes) {
    const AddFixedLayout L = add_fixed_layout(A->nrows, A->nnz, C->nnz);
    TR_REQUIRE(L.total() <= bytes, "spadd: fixed scratch region too small");
    char *q = base;
    auto carve = [&](size_t b) {
        char *p = q;
        q += align_up(b);
        return p;
    };
    AddPlan ap;
    AddTaskDev t{};
    t.xp = A->ptr_vec;
    t.xi = A->idx_vec;
    t.xv = A->val;
    t.yp = C->ptr_vec;
    t.yi = C->idx_vec;
    t.yv = C->val;
    t.m = A->nrows;
    t.optr = optr;
    t.row_base = 0;
    ap.nt = 1;
    ap.h = {t};
    ap.rows_total = A->nrows;
    ap.chunk_cap = L.chunk_cap;
    ap.d = (AddTaskDev *)carve(L.piece[0]);
    ap.nch = (int32_t *)carve(L.piece[1]);
    ap.cstart = (int64_t *)carve(L.piece[2]);
    ap.crow = (int32_t *)carve(L.piece[3]);
    ap.ccount = (int32_t *)carve(L.piece[4]);
    ap.coff = (int64_t *)carve(L.piece[5]);
    ap.scan_tmp = (int64_t *)carve(L.piece[6]);
    TR_LAUNCH(ctx, "put_add_task", k_put_add_task, 1, 64, 0, ap.d, t);
    add_count(ctx, ap, false);
    const int64_t nnz = add_total(ctx, ap);
    if (nnz > capacity) return nnz;
    add_write(ctx, ap, oi, ov);
    return nnz;
}

// out = A + C (both CSR, sorted), out context-owned.
void spadd(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *C, tr_csx_t *out) {
    Workspace ws(ctx, false);
    AddPlan ap;
    AddTaskDev t{};
    t.xp = A->ptr_vec;
    t.xi = A->idx_vec;
    t.xv = A->val;
    t.yp = C->ptr_vec;
    t.yi = C->idx_vec;
    t.yv = C->val;
    t.m = A->nrows;
    // the output row pointer is written into the tile allocated below;
    // point the task at it after allocation
    add_setup(ctx, ws, ap, {t}, A->nnz, C->nnz);
    add_count(ctx, ap, false);
    const int64_t nnz = add_total(ctx, ap);
    if (nnz > INT32_MAX) fail(TR_ERR_OVERFLOW, "spadd: nnz(A + C) exceeds INT32_MAX");
    alloc_tile(ctx, TR_ROWS, A->nrows, A->ncols, nnz, out);
    ap.h[0].optr = out->ptr_vec;
    TR_HIP(hipMemcpyAsync(ap.d, ap.h.data(), sizeof(AddTaskDev), hipMemcpyHostToDevice, ctx->stream));
    add_write(ctx, ap, out->idx_vec, out->val);
}

}  // namespace tr
