// Sorted sparse add C = A + B on gfx950 (partial-C accumulation).
//
// Replaces the spadd of LocalMatrix::spadd / sp_mma
// (include/KokkosWrap.hpp:341-352, :394-439: KokkosSparse spadd with
// alpha = beta = 1) and cusparse_spgeam (include/cusparse_helpers.cuh:
// 799-905: csrgeam2 bufferSize / Nnz / compute with a host sync after each).
//
// Merge-path formulation: every row's union is cut into chunks of kChunk
// merged positions; a wave owns a chunk and finds, per lane, the merge-path
// split of its position by binary search (ties take A's entry first, so an
// equal column appears as an adjacent (A, B) pair that collapses into one
// output entry a + b).  Pass 1 counts the kept positions per chunk, one scan
// over all chunks gives both the row pointer and every chunk's output offset,
// pass 2 writes.  Long rows are spread over many waves; there is no host
// round trip except the single read of the output size.
#include "dev_common.h"
#include "tr_internal.h"

namespace tr {
namespace {

using namespace dev;

constexpr int kChunk = 256;

__global__ void k_add_chunks(const int32_t *__restrict__ ap, const int32_t *__restrict__ bp, int32_t m,
                             int32_t *__restrict__ nch) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int64_t len = (int64_t)(ap[i + 1] - ap[i]) + (int64_t)(bp[i + 1] - bp[i]);
    nch[i] = (int32_t)((len + kChunk - 1) / kChunk);
}

__global__ void k_add_chunk_rows(const int64_t *__restrict__ cstart, int32_t m, int32_t *__restrict__ crow) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    for (int64_t c = cstart[i]; c < cstart[i + 1]; ++c) crow[c] = i;
}

// merged element at position t of row (A_i, B_i); returns key, value and
// whether the position is kept (a B entry equal to the preceding A entry is
// folded into it)
__device__ __forceinline__ bool merge_at(const int32_t *__restrict__ ai, const double *__restrict__ av, int32_t la,
                                         const int32_t *__restrict__ bi, const double *__restrict__ bv, int32_t lb,
                                         int32_t t, int32_t &key, double &val, bool want_val) {
    int32_t lo = t - lb > 0 ? t - lb : 0;
    int32_t hi = t < la ? t : la;
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if (ai[mid] <= bi[t - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    const int32_t i = lo, j = t - lo;
    bool fromA = (i < la) && (j >= lb || ai[i] <= bi[j]);
    if (fromA) {
        key = ai[i];
        if (want_val) {
            val = av[i];
            if (j < lb && bi[j] == key) val = val + bv[j];
        }
        return true;
    }
    key = bi[j];
    if (i > 0 && ai[i - 1] == key) return false;  // folded into the A entry
    if (want_val) val = bv[j];
    return true;
}

template <bool kWrite>
__global__ __launch_bounds__(256) void k_add_pass(const int32_t *__restrict__ ap, const int32_t *__restrict__ ai,
                                                  const double *__restrict__ av, const int32_t *__restrict__ bp,
                                                  const int32_t *__restrict__ bi, const double *__restrict__ bv,
                                                  const int64_t *__restrict__ cstart,
                                                  const int32_t *__restrict__ crow, int64_t nchunks,
                                                  int32_t *__restrict__ ccount, const int64_t *__restrict__ coff,
                                                  int32_t *__restrict__ oi, double *__restrict__ ov) {
    const int l = lane_id();
    const int64_t c = (int64_t)blockIdx.x * 4 + threadIdx.x / kWave;
    if (c >= nchunks) return;
    const int32_t row = crow[c];
    const int32_t a0 = ap[row], la = ap[row + 1] - a0;
    const int32_t b0 = bp[row], lb = bp[row + 1] - b0;
    const int32_t t0 = (int32_t)(c - cstart[row]) * kChunk;
    const int32_t tend = min(t0 + kChunk, la + lb);
    int64_t out = kWrite ? coff[c] : 0;
    int32_t kept_total = 0;
    for (int32_t tb = t0; tb < tend; tb += kWave) {
        const int32_t t = tb + l;
        int32_t key = 0;
        double val = 0.0;
        bool keep = false;
        if (t < tend) keep = merge_at(ai + a0, av + a0, la, bi + b0, bv + b0, lb, t, key, val, kWrite);
        const uint64_t msk = ballot(keep);
        if (kWrite && keep) {
            const int64_t q = out + rank_below(msk);
            oi[q] = key;
            ov[q] = val;
        }
        out += __popcll(msk);
        kept_total += __popcll(msk);
    }
    if (!kWrite && l == 0) ccount[c] = kept_total;
}

__global__ void k_add_rowptr(const int64_t *__restrict__ cstart, const int64_t *__restrict__ coff, int32_t m,
                             int32_t *__restrict__ optr) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > m) return;
    optr[i] = (int32_t)coff[cstart[i]];
}

}  // namespace

// out = A + C (both CSR, sorted), out context-owned.
void spadd(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *C, tr_csx_t *out) {
    const int32_t m = A->nrows;
    int32_t *nch = (int32_t *)scratch(ctx, S_MISC, sizeof(int32_t) * ((size_t)m + 1));
    int64_t *cstart = (int64_t *)scratch(ctx, S_ROWPTR64, sizeof(int64_t) * ((size_t)m + 1));
    TR_LAUNCH(ctx, "add_chunks", k_add_chunks, div_up((uint64_t)m, 256), 256, 0, A->ptr_vec, C->ptr_vec, m, nch);
    launch_exclusive_scan_i32_to_i64(ctx, nch, cstart, m);
    // chunk count bound: m + (nnzA + nnzC) / kChunk
    const int64_t cap = (int64_t)m + (A->nnz + C->nnz) / kChunk + 1;
    int32_t *crow = (int32_t *)scratch(ctx, S_BINS, sizeof(int32_t) * (size_t)cap);
    int32_t *ccount = (int32_t *)scratch(ctx, S_MISC2, sizeof(int32_t) * ((size_t)cap + 1));
    int64_t *coff = (int64_t *)scratch(ctx, S_FLOPS, sizeof(int64_t) * ((size_t)cap + 1));
    TR_LAUNCH(ctx, "add_chunk_rows", k_add_chunk_rows, div_up((uint64_t)m, 256), 256, 0, cstart, m, crow);
    const int64_t nchunks = read_scalar<int64_t>(ctx, cstart + m);
    TR_REQUIRE(nchunks <= cap, "spadd: chunk bound violated");
    TR_LAUNCH(ctx, "add_count", k_add_pass<false>, div_up((uint64_t)nchunks, 4), 256, 0, A->ptr_vec, A->idx_vec,
              A->val, C->ptr_vec, C->idx_vec, C->val, cstart, crow, nchunks, ccount, (const int64_t *)nullptr,
              (int32_t *)nullptr, (double *)nullptr);
    launch_exclusive_scan_i32_to_i64(ctx, ccount, coff, nchunks);
    const int64_t nnz = read_scalar<int64_t>(ctx, coff + nchunks);
    if (nnz > INT32_MAX) fail(TR_ERR_OVERFLOW, "spadd: nnz(A + C) exceeds INT32_MAX");
    alloc_tile(ctx, TR_ROWS, m, A->ncols, nnz, out);
    TR_LAUNCH(ctx, "add_rowptr", k_add_rowptr, div_up((uint64_t)m + 1, 256), 256, 0, cstart, coff, m, out->ptr_vec);
    TR_LAUNCH(ctx, "add_write", k_add_pass<true>, div_up((uint64_t)nchunks, 4), 256, 0, A->ptr_vec, A->idx_vec,
              A->val, C->ptr_vec, C->idx_vec, C->val, cstart, crow, nchunks, (int32_t *)nullptr, coff, out->idx_vec,
              out->val);
}

}  // namespace tr
