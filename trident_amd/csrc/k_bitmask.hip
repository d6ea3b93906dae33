// Bitmask tile compression on gfx950.
//
// Replaces SpaComm::gen_bitmask (include/spacomm.cuh:29-117),
// intersect_bitmasks (:126-148) and SpaCommHandler::Compress (:962-1091,
// with select_entries_cuda :808-889 and select_ptrs :556-623).
//
// Mask words are 64 bits: one wave64 ballot over 64 consecutive segments
// produces one word directly (the reference builds 8-bit words with a
// transform, a segmented reduction and a host round trip).  Compression is
// segment-parallel: a wave owns 64 consecutive segments, whose entries are
// contiguous in the input, so reads and writes stream; there is no per-entry
// binary search over the pointer array and values and indices move in the
// same pass.
#include "dev_common.h"
#include "tr_internal.h"

namespace tr {
namespace {

using namespace dev;

__global__ void k_bitmask_gen(const int32_t *__restrict__ ptr, int32_t n, uint64_t *__restrict__ mask) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool nonempty = false;
    if (i < n) nonempty = ptr[i] != ptr[i + 1];
    uint64_t word = ballot(nonempty);
    if (lane_id() == 0 && i < n) mask[i >> 6] = word;
}

__global__ void k_bitmask_and(const uint64_t *__restrict__ a, const uint64_t *__restrict__ b, int32_t words,
                              uint64_t *__restrict__ out) {
    int32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < words) out[w] = a[w] & b[w];
}

// kept[i] = bit(i) ? len(i) : 0
__global__ void k_compress_count(const int32_t *__restrict__ ptr, int32_t n, const uint64_t *__restrict__ mask,
                                 int32_t *__restrict__ kept) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool keep = (mask[i >> 6] >> (i & 63)) & 1ull;
    kept[i] = keep ? ptr[i + 1] - ptr[i] : 0;
}

// one wave per 64 consecutive segments; flattened over their entries.
// Reads stay in [0, nnz_in) and writes in [0, cap) whatever the pointer
// array holds (a malformed one gives a wrong tile, never a stray access).
__global__ void k_compress_write(const int32_t *__restrict__ ptr, const int32_t *__restrict__ idx,
                                 const double *__restrict__ val, int32_t n, const uint64_t *__restrict__ mask,
                                 const int64_t *__restrict__ newptr64, int32_t *__restrict__ optr,
                                 int32_t *__restrict__ oidx, double *__restrict__ oval, int64_t nnz_in,
                                 int64_t cap) {
    const int l = lane_id();
    const int64_t g = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int64_t s0 = g * kWave;
    if (s0 >= n) return;
    const int64_t s = s0 + l;
    const bool valid = s < n;
    // segment start (as exclusive prefix over the group) and keep flag
    int32_t beg = valid ? ptr[s] : ptr[n];
    bool keep = valid && ((mask[s >> 6] >> (s & 63)) & 1ull);
    int64_t nbeg = valid ? newptr64[s] : 0;
    if (valid) optr[s] = (int32_t)nbeg;
    if (s == n - 1) optr[n] = (int32_t)newptr64[n];
    const int32_t gbeg = __shfl(beg, 0, kWave);
    int32_t last = (s0 + kWave < n) ? ptr[s0 + kWave] : ptr[n];
    last = __shfl(last, 0, kWave);
    const int64_t rel = (int64_t)(beg - gbeg);
    for (int64_t x0 = 0; x0 < (int64_t)(last - gbeg); x0 += kWave) {
        int64_t x = x0 + l;
        int o = owner_lane(rel, x);
        // owner_lane may land on a trailing invalid lane only if x is past
        // the group's entries, which the bound below excludes
        int32_t ob = __shfl(beg, o, kWave);
        int keep_o = __shfl((int)keep, o, kWave);
        int64_t nb_o = __shfl(nbeg, o, kWave);
        if (x < (int64_t)(last - gbeg) && keep_o) {
            int64_t p = (int64_t)gbeg + x;
            int64_t q = nb_o + (p - ob);
            if (p >= 0 && p < nnz_in && q >= 0 && q < cap) {
                oidx[q] = idx[p];
                oval[q] = val[p];
            }
        }
    }
}

}  // namespace

void launch_bitmask_gen(tr_ctx *ctx, const int32_t *ptr, int32_t n, uint64_t *mask) {
    if (n <= 0) return;
    TR_LAUNCH(ctx, "bitmask_gen", k_bitmask_gen, div_up((uint64_t)n, 256), 256, 0, ptr, n, mask);
}

void launch_bitmask_and(tr_ctx *ctx, const uint64_t *a, const uint64_t *b, int32_t words, uint64_t *out) {
    if (words <= 0) return;
    TR_LAUNCH(ctx, "bitmask_and", k_bitmask_and, div_up((uint64_t)words, 256), 256, 0, a, b, words, out);
}

// Segment compression into caller-provided arrays with capacity for the
// input's nnz; no host round trip (usable inside a captured graph).
void launch_compress_fixed(tr_ctx *ctx, int32_t n, const int32_t *ptr, const int32_t *idx, const double *val,
                           int64_t nnz_in, const uint64_t *mask, int32_t *kept_tmp, int64_t *np64, int64_t *scan_tmp,
                           int32_t *optr, int32_t *oidx, double *oval) {
    if (n <= 0) return;
    TR_LAUNCH(ctx, "compress_count", k_compress_count, div_up((uint64_t)n, 256), 256, 0, ptr, n, mask, kept_tmp);
    launch_exclusive_scan_i32_to_i64(ctx, kept_tmp, np64, n, scan_tmp);
    const int wpb = 4;
    const uint32_t groups = div_up((uint64_t)n, kWave);
    TR_LAUNCH(ctx, "compress_write", k_compress_write, div_up(groups, wpb), wpb * kWave, 0, ptr, idx, val, n, mask,
              np64, optr, oidx, oval, nnz_in, nnz_in);
}

int64_t compress(tr_ctx *ctx, const tr_csx_t *m, const uint64_t *mask, tr_csx_t *out) {
    const int32_t n = ptrdim(m);
    int32_t *kept = (int32_t *)scratch(ctx, S_MISC, sizeof(int32_t) * ((size_t)n + 1));
    int64_t *np64 = (int64_t *)scratch(ctx, S_ROWPTR64, sizeof(int64_t) * ((size_t)n + 1));
    TR_LAUNCH(ctx, "compress_count", k_compress_count, div_up((uint64_t)n, 256), 256, 0, m->ptr_vec, n, mask, kept);
    launch_exclusive_scan_i32_to_i64(ctx, kept, np64, n);
    int64_t nnz = read_scalar<int64_t>(ctx, np64 + n);
    alloc_tile(ctx, m->majordim, m->nrows, m->ncols, nnz, out);
    if (n == 0) {
        TR_HIP(hipMemsetAsync(out->ptr_vec, 0, sizeof(int32_t), ctx->stream));
        return nnz;
    }
    const int wpb = 4;
    uint32_t groups = div_up((uint64_t)n, kWave);
    TR_LAUNCH(ctx, "compress_write", k_compress_write, div_up(groups, wpb), wpb * kWave, 0, m->ptr_vec, m->idx_vec,
              m->val, n, mask, np64, out->ptr_vec, out->idx_vec, out->val, m->nnz, nnz);
    return nnz;
}

}  // namespace tr
