// Utility kernels of libtrident_hip: scans, CSR validation, transposition
// and per-row sorting.  These are not the hot path; they are written for
// correctness with coalesced access and no host round trips.
#include "dev_common.h"
#include "tr_internal.h"

namespace tr {
namespace {

using namespace dev;

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

template <typename Tin>
__global__ void k_scan_reduce(const Tin *__restrict__ in, int64_t n, int64_t *__restrict__ tile_sums) {
    __shared__ int64_t wsum[kScanThreads / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    int64_t s = 0;
    for (int t = 0; t < kScanItems; ++t) {
        int64_t i = base + (int64_t)t * kScanThreads + threadIdx.x;
        if (i < n) s += (int64_t)in[i];
    }
    s = wave_sum(s);
    if (lane_id() == 0) wsum[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t tot = 0;
        for (int w = 0; w < kScanThreads / kWave; ++w) tot += wsum[w];
        tile_sums[blockIdx.x] = tot;
    }
}

// Exclusive scan of the tile sums in place, single block; writes the grand
// total to *total.
__global__ void k_scan_tiles(int64_t *__restrict__ sums, int64_t ntiles, int64_t *__restrict__ total) {
    __shared__ int64_t wsum[kScanThreads / kWave];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < ntiles; base += kScanThreads) {
        int64_t i = base + threadIdx.x;
        int64_t v = i < ntiles ? sums[i] : 0;
        int64_t inc = wave_incl_scan(v);
        if (lane_id() == kWave - 1) wsum[threadIdx.x / kWave] = inc;
        __syncthreads();
        int64_t off = carry;
        for (int w = 0; w < (int)(threadIdx.x / kWave); ++w) off += wsum[w];
        if (i < ntiles) sums[i] = off + inc - v;
        __syncthreads();
        if (threadIdx.x == kScanThreads - 1) carry = off + inc;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

template <typename Tin>
__global__ void k_scan_apply(const Tin *__restrict__ in, int64_t n, const int64_t *__restrict__ tile_off,
                             int64_t *__restrict__ out) {
    __shared__ int64_t wsum[kScanThreads / kWave];
    __shared__ int64_t carry;
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    if (threadIdx.x == 0) carry = tile_off[blockIdx.x];
    __syncthreads();
    for (int t = 0; t < kScanItems; ++t) {
        int64_t i = base + (int64_t)t * kScanThreads + threadIdx.x;
        int64_t v = i < n ? (int64_t)in[i] : 0;
        int64_t inc = wave_incl_scan(v);
        if (lane_id() == kWave - 1) wsum[threadIdx.x / kWave] = inc;
        __syncthreads();
        int64_t off = carry;
        for (int w = 0; w < (int)(threadIdx.x / kWave); ++w) off += wsum[w];
        if (i < n) out[i] = off + inc - v;
        __syncthreads();
        if (threadIdx.x == kScanThreads - 1) carry = off + inc;
        __syncthreads();
    }
}

__global__ void k_store0_i64(int64_t *p) { *p = 0; }

template <typename Tin>
void scan_impl(tr_ctx *ctx, const Tin *in, int64_t *out, int64_t n, int64_t *tmp) {
    // out[0..n] = exclusive prefix, out[n] = total
    int64_t ntiles = (n + kScanTile - 1) / kScanTile;
    if (ntiles == 0) {
        TR_LAUNCH(ctx, "scan_empty", k_store0_i64, 1, 1, 0, out);
        return;
    }
    int64_t *sums = tmp ? tmp : (int64_t *)scratch(ctx, S_SCAN_TMP, sizeof(int64_t) * (size_t)ntiles);
    TR_LAUNCH(ctx, "scan_reduce", k_scan_reduce<Tin>, (uint32_t)ntiles, kScanThreads, 0, in, n, sums);
    TR_LAUNCH(ctx, "scan_tiles", k_scan_tiles, 1, kScanThreads, 0, sums, ntiles, out + n);
    TR_LAUNCH(ctx, "scan_apply", k_scan_apply<Tin>, (uint32_t)ntiles, kScanThreads, 0, in, n, sums, out);
}

// plain streaming copy, 16 bytes per lane per iteration (counter calibration)
__global__ void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n16) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// ---- validation ----
// status bits: 1 ptr malformed, 2 idx out of range, 4 unsorted/duplicate.
__global__ void k_validate_ptr(const int32_t *__restrict__ ptr, int32_t P, int64_t nnz,
                               uint32_t *__restrict__ starts, int32_t *__restrict__ status) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > P) return;
    int32_t v = ptr[i];
    bool bad = (v < 0) || ((int64_t)v > nnz);
    if (i == 0 && v != 0) bad = true;
    if (i == P && (int64_t)v != nnz) bad = true;
    if (i < P) {
        int32_t w = ptr[i + 1];
        if (w < v) bad = true;
        if (!bad && v < w && (int64_t)v < nnz) atomicOr(&starts[v >> 5], 1u << (v & 31));
    }
    if (bad) atomicOr(status, 1);
}

__global__ void k_validate_idx(const int32_t *__restrict__ idx, int64_t nnz, int32_t O,
                               const uint32_t *__restrict__ starts, int32_t *__restrict__ status) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int flags = 0;
    for (; p < nnz; p += stride) {
        int32_t c = idx[p];
        if (c < 0 || c >= O) flags |= 2;
        if (p + 1 < nnz) {
            int64_t q = p + 1;
            bool row_start = (starts[q >> 5] >> (q & 31)) & 1u;
            if (!row_start && idx[q] <= c) flags |= 4;
        }
    }
    if (flags) atomicOr(status, flags);
}

// ---- transpose ----
__global__ void k_count_cols(const int32_t *__restrict__ idx, int64_t nnz, int32_t *__restrict__ cnt) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; p < nnz; p += stride) atomicAdd(&cnt[idx[p]], 1);
}

__global__ void k_i64_to_i32(const int64_t *__restrict__ in, int64_t n, int32_t *__restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = (int32_t)in[i];
}

// one wave per source segment; scatter entries to their destination
// segments through per-segment cursors (order fixed afterwards by sorting)
__global__ void k_transpose_scatter(const int32_t *__restrict__ ptr, const int32_t *__restrict__ idx,
                                    const double *__restrict__ val, int32_t P, int32_t *__restrict__ cursor,
                                    int32_t *__restrict__ oidx, double *__restrict__ oval) {
    int32_t seg = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (seg >= P) return;
    for (int32_t p = ptr[seg] + lane_id(); p < ptr[seg + 1]; p += kWave) {
        int32_t c = idx[p];
        int32_t q = atomicAdd(&cursor[c], 1);
        oidx[q] = seg;
        oval[q] = val[p];
    }
}

__global__ void k_rowptr_rebase(const int64_t *__restrict__ rp, int32_t r0, int32_t n, int32_t *__restrict__ o) {
    int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= n) o[t] = (int32_t)(rp[r0 + t] - rp[r0]);
}

// ---- per-row sort (keys with original positions, values gathered after) ----
// wave per row for rows up to kSortWaveMax entries
constexpr int kSortWaveMax = 2048;
constexpr int kSortBlockMax = 16384;

constexpr int kSortWaves = 4;

__global__ void k_sort_rows_wave(const int32_t *__restrict__ ptr, int32_t *__restrict__ idx,
                                 double *__restrict__ val, const double *__restrict__ val_copy,
                                 const int32_t *__restrict__ rows, int32_t nrows_list) {
    __shared__ int32_t smem[kSortWaves * 2 * kSortWaveMax];
    const int w = threadIdx.x / kWave;
    const int l = lane_id();
    int32_t li = blockIdx.x * (blockDim.x / kWave) + w;
    if (li >= nrows_list) return;
    int32_t *keys = smem + (size_t)w * 2 * kSortWaveMax;
    int32_t *pos = keys + kSortWaveMax;
    const int32_t r = rows[li];
    const int32_t b = ptr[r];
    const int n = ptr[r + 1] - b;
    for (int t = l; t < n; t += kWave) {
        keys[t] = idx[b + t];
        pos[t] = b + t;
    }
    wave_sync();
    wave_sort_lds<int32_t>(keys, pos, n);
    for (int t = l; t < n; t += kWave) {
        idx[b + t] = keys[t];
        val[b + t] = val_copy[pos[t]];
    }
}

// block-level bitonic over (key, pos) pairs in LDS or in a global scratch
// region for longer rows
template <bool kGlobal>
__global__ void k_sort_rows_block(const int32_t *__restrict__ ptr, int32_t *__restrict__ idx,
                                  double *__restrict__ val, const double *__restrict__ val_copy,
                                  const int32_t *__restrict__ rows, int32_t nrows_list,
                                  int32_t *__restrict__ gkeys, int32_t *__restrict__ gpos, int64_t gcap) {
    __shared__ int32_t smem[kGlobal ? 1 : 2 * kSortBlockMax];
    const int32_t r = rows[blockIdx.x];
    const int32_t b = ptr[r];
    const int n = ptr[r + 1] - b;
    int N = 64;
    while (N < n) N <<= 1;
    int32_t *keys;
    int32_t *pos;
    if (kGlobal) {
        if ((int64_t)N > gcap) return;  // host sizes gcap for the longest row
        keys = gkeys;
        pos = gpos;
    } else {
        keys = smem;
        pos = smem + kSortBlockMax;
    }
    for (int t = threadIdx.x; t < N; t += blockDim.x) {
        if (t < n) {
            keys[t] = idx[b + t];
            pos[t] = b + t;
        } else {
            keys[t] = kKeyPad;
            pos[t] = 0;
        }
    }
    __syncthreads();
    for (int size = 2; size <= N; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < (N >> 1); i += blockDim.x) {
                int lo = ((i & ~(stride - 1)) << 1) | (i & (stride - 1));
                int hi = lo + stride;
                bool up = ((lo & size) == 0);
                int32_t ka = keys[lo], kb = keys[hi];
                if ((ka > kb) == up) {
                    int32_t pa = pos[lo];
                    keys[lo] = kb;
                    keys[hi] = ka;
                    pos[lo] = pos[hi];
                    pos[hi] = pa;
                }
            }
            if (kGlobal) __threadfence_block();
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        idx[b + t] = keys[t];
        val[b + t] = val_copy[pos[t]];
    }
}

// classify rows by length for the sort kernels
__global__ void k_sort_classify(const int32_t *__restrict__ ptr, int32_t m, int32_t *__restrict__ lists,
                                int32_t *__restrict__ counts, int32_t *__restrict__ maxlen) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int32_t n = ptr[i + 1] - ptr[i];
    if (n <= 1) return;
    int c = n <= kSortWaveMax ? 0 : (n <= kSortBlockMax ? 1 : 2);
    int32_t slot = atomicAdd(&counts[c], 1);
    lists[(int64_t)c * m + slot] = i;
    if (c == 2) atomicMax(maxlen, n);
}

}  // namespace

size_t scan_tmp_bytes(int64_t n) { return sizeof(int64_t) * (size_t)((n + kScanTile - 1) / kScanTile + 1); }
void launch_exclusive_scan_i64(tr_ctx *ctx, const int64_t *in, int64_t *out, int64_t n, int64_t *tmp) {
    scan_impl<int64_t>(ctx, in, out, n, tmp);
}
void launch_exclusive_scan_i32_to_i64(tr_ctx *ctx, const int32_t *in, int64_t *out, int64_t n, int64_t *tmp) {
    scan_impl<int32_t>(ctx, in, out, n, tmp);
}

void launch_copy16(tr_ctx *ctx, void *dst, const void *src, size_t bytes) {
    const int64_t n16 = (int64_t)(bytes / 16);
    if (n16 <= 0) return;
    const uint32_t g = (uint32_t)std::min<int64_t>((n16 + 255) / 256, (int64_t)ctx->cu_count * 32);
    TR_LAUNCH(ctx, "copy16", k_copy16, g, 256, 0, (const uint4 *)src, (uint4 *)dst, n16);
}

void launch_fill_i32(tr_ctx *ctx, int32_t *p, int32_t v, int64_t n) {
    if (n <= 0) return;
    uint32_t g = (uint32_t)std::min<int64_t>((n + 255) / 256, 65536);
    TR_LAUNCH(ctx, "fill_i32", k_fill_i32, g, 256, 0, p, v, n);
}

size_t validate_tmp_bytes(int64_t nnz) { return sizeof(uint32_t) * (size_t)(nnz / 32 + 2); }

void launch_validate(tr_ctx *ctx, const tr_csx_t *m, int32_t *d_status, uint32_t *starts) {
    const int32_t P = ptrdim(m), O = otherdim(m);
    const size_t words = validate_tmp_bytes(m->nnz) / sizeof(uint32_t);
    if (starts) {
        launch_fill_i32(ctx, (int32_t *)starts, 0, (int64_t)words);
    } else {
        starts = (uint32_t *)scratch(ctx, S_MISC2, words * sizeof(uint32_t));
        TR_HIP(hipMemsetAsync(starts, 0, words * sizeof(uint32_t), ctx->stream));
    }
    TR_LAUNCH(ctx, "validate_ptr", k_validate_ptr, div_up((uint64_t)P + 1, 256), 256, 0, m->ptr_vec, P, m->nnz,
              starts, d_status);
    if (m->nnz > 0) {
        uint32_t g = (uint32_t)std::min<int64_t>((m->nnz + 255) / 256, 16384);
        TR_LAUNCH(ctx, "validate_idx", k_validate_idx, g, 256, 0, m->idx_vec, m->nnz, O, starts, d_status);
    }
}

void launch_rowptr_i64_to_i32(tr_ctx *ctx, const int64_t *rp64, int32_t r0, int32_t r1, int32_t *out) {
    // out[t] = rp64[r0 + t] - rp64[r0], t in [0, r1-r0]
    int32_t n = r1 - r0;
    TR_LAUNCH(ctx, "rowptr_to_i32", k_rowptr_rebase, div_up((uint64_t)n + 1, 256), 256, 0, rp64, r0, n, out);
}

void launch_sort_rows(tr_ctx *ctx, tr_csx_t *m) {
    const int32_t P = ptrdim(m);
    if (m->nnz <= 1 || P == 0) return;
    int32_t *lists = (int32_t *)scratch(ctx, S_BINS, sizeof(int32_t) * 3 * (size_t)P);
    int32_t *cnt = (int32_t *)scratch(ctx, S_BINCNT, sizeof(int32_t) * 4);
    TR_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * 4, ctx->stream));
    TR_LAUNCH(ctx, "sort_classify", k_sort_classify, div_up((uint64_t)P, 256), 256, 0, m->ptr_vec, P, lists, cnt,
              cnt + 3);
    int32_t hc[4];
    read_array(ctx, hc, cnt, sizeof(hc));
    double *vcopy = (double *)scratch(ctx, S_TMP_CSX, sizeof(double) * (size_t)m->nnz);
    TR_HIP(hipMemcpyAsync(vcopy, m->val, sizeof(double) * (size_t)m->nnz, hipMemcpyDeviceToDevice, ctx->stream));
    if (hc[0] > 0) {
        TR_LAUNCH(ctx, "sort_rows_wave", k_sort_rows_wave, div_up((uint64_t)hc[0], kSortWaves), kSortWaves * kWave, 0,
                  m->ptr_vec, m->idx_vec, m->val, vcopy, lists, hc[0]);
    }
    if (hc[1] > 0) {
        TR_LAUNCH(ctx, "sort_rows_block", (k_sort_rows_block<false>), (uint32_t)hc[1], 1024, 0, m->ptr_vec,
                  m->idx_vec, m->val, vcopy, lists + P, hc[1], (int32_t *)nullptr, (int32_t *)nullptr, (int64_t)0);
    }
    if (hc[2] > 0) {
        int64_t N = 64;
        while (N < hc[3]) N <<= 1;
        int32_t *g = (int32_t *)scratch(ctx, S_MISC, sizeof(int32_t) * 2 * (size_t)N);
        // one row at a time: the global scratch is shared
        for (int32_t t = 0; t < hc[2]; ++t) {
            TR_LAUNCH(ctx, "sort_rows_global", (k_sort_rows_block<true>), 1, 1024, 0, m->ptr_vec, m->idx_vec, m->val,
                      vcopy, lists + 2 * (int64_t)P + t, 1, g, g + N, N);
        }
    }
}

// out receives the segments of m's other dimension (for a CSR m: the CSR of
// m^T); its descriptor and storage are prepared by the caller.
void launch_transpose(tr_ctx *ctx, const tr_csx_t *m, tr_csx_t *out) {
    const int32_t P = ptrdim(m), O = otherdim(m);
    int32_t *cnt = (int32_t *)scratch(ctx, S_MISC, sizeof(int32_t) * ((size_t)O + 1));
    TR_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * ((size_t)O + 1), ctx->stream));
    if (m->nnz > 0) {
        uint32_t g = (uint32_t)std::min<int64_t>((m->nnz + 255) / 256, 16384);
        TR_LAUNCH(ctx, "transpose_count", k_count_cols, g, 256, 0, m->idx_vec, m->nnz, cnt);
    }
    int64_t *off = (int64_t *)scratch(ctx, S_ROWPTR64, sizeof(int64_t) * ((size_t)O + 1));
    launch_exclusive_scan_i32_to_i64(ctx, cnt, off, O);
    TR_LAUNCH(ctx, "i64_to_i32", k_i64_to_i32, div_up((uint64_t)O + 1, 256), 256, 0, off, (int64_t)O + 1,
              out->ptr_vec);
    // cursors start at the segment offsets
    TR_HIP(hipMemcpyAsync(cnt, out->ptr_vec, sizeof(int32_t) * (size_t)O, hipMemcpyDeviceToDevice, ctx->stream));
    if (m->nnz > 0) {
        TR_LAUNCH(ctx, "transpose_scatter", k_transpose_scatter, div_up((uint64_t)P, 4), 256, 0, m->ptr_vec,
                  m->idx_vec, m->val, P, cnt, out->idx_vec, out->val);
    }
    launch_sort_rows(ctx, out);
}

}  // namespace tr
