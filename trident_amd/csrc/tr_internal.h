// Internal host-side definitions of libtrident_hip: context, device memory
// pool, error propagation and launch instrumentation.  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/trident_hip.h"
#include "arena.h"

namespace tr {

// Internal errors are thrown and converted to status codes at the C-ABI
// boundary (api.hip), so the library never exits the process.
struct Error {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string &msg);

#define TR_HIP(expr)                                                                              \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess)                                                                     \
            ::tr::fail(_e == hipErrorOutOfMemory ? TR_ERR_OOM : TR_ERR_HIP,                       \
                       std::string(#expr) + ": " + hipGetErrorString(_e));                         \
    } while (0)

#define TR_REQUIRE(cond, msg)                                                                     \
    do {                                                                                          \
        if (!(cond)) ::tr::fail(TR_ERR_INVALID, msg);                                             \
    } while (0)

// Caching device allocator.  All work of a context is ordered on one stream,
// so a block released by stream-ordered work can be handed to later work on
// the same stream without synchronising.
class Pool {
  public:
    Pool() = default;
    ~Pool();
    void *alloc(size_t bytes);
    void release(void *p);
    bool owns(void *p) const { return live_.count(p) != 0; }
    size_t size_of(void *p) const;
    void trim();
    void trim_to(size_t total);  // frees cached blocks, largest first, until in use + cached <= total
    size_t in_use() const { return in_use_; }
    size_t cached() const { return cached_; }
    size_t n_malloc = 0, n_free = 0;  // device allocations / releases made (not cache hits)
    void set_limit(size_t bytes) { limit_ = bytes; }
    // Before a new device allocation, cached blocks are freed (largest first)
    // until in use + cached + the new block fit `bytes` (0: no cap).  Idle
    // cached blocks still count against the device memory the driver places.
    void set_cache_cap(size_t bytes) { cache_cap_ = bytes; }
    // Reserves `bytes` of device memory once; later blocks are placed in it
    // (best fit, coalescing) before the cache and hipMalloc are tried.
    // Returns false if the reservation cannot be allocated.
    bool reserve(size_t bytes);

  private:
    bool in_arena(const void *p) const {
        return base_ && (const char *)p >= (const char *)base_ && (const char *)p < (const char *)base_ + arena_.size();
    }
    std::multimap<size_t, void *> free_;
    std::unordered_map<void *, size_t> live_;
    size_t in_use_ = 0, cached_ = 0, limit_ = 0, cache_cap_ = 0;
    void *base_ = nullptr;
    Arena arena_;
};

struct TimingRecord {
    const char *name;
    hipEvent_t start, stop;
};

struct KernelStat {
    int64_t launches = 0;
    double total_ms = 0.0;
};

}  // namespace tr

namespace tr {
// built-in multiply tuning (see tr_tuning_t) and its validation
tr_tuning_t tuning_builtin();
tr_tuning_t tuning_process_default();
void tuning_check(const tr_tuning_t &t);  // throws TR_ERR_INVALID
}  // namespace tr

struct tr_ctx {
    int device = 0;
    tr_tuning_t tune = tr::tuning_process_default();
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string last_error;
    tr::Pool pool;
    int cu_count = 256;

    // Grow-only scratch buffers keyed by slot id (stream ordered reuse).
    std::vector<void *> scratch_ptr;
    std::vector<size_t> scratch_size;

    // Pinned host staging for small device->host reads.
    void *pinned = nullptr;
    size_t pinned_size = 0;

    // Launch instrumentation.
    bool timing = false;
    std::vector<tr::TimingRecord> pending;
    std::vector<hipEvent_t> free_events;
    std::map<std::string, tr::KernelStat> stats;

    explicit tr_ctx(int dev) : device(dev) {}
};

namespace tr {

// Scratch slots used by the kernels of one call.
enum ScratchSlot : int {
    S_FLOPS = 0,
    S_ROWNNZ,
    S_ROWPTR64,
    S_BINS,
    S_BINCNT,
    S_SCAN_TMP,
    S_WINDOWS,
    S_ITEMS,
    S_CURSORS,
    S_MISC,
    S_MISC2,
    S_PANELS,
    S_TMP_CSX,
    S_ORDER,       // heavy items in column order
    S_ORDER_HIST,  // column-bucket histogram / cursors of the item order
    S_NUM_SLOTS
};

void *scratch(tr_ctx *ctx, int slot, size_t bytes);
// Bytes the pool sets aside for a request of `b` bytes (256-byte steps up to
// 1 MiB, 2 MiB granules above); scratch slots grow to 1.25x a request.
size_t pool_block_bytes(size_t b);
inline size_t scratch_block_bytes(size_t b) { return pool_block_bytes(b + b / 4); }
void *pinned(tr_ctx *ctx, size_t bytes);

// Launch bracketing for per-kernel timing.
struct TimerScope {
    tr_ctx *ctx;
    const char *name;
    hipEvent_t ev0 = nullptr;
    TimerScope(tr_ctx *c, const char *n);
    ~TimerScope();
};
void fold_timing(tr_ctx *ctx);

#define TR_LAUNCH(ctx, name, kernel, grid, block, shmem, ...)                                     \
    do {                                                                                          \
        dim3 _g(grid), _b(block);                                                                 \
        if ((uint64_t)_g.x * _b.x > 0xffffffffull)                                                \
            ::tr::fail(TR_ERR_HIP, std::string("launch of ") + (name) + ": grid exceeds 2^32 work-items"); \
        if (_g.x > 0 && _g.y > 0 && _g.z > 0) {                                                   \
            ::tr::TimerScope _ts((ctx), (name));                                                  \
            hipLaunchKernelGGL(kernel, _g, _b, (shmem), (ctx)->stream, __VA_ARGS__);               \
            TR_HIP(hipGetLastError());                                                            \
        }                                                                                         \
    } while (0)

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

inline int32_t ptrdim(const tr_csx_t *m) { return m->majordim == TR_ROWS ? m->nrows : m->ncols; }
inline int32_t otherdim(const tr_csx_t *m) { return m->majordim == TR_ROWS ? m->ncols : m->nrows; }

// Allocate a context-owned contiguous tile [val | idx | ptr] and fill the
// descriptor (ptr array not initialised).
void alloc_tile(tr_ctx *ctx, int32_t majordim, int32_t nrows, int32_t ncols, int64_t nnz, tr_csx_t *out);
void release_csx(tr_ctx *ctx, tr_csx_t *m);
void check_csx(const tr_csx_t *m, const char *what);

// ---- launchers implemented in the kernel translation units ----

// k_util.hip
// out[0..n] = exclusive prefix (out[n] = total).  tmp: caller-owned temp of
// scan_tmp_bytes(n) bytes, or nullptr to use the context scratch.
size_t scan_tmp_bytes(int64_t n);
void launch_exclusive_scan_i64(tr_ctx *ctx, const int64_t *in, int64_t *out, int64_t n, int64_t *tmp = nullptr);
void launch_exclusive_scan_i32_to_i64(tr_ctx *ctx, const int32_t *in, int64_t *out, int64_t n,
                                      int64_t *tmp = nullptr);
// starts: caller-owned temp of validate_tmp_bytes(nnz) bytes, or nullptr to
// use the context scratch.
size_t validate_tmp_bytes(int64_t nnz);
void launch_validate(tr_ctx *ctx, const tr_csx_t *m, int32_t *d_status, uint32_t *starts = nullptr);
void launch_fill_i32(tr_ctx *ctx, int32_t *p, int32_t v, int64_t n);
void launch_copy16(tr_ctx *ctx, void *dst, const void *src, size_t bytes);
void launch_transpose(tr_ctx *ctx, const tr_csx_t *m, tr_csx_t *out);
void launch_sort_rows(tr_ctx *ctx, tr_csx_t *m);
void launch_rowptr_i64_to_i32(tr_ctx *ctx, const int64_t *rp64, int32_t r0, int32_t r1, int32_t *out);

// k_bitmask.hip
void launch_bitmask_gen(tr_ctx *ctx, const int32_t *ptr, int32_t n, uint64_t *mask);
void launch_bitmask_and(tr_ctx *ctx, const uint64_t *a, const uint64_t *b, int32_t words, uint64_t *out);
int64_t compress(tr_ctx *ctx, const tr_csx_t *m, const uint64_t *mask, tr_csx_t *out);
void launch_compress_fixed(tr_ctx *ctx, int32_t n, const int32_t *ptr, const int32_t *idx, const double *val,
                           int64_t nnz_in, const uint64_t *mask, int32_t *kept_tmp, int64_t *np64, int64_t *scan_tmp,
                           int32_t *optr, int32_t *oidx, double *oval);

// k_spgemm.hip
void validate_operands(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *B);
void check_operand_shapes(const tr_csx_t *A, const tr_csx_t *B);  // host-side shape checks only
// internal batch flag (not part of the C-ABI): a single run, workspace in
// the context's scratch slots (valid only without accumulators)
constexpr int32_t kBatchSingleRun = 1 << 30;
int batch_create(tr_ctx_t ctx, int32_t ntasks, const tr_csx_t *A, const tr_csx_t *B, const tr_csx_t *C_in,
                 int32_t flags, bool validate, tr_batch_t *out);
size_t batch_retained_bytes(tr_batch_t b);  // device bytes a finished batch keeps
int64_t spgemm_flops(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *B);
// Runs the symbolic phase and returns nnz(C) (and bytes needed).
void spgemm_symbolic_only(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *B, int64_t *nnz, size_t *bytes);
// Full product into panels.
int spgemm_panels(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *B, int64_t max_panel_nnz,
                  tr_csx_t *panels, int32_t *row_begin, int32_t max_panels, int32_t *npanels);

// k_spadd.hip
void spadd(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *C, tr_csx_t *out);
int64_t spadd_into(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *C, int32_t *optr, int32_t *oi, double *ov,
                   int64_t capacity);
// The same add with every temporary carved from the device region
// [base, base + bytes) instead of the context's pool: no allocation.
// spadd_fixed_bytes gives the bytes it needs (256-byte aligned pieces).
size_t spadd_fixed_bytes(int32_t rows, int64_t nnz_x, int64_t nnz_y);
int64_t spadd_into_fixed(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *C, int32_t *optr, int32_t *oi, double *ov,
                         int64_t capacity, char *base, size_t bytes);

// noalloc.hip: the per-device static scratch of the allocation-free entry
// points; noalloc_init runs once per device (from tr_ctx_create, with the
// device current) and noalloc_spadd_into is the body of
// tr_spadd_into_noalloc.
void noalloc_init(int device);
int64_t noalloc_spadd_into(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *C, tr_csx_t *out, int64_t capacity);

// Read a device value through pinned staging (synchronises the stream).
template <typename T>
T read_scalar(tr_ctx *ctx, const T *dptr) {
    T *h = (T *)pinned(ctx, sizeof(T));
    TR_HIP(hipMemcpyAsync(h, dptr, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    TR_HIP(hipStreamSynchronize(ctx->stream));
    return *h;
}
void read_array(tr_ctx *ctx, void *host, const void *dptr, size_t bytes);

}  // namespace tr
