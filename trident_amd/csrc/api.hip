// C-ABI entry points of libtrident_hip (declared in include/trident_hip.h),
// the context, its caching device pool and launch instrumentation.
#include "dev_common.h"
#include "tr_internal.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>

namespace tr {

[[noreturn]] void fail(int code, const std::string &msg) { throw Error{code, msg}; }

// ------------------------------------------------------------------ pool
size_t pool_block_bytes(size_t b) {
    if (b == 0) b = 1;
    if (b <= (1u << 20)) return (b + 255) & ~(size_t)255;
    return (b + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);  // 2 MiB granules
}

Pool::~Pool() {
    for (auto &kv : free_) (void)hipFree(kv.second);
    for (auto &kv : live_)
        if (!in_arena(kv.first)) (void)hipFree(kv.first);
    if (base_) (void)hipFree(base_);
}

bool Pool::reserve(size_t bytes) {
    if (base_ || bytes == 0) return base_ != nullptr;
    const size_t b = pool_block_bytes(bytes);
    if (hipMalloc(&base_, b) != hipSuccess) {
        (void)hipGetLastError();
        base_ = nullptr;
        return false;
    }
    ++n_malloc;
    arena_.init(b);
    return true;
}

void *Pool::alloc(size_t bytes) {
    size_t b = pool_block_bytes(bytes);
    // the test limit applies to every block handed out, cached ones included
    if (limit_ && in_use_ + b > limit_) fail(TR_ERR_OOM, "device pool limit reached");
    if (base_) {
        size_t off = 0;
        if (arena_.alloc(b, off)) {
            void *p = (char *)base_ + off;
            if (limit_ && in_use_ + b > limit_) {
                arena_.free(off, b);
                fail(TR_ERR_OOM, "device pool limit reached");
            }
            live_[p] = b;
            in_use_ += b;
            return p;
        }
    }
    auto it = free_.lower_bound(b);
    if (it != free_.end() && it->first <= 2 * b + (4u << 20) && !(limit_ && in_use_ + it->first > limit_)) {
        void *p = it->second;
        size_t sz = it->first;
        free_.erase(it);
        cached_ -= sz;
        live_[p] = sz;
        in_use_ += sz;
        return p;
    }
    if (limit_ && in_use_ + b > limit_) fail(TR_ERR_OOM, "device pool limit reached");
    if (cache_cap_ && in_use_ + cached_ + b > cache_cap_) trim_to(cache_cap_ > b ? cache_cap_ - b : 0);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim();
        e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail(TR_ERR_OOM, "hipMalloc of " + std::to_string(b) + " bytes failed");
        }
    }
    ++n_malloc;
    live_[p] = b;
    in_use_ += b;
    return p;
}

void Pool::release(void *p) {
    if (!p) return;
    auto it = live_.find(p);
    if (it == live_.end()) fail(TR_ERR_INVALID, "release of a pointer the context does not own");
    size_t sz = it->second;
    live_.erase(it);
    in_use_ -= sz;
    if (in_arena(p)) {
        arena_.free((size_t)((char *)p - (char *)base_), sz);
        return;
    }
    free_.emplace(sz, p);
    cached_ += sz;
}

size_t Pool::size_of(void *p) const {
    auto it = live_.find(p);
    return it == live_.end() ? 0 : it->second;
}

void Pool::trim() {
    n_free += free_.size();
    for (auto &kv : free_) (void)hipFree(kv.second);
    free_.clear();
    cached_ = 0;
}

void Pool::trim_to(size_t total) {
    while (!free_.empty() && in_use_ + cached_ > total) {
        auto it = std::prev(free_.end());
        (void)hipFree(it->second);
        ++n_free;
        cached_ -= it->first;
        free_.erase(it);
    }
}

// ------------------------------------------------------------------ scratch / staging
void *scratch(tr_ctx *ctx, int slot, size_t bytes) {
    if ((int)ctx->scratch_ptr.size() < S_NUM_SLOTS) {
        ctx->scratch_ptr.resize(S_NUM_SLOTS, nullptr);
        ctx->scratch_size.resize(S_NUM_SLOTS, 0);
    }
    if (bytes == 0) bytes = 1;
    if (ctx->scratch_size[slot] < bytes) {
        if (ctx->scratch_ptr[slot]) ctx->pool.release(ctx->scratch_ptr[slot]);
        ctx->scratch_ptr[slot] = nullptr;
        ctx->scratch_size[slot] = 0;
        size_t want = bytes + bytes / 4;
        ctx->scratch_ptr[slot] = ctx->pool.alloc(want);
        ctx->scratch_size[slot] = want;
    }
    return ctx->scratch_ptr[slot];
}

void *pinned(tr_ctx *ctx, size_t bytes) {
    if (ctx->pinned_size < bytes) {
        if (ctx->pinned) {
            TR_HIP(hipStreamSynchronize(ctx->stream));
            (void)hipHostFree(ctx->pinned);
            ctx->pinned = nullptr;
            ctx->pinned_size = 0;
        }
        size_t want = std::max<size_t>(bytes, 4096);
        TR_HIP(hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault));
        ctx->pinned_size = want;
    }
    return ctx->pinned;
}

void read_array(tr_ctx *ctx, void *host, const void *dptr, size_t bytes) {
    if (bytes == 0) return;
    TR_HIP(hipMemcpyAsync(host, dptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TR_HIP(hipStreamSynchronize(ctx->stream));
}

// ------------------------------------------------------------------ timing
TimerScope::TimerScope(tr_ctx *c, const char *n) : ctx(c), name(n) {
    if (!ctx->timing) return;
    if (ctx->free_events.empty()) {
        hipEvent_t e;
        TR_HIP(hipEventCreate(&e));
        ctx->free_events.push_back(e);
    }
    ev0 = ctx->free_events.back();
    ctx->free_events.pop_back();
    TR_HIP(hipEventRecord(ev0, ctx->stream));
}

TimerScope::~TimerScope() {
    if (!ev0) return;
    hipEvent_t ev1;
    if (ctx->free_events.empty()) {
        if (hipEventCreate(&ev1) != hipSuccess) return;
    } else {
        ev1 = ctx->free_events.back();
        ctx->free_events.pop_back();
    }
    if (hipEventRecord(ev1, ctx->stream) != hipSuccess) return;
    ctx->pending.push_back(TimingRecord{name, ev0, ev1});
    if (ctx->pending.size() > 4096) {
        try {
            fold_timing(ctx);
        } catch (...) {
        }
    }
}

void fold_timing(tr_ctx *ctx) {
    if (ctx->pending.empty()) return;
    TR_HIP(hipStreamSynchronize(ctx->stream));
    for (auto &r : ctx->pending) {
        float ms = 0.f;
        TR_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
        auto &st = ctx->stats[r.name];
        st.launches += 1;
        st.total_ms += ms;
        ctx->free_events.push_back(r.start);
        ctx->free_events.push_back(r.stop);
    }
    ctx->pending.clear();
}

// ------------------------------------------------------------------ tiles
void check_csx(const tr_csx_t *m, const char *what) {
    std::string w(what);
    TR_REQUIRE(m != nullptr, w + ": null matrix");
    TR_REQUIRE(m->majordim == TR_ROWS || m->majordim == TR_COLS, w + ": majordim must be TR_ROWS or TR_COLS");
    TR_REQUIRE(m->nrows >= 0 && m->ncols >= 0, w + ": negative dimension");
    TR_REQUIRE(m->nnz >= 0 && m->nnz <= INT32_MAX, w + ": nnz must be in [0, INT32_MAX]");
    TR_REQUIRE(m->ptr_vec != nullptr, w + ": uninitialised matrix (ptr_vec == NULL)");
    TR_REQUIRE(m->nnz == 0 || (m->idx_vec != nullptr && m->val != nullptr), w + ": null idx_vec/val");
}

void alloc_tile(tr_ctx *ctx, int32_t majordim, int32_t nrows, int32_t ncols, int64_t nnz, tr_csx_t *out) {
    if (nnz > INT32_MAX) fail(TR_ERR_OVERFLOW, "tile nnz exceeds INT32_MAX");
    const int32_t P = majordim == TR_ROWS ? nrows : ncols;
    size_t bytes = (size_t)nnz * (sizeof(double) + sizeof(int32_t)) + ((size_t)P + 1) * sizeof(int32_t);
    char *buf = (char *)ctx->pool.alloc(bytes);
    out->majordim = majordim;
    out->nrows = nrows;
    out->ncols = ncols;
    out->nnz = nnz;
    out->contig = 1;
    out->buf = buf;
    out->buf_size = bytes;
    out->val = (double *)buf;
    out->idx_vec = (int32_t *)(buf + (size_t)nnz * sizeof(double));
    out->ptr_vec = (int32_t *)(buf + (size_t)nnz * (sizeof(double) + sizeof(int32_t)));
}

void release_csx(tr_ctx *ctx, tr_csx_t *m) {
    if (m->contig && m->buf && ctx->pool.owns(m->buf)) ctx->pool.release(m->buf);
    std::memset(m, 0, sizeof(*m));
}

// Device validation of one operand; fails with TR_ERR_INVALID on a malformed
// pointer or an out-of-range index (and on unsorted rows if need_sorted).
static void require_valid(tr_ctx *ctx, const tr_csx_t *m, const char *what, bool need_sorted) {
    int32_t *st = (int32_t *)scratch(ctx, S_PANELS, 64);
    TR_HIP(hipMemsetAsync(st, 0, sizeof(int32_t), ctx->stream));
    launch_validate(ctx, m, st);
    int32_t hs = 0;
    read_array(ctx, &hs, st, sizeof(hs));
    if (hs & 1) fail(TR_ERR_INVALID, std::string(what) + ": malformed pointer array");
    if (hs & 2) fail(TR_ERR_INVALID, std::string(what) + ": index out of range");
    if (need_sorted && (hs & 4)) fail(TR_ERR_INVALID, std::string(what) + ": unsorted or duplicate indices");
}

static bool ctx_owned(tr_ctx *ctx, const tr_csx_t *m) {
    return m->contig && m->buf && ctx->pool.owns(m->buf);
}

static void copy_tile(tr_ctx *ctx, const tr_csx_t *src, tr_csx_t *dst) {
    tr_csx_t t;
    alloc_tile(ctx, src->majordim, src->nrows, src->ncols, src->nnz, &t);
    const int32_t P = ptrdim(src);
    TR_HIP(hipMemcpyAsync(t.ptr_vec, src->ptr_vec, sizeof(int32_t) * ((size_t)P + 1), hipMemcpyDeviceToDevice,
                          ctx->stream));
    if (src->nnz) {
        TR_HIP(hipMemcpyAsync(t.idx_vec, src->idx_vec, sizeof(int32_t) * (size_t)src->nnz, hipMemcpyDeviceToDevice,
                              ctx->stream));
        TR_HIP(hipMemcpyAsync(t.val, src->val, sizeof(double) * (size_t)src->nnz, hipMemcpyDeviceToDevice,
                              ctx->stream));
    }
    *dst = t;
}



// ---------------------------------------------------------------- tuning
tr_tuning_t tuning_builtin() {
    tr_tuning_t t;
    t.light_max_flops = 1024;
    t.window_entries = 1152;
    t.item_windows = 16;
    t.heavy_table = 2048;
    t.dense_width = 2048;
    return t;
}

namespace {
std::mutex g_tune_mu;
tr_tuning_t g_tune = tuning_builtin();
}  // namespace

tr_tuning_t tuning_process_default() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return g_tune;
}

void tuning_check(const tr_tuning_t &t) {
    TR_REQUIRE(t.light_max_flops >= 64 && t.light_max_flops <= 1024,
               "tuning: light_max_flops must be in [64, 1024]");
    TR_REQUIRE(t.heavy_table == 2048 || t.heavy_table == 4096, "tuning: heavy_table must be 2048 or 4096");
    TR_REQUIRE(t.window_entries >= 64 && t.window_entries <= 3 * t.heavy_table / 4 - 127,
               "tuning: window_entries must be in [64, 3 * heavy_table / 4 - 127]");
    TR_REQUIRE(t.item_windows >= 1 && t.item_windows <= 64, "tuning: item_windows must be in [1, 64]");
    TR_REQUIRE(t.dense_width >= 0 && t.dense_width <= t.heavy_table, "tuning: dense_width must be in [0, heavy_table]");
}

}  // namespace tr

using namespace tr;

// Runs `body`, converting internal errors into status codes.
template <typename F>
static int guarded(tr_ctx *ctx, F &&body) {
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

#define TR_CTX_REQUIRED(ctx)                                                                                      \
    if (!(ctx)) return TR_ERR_INVALID

extern "C" {

const char *tr_version(void) {
    return "trident_hip gfx950: light rows wave hash, heavy rows bitmap windows + 2048-slot workgroup tables, merge-path add";
}

int tr_ctx_create(tr_ctx_t *out, int device, void *hip_stream) {
    if (!out) return TR_ERR_INVALID;
    *out = nullptr;
    tr_ctx *ctx = nullptr;
    int rc = guarded(nullptr, [&] {
        int ndev = 0;
        TR_HIP(hipGetDeviceCount(&ndev));
        TR_REQUIRE(device >= 0 && device < ndev, "tr_ctx_create: no such device");
        TR_HIP(hipSetDevice(device));
        ctx = new tr_ctx(device);
        if (hip_stream) {
            ctx->stream = (hipStream_t)hip_stream;
            ctx->own_stream = false;
        } else {
            TR_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            ctx->own_stream = true;
        }
        hipDeviceProp_t prop;
        TR_HIP(hipGetDeviceProperties(&prop, device));
        ctx->cu_count = prop.multiProcessorCount;
        (void)pinned(ctx, 4096);  // staging of scalar reads, so later calls need no host allocation
        noalloc_init(device);     // device address of the allocation-free entry points' scratch
    });
    if (rc != TR_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return TR_OK;
}

int tr_ctx_destroy(tr_ctx_t ctx) {
    if (!ctx) return TR_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &r : ctx->pending) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    for (auto e : ctx->free_events) (void)hipEventDestroy(e);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;  // pool destructor frees device memory
    return TR_OK;
}

const char *tr_last_error(tr_ctx_t ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int tr_ctx_synchronize(tr_ctx_t ctx) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] { TR_HIP(hipStreamSynchronize(ctx->stream)); });
}

void *tr_ctx_stream(tr_ctx_t ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int tr_spgemm(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, tr_csx_t *C_out) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_REQUIRE(C_out != nullptr, "tr_spgemm: null output");
        int32_t rb[2], np = 0;
        spgemm_panels(ctx, A, B, INT32_MAX, C_out, rb, 1, &np);
    });
}

int tr_spgemm_panels(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, int64_t max_panel_nnz, tr_csx_t *panels,
                     int32_t *row_begin, int32_t max_panels, int32_t *npanels) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] { spgemm_panels(ctx, A, B, max_panel_nnz, panels, row_begin, max_panels, npanels); });
}

int tr_sp_mma(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, tr_csx_t *C) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_REQUIRE(C != nullptr, "tr_sp_mma: null accumulator");
        tr_csx_t prod;
        std::memset(&prod, 0, sizeof(prod));
        int32_t rb[2], np = 0;
        spgemm_panels(ctx, A, B, INT32_MAX, &prod, rb, 1, &np);
        if (C->ptr_vec == nullptr) {  // uninitialised: C := A*B
            *C = prod;
            return;
        }
        if (prod.nnz == 0) {  // sp_mma_hybrid returns early on an empty product
            release_csx(ctx, &prod);
            return;
        }
        check_csx(C, "C");
        TR_REQUIRE(C->majordim == TR_ROWS && C->nrows == prod.nrows && C->ncols == prod.ncols,
                   "tr_sp_mma: accumulator shape mismatch");
        tr_csx_t acc;
        spadd(ctx, &prod, C, &acc);
        release_csx(ctx, &prod);
        if (ctx_owned(ctx, C)) release_csx(ctx, C);
        *C = acc;
    });
}

// Both operands of a sorted add: CSR, same shape, well formed, sorted rows.
static void require_add_operands(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *C, const char *what) {
    const std::string w(what);
    check_csx(C, "C");
    TR_REQUIRE(A->majordim == TR_ROWS && C->majordim == TR_ROWS, w + ": operands must be CSR");
    TR_REQUIRE(A->nrows == C->nrows && A->ncols == C->ncols, w + ": shape mismatch");
    int32_t *st = (int32_t *)scratch(ctx, S_PANELS, 64);
    TR_HIP(hipMemsetAsync(st, 0, 2 * sizeof(int32_t), ctx->stream));
    launch_validate(ctx, A, st);
    launch_validate(ctx, C, st + 1);
    int32_t hs[2];
    read_array(ctx, hs, st, sizeof(hs));
    if ((hs[0] | hs[1]) & 3) fail(TR_ERR_INVALID, w + ": malformed CSR operand");
    if ((hs[0] | hs[1]) & 4) fail(TR_ERR_INVALID, w + ": operands need sorted unique columns");
}

int tr_spadd(tr_ctx_t ctx, const tr_csx_t *A, tr_csx_t *C) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_REQUIRE(C != nullptr, "tr_spadd: null accumulator");
        check_csx(A, "A");
        if (A->nnz == 0) return;  // KokkosWrap.hpp:400-402
        if (C->ptr_vec == nullptr) {
            copy_tile(ctx, A, C);  // KokkosWrap.hpp:403-407 (deep copy here)
            return;
        }
        require_add_operands(ctx, A, C, "tr_spadd");
        tr_csx_t acc;
        spadd(ctx, A, C, &acc);
        if (ctx_owned(ctx, C)) release_csx(ctx, C);
        *C = acc;
    });
}

int tr_spadd_into(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, tr_csx_t *out, int64_t capacity,
                  int64_t *nnz_needed) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_REQUIRE(out != nullptr && nnz_needed != nullptr, "tr_spadd_into: null output");
        check_csx(A, "A");
        require_add_operands(ctx, A, B, "tr_spadd_into");
        TR_REQUIRE(capacity >= 0 && capacity <= INT32_MAX, "tr_spadd_into: capacity must be in [0, INT32_MAX]");
        TR_REQUIRE(out->ptr_vec != nullptr && (capacity == 0 || (out->idx_vec && out->val)),
                   "tr_spadd_into: output arrays missing");
        for (const tr_csx_t *in : {A, B})
            TR_REQUIRE(out->ptr_vec != in->ptr_vec && out->idx_vec != in->idx_vec && out->val != in->val,
                       "tr_spadd_into: output aliases an operand");
        const int64_t nnz = spadd_into(ctx, A, B, out->ptr_vec, out->idx_vec, out->val, capacity);
        *nnz_needed = nnz;
        if (nnz > capacity) fail(TR_ERR_OVERFLOW, "tr_spadd_into: result needs " + std::to_string(nnz) +
                                                      " entries, capacity " + std::to_string(capacity));
        out->majordim = TR_ROWS;
        out->nrows = A->nrows;
        out->ncols = A->ncols;
        out->nnz = nnz;
    });
}

int tr_calibration_copy(tr_ctx_t ctx, void *dst, const void *src, size_t bytes) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_REQUIRE(bytes % 16 == 0 && ((uintptr_t)dst | (uintptr_t)src) % 16 == 0,
                   "tr_calibration_copy: 16-byte aligned pointers and a multiple of 16 bytes");
        TR_REQUIRE(bytes == 0 || (dst && src), "tr_calibration_copy: null pointer");
        launch_copy16(ctx, dst, src, bytes);
    });
}

int tr_spadd_into_noalloc(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, tr_csx_t *out, int64_t capacity,
                          int64_t *nnz_needed) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_REQUIRE(out != nullptr && nnz_needed != nullptr, "tr_spadd_into_noalloc: null output");
        check_csx(A, "A");
        check_csx(B, "B");
        TR_REQUIRE(A->majordim == TR_ROWS && B->majordim == TR_ROWS, "tr_spadd_into_noalloc: operands must be CSR");
        TR_REQUIRE(A->nrows == B->nrows && A->ncols == B->ncols, "tr_spadd_into_noalloc: shape mismatch");
        TR_REQUIRE(capacity >= 0 && capacity <= INT32_MAX,
                   "tr_spadd_into_noalloc: capacity must be in [0, INT32_MAX]");
        TR_REQUIRE(out->ptr_vec != nullptr && (capacity == 0 || (out->idx_vec && out->val)),
                   "tr_spadd_into_noalloc: output arrays missing");
        for (const tr_csx_t *in : {A, B})
            TR_REQUIRE(out->ptr_vec != in->ptr_vec && out->idx_vec != in->idx_vec && out->val != in->val,
                       "tr_spadd_into_noalloc: output aliases an operand");
        const int64_t nnz = noalloc_spadd_into(ctx, A, B, out, capacity);
        *nnz_needed = nnz;
        if (nnz > capacity) fail(TR_ERR_OVERFLOW, "tr_spadd_into_noalloc: result needs " + std::to_string(nnz) +
                                                      " entries, capacity " + std::to_string(capacity));
        out->majordim = TR_ROWS;
        out->nrows = A->nrows;
        out->ncols = A->ncols;
        out->nnz = nnz;
    });
}

int tr_symbolic(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, int64_t *nnz_C, size_t *bytes_needed) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] { spgemm_symbolic_only(ctx, A, B, nnz_C, bytes_needed); });
}

int tr_spgemm_flops(tr_ctx_t ctx, const tr_csx_t *A, const tr_csx_t *B, int64_t *flops) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_REQUIRE(flops != nullptr, "tr_spgemm_flops: null output");
        *flops = spgemm_flops(ctx, A, B);
    });
}

int tr_tuning_default(tr_tuning_t *t) {
    if (!t) return TR_ERR_INVALID;
    *t = tuning_builtin();
    return TR_OK;
}

int tr_set_default_tuning(const tr_tuning_t *t) {
    return guarded(nullptr, [&] {
        const tr_tuning_t v = t ? *t : tuning_builtin();
        tuning_check(v);
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune = v;
    });
}

int tr_ctx_set_tuning(tr_ctx_t ctx, const tr_tuning_t *t) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        const tr_tuning_t v = t ? *t : tuning_process_default();
        tuning_check(v);
        ctx->tune = v;
    });
}

int tr_ctx_get_tuning(tr_ctx_t ctx, tr_tuning_t *t) {
    TR_CTX_REQUIRED(ctx);
    if (!t) return TR_ERR_INVALID;
    *t = ctx->tune;
    return TR_OK;
}

int tr_bitmask_gen(tr_ctx_t ctx, const int32_t *ptr, int32_t n, uint64_t *mask) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_REQUIRE(n >= 0, "tr_bitmask_gen: negative length");
        TR_REQUIRE(n == 0 || (ptr && mask), "tr_bitmask_gen: null pointer");
        launch_bitmask_gen(ctx, ptr, n, mask);
    });
}

int tr_bitmask_and(tr_ctx_t ctx, const uint64_t *a, const uint64_t *b, int32_t words, uint64_t *out) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_REQUIRE(words >= 0, "tr_bitmask_and: negative length");
        TR_REQUIRE(words == 0 || (a && b && out), "tr_bitmask_and: null pointer");
        launch_bitmask_and(ctx, a, b, words, out);
    });
}

int tr_compress(tr_ctx_t ctx, const tr_csx_t *M, const uint64_t *mask, tr_csx_t *out) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        check_csx(M, "M");
        TR_REQUIRE(out != nullptr && (mask != nullptr || ptrdim(M) == 0), "tr_compress: null pointer");
        require_valid(ctx, M, "tr_compress", false);
        compress(ctx, M, mask, out);
    });
}

int tr_csx_transpose(tr_ctx_t ctx, const tr_csx_t *M, tr_csx_t *out) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        check_csx(M, "M");
        TR_REQUIRE(out != nullptr, "tr_csx_transpose: null output");
        require_valid(ctx, M, "tr_csx_transpose", false);
        // segments of the result = the other dimension of M
        tr_csx_t t;
        if (M->majordim == TR_ROWS) alloc_tile(ctx, TR_ROWS, M->ncols, M->nrows, M->nnz, &t);  // CSR of M^T
        else alloc_tile(ctx, TR_COLS, M->ncols, M->nrows, M->nnz, &t);                        // CSC of M^T
        launch_transpose(ctx, M, &t);
        *out = t;
    });
}

int tr_csc_to_csr(tr_ctx_t ctx, const tr_csx_t *M, tr_csx_t *out) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        check_csx(M, "M");
        TR_REQUIRE(M->majordim == TR_COLS, "tr_csc_to_csr: input must be CSC (TR_COLS)");
        TR_REQUIRE(out != nullptr, "tr_csc_to_csr: null output");
        require_valid(ctx, M, "tr_csc_to_csr", false);
        tr_csx_t t;
        alloc_tile(ctx, TR_ROWS, M->nrows, M->ncols, M->nnz, &t);
        launch_transpose(ctx, M, &t);
        *out = t;
    });
}

int tr_csc_to_csr_into(tr_ctx_t ctx, const tr_csx_t *M, int32_t *ptr, int32_t *idx, double *val) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        check_csx(M, "M");
        TR_REQUIRE(M->majordim == TR_COLS, "tr_csc_to_csr_into: input must be CSC (TR_COLS)");
        TR_REQUIRE(ptr != nullptr && (M->nnz == 0 || (idx && val)), "tr_csc_to_csr_into: null output");
        TR_REQUIRE(ptr != M->ptr_vec && (M->nnz == 0 || (idx != M->idx_vec && val != M->val)),
                   "tr_csc_to_csr_into: output aliases the input");
        require_valid(ctx, M, "tr_csc_to_csr_into", false);
        tr_csx_t t;
        std::memset(&t, 0, sizeof(t));
        t.majordim = TR_ROWS;
        t.nrows = M->nrows;
        t.ncols = M->ncols;
        t.nnz = M->nnz;
        t.ptr_vec = ptr;
        t.idx_vec = idx;
        t.val = val;
        launch_transpose(ctx, M, &t);
    });
}

int tr_csx_sort_rows(tr_ctx_t ctx, tr_csx_t *M) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        check_csx(M, "M");
        require_valid(ctx, M, "tr_csx_sort_rows", false);
        launch_sort_rows(ctx, M);
    });
}

int tr_csx_validate(tr_ctx_t ctx, const tr_csx_t *M, int32_t *status) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        check_csx(M, "M");
        TR_REQUIRE(status != nullptr, "tr_csx_validate: null status");
        int32_t *st = (int32_t *)scratch(ctx, S_PANELS, 64);
        TR_HIP(hipMemsetAsync(st, 0, sizeof(int32_t), ctx->stream));
        launch_validate(ctx, M, st);
        read_array(ctx, status, st, sizeof(int32_t));
    });
}

int tr_csx_free(tr_ctx_t ctx, tr_csx_t *M) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        if (!M) return;
        if (ctx_owned(ctx, M)) release_csx(ctx, M);
        else std::memset(M, 0, sizeof(*M));
    });
}

int tr_csx_upload(tr_ctx_t ctx, int32_t majordim, int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *h_ptr,
                  const int32_t *h_idx, const double *h_val, tr_csx_t *out) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_REQUIRE(out != nullptr && h_ptr != nullptr, "tr_csx_upload: null pointer");
        TR_REQUIRE(majordim == TR_ROWS || majordim == TR_COLS, "tr_csx_upload: bad majordim");
        TR_REQUIRE(nrows >= 0 && ncols >= 0 && nnz >= 0 && nnz <= INT32_MAX, "tr_csx_upload: bad sizes");
        TR_REQUIRE(nnz == 0 || (h_idx && h_val), "tr_csx_upload: null arrays");
        tr_csx_t t;
        alloc_tile(ctx, majordim, nrows, ncols, nnz, &t);
        const int32_t P = ptrdim(&t);
        TR_HIP(hipMemcpyAsync(t.ptr_vec, h_ptr, sizeof(int32_t) * ((size_t)P + 1), hipMemcpyHostToDevice,
                              ctx->stream));
        if (nnz) {
            TR_HIP(hipMemcpyAsync(t.idx_vec, h_idx, sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice,
                                  ctx->stream));
            TR_HIP(hipMemcpyAsync(t.val, h_val, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice, ctx->stream));
        }
        TR_HIP(hipStreamSynchronize(ctx->stream));
        *out = t;
    });
}

int tr_csx_download(tr_ctx_t ctx, const tr_csx_t *M, int32_t *h_ptr, int32_t *h_idx, double *h_val) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        check_csx(M, "M");
        const int32_t P = ptrdim(M);
        if (h_ptr)
            TR_HIP(hipMemcpyAsync(h_ptr, M->ptr_vec, sizeof(int32_t) * ((size_t)P + 1), hipMemcpyDeviceToHost,
                                  ctx->stream));
        if (M->nnz && h_idx)
            TR_HIP(hipMemcpyAsync(h_idx, M->idx_vec, sizeof(int32_t) * (size_t)M->nnz, hipMemcpyDeviceToHost,
                                  ctx->stream));
        if (M->nnz && h_val)
            TR_HIP(hipMemcpyAsync(h_val, M->val, sizeof(double) * (size_t)M->nnz, hipMemcpyDeviceToHost,
                                  ctx->stream));
        TR_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int tr_malloc(tr_ctx_t ctx, void **ptr, size_t bytes) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_REQUIRE(ptr != nullptr, "tr_malloc: null output");
        *ptr = ctx->pool.alloc(bytes);
    });
}

int tr_free(tr_ctx_t ctx, void *ptr) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] { ctx->pool.release(ptr); });
}

int tr_memcpy_h2d(tr_ctx_t ctx, void *dst, const void *src, size_t bytes) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        TR_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int tr_memcpy_d2h(tr_ctx_t ctx, void *dst, const void *src, size_t bytes) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] { read_array(ctx, dst, src, bytes); });
}

int tr_memcpy_d2d(tr_ctx_t ctx, void *dst, const void *src, size_t bytes) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] { TR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream)); });
}

int tr_memset(tr_ctx_t ctx, void *dst, int value, size_t bytes) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_REQUIRE(dst != nullptr || bytes == 0, "tr_memset: null pointer");
        if (bytes) TR_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
    });
}

int tr_device_malloc(int device, void **ptr, size_t bytes) {
    return guarded(nullptr, [&] {
        TR_REQUIRE(ptr != nullptr, "tr_device_malloc: null output");
        *ptr = nullptr;
        TR_HIP(hipSetDevice(device));
        TR_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    });
}

int tr_device_free(void *ptr) {
    return guarded(nullptr, [&] {
        if (ptr) TR_HIP(hipFree(ptr));
    });
}

int tr_event_create(tr_event_t *ev) {
    if (!ev) return TR_ERR_INVALID;
    *ev = nullptr;
    return guarded(nullptr, [&] {
        hipEvent_t e;
        TR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        *ev = (tr_event_t)e;
    });
}

int tr_event_record(tr_ctx_t ctx, tr_event_t ev) {
    TR_CTX_REQUIRED(ctx);
    if (!ev) return TR_ERR_INVALID;
    return guarded(ctx, [&] { TR_HIP(hipEventRecord((hipEvent_t)ev, ctx->stream)); });
}

int tr_ctx_wait_event(tr_ctx_t ctx, tr_event_t ev) {
    TR_CTX_REQUIRED(ctx);
    if (!ev) return TR_ERR_INVALID;
    return guarded(ctx, [&] { TR_HIP(hipStreamWaitEvent(ctx->stream, (hipEvent_t)ev, 0)); });
}

int tr_event_synchronize(tr_event_t ev) {
    if (!ev) return TR_ERR_INVALID;
    return guarded(nullptr, [&] { TR_HIP(hipEventSynchronize((hipEvent_t)ev)); });
}

int tr_event_destroy(tr_event_t ev) {
    if (!ev) return TR_OK;
    return guarded(nullptr, [&] { TR_HIP(hipEventDestroy((hipEvent_t)ev)); });
}

int tr_pool_trim(tr_ctx_t ctx) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        TR_HIP(hipStreamSynchronize(ctx->stream));
        // drop internal scratch too
        for (size_t s = 0; s < ctx->scratch_ptr.size(); ++s) {
            if (ctx->scratch_ptr[s]) ctx->pool.release(ctx->scratch_ptr[s]);
            ctx->scratch_ptr[s] = nullptr;
            ctx->scratch_size[s] = 0;
        }
        ctx->pool.trim();
    });
}

int tr_pool_stats(tr_ctx_t ctx, size_t *in_use, size_t *cached) {
    TR_CTX_REQUIRED(ctx);
    if (in_use) *in_use = ctx->pool.in_use();
    if (cached) *cached = ctx->pool.cached();
    return TR_OK;
}

int tr_pool_set_limit(tr_ctx_t ctx, size_t bytes) {
    TR_CTX_REQUIRED(ctx);
    ctx->pool.set_limit(bytes);
    return TR_OK;
}

int tr_ctx_set_timing(tr_ctx_t ctx, int enable) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        if (!enable) fold_timing(ctx);
        ctx->timing = enable != 0;
    });
}

int tr_ctx_reset_timing(tr_ctx_t ctx) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        fold_timing(ctx);
        ctx->stats.clear();
    });
}

int tr_ctx_timing_count(tr_ctx_t ctx) {
    if (!ctx) return 0;
    try {
        fold_timing(ctx);
    } catch (...) {
    }
    return (int)ctx->stats.size();
}

int tr_ctx_timing_get(tr_ctx_t ctx, int i, char *name, int name_len, int64_t *launches, double *total_ms) {
    TR_CTX_REQUIRED(ctx);
    return guarded(ctx, [&] {
        fold_timing(ctx);
        TR_REQUIRE(i >= 0 && i < (int)ctx->stats.size(), "tr_ctx_timing_get: index out of range");
        auto it = ctx->stats.begin();
        std::advance(it, i);
        if (name && name_len > 0) {
            std::strncpy(name, it->first.c_str(), (size_t)name_len - 1);
            name[name_len - 1] = 0;
        }
        if (launches) *launches = it->second.launches;
        if (total_ms) *total_ms = it->second.total_ms;
    });
}

}  // extern "C"
