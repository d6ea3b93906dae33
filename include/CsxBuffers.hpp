// CsxBuffers.hpp — growable device storage of one CSR matrix (the partial C
// of a stage loop, conversion and gather buffers), with the fields and
// methods of the reference's CsxBuffers (include/utils.cuh:527-784).
//
// Storage comes from tr_device_malloc, outside any library context, because
// the reference grows these buffers from the accumulation thread while the
// main thread owns the multiply (src/hns_spgemm.cu:34-65): a context may not
// be shared between host threads, device allocations may.
//
// Differences to the reference, all deliberate:
//  - the library keeps its own scratch, so there are no cub temporary
//    buffers: ensure_tmp*() only record the request;
//  - growth reallocates without copying, as the reference's does;
//  - failures throw KokkosWrap::Error instead of exit(EXIT_FAILURE).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <new>
#include <stdexcept>
#include <string>

#include "mmio_csx.hpp"
#include "trident_hip.h"

namespace KokkosWrap {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

namespace detail {

struct ThreadCtx {
    tr_ctx_t ctx = nullptr;
    int device = 0;
    bool borrowed = false;  // bound with bind_context(): owned elsewhere
    ~ThreadCtx() {
        if (ctx && !borrowed) tr_ctx_destroy(ctx);
    }
};

inline ThreadCtx &thread_ctx() {
    static thread_local ThreadCtx tc;
    return tc;
}

inline tr_ctx_t ctx() {
    ThreadCtx &tc = thread_ctx();
    if (!tc.ctx) {
        int rc = tr_ctx_create(&tc.ctx, tc.device, nullptr);
        if (rc != TR_OK) throw Error(rc, "KokkosWrap: tr_ctx_create failed");
    }
    return tc.ctx;
}

inline void check(int rc, const char *what) {
    if (rc != TR_OK) throw Error(rc, std::string(what) + ": " + tr_last_error(ctx()));
}

inline void check_plain(int rc, const char *what) {
    if (rc != TR_OK) throw Error(rc, std::string(what) + " failed");
}

}  // namespace detail

// Select the GPU of the calling thread's context (before its first use).
inline void set_device(int device) {
    detail::ThreadCtx &tc = detail::thread_ctx();
    if (tc.ctx) throw Error(TR_ERR_INVALID, "KokkosWrap::set_device: context already created on this thread");
    tc.device = device;
}

// The calling thread's library context (for direct C-ABI calls).
inline tr_ctx_t context() { return detail::ctx(); }

// Make the calling thread work on an existing context (owned and destroyed
// by the caller, after this thread's last call), instead of its own.
inline void bind_context(tr_ctx_t c) {
    detail::ThreadCtx &tc = detail::thread_ctx();
    if (tc.ctx && !tc.borrowed) throw Error(TR_ERR_INVALID, "KokkosWrap::bind_context: thread already has a context");
    tc.ctx = c;
    tc.borrowed = true;
}

// The handle the reference's signatures take (a cusparseHandle_t there).
// The library orders work on the calling thread's context, so the handle
// carries nothing; it exists so that call sites keep their shape.
struct Handle {};

}  // namespace KokkosWrap

template <typename IT, typename VT>
struct CsxBuffers {
    static_assert(sizeof(IT) == 4 && sizeof(VT) == 8, "int32 indices, fp64 values");

    uint64_t nnz = 0;
    uint64_t max_nnz = 0;
    uint64_t ptr_dim = 0;
    uint64_t other_dim = 0;
    uint64_t nbufs = 0;
    int initialized = 0;
    VT *d_node_vals = nullptr;
    IT *d_node_colinds = nullptr;
    IT *d_node_rowptrs = nullptr;
    void *stream = nullptr;  // the caller's stream handle (hipStream_t*), kept for the interface
    bool async_csx = false;
    uint64_t tmp_request = 0;  // largest ensure_tmp*() request (informational)
    int device = 0;

    // input_nnz entries, input_ptr_dim row pointers (zeroed), other_dim
    // columns (include/utils.cuh:544-584)
    CsxBuffers(uint64_t input_nnz, uint64_t input_ptr_dim, uint64_t _other_dim, void *_stream = nullptr,
               uint64_t _nbufs = 1, bool _async_csx = false) {
        stream = _stream;
        nbufs = _nbufs;
        async_csx = _async_csx;
        allocate(input_nnz, input_ptr_dim, _other_dim);
    }

    CsxBuffers() {}

    CsxBuffers(const CsxBuffers &) = delete;
    CsxBuffers &operator=(const CsxBuffers &) = delete;

    size_t total_size() const { return (sizeof(VT) + sizeof(IT)) * max_nnz + sizeof(IT) * ptr_dim; }

    // header over the buffers, malloc'd like the reference's (:609-619)
    mmio::CSX<IT, VT> *to_mmio_csx() const {
        auto *csx = (mmio::CSX<IT, VT> *)std::malloc(sizeof(mmio::CSX<IT, VT>));
        csx->majordim = mmio::MajorDim::ROWS;
        csx->nnz = (IT)nnz;
        csx->nrows = (IT)(ptr_dim - 1);
        csx->ncols = (IT)other_dim;
        csx->val = d_node_vals;
        csx->ptr_vec = d_node_rowptrs;
        csx->idx_vec = d_node_colinds;
        csx->contig = false;
        csx->buf = nullptr;
        csx->buf_size = 0;
        return csx;
    }

    // tr_csx_t view of the current contents (nnz entries, ptr_dim - 1 rows)
    tr_csx_t view() const {
        tr_csx_t t{};
        t.majordim = TR_ROWS;
        t.nrows = (int32_t)(ptr_dim ? ptr_dim - 1 : 0);
        t.ncols = (int32_t)other_dim;
        t.nnz = (int64_t)nnz;
        t.ptr_vec = d_node_rowptrs;
        t.idx_vec = d_node_colinds;
        t.val = d_node_vals;
        return t;
    }

    // Room for input_nnz entries and input_ptr_dim row pointers; sets nnz
    // (include/utils.cuh:621-667).  Growth reallocates without copying.
    void ensure(uint64_t input_nnz, uint64_t input_ptr_dim, uint64_t _other_dim) {
        if (!initialized) {
            allocate(input_nnz, input_ptr_dim, _other_dim);
            nnz = input_nnz;
            return;
        }
        nnz = input_nnz;
        other_dim = _other_dim;
        if (input_nnz > max_nnz) {
            detail_free(d_node_vals);
            detail_free(d_node_colinds);
            d_node_vals = nullptr;
            d_node_colinds = nullptr;
            max_nnz = input_nnz;
            d_node_vals = (VT *)detail_alloc(sizeof(VT) * max_nnz);
            d_node_colinds = (IT *)detail_alloc(sizeof(IT) * max_nnz);
        }
        if (input_ptr_dim > ptr_dim) {
            detail_free(d_node_rowptrs);
            d_node_rowptrs = nullptr;
            ptr_dim = input_ptr_dim;
            d_node_rowptrs = (IT *)detail_alloc(sizeof(IT) * ptr_dim);
            zero_rowptrs();
        }
    }

    // same contract, the stream-ordered variant of the reference (:669-716)
    void ensure_async(uint64_t input_nnz, uint64_t input_ptr_dim, uint64_t _other_dim) {
        ensure(input_nnz, input_ptr_dim, _other_dim);
    }

    void ensure_tmp(uint64_t required_size, uint64_t = 0) {
        if (required_size > tmp_request) tmp_request = required_size;
    }
    void ensure_tmp_async(uint64_t required_size, uint64_t idx = 0) { ensure_tmp(required_size, idx); }
    void free_tmp_async(uint64_t = 0) {}

    void explicitFree() {
        if (!initialized) return;
        detail_free(d_node_colinds);
        detail_free(d_node_rowptrs);
        detail_free(d_node_vals);
        d_node_colinds = nullptr;
        d_node_rowptrs = nullptr;
        d_node_vals = nullptr;
        initialized = 0;
        ptr_dim = 0;
        nnz = 0;
        max_nnz = 0;
        nbufs = 0;
    }
    void explicitFreeAsync() { explicitFree(); }

    ~CsxBuffers() {
        try {
            explicitFree();
        } catch (...) {
        }
    }

  private:
    void allocate(uint64_t input_nnz, uint64_t input_ptr_dim, uint64_t _other_dim) {
        device = KokkosWrap::detail::thread_ctx().device;
        nnz = 0;
        max_nnz = input_nnz;
        ptr_dim = input_ptr_dim;
        other_dim = _other_dim;
        d_node_vals = (VT *)detail_alloc(sizeof(VT) * (max_nnz ? max_nnz : 1));
        d_node_colinds = (IT *)detail_alloc(sizeof(IT) * (max_nnz ? max_nnz : 1));
        d_node_rowptrs = (IT *)detail_alloc(sizeof(IT) * (ptr_dim ? ptr_dim : 1));
        initialized = 1;
        zero_rowptrs();
    }
    void zero_rowptrs() {
        if (!ptr_dim) return;
        tr_ctx_t c = KokkosWrap::detail::ctx();
        KokkosWrap::detail::check(tr_memset(c, d_node_rowptrs, 0, sizeof(IT) * ptr_dim), "CsxBuffers: memset");
        KokkosWrap::detail::check(tr_ctx_synchronize(c), "CsxBuffers: synchronize");
    }
    void *detail_alloc(size_t bytes) {
        void *p = nullptr;
        KokkosWrap::detail::check_plain(tr_device_malloc(device, &p, bytes), "CsxBuffers: tr_device_malloc");
        return p;
    }
    static void detail_free(void *p) {
        if (p) KokkosWrap::detail::check_plain(tr_device_free(p), "CsxBuffers: tr_device_free");
    }
};
