// KokkosWrap.hpp — the reference's KokkosWrap::LocalMatrix interface
// (include/KokkosWrap.hpp:43-77 and :315-439 of the reference), rebuilt on
// the C-ABI of libtrident_hip (trident_hip.h) for MI355X.
//
// Callers keep the same calls:
//     LocalMatrix<int32_t,int32_t,double> A(csx_a), B(csx_b), C;
//     auto P = LocalMatrix<...>::spgemm(A, B);   // C = A*B
//     LocalMatrix<...>::sp_mma(A, B, C);          // C += A*B
//     LocalMatrix<...>::spadd(A, C);              // C += A
//     mmio::CSX<int32_t,double>* c = C.get_csx();
//     C.freeBuffers();
//
// Differences to the reference, all deliberate:
//  - instantiated for int32 indices and fp64 values (the reference's
//    instantiation is <int32_t,float>);
//  - storage is a tr_csx_t (device pointers); the product owns library
//    storage released by freeBuffers(); views built from caller arrays stay
//    borrowed, as the reference's unmanaged views (KokkosWrap.hpp:216-230);
//  - failures throw KokkosWrap::Error (status code + message) instead of
//    exit(EXIT_FAILURE);
//  - spadd on an uninitialised C deep-copies A (the reference aliases it).
//
// All operations run on a per-thread library context: one HIP stream per
// host thread, as the reference's per-thread default stream
// (include/common.h:13).  Use set_device() before the first call on a thread
// to pick a GPU other than 0.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "KokkosWrapTriple.hpp"
#include "mmio_csx.hpp"
#include "trident_hip.h"

using MajorDim = mmio::MajorDim;

namespace KokkosWrap {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

namespace detail {

struct ThreadCtx {
    tr_ctx_t ctx = nullptr;
    int device = 0;
    ~ThreadCtx() {
        if (ctx) tr_ctx_destroy(ctx);
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

}  // namespace detail

// Select the GPU of the calling thread's context (before its first use).
inline void set_device(int device) {
    detail::ThreadCtx &tc = detail::thread_ctx();
    if (tc.ctx) throw Error(TR_ERR_INVALID, "KokkosWrap::set_device: context already created on this thread");
    tc.device = device;
}

// The calling thread's library context (for direct C-ABI calls).
inline tr_ctx_t context() { return detail::ctx(); }

template <typename KIT, typename DIT, typename VT>
struct LocalMatrix {
    static_assert(sizeof(KIT) == 4 && sizeof(DIT) == 4, "int32 indices");
    static_assert(sizeof(VT) == 8, "fp64 values");

    bool initialized;
    tr_csx_t storage;

    LocalMatrix() : initialized(false), storage{} {}

    // View over a caller-owned device CSR (or CSC, converted to CSR like the
    // reference's csc_to_csr, KokkosWrap.hpp:441-531).
    explicit LocalMatrix(mmio::CSX<DIT, VT> *csx) : initialized(true), storage{} {
        storage.majordim = csx->majordim == MajorDim::ROWS ? TR_ROWS : TR_COLS;
        storage.nrows = csx->nrows;
        storage.ncols = csx->ncols;
        storage.nnz = csx->nnz;
        storage.ptr_vec = (int32_t *)csx->ptr_vec;
        storage.idx_vec = (int32_t *)csx->idx_vec;
        storage.val = (double *)csx->val;
        storage.contig = 0;
        if (storage.majordim == TR_COLS) {
            tr_csx_t csr;
            detail::check(tr_csc_to_csr(detail::ctx(), &storage, &csr), "LocalMatrix(csc)");
            storage = csr;
        }
    }

    explicit LocalMatrix(mmio::CSR<DIT, VT> *csr) : initialized(true), storage{} {
        storage.majordim = TR_ROWS;
        storage.nrows = csr->nrows;
        storage.ncols = csr->ncols;
        storage.nnz = csr->nnz;
        storage.ptr_vec = (int32_t *)csr->row_ptr;
        storage.idx_vec = (int32_t *)csr->col_idx;
        storage.val = (double *)csr->val;
    }

    LocalMatrix(DIT nrows, DIT ncols, DIT nnz, DIT *ptrvec, DIT *idxvec, VT *values, MajorDim layout)
        : initialized(false), storage{} {
        mmio::CSX<DIT, VT> c{layout, nrows, ncols, nnz, ptrvec, idxvec, values, false, nullptr, 0};
        *this = LocalMatrix(&c);
    }

    // C = C + A*B (KokkosWrap.hpp:315-364)
    static void sp_mma(const LocalMatrix &A, const LocalMatrix &B, LocalMatrix &C) {
        tr_csx_t c = C.initialized ? C.storage : tr_csx_t{};
        detail::check(tr_sp_mma(detail::ctx(), &A.storage, &B.storage, &c), "LocalMatrix::sp_mma");
        C.storage = c;
        C.initialized = true;
    }

    // C = A*B (KokkosWrap.hpp:367-391)
    static LocalMatrix spgemm(const LocalMatrix &A, const LocalMatrix &B) {
        LocalMatrix r;
        detail::check(tr_spgemm(detail::ctx(), &A.storage, &B.storage, &r.storage), "LocalMatrix::spgemm");
        r.initialized = true;
        return r;
    }

    // C = A + C (KokkosWrap.hpp:394-439)
    static void spadd(const LocalMatrix &A, LocalMatrix &C) {
        tr_csx_t c = C.initialized ? C.storage : tr_csx_t{};
        detail::check(tr_spadd(detail::ctx(), &A.storage, &c), "LocalMatrix::spadd");
        C.storage = c;
        C.initialized = c.ptr_vec != nullptr;
    }

    // header over the storage, malloc'd like the reference's rawptr_get
    // (KokkosWrap.hpp:89-105); the caller frees the header with free().
    mmio::CSX<DIT, VT> *get_csx() {
        auto *c = (mmio::CSX<DIT, VT> *)std::malloc(sizeof(mmio::CSX<DIT, VT>));
        c->majordim = MajorDim::ROWS;
        c->nrows = storage.nrows;
        c->ncols = storage.ncols;
        c->nnz = (DIT)storage.nnz;
        c->ptr_vec = (DIT *)storage.ptr_vec;
        c->idx_vec = (DIT *)storage.idx_vec;
        c->val = (VT *)storage.val;
        c->contig = storage.contig != 0;
        c->buf = (char *)storage.buf;
        c->buf_size = storage.buf_size;
        return c;
    }

    // Release library-owned storage (KokkosWrap.hpp:294-302); borrowed views
    // are left to their owner.
    void freeBuffers() {
        if (initialized) {
            detail::check(tr_csx_free(detail::ctx(), &storage), "LocalMatrix::freeBuffers");
            initialized = false;
        }
    }

    ~LocalMatrix() {}  // views are not freed implicitly (KokkosWrap.hpp:304-308)
};

template <typename KIT, typename VT>
mmio::CSX<KIT, VT> *rawptr_get(LocalMatrix<KIT, KIT, VT> &m) {
    return m.get_csx();
}

}  // namespace KokkosWrap
