// KokkosWrap.hpp — the reference's KokkosWrap::LocalMatrix interface
// (include/KokkosWrap.hpp:43-77 and :315-439 of the reference), rebuilt on
// the C-ABI of libtrident_hip (trident_hip.h) for MI355X.
//
// Callers keep the same calls:
//     LocalMatrix<int32_t,int32_t,double> A(csx_a), B(csx_b), C;
//     LocalMatrix<int32_t,int32_t,double> At(handle, csc_a, conv_buffs);  // CSC -> CSR
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

#include "CsxBuffers.hpp"
#include "KokkosWrapTriple.hpp"
#include "mmio_csx.hpp"
#include "trident_hip.h"

using MajorDim = mmio::MajorDim;

namespace KokkosWrap {

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

    // View over a caller-owned device CSR; a CSC tile is converted to CSR,
    // into `buffs` when given (no allocation once they are large enough),
    // as the reference's constructor does (KokkosWrap.hpp:210-237).  The
    // handle is accepted for the call-site shape and not used.
    template <typename H>
    LocalMatrix(H &handle, mmio::CSX<DIT, VT> *csx, CsxBuffers<DIT, VT> *buffs = nullptr)
        : initialized(true), storage{} {
        if (csx->majordim == MajorDim::ROWS) {
            *this = LocalMatrix(csx);
        } else {
            storage = csc_to_csr(handle, csx, buffs);
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

    // CSR of a CSC tile (KokkosWrap.hpp:441-531).  With `buffers` the result
    // lives in them (grown with ensure(nnz, nrows + 1, ncols) like the
    // reference) and stays borrowed; without, it is library-owned storage
    // that freeBuffers() releases.
    template <typename H>
    tr_csx_t csc_to_csr(H &, mmio::CSX<DIT, VT> *csc, CsxBuffers<DIT, VT> *buffers = nullptr) {
        if (csc->majordim != MajorDim::COLS)
            throw KokkosWrap::Error(TR_ERR_INVALID, "LocalMatrix::csc_to_csr: input is not CSC");
        tr_csx_t in{};
        in.majordim = TR_COLS;
        in.nrows = csc->nrows;
        in.ncols = csc->ncols;
        in.nnz = csc->nnz;
        in.ptr_vec = (int32_t *)csc->ptr_vec;
        in.idx_vec = (int32_t *)csc->idx_vec;
        in.val = (double *)csc->val;
        tr_csx_t out{};
        if (buffers == nullptr) {
            detail::check(tr_csc_to_csr(detail::ctx(), &in, &out), "LocalMatrix::csc_to_csr");
            return out;
        }
        buffers->ensure((uint64_t)csc->nnz, (uint64_t)csc->nrows + 1, (uint64_t)csc->ncols);
        detail::check(tr_csc_to_csr_into(detail::ctx(), &in, buffers->d_node_rowptrs, buffers->d_node_colinds,
                                         buffers->d_node_vals),
                      "LocalMatrix::csc_to_csr");
        out = buffers->view();
        return out;
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
