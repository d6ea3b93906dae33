// AccumCSX.hpp — the partial-C wrapper of a stage loop and sp_mma_hybrid,
// the call with which the reference's two stage loops multiply a pair of
// tiles and add the product into the rank's partial C
// (include/cusparse_helpers.cuh:134-484 and :708-739; called at
// src/hns_spgemm.cu:969 and src/sparse_summa.cu:99).
//
// AccumCSX keeps the reference wrapper's states and constructors: a matrix
// header (Mat), growable buffers (Buffs), both (Both, a view of a
// LocalMatrix) or nothing (Null).  The reference calls the type CusparseCSX;
// that name is kept below as an alias so the stage loops compile unchanged.
//
//     CusparseCSX<IT,VT> *C_accum = new CusparseCSX<IT,VT>(C_accum_buffs);
//     CusparseCSX<IT,VT> *C_local = new CusparseCSX<IT,VT>(C_local_buffs);
//     ...
//     LocalMatrix<IT,IT,VT> A_p(A_remote->mat), B_p(B_node->mat);
//     sp_mma_hybrid(&handle, A_p, B_p, &C_local, &C_accum, did_one_spgemm);
//
// sp_mma_hybrid: C_prod = A*B; an empty product returns at once (did_first
// untouched); otherwise C_accum = C_prod + C_local (tr_spadd_into, growing
// C_accum's buffers when the result does not fit, as cusparse_spgeam does,
// :866-867), did_first = true, and the two buffers swap so C_local holds
// the running sum.  Values are fp64 and every add is exact-order
// deterministic (c = a + b per entry), so the partial C is bit-identical to
// the oracle's sum of the stage products in stage order.
#pragma once

#include <cstdlib>
#include <utility>

#include "CsxBuffers.hpp"
#include "KokkosWrap.hpp"

template <typename IT, typename VT>
struct AccumCSX {
    enum class State { Mat, Buffs, Null, Both };

    mmio::CSX<IT, VT> *mat = nullptr;
    CsxBuffers<IT, VT> *buffers = nullptr;
    State state = State::Null;
    bool owns_mat_storage = false;  // mat converted from CSC into library-owned storage
    tr_csx_t owned{};               // that storage
    bool own_header = false;        // mat was malloc'd here

    AccumCSX() {}

    explicit AccumCSX(mmio::CSX<IT, VT> *m) : mat(m), state(State::Mat) {}

    // view of a LocalMatrix's storage (cusparse_helpers.cuh:182-201)
    explicit AccumCSX(KokkosWrap::LocalMatrix<IT, IT, VT> &lm) : state(State::Both) {
        mat = lm.get_csx();
        own_header = true;
        buffers = new CsxBuffers<IT, VT>();
        buffers->d_node_vals = mat->val;
        buffers->d_node_colinds = mat->idx_vec;
        buffers->d_node_rowptrs = mat->ptr_vec;
        buffers->nnz = (uint64_t)mat->nnz;
        buffers->ptr_dim = (uint64_t)mat->nrows + 1;
        buffers->other_dim = (uint64_t)mat->ncols;
    }

    // a tile header; a CSC tile is converted to CSR, into the conversion
    // buffers when given (cusparse_helpers.cuh:204-228)
    template <typename H>
    AccumCSX(H *handle, mmio::CSX<IT, VT> *m, CsxBuffers<IT, VT> *conversion_buffers) : state(State::Mat) {
        if (m->majordim == mmio::MajorDim::COLS) {
            KokkosWrap::LocalMatrix<IT, IT, VT> tmp;
            tr_csx_t csr = tmp.csc_to_csr(*handle, m, conversion_buffers);
            owns_mat_storage = conversion_buffers == nullptr;
            if (owns_mat_storage) owned = csr;
            mat = (mmio::CSX<IT, VT> *)std::malloc(sizeof(mmio::CSX<IT, VT>));
            own_header = true;
            mat->majordim = mmio::MajorDim::ROWS;
            mat->nrows = csr.nrows;
            mat->ncols = csr.ncols;
            mat->nnz = (IT)csr.nnz;
            mat->ptr_vec = csr.ptr_vec;
            mat->idx_vec = csr.idx_vec;
            mat->val = csr.val;
            mat->contig = false;
            mat->buf = nullptr;
            mat->buf_size = 0;
        } else {
            mat = m;
        }
    }

    explicit AccumCSX(CsxBuffers<IT, VT> *b) : buffers(b), state(State::Buffs) {}

    AccumCSX(const AccumCSX &) = delete;
    AccumCSX &operator=(const AccumCSX &) = delete;

    bool is_mat() const { return state == State::Mat || state == State::Both; }
    bool is_buffs() const { return state == State::Buffs || state == State::Both; }
    bool is_null() const { return state == State::Null; }

    IT nrows() const { return is_mat() ? mat->nrows : is_buffs() ? (IT)(buffers->ptr_dim - 1) : 0; }
    IT ncols() const { return is_mat() ? mat->ncols : is_buffs() ? (IT)buffers->other_dim : 0; }
    IT nnz() const { return is_mat() ? mat->nnz : is_buffs() ? (IT)buffers->nnz : 0; }

    // tr_csx_t view of the current contents
    tr_csx_t view() const {
        if (is_buffs()) return buffers->view();
        tr_csx_t t{};
        if (!is_mat()) return t;
        t.majordim = mat->majordim == mmio::MajorDim::ROWS ? TR_ROWS : TR_COLS;
        t.nrows = mat->nrows;
        t.ncols = mat->ncols;
        t.nnz = mat->nnz;
        t.ptr_vec = mat->ptr_vec;
        t.idx_vec = mat->idx_vec;
        t.val = mat->val;
        return t;
    }

    void assert_buffs(const char *name) const {
        if (!is_buffs())
            throw KokkosWrap::Error(TR_ERR_INVALID, std::string("expected ") + name + " to hold CsxBuffers");
    }

    void explicit_free() {
        if (owns_mat_storage) {
            (void)tr_csx_free(KokkosWrap::context(), &owned);
            owns_mat_storage = false;
        }
        if (state == State::Buffs && buffers) buffers->explicitFree();
    }

    ~AccumCSX() {
        try {
            if (owns_mat_storage) (void)tr_csx_free(KokkosWrap::context(), &owned);
        } catch (...) {
        }
        // a Both view owns neither the LocalMatrix storage nor the buffers'
        // arrays, only the descriptor objects
        if (state == State::Both) delete buffers;
        if (own_header) std::free(mat);
    }
};

// The reference's name for this wrapper (include/cusparse_helpers.cuh:134).
template <typename IT, typename VT>
using CusparseCSX = AccumCSX<IT, VT>;

// C_accum = C_prod + C_local into C_accum's buffers (growing them when the
// result does not fit), then C_local and C_accum swap.  The reference's
// cusparse_spgeam (include/cusparse_helpers.cuh:799-905) plus the swap of
// sp_mma_hybrid (:732).
template <typename IT, typename VT, typename H>
void accumulate_into(H *, const tr_csx_t &prod, AccumCSX<IT, VT> **C_local, AccumCSX<IT, VT> **C_accum) {
    (*C_local)->assert_buffs("C_local");
    (*C_accum)->assert_buffs("C_accum");
    tr_ctx_t c = KokkosWrap::context();
    CsxBuffers<IT, VT> *acc = (*C_accum)->buffers;
    const tr_csx_t local = (*C_local)->view();
    if (acc->ptr_dim < (uint64_t)prod.nrows + 1 || !acc->initialized)
        acc->ensure(acc->max_nnz, (uint64_t)prod.nrows + 1, (uint64_t)prod.ncols);
    for (int attempt = 0; attempt < 2; ++attempt) {
        tr_csx_t out = acc->view();
        int64_t need = 0;
        const int rc = tr_spadd_into(c, &prod, &local, &out, (int64_t)acc->max_nnz, &need);
        if (rc == TR_OK) {
            acc->nnz = (uint64_t)need;
            acc->other_dim = (uint64_t)prod.ncols;
            std::swap(*C_local, *C_accum);
            return;
        }
        if (rc != TR_ERR_OVERFLOW || attempt == 1) KokkosWrap::detail::check(rc, "sp_mma_hybrid: add");
        // grow like the reference's ensure_async (no copy: the add rewrites everything)
        acc->ensure((uint64_t)need + (uint64_t)need / 4, (uint64_t)prod.nrows + 1, (uint64_t)prod.ncols);
    }
}

// include/cusparse_helpers.cuh:708-739
template <typename IT, typename VT, typename H>
void sp_mma_hybrid(H *handle, const KokkosWrap::LocalMatrix<IT, IT, VT> &A, const KokkosWrap::LocalMatrix<IT, IT, VT> &B,
                   AccumCSX<IT, VT> **C_local, AccumCSX<IT, VT> **C_accum, bool &did_first) {
    KokkosWrap::LocalMatrix<IT, IT, VT> C_prod = KokkosWrap::LocalMatrix<IT, IT, VT>::spgemm(A, B);
    if (C_prod.storage.nnz == 0) {
        C_prod.freeBuffers();
        return;
    }
    did_first = true;
    try {
        accumulate_into(handle, C_prod.storage, C_local, C_accum);
    } catch (...) {
        C_prod.freeBuffers();
        throw;
    }
    C_prod.freeBuffers();
}
