// SpaCommHandler.hpp — bitmask tile compression with the reference's
// per-stage filter selection (SpaCommHandler, include/spacomm.cuh:918-1091),
// for a grid whose tiles all live on this GPU.
//
// Filters.  On rank (r, c) of a P x P grid the reference's spcomm_2D
// (include/spacomm.cuh:165-352) leaves
//     A_column_filters[j] = colmask(A(r,c)) & rowmask(B(c,j))   j = 0..P-1
//     B_rows_filters[i]   = colmask(A(i,r)) & rowmask(B(r,c))   i = 0..P-1
// where colmask(A) has bit k set iff column k of the (CSC) A tile is
// non-empty and rowmask(B) bit k iff row k of the (CSR) B tile is: the
// allgathers along the process row and column, the intersection and the
// all-to-all of that function, evaluated here with tr_bitmask_gen and
// tr_bitmask_and on the tiles themselves (node_size 1).
//
// Compress(M, iteration_number, buffs, stream) selects the slice the way the
// reference does (:979-987): a CSR (B) tile takes B_rows_filters +
// mask_len * iteration_number, a CSC (A) tile A_column_filters +
// mask_len * iteration_number; then it keeps the segments whose bit is set
// (tr_compress) and returns a malloc'd header over a contiguous tile owned
// by `buffs`.
//
// Masks are 64-bit words (mask_len = ceil(K / 64)); on little-endian hosts
// their bytes equal the reference's MASK_SIZE = 8 masks.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <vector>

#include "CsxBuffers.hpp"
#include "mmio_csx.hpp"
#include "trident_hip.h"

// The process grid of one GPU: P x P tiles, this rank's position, and every
// rank's local A tile (CSC) and B tile (CSR), row-major by (row, col).
template <typename IT, typename VT>
struct LocalGrid {
    int row_size = 0, col_size = 0;
    int row_rank = 0, col_rank = 0;  // this rank's position (r, c)
    int node_size = 1, node_rank = 0;
    std::vector<mmio::CSX<IT, VT> *> A_tiles, B_tiles;
};

// Output storage of Compress (the reference's SpaCommBuffers,
// include/spacomm.cuh:635-700): the last compressed tile.
template <typename IT, typename VT>
struct SpaCommBuffers {
    tr_csx_t tile{};
    bool initialized = true;
    explicit SpaCommBuffers(const mmio::CSX<IT, VT> * = nullptr) {}
    SpaCommBuffers(const SpaCommBuffers &) = delete;
    SpaCommBuffers &operator=(const SpaCommBuffers &) = delete;
    void explicitFree() {
        if (tile.ptr_vec) (void)tr_csx_free(KokkosWrap::context(), &tile);
    }
    ~SpaCommBuffers() {
        try {
            explicitFree();
        } catch (...) {
        }
    }
};

template <typename IT, typename VT>
struct SpaCommHandler {
    int nfilters = 0;
    int32_t mask_len = 0;      // 64-bit words per filter
    int32_t sub_mask_len = 0;  // mask_len / node_size
    int node_rank = 0;
    uint64_t *A_column_filters = nullptr;  // device, nfilters * mask_len words
    uint64_t *B_rows_filters = nullptr;
    bool initialized = false;

    // Filters of rank (grid->row_rank, grid->col_rank) from the grid's tiles;
    // csx_A / csx_B are this rank's local tiles (checked against the grid).
    SpaCommHandler(mmio::CSX<IT, VT> *csx_A, mmio::CSX<IT, VT> *csx_B, const LocalGrid<IT, VT> *grid) {
        using KokkosWrap::Error;
        if (grid->row_size != grid->col_size) throw Error(TR_ERR_INVALID, "SpaCommHandler: grid must be square");
        const int P = grid->row_size, r = grid->row_rank, c = grid->col_rank;
        if ((int)grid->A_tiles.size() != P * P || (int)grid->B_tiles.size() != P * P)
            throw Error(TR_ERR_INVALID, "SpaCommHandler: grid tiles missing");
        if (grid->A_tiles[r * P + c] != csx_A || grid->B_tiles[r * P + c] != csx_B)
            throw Error(TR_ERR_INVALID, "SpaCommHandler: local tiles differ from the grid's");
        if (csx_A->majordim != mmio::MajorDim::COLS || csx_B->majordim != mmio::MajorDim::ROWS)
            throw Error(TR_ERR_INVALID, "SpaCommHandler: A must be CSC and B CSR");
        if (csx_A->ncols != csx_B->nrows)
            throw Error(TR_ERR_INVALID, "SpaCommHandler: A columns must equal B rows");
        nfilters = P;
        node_rank = grid->node_rank;
        const IT K = csx_A->ncols;
        mask_len = (int32_t)((K + 63) / 64);
        sub_mask_len = mask_len / grid->node_size;
        tr_ctx_t ctx = KokkosWrap::context();
        const size_t fbytes = sizeof(uint64_t) * (size_t)nfilters * (size_t)mask_len;
        const int dev = KokkosWrap::detail::thread_ctx().device;
        void *pa = nullptr, *pb = nullptr, *tmp = nullptr;
        KokkosWrap::detail::check_plain(tr_device_malloc(dev, &pa, fbytes), "SpaCommHandler: filters");
        KokkosWrap::detail::check_plain(tr_device_malloc(dev, &pb, fbytes), "SpaCommHandler: filters");
        KokkosWrap::detail::check_plain(tr_device_malloc(dev, &tmp, 2 * sizeof(uint64_t) * (size_t)mask_len),
                                        "SpaCommHandler: masks");
        A_column_filters = (uint64_t *)pa;
        B_rows_filters = (uint64_t *)pb;
        uint64_t *ma = (uint64_t *)tmp, *mb = ma + mask_len;
        auto gen = [&](const mmio::CSX<IT, VT> *t, uint64_t *out) {
            KokkosWrap::detail::check(tr_memset(ctx, out, 0, sizeof(uint64_t) * (size_t)mask_len),
                                      "SpaCommHandler: memset");
            const IT n = t->majordim == mmio::MajorDim::ROWS ? t->nrows : t->ncols;
            if (n != K) throw Error(TR_ERR_INVALID, "SpaCommHandler: tile dimension differs from K");
            KokkosWrap::detail::check(tr_bitmask_gen(ctx, (const int32_t *)t->ptr_vec, (int32_t)n, out),
                                      "SpaCommHandler: gen_bitmask");
        };
        for (int j = 0; j < P; ++j) {  // A(r,c) for receiver (r,j), which multiplies it by B(c,j)
            gen(grid->A_tiles[r * P + c], ma);
            gen(grid->B_tiles[c * P + j], mb);
            KokkosWrap::detail::check(tr_bitmask_and(ctx, ma, mb, mask_len, A_column_filters + (size_t)j * mask_len),
                                      "SpaCommHandler: intersect_bitmasks");
        }
        for (int i = 0; i < P; ++i) {  // B(r,c) for receiver (i,c), which multiplies A(i,r) by it
            gen(grid->A_tiles[i * P + r], ma);
            gen(grid->B_tiles[r * P + c], mb);
            KokkosWrap::detail::check(tr_bitmask_and(ctx, ma, mb, mask_len, B_rows_filters + (size_t)i * mask_len),
                                      "SpaCommHandler: intersect_bitmasks");
        }
        KokkosWrap::detail::check(tr_ctx_synchronize(ctx), "SpaCommHandler: synchronize");
        (void)tr_device_free(tmp);
        initialized = true;
    }

    SpaCommHandler(const SpaCommHandler &) = delete;
    SpaCommHandler &operator=(const SpaCommHandler &) = delete;

    // include/spacomm.cuh:962-1091
    mmio::CSX<IT, VT> *Compress(const mmio::CSX<IT, VT> *M, int iteration_number, SpaCommBuffers<IT, VT> *buffs,
                                void * /*stream*/ = nullptr) {
        using KokkosWrap::Error;
        if (iteration_number < 0 || iteration_number >= nfilters)
            throw Error(TR_ERR_INVALID, "SpaCommHandler::Compress: invalid iteration number");
        const bool rows = M->majordim == mmio::MajorDim::ROWS;
        const uint64_t *mask = rows ? B_rows_filters + (size_t)mask_len * iteration_number +
                                          (size_t)node_rank * sub_mask_len
                                    : A_column_filters + (size_t)mask_len * iteration_number;
        const int32_t considered = rows ? sub_mask_len : mask_len;
        const IT ptr_segments = rows ? M->nrows : M->ncols;
        if ((int64_t)ptr_segments > (int64_t)considered * 64)
            throw Error(TR_ERR_INVALID, "SpaCommHandler::Compress: tile has more segments than its mask");
        tr_csx_t in{};
        in.majordim = rows ? TR_ROWS : TR_COLS;
        in.nrows = M->nrows;
        in.ncols = M->ncols;
        in.nnz = M->nnz;
        in.ptr_vec = (int32_t *)M->ptr_vec;
        in.idx_vec = (int32_t *)M->idx_vec;
        in.val = (double *)M->val;
        buffs->explicitFree();
        KokkosWrap::detail::check(tr_compress(KokkosWrap::context(), &in, mask, &buffs->tile),
                                  "SpaCommHandler::Compress");
        auto *out = (mmio::CSX<IT, VT> *)std::malloc(sizeof(mmio::CSX<IT, VT>));
        out->majordim = M->majordim;
        out->nnz = (IT)buffs->tile.nnz;
        out->nrows = M->nrows;
        out->ncols = M->ncols;
        out->contig = true;
        out->buf = (char *)buffs->tile.buf;
        out->buf_size = buffs->tile.buf_size;
        out->ptr_vec = buffs->tile.ptr_vec;
        out->idx_vec = buffs->tile.idx_vec;
        out->val = buffs->tile.val;
        return out;
    }

    void explicitFree() {
        if (A_column_filters) (void)tr_device_free(A_column_filters);
        if (B_rows_filters) (void)tr_device_free(B_rows_filters);
        A_column_filters = B_rows_filters = nullptr;
        initialized = false;
    }

    ~SpaCommHandler() { explicitFree(); }
};
