// TileHolder.hpp — the CSR/CSC tile buffer of a stage loop, with the layout
// and the local methods of the reference's TileHolder
// (include/tile_holder.cuh:16-233), without MPI.
//
// A tile lives in one device buffer d_buf in the order the reference sends
// it (set_csx_ptrs, :79-89):
//     [ VT vals[nnz] | IT inds[nnz] | IT ptrs[ptr_size + 1] ]
// and is exchanged with the wire header IT[2] = {nnz, buf_size} (:121).
// The reference asserts sizeof(IT) == sizeof(VT) (:24, :81), which rules
// out its fp64 instantiation with int32 indices; the layout only needs the
// value array first and sizeof(VT) a multiple of sizeof(IT) for every array
// to stay aligned, which is what is asserted here.
//
// The MPI transport (send_tile_contig / recv_tile_contig / tile_bcast /
// node_allgather, :92-190, :236-511) is out of scope on one GPU;
// load_tile_contig() takes the place of recv_tile_contig: it copies a tile
// that already sits in device memory in the wire layout (what a receive
// would have delivered into d_buf).
#pragma once

#include <cstdint>
#include <cstdlib>

#include "CsxBuffers.hpp"
#include "mmio_csx.hpp"
#include "trident_hip.h"

template <typename IT, typename VT>
struct TileHolder {
    static_assert(sizeof(VT) % sizeof(IT) == 0, "value size must be a multiple of the index size");
    static_assert(sizeof(IT) == 4 && sizeof(VT) == 8, "the library works on int32 indices and fp64 values");

    char *d_buf = nullptr;
    size_t d_buf_size = 0;
    IT max_nnz = 0;
    IT ptr_size = 0;  // segments of a tile (rows for CSR, columns for CSC)
    IT current_nnz = 0;
    VT *d_vals_buf = nullptr;
    IT *d_inds_buf = nullptr;
    IT *d_ptrs_buf = nullptr;

    // _buf_size bytes (mmio::CSX_buf_size of the largest tile), tiles of
    // _ptr_size segments and at most nnz_size entries (:22-36); the buffer
    // starts zeroed like the reference's cudaMemset
    TileHolder(const size_t _buf_size, const IT _ptr_size, const IT nnz_size)
        : d_buf_size(_buf_size), max_nnz(nnz_size), ptr_size(_ptr_size) {
        if (_buf_size < mmio::CSX_buf_size<IT, VT>(_ptr_size, 0, nnz_size, mmio::MajorDim::ROWS))
            throw KokkosWrap::Error(TR_ERR_INVALID, "TileHolder: buffer smaller than the tile it must hold");
        const int dev = KokkosWrap::detail::thread_ctx().device;
        void *p = nullptr;
        KokkosWrap::detail::check_plain(tr_device_malloc(dev, &p, d_buf_size), "TileHolder: tr_device_malloc");
        d_buf = (char *)p;
        tr_ctx_t c = KokkosWrap::context();
        KokkosWrap::detail::check(tr_memset(c, d_buf, 0, d_buf_size), "TileHolder: memset");
        KokkosWrap::detail::check(tr_ctx_synchronize(c), "TileHolder: synchronize");
    }

    TileHolder() {}
    TileHolder(const TileHolder &) = delete;
    TileHolder &operator=(const TileHolder &) = delete;

    ~TileHolder() {
        if (d_buf) (void)tr_device_free(d_buf);
    }

    void set_csx_ptrs() { set_csx_ptrs(current_nnz); }

    // array starts for a tile of nnz entries (:79-89)
    void set_csx_ptrs(const IT nnz) {
        if (nnz < 0 || nnz > max_nnz)
            throw KokkosWrap::Error(TR_ERR_OVERFLOW, "TileHolder: tile nnz exceeds the holder's capacity");
        current_nnz = nnz;
        size_t offset = 0;
        d_vals_buf = (VT *)d_buf;
        offset += sizeof(VT) * (size_t)nnz;
        d_inds_buf = (IT *)(d_buf + offset);
        offset += sizeof(IT) * (size_t)nnz;
        d_ptrs_buf = (IT *)(d_buf + offset);
    }

    // wire header of the current tile: {nnz, buf_size} (:121)
    void wire_header(IT payload[2]) const {
        payload[0] = current_nnz;
        payload[1] = (IT)mmio::CSX_buf_size<IT, VT>(ptr_size, 0, current_nnz, mmio::MajorDim::ROWS);
    }

    // Copy a tile from device memory already in the wire layout (the bytes a
    // receive would deliver, :163-188): `nnz` entries, `buf_size` bytes.  An
    // empty tile gets a zero pointer array, as the reference's receive does.
    IT load_tile_contig(const void *d_src, const IT nnz, const size_t buf_size) {
        set_csx_ptrs(nnz);
        tr_ctx_t c = KokkosWrap::context();
        if (nnz > 0) {
            if (buf_size > d_buf_size)
                throw KokkosWrap::Error(TR_ERR_OVERFLOW, "TileHolder: tile larger than the buffer");
            KokkosWrap::detail::check(tr_memcpy_d2d(c, d_buf, d_src, buf_size), "TileHolder::load_tile_contig");
        } else {
            KokkosWrap::detail::check(tr_memset(c, d_ptrs_buf, 0, sizeof(IT) * ((size_t)ptr_size + 1)),
                                      "TileHolder::load_tile_contig");
        }
        return nnz;
    }

    // Copy this rank's own tile into the holder (:192-209).
    IT copy_device_local_csx(mmio::CSX<IT, VT> *input, void * /*stream*/ = nullptr) {
        set_csx_ptrs(input->nnz);
        tr_ctx_t c = KokkosWrap::context();
        if (input->nnz > 0) {
            KokkosWrap::detail::check(tr_memcpy_d2d(c, d_vals_buf, input->val, sizeof(VT) * (size_t)input->nnz),
                                      "TileHolder::copy_device_local_csx");
            KokkosWrap::detail::check(
                tr_memcpy_d2d(c, d_inds_buf, input->idx_vec, sizeof(IT) * (size_t)input->nnz),
                "TileHolder::copy_device_local_csx");
        }
        KokkosWrap::detail::check(
            tr_memcpy_d2d(c, d_ptrs_buf, input->ptr_vec, sizeof(IT) * ((size_t)ptr_size + 1)),
            "TileHolder::copy_device_local_csx");
        return input->nnz;
    }

    // malloc'd header over given device arrays (:212-227)
    mmio::CSX<IT, VT> *form_mmiocsx(const IT nrows, const IT ncols, const IT nnz, mmio::MajorDim layout, VT *d_vals,
                                    IT *d_inds, IT *d_ptrs) {
        auto *csx = (mmio::CSX<IT, VT> *)std::malloc(sizeof(mmio::CSX<IT, VT>));
        csx->majordim = layout;
        csx->nrows = nrows;
        csx->ncols = ncols;
        csx->nnz = nnz;
        csx->contig = false;
        csx->buf = nullptr;
        csx->buf_size = 0;
        csx->ptr_vec = d_ptrs;
        csx->idx_vec = d_inds;
        csx->val = d_vals;
        return csx;
    }

    // header over the holder's current tile (:230-233)
    mmio::CSX<IT, VT> *form_mmiocsx(const IT nrows, const IT ncols, const IT nnz, mmio::MajorDim layout) {
        return form_mmiocsx(nrows, ncols, nnz, layout, d_vals_buf, d_inds_buf, d_ptrs_buf);
    }
};
