// mmio_csx.hpp — the subset of dmmio's mmio::CSX that the KokkosWrap /
// TileHolder interfaces use, for callers that build without dmmio.
//
// dmmio is an un-vendored submodule of the reference (.gitmodules); the
// fields below are the ones the reference reads and writes
// (include/KokkosWrap.hpp:95-104, include/tile_holder.cuh:215-226,
// include/utils.cuh:455-466).  When dmmio is available its own header can
// be included instead, as long as the field names match.
#pragma once

#include <cstddef>
#include <cstdint>

namespace mmio {

enum class MajorDim : int32_t { ROWS = 0, COLS = 1 };

template <typename IT, typename VT>
struct CSX {
    MajorDim majordim;
    IT nrows;
    IT ncols;
    IT nnz;
    IT *ptr_vec;
    IT *idx_vec;
    VT *val;
    bool contig;
    char *buf;
    size_t buf_size;
};

template <typename IT, typename VT>
struct CSR {
    IT nrows;
    IT ncols;
    IT nnz;
    IT *row_ptr;
    IT *col_idx;
    VT *val;
};

// Bytes of the contiguous tile [val | idx | ptr] (tile_holder.cuh:80-89).
template <typename IT, typename VT>
inline size_t CSX_buf_size(IT nrows, IT ncols, IT nnz, MajorDim md) {
    const size_t P = (md == MajorDim::ROWS ? (size_t)nrows : (size_t)ncols) + 1;
    return (size_t)nnz * (sizeof(VT) + sizeof(IT)) + P * sizeof(IT);
}

}  // namespace mmio
