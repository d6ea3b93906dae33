// KokkosWrapTriple.hpp — COO triple used for sorting coordinate data
// (same layout as the reference's include/KokkosWrapTriple.hpp:5-10).
#pragma once

namespace KokkosWrap {

template <typename IT, typename VT>
struct Triple {
    IT row;
    IT col;
    VT val;
};

}  // namespace KokkosWrap
