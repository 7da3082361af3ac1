// Conversions between the NeighborList container and the CSR arrays of the
// C ABI, plus the status-to-exception bridge used by the drop-in headers.
#ifndef UMAPPP_AMD_DETAIL_CSR_HPP
#define UMAPPP_AMD_DETAIL_CSR_HPP

#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../umappp_amd.h"
#include "NeighborList.hpp"

namespace umappp {
namespace detail {

template <typename Float_>
constexpr int dtype_of() {
    static_assert(std::is_same<Float_, float>::value || std::is_same<Float_, double>::value,
                  "Float_ must be float or double");
    return std::is_same<Float_, float>::value ? UMX_F32 : UMX_F64;
}

inline void check(int rc) {
    if (rc == UMX_OK) return;
    const std::string msg = umx_last_error();
    if (rc == UMX_ERR_ARGUMENT || rc == UMX_ERR_GRAPH) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

template <typename Index_, typename Float_>
struct Csr {
    std::vector<int64_t> row_ptr;
    std::vector<int32_t> idx;
    std::vector<Float_> val;
};

template <typename Index_, typename Float_>
Csr<Index_, Float_> to_csr(const NeighborList<Index_, Float_>& x) {
    Csr<Index_, Float_> c;
    c.row_ptr.resize(x.size() + 1);
    c.row_ptr[0] = 0;
    for (size_t i = 0; i < x.size(); ++i) c.row_ptr[i + 1] = c.row_ptr[i] + static_cast<int64_t>(x[i].size());
    c.idx.reserve(static_cast<size_t>(c.row_ptr.back()));
    c.val.reserve(static_cast<size_t>(c.row_ptr.back()));
    for (const auto& row : x) {
        for (const auto& p : row) {
            if (static_cast<int64_t>(p.first) < 0 || static_cast<int64_t>(p.first) > std::numeric_limits<int32_t>::max())
                throw std::invalid_argument("neighbour index does not fit a 32-bit index");
            c.idx.push_back(static_cast<int32_t>(p.first));
            c.val.push_back(p.second);
        }
    }
    return c;
}

template <typename Index_, typename Float_>
void from_csr(const int64_t* rp, const int32_t* idx, const Float_* val, size_t n, NeighborList<Index_, Float_>& x) {
    x.resize(n);
    for (size_t i = 0; i < n; ++i) {
        auto& row = x[i];
        row.clear();
        for (int64_t e = rp[i]; e < rp[i + 1]; ++e) row.emplace_back(static_cast<Index_>(idx[e]), val[e]);
    }
}

}  // namespace detail
}  // namespace umappp

#endif
