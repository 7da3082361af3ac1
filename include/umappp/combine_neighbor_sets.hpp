// Drop-in for umappp/combine_neighbor_sets.hpp (reference :15-100): fuzzy-set
// union / intersection / mix of the directed kNN weights, on the GPU.  Rows
// come back sorted by index, exactly as the reference leaves them.
#ifndef UMAPPP_AMD_COMBINE_NEIGHBOR_SETS_HPP
#define UMAPPP_AMD_COMBINE_NEIGHBOR_SETS_HPP

#include "NeighborList.hpp"
#include "detail_csr.hpp"

namespace umappp {

template <typename Index_, typename Float_>
void combine_neighbor_sets(NeighborList<Index_, Float_>& x, const Float_ mix_ratio) {
    auto c = detail::to_csr(x);
    const int64_t cap = 2 * c.row_ptr.back() + 1;
    std::vector<int64_t> rp(x.size() + 1);
    std::vector<int32_t> idx(static_cast<size_t>(cap));
    std::vector<Float_> val(static_cast<size_t>(cap));
    int64_t nnz = 0;
    detail::check(umx_combine_neighbor_sets(detail::dtype_of<Float_>(), static_cast<int64_t>(x.size()),
                                            c.row_ptr.data(), c.idx.data(), c.val.data(), mix_ratio, cap, rp.data(),
                                            idx.data(), val.data(), &nnz));
    detail::from_csr(rp.data(), idx.data(), val.data(), x.size(), x);
}

}  // namespace umappp

#endif
