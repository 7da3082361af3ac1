// Drop-in for umappp/neighbor_similarities.hpp (reference :38-207): distances
// become fuzzy-set weights in place, computed on the GPU in double precision.
#ifndef UMAPPP_AMD_NEIGHBOR_SIMILARITIES_HPP
#define UMAPPP_AMD_NEIGHBOR_SIMILARITIES_HPP

#include "NeighborList.hpp"
#include "detail_csr.hpp"

namespace umappp {

template <typename Float_>
struct NeighborSimilaritiesOptions {
    Float_ local_connectivity = 1.0;
    Float_ bandwidth = 1.0;
    Float_ min_k_dist_scale = 1e-3;
    int num_threads = 1;  // accepted for compatibility; the GPU ignores it
};

template <bool use_newton_ = true, typename Index_, typename Float_>
void neighbor_similarities(NeighborList<Index_, Float_>& x, const NeighborSimilaritiesOptions<Float_>& options) {
    auto c = detail::to_csr(x);
    detail::check(umx_neighbor_similarities(detail::dtype_of<Float_>(), static_cast<int64_t>(x.size()),
                                            c.row_ptr.data(), c.val.data(), options.local_connectivity,
                                            options.bandwidth, options.min_k_dist_scale, use_newton_ ? 1 : 0,
                                            nullptr, nullptr));
    size_t q = 0;
    for (auto& row : x)
        for (auto& p : row) p.second = c.val[q++];
}

}  // namespace umappp

#endif
