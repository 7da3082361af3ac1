// Drop-in for the EpochData / similarities_to_epochs part of
// umappp/optimize_layout.hpp (reference :27-92).  The SGD epochs themselves
// run on the GPU behind Status::run(); EpochData here is a host snapshot of
// the GPU schedule.
#ifndef UMAPPP_AMD_OPTIMIZE_LAYOUT_HPP
#define UMAPPP_AMD_OPTIMIZE_LAYOUT_HPP

#include <cstddef>
#include <vector>

#include "NeighborList.hpp"
#include "detail_csr.hpp"

namespace umappp {

template <typename Index_, typename Float_>
struct EpochData {
    EpochData() = default;
    explicit EpochData(const Index_ nobs) : cumulative_num_edges(static_cast<size_t>(nobs) + 1) {}

    int total_epochs = 0;
    int current_epoch = 0;
    std::vector<std::size_t> cumulative_num_edges;
    std::vector<Index_> edge_targets;
    std::vector<Float_> epochs_per_sample;
    std::vector<Float_> epoch_of_next_sample;
    std::vector<Float_> epoch_of_next_negative_sample;
    Float_ negative_sample_rate = 5;
};

namespace detail {

template <typename Index_, typename Float_>
EpochData<Index_, Float_> export_epochs(const umx_layout* L, Float_ nsr) {
    int64_t n = 0, nnz = 0;
    int total = 0, current = 0;
    check(umx_layout_info(L, &n, &nnz, &total, &current));
    EpochData<Index_, Float_> out(static_cast<Index_>(n));
    std::vector<int64_t> rp(static_cast<size_t>(n) + 1);
    std::vector<int32_t> tg(static_cast<size_t>(nnz));
    out.epochs_per_sample.resize(static_cast<size_t>(nnz));
    out.epoch_of_next_sample.resize(static_cast<size_t>(nnz));
    out.epoch_of_next_negative_sample.resize(static_cast<size_t>(nnz));
    check(umx_layout_export_epochs(L, rp.data(), tg.data(), out.epochs_per_sample.data(),
                                   out.epoch_of_next_sample.data(), out.epoch_of_next_negative_sample.data()));
    for (size_t i = 0; i < rp.size(); ++i) out.cumulative_num_edges[i] = static_cast<std::size_t>(rp[i]);
    out.edge_targets.assign(tg.begin(), tg.end());
    out.total_epochs = total;
    out.current_epoch = current;
    out.negative_sample_rate = nsr;
    return out;
}

}  // namespace detail

template <typename Index_, typename Float_>
EpochData<Index_, Float_> similarities_to_epochs(const NeighborList<Index_, Float_>& p, const int num_epochs,
                                                 const Float_ negative_sample_rate) {
    auto c = detail::to_csr(p);
    umx_layout* L = nullptr;
    detail::check(umx_layout_create_csr(detail::dtype_of<Float_>(), static_cast<int64_t>(p.size()),
                                        c.row_ptr.data(), c.idx.data(), c.val.data(), 0, num_epochs,
                                        negative_sample_rate, &L));
    try {
        auto out = detail::export_epochs<Index_, Float_>(L, negative_sample_rate);
        umx_layout_destroy(L);
        return out;
    } catch (...) {
        umx_layout_destroy(L);
        throw;
    }
}

}  // namespace umappp

#endif
