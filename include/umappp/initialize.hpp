// Drop-in for umappp/initialize.hpp (reference :27-119): similarities, fuzzy
// union, spectral initialisation and schedule are built on the GPU; random
// initial coordinates, (a, b) and the epoch count are host steps as in the
// reference.
#ifndef UMAPPP_AMD_INITIALIZE_HPP
#define UMAPPP_AMD_INITIALIZE_HPP

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <vector>

#include "NeighborList.hpp"
#include "Options.hpp"
#include "Status.hpp"
#include "detail_csr.hpp"
#include "find_ab.hpp"

namespace umappp {

// 500 epochs up to 10,000 observations, else 200 + ceil(300 * 10^4 / N)
// (reference initialize.hpp:27-47).
template <typename Index_>
int choose_num_epochs(const std::optional<int> num_epochs, const Index_ size) {
    if (num_epochs.has_value()) return *num_epochs;
    constexpr double limit = 10000;
    if (static_cast<double>(size) <= limit) return 500;
    return 200 + static_cast<int>(std::ceil(300.0 * limit / static_cast<double>(size)));
}

// Uniform draws in [-scale, scale) from std::mt19937_64(seed): 53 random
// bits per draw scaled to [0, 1).  Same role as the reference's random_init
// (spectral_init.hpp:209-225); its aarand mapping is not reproduced.
template <typename Index_, typename Float_>
void random_init(const Index_ num_obs, const std::size_t num_dim, Float_* const vals,
                 const RngEngine::result_type seed, const double scale) {
    RngEngine rng(seed);
    const Float_ mult = static_cast<Float_>(scale * 2);
    const Float_ shift = static_cast<Float_>(scale);
    const std::size_t total = num_dim * static_cast<std::size_t>(num_obs);
    for (std::size_t i = 0; i < total; ++i) {
        const double u = static_cast<double>(rng() >> 11) * 0x1.0p-53;
        vals[i] = static_cast<Float_>(u) * mult - shift;
    }
}

template <typename Index_, typename Float_>
Status<Index_, Float_> initialize(NeighborList<Index_, Float_> x, const std::size_t num_dim, Float_* const embedding,
                                  Options options) {
    auto c = detail::to_csr(x);
    const int64_t n = static_cast<int64_t>(x.size());
    umx_graph* g = nullptr;
    detail::check(umx_graph_create(detail::dtype_of<Float_>(), n, c.row_ptr.data(), c.idx.data(), c.val.data(), 0, &g));
    std::unique_ptr<umx_graph, void (*)(umx_graph*)> graph(g, &umx_graph_destroy);
    detail::check(umx_graph_neighbor_similarities(g, options.local_connectivity, options.bandwidth, 1e-3, 1));
    detail::check(umx_graph_combine(g, options.mix_ratio));

    bool use_random = options.initialize_method == InitializeMethod::RANDOM;
    if (options.initialize_method == InitializeMethod::SPECTRAL) {
        // GPU spectral_init on the symmetrised graph (reference initialize.hpp:79-93)
        int ok = 0;
        detail::check(umx_graph_spectral_init(g, static_cast<int>(num_dim), options.initialize_spectral_scale,
                                              options.initialize_spectral_jitter ? 1 : 0,
                                              options.initialize_spectral_jitter_sd, options.initialize_seed,
                                              embedding, 0, &ok, nullptr));
        use_random = options.initialize_random_on_spectral_fail && !ok;
    }
    if (use_random) random_init<Index_, Float_>(static_cast<Index_>(n), num_dim, embedding, options.initialize_seed,
                                                options.initialize_random_scale);

    if (!options.a.has_value() || !options.b.has_value()) {
        const auto found = find_ab<double>(options.spread, options.min_dist);
        options.a = found.first;
        options.b = found.second;
    }
    options.num_epochs = choose_num_epochs<Index_>(options.num_epochs, static_cast<Index_>(n));

    umx_layout* L = nullptr;
    detail::check(umx_layout_create(g, *options.num_epochs, options.negative_sample_rate, &L));
    return Status<Index_, Float_>(L, std::move(options), num_dim);
}

// Data overload (reference initialize.hpp:165-177, which takes a
// knncolle::Builder): the exact options.num_neighbors nearest neighbours of
// every observation are found on the GPU (umx_knn_bruteforce), then as
// above.  data: num_obs observations of data_dim values each
// ("data_dim x num_obs" column-major, as in the reference).
template <typename Index_, typename Float_>
Status<Index_, Float_> initialize(const std::size_t data_dim, const Index_ num_obs, const Float_* const data,
                                  const std::size_t num_dim, Float_* const embedding, Options options) {
    const int k = options.num_neighbors;
    std::vector<int32_t> idx(static_cast<std::size_t>(num_obs) * static_cast<std::size_t>(k));
    std::vector<Float_> dist(idx.size());
    detail::check(umx_knn_bruteforce(detail::dtype_of<Float_>(), static_cast<int64_t>(num_obs),
                                     static_cast<int>(data_dim), data, k, idx.data(), dist.data(), 0));
    NeighborList<Index_, Float_> nl(static_cast<std::size_t>(num_obs));
    for (Index_ i = 0; i < num_obs; ++i) {
        auto& row = nl[static_cast<std::size_t>(i)];
        row.reserve(static_cast<std::size_t>(k));
        for (int q = 0; q < k; ++q) {
            const std::size_t at = static_cast<std::size_t>(i) * static_cast<std::size_t>(k) + static_cast<std::size_t>(q);
            row.emplace_back(static_cast<Index_>(idx[at]), dist[at]);
        }
    }
    return initialize<Index_, Float_>(std::move(nl), num_dim, embedding, std::move(options));
}

}  // namespace umappp

#endif
