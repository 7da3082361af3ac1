// Drop-in for umappp/Status.hpp (reference :25-142).  The schedule and the
// SGD epochs live on the GPU (a umx_layout handle); run() reads the caller's
// embedding, runs the epochs there and writes the result back, as the
// reference's run() does on the host.  Copies are deep, as the reference's
// (its members are values): a copy owns a clone of the device state
// (umx_layout_clone) and runs independently of the original.
#ifndef UMAPPP_AMD_STATUS_HPP
#define UMAPPP_AMD_STATUS_HPP

#include <cstddef>
#include <memory>

#include "Options.hpp"
#include "detail_csr.hpp"
#include "optimize_layout.hpp"

namespace umappp {

template <typename Index_, typename Float_>
class Status {
public:
    // Takes ownership of a layout handle whose schedule is already built.
    Status(umx_layout* layout, Options options, const std::size_t num_dim)
        : my_layout(layout, &umx_layout_destroy), my_options(std::move(options)), my_num_dim(num_dim) {
        detail::check(umx_layout_set_params(my_layout.get(), static_cast<int>(num_dim), *my_options.a,
                                            *my_options.b, my_options.repulsion_strength, my_options.learning_rate,
                                            my_options.optimize_seed));
    }

    Status(const Status& other)
        : my_layout(clone(other.my_layout.get()), &umx_layout_destroy),
          my_options(other.my_options),
          my_num_dim(other.my_num_dim) {}
    Status& operator=(const Status& other) {
        if (this != &other) {
            Status copy(other);
            *this = std::move(copy);
        }
        return *this;
    }
    Status(Status&&) = default;
    Status& operator=(Status&&) = default;

    // Host snapshot of the GPU schedule (reference: testing only).
    EpochData<Index_, Float_> get_epoch_data() const {
        return detail::export_epochs<Index_, Float_>(my_layout.get(), static_cast<Float_>(my_options.negative_sample_rate));
    }

    std::size_t num_dimensions() const { return my_num_dim; }

    int epoch() const {
        int cur = 0;
        detail::check(umx_layout_info(my_layout.get(), nullptr, nullptr, nullptr, &cur));
        return cur;
    }

    int num_epochs() const {
        int tot = 0;
        detail::check(umx_layout_info(my_layout.get(), nullptr, nullptr, &tot, nullptr));
        return tot;
    }

    Index_ num_observations() const {
        int64_t n = 0;
        detail::check(umx_layout_info(my_layout.get(), &n, nullptr, nullptr, nullptr));
        return static_cast<Index_>(n);
    }

    void run(Float_* const embedding, int epoch_limit) {
        detail::check(umx_layout_set_embedding(my_layout.get(), embedding, 0));
        detail::check(umx_layout_run(my_layout.get(), epoch_limit));
        detail::check(umx_layout_get_embedding(my_layout.get(), embedding, 0));
    }

    void run(Float_* const embedding) { run(embedding, num_epochs()); }

private:
    static umx_layout* clone(const umx_layout* layout) {
        umx_layout* copy = nullptr;
        detail::check(umx_layout_clone(layout, &copy));
        return copy;
    }

    std::unique_ptr<umx_layout, void (*)(umx_layout*)> my_layout;
    Options my_options;
    std::size_t my_num_dim;
};

}  // namespace umappp

#endif
