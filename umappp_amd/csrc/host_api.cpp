// Host-only entry points of the C ABI: find_ab and choose_num_epochs.
#include <cstdint>
#include <optional>

#include "umappp_amd.h"
#include "umappp/find_ab.hpp"
#include "umappp/initialize.hpp"

extern "C" {

int umx_find_ab(double spread, double min_dist, double* a, double* b) {
    if (!a || !b) return UMX_ERR_ARGUMENT;
    const auto ab = umappp::find_ab<double>(spread, min_dist);
    *a = ab.first;
    *b = ab.second;
    return UMX_OK;
}

int umx_choose_num_epochs(int num_epochs, int64_t size) {
    return umappp::choose_num_epochs<int64_t>(num_epochs >= 0 ? std::optional<int>(num_epochs) : std::nullopt, size);
}

}
