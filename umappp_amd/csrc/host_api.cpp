// Host-only code of the C ABI: find_ab, choose_num_epochs and the symmetric
// tridiagonal eigen-solver of the spectral initialisation.
#include <cmath>
#include <cstdint>
#include <limits>
#include <optional>
#include <vector>

#include "umappp_amd.h"
#include "umappp/find_ab.hpp"
#include "umappp/initialize.hpp"

namespace umx {

// Symmetric tridiagonal eigenproblem (diagonal d[0..k), off-diagonal
// e[0..k-1)) by the QL algorithm with implicit Wilkinson-type shifts and
// plane rotations accumulated into z (as in LAPACK's xSTEQR / the EISPACK
// tql2 routine); eigenvalues into d, eigenvectors into the columns of z
// (k x k, row-major, z[r * k + c]).  Returns false if an eigenvalue needs more
// than 60 iterations.
bool tridiagonal_eigen(int k, std::vector<double>& d, const std::vector<double>& e_in, std::vector<double>& z) {
    std::vector<double> e(e_in);
    z.assign(static_cast<size_t>(k) * k, 0.0);
    for (int r = 0; r < k; ++r) z[static_cast<size_t>(r) * k + r] = 1.0;
    e.resize(static_cast<size_t>(k), 0.0);
    if (k > 0) e[k - 1] = 0.0;
    for (int l = 0; l < k; ++l) {
        int iter = 0;
        int m;
        do {
            for (m = l; m < k - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= std::numeric_limits<double>::epsilon() * dd) break;
            }
            if (m != l) {
                if (++iter > 60) return false;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    for (int q = 0; q < k; ++q) {
                        f = z[static_cast<size_t>(q) * k + i + 1];
                        z[static_cast<size_t>(q) * k + i + 1] = s * z[static_cast<size_t>(q) * k + i] + c * f;
                        z[static_cast<size_t>(q) * k + i] = c * z[static_cast<size_t>(q) * k + i] - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    return true;
}


}  // namespace umx

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

int umx_tridiagonal_eigen(int k, double* d, const double* e, double* z) {
    if (k < 0 || (k > 0 && (!d || !z || (k > 1 && !e)))) return UMX_ERR_ARGUMENT;
    std::vector<double> dd(d, d + k), ee(e ? e : d, e ? e + (k > 0 ? k - 1 : 0) : d), zz;
    if (!umx::tridiagonal_eigen(k, dd, ee, zz)) return UMX_ERR_STATE;
    for (int q = 0; q < k; ++q) d[q] = dd[q];
    for (size_t q = 0; q < zz.size(); ++q) z[q] = zz[q];
    return UMX_OK;
}

}
