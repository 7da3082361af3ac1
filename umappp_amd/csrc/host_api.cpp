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

// Symmetric tridiagonal eigenproblem: diagonal d[0..k), off-diagonal
// e[0..k-1) (e[i] couples i and i + 1).  The implicit symmetric QR algorithm
// with Wilkinson shifts (Golub & Van Loan, Matrix Computations, 4th ed.,
// algorithms 8.3.2 and 8.3.3): off-diagonals below eps * (|d_i| + |d_i+1|)
// are set to zero, the trailing unreduced block [lo, hi] gets one shifted QR
// step, carried out as a chase of the bulge by plane rotations, and every
// rotation is applied to the columns of z.  Eigenvalues into d (unsorted),
// eigenvectors into the columns of z (k x k, row-major, z[r * k + c]).
// Returns false if 30 * k steps do not diagonalise the matrix.
namespace {

// One implicit QR step with a Wilkinson shift on the unreduced block
// [lo, hi] (hi > lo).
void qr_step(int k, int lo, int hi, std::vector<double>& d, std::vector<double>& e, std::vector<double>& z) {
    // shift: the eigenvalue of the trailing 2 x 2 block nearer to d[hi]
    const double half = (d[hi - 1] - d[hi]) / 2;
    const double off = e[hi - 1];
    const double root = std::hypot(half, off);
    const double mu = d[hi] - off * off / (half + (half >= 0 ? root : -root));
    double x = d[lo] - mu, bulge = e[lo];
    for (int q = lo; q < hi; ++q) {
        // rotation in plane (q, q + 1) taking (x, bulge) to (r, 0)
        const double r = std::hypot(x, bulge);
        const double c = r > 0 ? x / r : 1.0, s = r > 0 ? -bulge / r : 0.0;
        if (q > lo) e[q - 1] = r;
        const double a = d[q], b = d[q + 1], f = e[q];
        d[q] = c * c * a - 2 * c * s * f + s * s * b;
        d[q + 1] = s * s * a + 2 * c * s * f + c * c * b;
        e[q] = c * s * (a - b) + (c * c - s * s) * f;
        if (q + 1 < hi) {
            const double g = e[q + 1];
            bulge = -s * g;
            e[q + 1] = c * g;
            x = e[q];
        }
        for (int row = 0; row < k; ++row) {
            double* zr = &z[static_cast<size_t>(row) * k];
            const double zq = zr[q], zq1 = zr[q + 1];
            zr[q] = c * zq - s * zq1;
            zr[q + 1] = s * zq + c * zq1;
        }
    }
}

}  // namespace

bool tridiagonal_eigen(int k, std::vector<double>& d, const std::vector<double>& e_in, std::vector<double>& z) {
    z.assign(static_cast<size_t>(k) * k, 0.0);
    for (int r = 0; r < k; ++r) z[static_cast<size_t>(r) * k + r] = 1.0;
    if (k < 2) return true;
    std::vector<double> e(e_in.begin(), e_in.begin() + (k - 1));
    const double eps = std::numeric_limits<double>::epsilon();
    for (int steps = 0;; ++steps) {
        for (int q = 0; q + 1 < k; ++q)
            if (std::fabs(e[q]) <= eps * (std::fabs(d[q]) + std::fabs(d[q + 1]))) e[q] = 0.0;
        int hi = k - 1;
        while (hi > 0 && e[hi - 1] == 0.0) --hi;
        if (hi == 0) return true;
        int lo = hi - 1;
        while (lo > 0 && e[lo - 1] != 0.0) --lo;
        if (steps >= 30 * k) return false;
        qr_step(k, lo, hi, d, e, z);
    }
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
