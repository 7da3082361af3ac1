// Drop-in for umappp/find_ab.hpp (host code; no GPU work).
//
// The problem is the reference's (find_ab.hpp:44-168): the least-squares fit
// of the curve 1 / (1 + a x^(2b)) to the target pmin(1, exp(-(x - min_dist) /
// spread)) over the 300 grid points x_g = g * 3 * spread / 300, g = 1..300.
//
// The method is this project's own.  With u = log a and l = log x the curve
// is a logistic function of a linear predictor,
//   f(x) = 1 / (1 + exp(eta)),   eta = u + 2 b l,
// so (u, b) is fitted as a logistic-curve least-squares problem:
//   start:  ordinary least squares of the target's logit log(1/y - 1) on
//           (1, 2 l) over the grid points with 0 < y < 1 (where the logit is
//           finite); a logistic curve fits its own logits exactly, so this
//           lands near the optimum;
//   step:   Gauss-Newton in (u, b): the residual's Jacobian is
//           f (1 - f) (1, 2 l); the 2 x 2 normal equations are solved
//           directly and the step is halved until the sum of squares falls;
//   stop:   when no halving of the step lowers the sum of squares, when the
//           relative decrease is at most 1e-14, or after 100 steps.
// Working in log a keeps a positive without a constraint.  When the target
// is 1 at every grid point (min_dist >= 3 * spread) the sum of squares is 0
// exactly at a = 0, which is returned (with b = 1).
#ifndef UMAPPP_AMD_FIND_AB_HPP
#define UMAPPP_AMD_FIND_AB_HPP

#include <cmath>
#include <cstddef>
#include <utility>
#include <vector>

namespace umappp {

template <typename Float_>
std::pair<Float_, Float_> find_ab(const Float_ spread, const Float_ min_dist) {
    constexpr std::size_t npts = 300;
    const double sp = static_cast<double>(spread), md = static_cast<double>(min_dist);
    std::vector<double> l(npts), y(npts);
    for (std::size_t g = 0; g < npts; ++g) {
        const double x = static_cast<double>(g + 1) * (sp * 3 / npts);
        l[g] = std::log(x);
        y[g] = x <= md ? 1.0 : std::exp(-(x - md) / sp);
    }

    // Start: least squares of logit(y) = u + 2 b l over the points with a
    // finite logit (2 x 2 normal equations in (u, b)).
    double s0 = 0, s1 = 0, s2 = 0, t0 = 0, t1 = 0;
    for (std::size_t g = 0; g < npts; ++g) {
        if (!(y[g] > 0 && y[g] < 1)) continue;
        const double z = std::log(1 / y[g] - 1), w = 2 * l[g];
        s0 += 1;
        s1 += w;
        s2 += w * w;
        t0 += z;
        t1 += w * z;
    }
    if (s0 == 0) return std::make_pair(static_cast<Float_>(0), static_cast<Float_>(1));
    double u = 0, b = 1;
    const double det0 = s0 * s2 - s1 * s1;
    if (s0 >= 2 && det0 > 0) {
        u = (s2 * t0 - s1 * t1) / det0;
        b = (s0 * t1 - s1 * t0) / det0;
    }

    std::vector<double> f(npts), f_try(npts);
    auto sum_sq = [&](double uu, double bb, std::vector<double>& fit) {
        double ss = 0;
        for (std::size_t g = 0; g < npts; ++g) {
            fit[g] = 1 / (1 + std::exp(uu + 2 * bb * l[g]));
            const double r = y[g] - fit[g];
            ss += r * r;
        }
        return ss;
    };
    double ss = sum_sq(u, b, f);

    for (int it = 0; it < 100; ++it) {
        // Gauss-Newton: residual r = y - f, dr/d(u, b) = f (1 - f) (1, 2 l).
        double h00 = 0, h01 = 0, h11 = 0, g0 = 0, g1 = 0;
        for (std::size_t g = 0; g < npts; ++g) {
            const double w = f[g] * (1 - f[g]);
            const double j0 = w, j1 = 2 * l[g] * w;
            const double r = y[g] - f[g];
            h00 += j0 * j0;
            h01 += j0 * j1;
            h11 += j1 * j1;
            g0 += j0 * r;
            g1 += j1 * r;
        }
        const double det = h00 * h11 - h01 * h01;
        if (!(det > 0)) break;
        const double du = -(h11 * g0 - h01 * g1) / det;
        const double db = -(h00 * g1 - h01 * g0) / det;
        double t = 1, ss_try = ss;
        bool lower = false;
        for (int h = 0; h < 60; ++h, t /= 2) {
            ss_try = sum_sq(u + t * du, b + t * db, f_try);
            if (ss_try < ss) {
                lower = true;
                break;
            }
        }
        if (!lower) break;
        u += t * du;
        b += t * db;
        const double drop = ss - ss_try;
        ss = ss_try;
        f.swap(f_try);
        if (drop <= 1e-14 * ss) break;
    }
    return std::make_pair(static_cast<Float_>(std::exp(u)), static_cast<Float_>(b));
}

}  // namespace umappp

#endif
