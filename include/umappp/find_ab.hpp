// Drop-in for umappp/find_ab.hpp (host code; no GPU work).
//
// Fits y ~ 1 / (1 + a * x^(2b)) to the target curve pmin(1, exp(-(x - min_dist)
// / spread)) at 300 grid points x_g = g * 3 * spread / 300, g = 1..300, by
// Gauss-Newton with Levenberg-Marquardt damping, as umappp does
// (find_ab.hpp:44-168 of the reference):
//   start:  x_half = -spread * log(0.5) + min_dist, slope d_half = -0.5 / spread,
//           b0 = -d_half * x_half / (1/0.5 - 1) / (2 * 0.5^2), a0 = (1/0.5 - 1) / x_half^(2 b0)
//   step:   solve the damped 2x2 normal equations (J^T J with its diagonal
//           scaled by 1 + damping) for (delta_a, delta_b); accept if the sum
//           of squares falls, halving the damping; otherwise grow the damping
//           (0 -> machine epsilon -> doubling) until it reaches 1024, then stop.
//   stop:   at 50 accepted iterations or when the relative decrease of the
//           sum of squares is <= 1e-6.
#ifndef UMAPPP_AMD_FIND_AB_HPP
#define UMAPPP_AMD_FIND_AB_HPP

#include <cmath>
#include <cstddef>
#include <limits>
#include <utility>
#include <vector>

namespace umappp {

template <typename Float_>
std::pair<Float_, Float_> find_ab(const Float_ spread, const Float_ min_dist) {
    constexpr std::size_t npts = 300;
    std::vector<Float_> xs(npts), logx(npts), target(npts);
    const Float_ step = spread * 3 / npts;
    for (std::size_t g = 0; g < npts; ++g) {
        xs[g] = (g + 1) * step;
        logx[g] = std::log(xs[g]);
        target[g] = xs[g] <= min_dist ? 1 : std::exp(-(xs[g] - min_dist) / spread);
    }

    constexpr Float_ half = 0.5;
    const Float_ x_half = std::log(half) * -spread + min_dist;
    const Float_ d_half = half / -spread;
    Float_ b = -d_half * x_half / (1 / half - 1) / (2 * half * half);
    Float_ a = (1 / half - 1) / std::pow(x_half, 2 * b);

    std::vector<Float_> powx(npts), fitted(npts), resid(npts);
    auto sum_sq = [&](Float_ A, Float_ B) {
        Float_ ss = 0;
        for (std::size_t g = 0; g < npts; ++g) {
            powx[g] = std::pow(xs[g], 2 * B);
            fitted[g] = 1 / (1 + A * powx[g]);
            resid[g] = target[g] - fitted[g];
            ss += resid[g] * resid[g];
        }
        return ss;
    };
    Float_ ss = sum_sq(a, b);
    Float_ damping = 0;

    for (int it = 0; it < 50; ++it) {
        // Normal equations from the residuals of the last sum_sq() call.
        Float_ jaa = 0, jbb = 0, jab = 0, ra = 0, rb = 0;
        for (std::size_t g = 0; g < npts; ++g) {
            const Float_ fy = fitted[g];
            const Float_ da = powx[g] * fy * fy;
            const Float_ db = a * (logx[g] * 2) * da;
            jaa += da * da;
            jbb += db * db;
            jab += da * db;
            ra += da * resid[g];
            rb += db * resid[g];
        }

        bool accepted = false;
        Float_ next_a = a, next_b = b, next_ss = ss;
        while (damping < 1024) {
            const Float_ scale = 1 + damping;
            const Float_ daa = jaa * scale, dbb = jbb * scale;
            const Float_ det = daa * dbb - jab * jab;
            const Float_ step_a = -(ra * dbb - jab * rb) / det;
            const Float_ step_b = -(-ra * jab + daa * rb) / det;
            next_a = a + step_a;
            next_b = b + step_b;
            next_ss = sum_sq(next_a, next_b);
            if (next_ss < ss) {
                accepted = true;
                damping /= 2;
                break;
            }
            damping = damping == 0 ? std::numeric_limits<Float_>::epsilon() : damping * 2;
        }
        if (!accepted) break;
        if (ss - next_ss <= ss * static_cast<Float_>(1e-6)) break;
        a = next_a;
        b = next_b;
        ss = next_ss;
    }
    return std::make_pair(a, b);
}

}  // namespace umappp

#endif
