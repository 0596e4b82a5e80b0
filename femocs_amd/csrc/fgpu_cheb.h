// Chebyshev-Jacobi preconditioner of the FP64 CG: the Lanczos estimate of lambda_max(D^-1 A)
// from recorded CG coefficients, the polynomial's recurrence coefficients and its degree.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "fgpu_internal.h"

namespace fgpu {

// largest eigenvalue of the symmetric tridiagonal matrix (diagonal a, off-diagonal b) by
// bisection on the Sturm count
inline double tridiag_max_eig(const std::vector<double> &a, const std::vector<double> &b) {
    const int m = (int)a.size();
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < m; ++i) {
        const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    auto count_below = [&](double x) {   // eigenvalues < x
        int cnt = 0;
        double q = 1.0;
        for (int i = 0; i < m; ++i) {
            q = a[i] - x - (i > 0 ? b[i - 1] * b[i - 1] / q : 0.0);
            if (q == 0.0) q = -1e-300;
            if (q < 0) ++cnt;
        }
        return cnt;
    };
    for (int it = 0; it < 200 && hi - lo > 1e-12 * std::max(1.0, std::fabs(hi)); ++it) {
        const double mid = 0.5 * (lo + hi);
        if (count_below(mid) == m) hi = mid;
        else lo = mid;
    }
    return hi;
}

// Chebyshev coefficients for the interval [lmax / kChebRatio, lmax] of D^-1 A: 1/theta, then
// per step j = 1..deg-1 the three-term recurrence d_j = c1 d_{j-1} + c2 D^-1 res_j with
// c1 = rho_j rho_{j-1}, c2 = 2 rho_j / delta, rho_0 = delta / theta, rho_j = 1/(2 theta/delta - rho_{j-1})
constexpr double kChebRatio = 100.0;
constexpr int kLanczosSteps = 16;
inline void cheb_coefficients(double lmax, int deg, std::vector<double> &c) {
    const double lmin = lmax / kChebRatio;
    const double th = 0.5 * (lmax + lmin), de = 0.5 * (lmax - lmin), s1 = th / de;
    c.assign(2 * kChebMax + 1, 0.0);
    c[0] = 1.0 / th;
    double rho = 1.0 / s1;
    for (int j = 1; j < deg; ++j) {
        const double rn = 1.0 / (2.0 * s1 - rho);
        c[2 * j - 1] = rn * rho;
        c[2 * j] = 2.0 * rn / de;
        rho = rn;
    }
}

// Chebyshev degree for a solve whose largest system has n dofs: a CG step costs about deg
// operator applies (each about 125 bytes per dof in FP64) plus a fixed cost per step (the
// scalar kernels and reductions, about 25 us); the degree balances both
inline int cheb_auto_degree(int n) {
    const double t_apply = 125.0 * n / 4e12;
    const int k = (int)std::lround(std::sqrt(50.0 * 25e-6 / std::max(t_apply, 1e-9)));
    return std::max(2, std::min(8, k));
}

// lambda_max of D^-1 A from the first m CG steps' alpha_j, beta_j (h: 2 * kHist doubles,
// alpha_j at 2j, beta_j at 2j + 1): the largest eigenvalue of the Lanczos matrix, times 1.1
inline double lanczos_lmax(const double *h, int m) {
    m = std::min(m, kHist);
    std::vector<double> a(m), b(std::max(0, m - 1));
    for (int j = 0; j < m; ++j) {
        a[j] = 1.0 / h[2 * j] + (j > 0 ? h[2 * j - 1] / h[2 * (j - 1)] : 0.0);
        if (j + 1 < m) b[j] = std::sqrt(std::max(0.0, h[2 * j + 1])) / h[2 * j];
    }
    return 1.1 * tridiag_max_eig(a, b);
}

}  // namespace fgpu
