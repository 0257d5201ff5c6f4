// FP64 special functions of the SHORE basis (device side), shared by the
// shared-design set-up and the per-voxel (gradient-deviation) design builds.
#pragma once

#include <hip/hip_runtime.h>

namespace shorefit {

constexpr double kPiD = 3.14159265358979323846264338327950288;

__device__ inline double factorial_d(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= double(i);
    return f;
}

// Gamma(n + 3/2) = (2n+1)!! / 2^(n+1) * sqrt(pi), exact product form.
__device__ inline double gamma_n_three_halves(int n) {
    double g = sqrt(kPiD) * 0.5;  // Gamma(3/2)
    for (int i = 1; i <= n; ++i) g *= (double(i) + 0.5);
    return g;
}

// Generalised Laguerre L_n^a(x) by its three-term recurrence.
__device__ inline double laguerre_d(int n, double a, double x) {
    double lm1 = 1.0;
    if (n == 0) return lm1;
    double l = 1.0 + a - x;
    for (int k = 1; k < n; ++k) {
        const double lp1 = ((2.0 * k + 1.0 + a - x) * l - (k + a) * lm1) / (k + 1.0);
        lm1 = l;
        l = lp1;
    }
    return l;
}

// Associated Legendre P_l^m(x), Condon-Shortley phase, upward recurrence.
__device__ inline double legendre_d(int l, int m, double x) {
    double pmm = 1.0;
    if (m > 0) {
        const double s = sqrt(fmax((1.0 - x) * (1.0 + x), 0.0));
        double f = 1.0;
        for (int i = 0; i < m; ++i) {
            pmm *= -f * s;
            f += 2.0;
        }
    }
    if (l == m) return pmm;
    double p1 = x * (2.0 * m + 1.0) * pmm;
    if (l == m + 1) return p1;
    double pl = 0.0;
    for (int ll = m + 2; ll <= l; ++ll) {
        pl = ((2.0 * ll - 1.0) * x * p1 - (ll + m - 1.0) * pmm) / double(ll - m);
        pmm = p1;
        p1 = pl;
    }
    return pl;
}

}  // namespace shorefit
