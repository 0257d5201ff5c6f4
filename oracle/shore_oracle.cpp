// SHORE-fit oracle: an FP64 CPU restatement of cs_dmri's per-voxel 3D-SHORE
// fit path (PennLINC/cs_dmri 0.3.1).
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or
// calls this file; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, and only as the checker / CPU baseline.
//
// Every function names the reference statement it restates (file:line in the
// reference repository). Arithmetic is FP64 throughout, inputs f32 volumes,
// outputs f32 maps, exactly as src/fit.rs does.
//
// Parity pins (see tests/test_oracle_*.py):
//   math KATs ............ src/math.rs:168-247 + scipy.special closed forms
//   design matrix ........ src/basis/shore.rs:205-273 KATs + scipy golden
//   FISTA ................ 8 sklearn fixtures (tests/data/lasso_fixtures),
//                          KKT, limits (tests/optimization_{kkt,limits}.rs)
//   Lipschitz ............ numpy.linalg.svd
//   Tikhonov ............. normal-equation residual + scipy cho_solve
//   alpha machinery ...... closed forms (alpha_max zero solution, log path
//                          endpoints, BIC formula) and brute-force argmin
//   volume driver ........ per-voxel equality with the single-voxel entry
//                          points, R^2 closed form, scatter layout
//   graddev table ........ src/graddev.rs:218-253 KATs
//   post-fit ............. src/odf.rs KATs (scipy radial factors, index
//                          examples, projection vs direct SH evaluation),
//                          scipy hyp2f1 / binom, quadrature of the SHORE
//                          signal for RTOP, QIV, RTAP, RTPP and a finite-
//                          difference Laplacian for MSD; NG closed cases;
//                          src/io/microstructure.rs rejection KATs
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

// ---------------------------------------------------------------- math ----

// src/math.rs:8-37 — exact table up to 20!, gamma beyond.
double gamma_fn(double x);
double factorial(unsigned n) {
    static const double table[21] = {1.0,
                                     1.0,
                                     2.0,
                                     6.0,
                                     24.0,
                                     120.0,
                                     720.0,
                                     5040.0,
                                     40320.0,
                                     362880.0,
                                     3628800.0,
                                     39916800.0,
                                     479001600.0,
                                     6227020800.0,
                                     87178291200.0,
                                     1307674368000.0,
                                     20922789888000.0,
                                     355687428096000.0,
                                     6402373705728000.0,
                                     121645100408832000.0,
                                     2432902008176640000.0};
    if (n < 21) return table[n];
    return gamma_fn(double(n) + 1.0);
}

// src/math.rs:40-64 — Lanczos approximation, g = 7, nine terms (the published
// Godfrey coefficient set), with the reflection formula below 1/2.
double gamma_fn(double x) {
    if (x < 0.5) return kPi / (std::sin(kPi * x) * gamma_fn(1.0 - x));
    static const double c[9] = {0.99999999999980993,  676.5203681218851,     -1259.1392167224028,
                                771.32342877765313,   -176.61502916214059,   12.507343278686905,
                                -0.13857109526572012, 9.9843695780195716e-6, 1.5056327351493116e-7};
    const double g = 7.0;
    x -= 1.0;
    double a = c[0];
    const double t = x + g + 0.5;
    for (int i = 1; i < 9; ++i) a += c[i] / (x + double(i));
    return std::sqrt(2.0 * kPi) * std::pow(t, x + 0.5) * std::exp(-t) * a;
}

// src/math.rs:70-101 — P_l^m(x) with the Condon-Shortley phase, upward
// recurrence in l from P_m^m.
double assoc_legendre(unsigned l, unsigned m, double x) {
    double pmm = 1.0;
    if (m > 0) {
        const double somx2 = std::sqrt(std::max((1.0 - x) * (1.0 + x), 0.0));
        double fact = 1.0;
        for (unsigned i = 0; i < m; ++i) {
            pmm *= -fact * somx2;
            fact += 2.0;
        }
    }
    if (l == m) return pmm;
    double pmmp1 = x * (2.0 * double(m) + 1.0) * pmm;
    if (l == m + 1) return pmmp1;
    double pll = 0.0;
    for (unsigned ll = m + 2; ll <= l; ++ll) {
        const double lf = double(ll), mf = double(m);
        pll = ((2.0 * lf - 1.0) * x * pmmp1 - (lf + mf - 1.0) * pmm) / (lf - mf);
        pmm = pmmp1;
        pmmp1 = pll;
    }
    return pll;
}

// src/math.rs:108-125 — L_n^alpha(x) by the three-term recurrence.
double gen_laguerre(unsigned n, double alpha, double x) {
    if (n == 0) return 1.0;
    double lkm1 = 1.0;
    double lk = 1.0 + alpha - x;
    if (n == 1) return lk;
    for (unsigned k = 1; k < n; ++k) {
        const double kf = double(k);
        const double lkp1 = ((2.0 * kf + 1.0 + alpha - x) * lk - (kf + alpha) * lkm1) / (kf + 1.0);
        lkm1 = lk;
        lk = lkp1;
    }
    return lk;
}

// src/math.rs:157-166 — (r, theta = azimuth, phi = polar).
void cart2sphere(double x, double y, double z, double* r, double* theta, double* phi) {
    *r = std::sqrt(x * x + y * y + z * z);
    *theta = std::atan2(y, x);
    *phi = (*r > 0.0) ? std::acos(std::min(1.0, std::max(-1.0, z / *r))) : 0.0;
}

// -------------------------------------------------------------- basis ----

struct Index3 {
    unsigned n, ell;
    int m;
};

// src/basis/shore.rs:29-44 — for n in 0..=R, even ell <= n, m in -ell..=ell.
std::vector<Index3> shore_indices(unsigned radial_order) {
    std::vector<Index3> idx;
    for (unsigned n = 0; n <= radial_order; ++n)
        for (unsigned ell = 0; ell <= n; ell += 2)
            for (int m = -int(ell); m <= int(ell); ++m) idx.push_back({n, ell, m});
    return idx;
}

// src/basis/shore.rs:56-61
double kappa(double zeta, unsigned n, unsigned ell) {
    const double num = 2.0 * factorial(n - ell);
    const double den = std::pow(zeta, 1.5) * gamma_fn(double(n) + 1.5);
    return std::sqrt(num / den);
}

// src/basis/shore.rs:64-70
double shore_radial(unsigned n, unsigned ell, double q, double zeta) {
    const double x = q * q / zeta;
    const double lag = gen_laguerre(n - ell, double(ell) + 0.5, x);
    const double env = std::exp(-x * 0.5);
    const double pw = std::pow(x, double(ell) / 2.0);
    return kappa(zeta, n, ell) * lag * env * pw;
}

// src/basis/shore.rs:76-115 — BrainSuite ordering: cos terms k = ell..1,
// then P_ell^0, then signed sin terms k = 1..ell.
void brainsuite_sh_block(unsigned ell, double theta, double phi, double* out) {
    if (ell == 0) {
        out[0] = 1.0 / std::sqrt(4.0 * kPi);
        return;
    }
    const double cphi = std::cos(phi);
    const double tl1 = 2.0 * double(ell) + 1.0;
    std::vector<double> p(ell + 1);
    for (unsigned k = 0; k <= ell; ++k) {
        const double norm = std::sqrt(tl1 / (4.0 * kPi) * factorial(ell - k) / factorial(ell + k));
        p[k] = norm * assoc_legendre(ell, k, cphi);
    }
    const double sqrt2 = std::sqrt(2.0);
    for (unsigned k = 1; k <= ell; ++k) out[ell - k] = sqrt2 * p[k] * std::cos(double(k) * theta);
    out[ell] = p[0];
    for (unsigned k = 1; k <= ell; ++k) {
        const double sign = ((k + 1) % 2 == 0) ? 1.0 : -1.0;
        out[ell + k] = sqrt2 * sign * p[k] * std::sin(double(k) * theta);
    }
}

// src/qspace.rs:130-149 — tau = Delta - delta/3; q = sqrt(b / (4 pi^2 tau)),
// zero at or below the b0 threshold.
void qvals(const double* bvals, int n, double big_delta, double small_delta, double b0_thr, double* q) {
    const double tau = big_delta - small_delta / 3.0;
    const double denom = 4.0 * kPi * kPi * tau;
    for (int i = 0; i < n; ++i) q[i] = (bvals[i] <= b0_thr) ? 0.0 : std::sqrt(bvals[i] / denom);
}

// src/basis/shore.rs:122-178 — row-major N x K output.
void design_matrix(unsigned radial_order, double zeta, const double* q, const double* bvecs, int n_rows,
                   double* A) {
    const auto idx = shore_indices(radial_order);
    const int K = int(idx.size());
    std::vector<double> r(n_rows), th(n_rows), ph(n_rows);
    for (int i = 0; i < n_rows; ++i) {
        double rr, t, p;
        cart2sphere(bvecs[3 * i] * q[i], bvecs[3 * i + 1] * q[i], bvecs[3 * i + 2] * q[i], &rr, &t, &p);
        if (std::isnan(t)) t = 0.0;
        r[i] = rr;
        th[i] = t;
        ph[i] = p;
    }
    std::vector<double> block(2 * radial_order + 1);
    int col = 0;
    for (unsigned n = 0; n <= radial_order; ++n) {
        for (unsigned ell = 0; ell <= n; ell += 2) {
            const int bs = int(2 * ell + 1);
            for (int i = 0; i < n_rows; ++i) {
                const double rad = shore_radial(n, ell, r[i], zeta);
                brainsuite_sh_block(ell, th[i], ph[i], block.data());
                for (int k = 0; k < bs; ++k) A[size_t(i) * K + col + k] = rad * block[k];
            }
            col += bs;
        }
    }
}

// ------------------------------------------------------------- linalg ----

// Largest eigenvalue of a symmetric K x K matrix by cyclic Jacobi rotations
// (textbook algorithm, Golub & Van Loan 8.5). Used for sigma_max(A)^2, the
// quantity the reference reads off nalgebra's SVD (src/solver/fista.rs:45-51).
double sym_max_eig(std::vector<double> a, int K) {
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int p = 0; p < K; ++p) {
            diag += a[size_t(p) * K + p] * a[size_t(p) * K + p];
            for (int q = p + 1; q < K; ++q) off += a[size_t(p) * K + q] * a[size_t(p) * K + q];
        }
        if (off <= 1e-32 * diag || off == 0.0) break;
        for (int p = 0; p < K; ++p) {
            for (int q = p + 1; q < K; ++q) {
                const double apq = a[size_t(p) * K + q];
                if (apq == 0.0) continue;
                const double app = a[size_t(p) * K + p], aqq = a[size_t(q) * K + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < K; ++k) {
                    const double akp = a[size_t(k) * K + p], akq = a[size_t(k) * K + q];
                    a[size_t(k) * K + p] = c * akp - s * akq;
                    a[size_t(k) * K + q] = s * akp + c * akq;
                }
                for (int k = 0; k < K; ++k) {
                    const double apk = a[size_t(p) * K + k], aqk = a[size_t(q) * K + k];
                    a[size_t(p) * K + k] = c * apk - s * aqk;
                    a[size_t(q) * K + k] = s * apk + c * aqk;
                }
            }
        }
    }
    double mx = -std::numeric_limits<double>::infinity();
    for (int p = 0; p < K; ++p) mx = std::max(mx, a[size_t(p) * K + p]);
    return mx;
}

// G = A^T A for row-major A (N x K).
std::vector<double> gram(const double* A, int N, int K) {
    std::vector<double> g(size_t(K) * K, 0.0);
    for (int i = 0; i < N; ++i)
        for (int p = 0; p < K; ++p) {
            const double ap = A[size_t(i) * K + p];
            for (int q = 0; q < K; ++q) g[size_t(p) * K + q] += ap * A[size_t(i) * K + q];
        }
    return g;
}

// src/solver/fista.rs:38-62 — L = max(sigma_max^2 / n, f64::EPSILON).
double lipschitz(const double* A, int N, int K) {
    const double smax2 = std::max(0.0, sym_max_eig(gram(A, N, K), K));
    return std::max(smax2 / double(N), std::numeric_limits<double>::epsilon());
}

double nrm2(const double* v, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += v[i] * v[i];
    return std::sqrt(s);
}

struct Diag {
    uint32_t iterations;
    double residual_l2;
    bool converged;
    double alpha;
    uint8_t kind;
};

// src/solver/fista.rs:70-143 — statement-for-statement FISTA with an optional
// warm start (used only when its length matches, :76-79).
Diag fista(const double* A, int N, int K, const double* s, double alpha, uint32_t max_iter, double tol,
           bool non_negative, double L, const double* warm, double* x_out) {
    std::vector<double> x(K, 0.0), y(K), my(N), g(K), xn(K), yn(K);
    if (warm) std::copy(warm, warm + K, x.begin());
    y = x;
    double t = 1.0;
    const double step = 1.0 / L;
    const double thr = step * alpha;
    const double ninv = 1.0 / double(N);
    uint32_t iters = 0;
    bool converged = false;
    for (uint32_t it = 0; it < max_iter; ++it) {
        for (int i = 0; i < N; ++i) {  // m_y = M y ; resid = m_y - s
            double acc = 0.0;
            for (int k = 0; k < K; ++k) acc += A[size_t(i) * K + k] * y[k];
            my[i] = acc - s[i];
        }
        for (int k = 0; k < K; ++k) {  // grad = M^T resid * (1/n)
            double acc = 0.0;
            for (int i = 0; i < N; ++i) acc += A[size_t(i) * K + k] * my[i];
            g[k] = acc * ninv;
        }
        for (int k = 0; k < K; ++k) {  // soft threshold (strict inequalities)
            const double v = y[k] - step * g[k];
            double sv = (v > thr) ? v - thr : (v < -thr) ? v + thr : 0.0;
            xn[k] = non_negative ? std::max(sv, 0.0) : sv;
        }
        const double tn = 0.5 * (1.0 + std::sqrt(1.0 + 4.0 * t * t));
        const double mom = (t - 1.0) / tn;
        double dx2 = 0.0;
        for (int k = 0; k < K; ++k) {
            yn[k] = xn[k] + mom * (xn[k] - x[k]);
            dx2 += (xn[k] - x[k]) * (xn[k] - x[k]);
        }
        const double dx = std::sqrt(dx2);
        const double den = std::max(nrm2(xn.data(), K), 1e-12);
        const double rel = dx / den;
        x.swap(xn);
        y.swap(yn);
        t = tn;
        iters = it + 1;
        if (rel < tol) {
            converged = true;
            break;
        }
    }
    double r2 = 0.0;
    for (int i = 0; i < N; ++i) {
        double acc = 0.0;
        for (int k = 0; k < K; ++k) acc += A[size_t(i) * K + k] * x[k];
        r2 += (s[i] - acc) * (s[i] - acc);
    }
    std::copy(x.begin(), x.end(), x_out);
    return Diag{iters, std::sqrt(r2), converged, alpha, 1};
}

// src/solver/tikhonov.rs:26-53 — lower Cholesky factor of
// A^T A + lambda_n D_n + lambda_l D_l; returns false when not PD (the
// reference panics at :44).
bool tikhonov_factor(const double* A, int N, int K, const double* dn, const double* dl, double lam_n, double lam_l,
                     std::vector<double>& Lf) {
    Lf = gram(A, N, K);
    for (int i = 0; i < K; ++i) {
        if (dn) Lf[size_t(i) * K + i] += lam_n * dn[i];
        if (dl) Lf[size_t(i) * K + i] += lam_l * dl[i];
    }
    for (int j = 0; j < K; ++j) {
        double d = Lf[size_t(j) * K + j];
        for (int k = 0; k < j; ++k) d -= Lf[size_t(j) * K + k] * Lf[size_t(j) * K + k];
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        Lf[size_t(j) * K + j] = d;
        for (int i = j + 1; i < K; ++i) {
            double v = Lf[size_t(i) * K + j];
            for (int k = 0; k < j; ++k) v -= Lf[size_t(i) * K + k] * Lf[size_t(j) * K + k];
            Lf[size_t(i) * K + j] = v / d;
        }
    }
    for (int i = 0; i < K; ++i)
        for (int j = i + 1; j < K; ++j) Lf[size_t(i) * K + j] = 0.0;
    return true;
}

// src/solver/tikhonov.rs:59-76 — c = chol.solve(M^T s); residual ||s - M c||.
Diag tikhonov_fit(const double* A, int N, int K, const std::vector<double>& Lf, const double* s, double* c) {
    std::vector<double> z(K);
    for (int k = 0; k < K; ++k) {
        double acc = 0.0;
        for (int i = 0; i < N; ++i) acc += A[size_t(i) * K + k] * s[i];
        z[k] = acc;
    }
    for (int i = 0; i < K; ++i) {  // L w = rhs
        double v = z[i];
        for (int k = 0; k < i; ++k) v -= Lf[size_t(i) * K + k] * z[k];
        z[i] = v / Lf[size_t(i) * K + i];
    }
    for (int i = K - 1; i >= 0; --i) {  // L^T c = w
        double v = z[i];
        for (int k = i + 1; k < K; ++k) v -= Lf[size_t(k) * K + i] * z[k];
        z[i] = v / Lf[size_t(i) * K + i];
    }
    double r2 = 0.0;
    for (int i = 0; i < N; ++i) {
        double acc = 0.0;
        for (int k = 0; k < K; ++k) acc += A[size_t(i) * K + k] * z[k];
        r2 += (s[i] - acc) * (s[i] - acc);
    }
    std::copy(z.begin(), z.end(), c);
    return Diag{1, std::sqrt(r2), true, 0.0, 2};
}

// ------------------------------------------------------- alpha machinery ----

// src/solver/alpha.rs:47-53
double alpha_max(const double* A, int N, int K, const double* s) {
    if (N == 0) return 0.0;
    double m = 0.0;
    for (int k = 0; k < K; ++k) {
        double acc = 0.0;
        for (int i = 0; i < N; ++i) acc += A[size_t(i) * K + k] * s[i];
        m = std::max(m, std::fabs(acc));
    }
    return m / double(N);
}

// src/solver/alpha.rs:57-69
void log_path(double amax, int n, double eps, double* out) {
    const double lmax = std::log(amax), lmin = std::log(amax * eps);
    for (int i = 0; i < n; ++i) {
        const double t = double(i) / double(n - 1);
        out[i] = std::exp(lmax + t * (lmin - lmax));
    }
}

// src/solver/alpha.rs:99-103
double bic(int n_grads, double rss, int k_nonzero) {
    const double n = double(n_grads);
    const double rc = std::max(rss, std::numeric_limits<double>::min());
    return n * std::log(rc / n) + std::log(n) * double(k_nonzero);
}

int count_nonzero(const double* c, int K) {
    int k = 0;
    for (int i = 0; i < K; ++i) k += (c[i] != 0.0);
    return k;
}

// Strategy tags mirror AlphaStrategy (src/solver/alpha.rs:195-204).
enum { kFixed = 0, kAlphaRatio = 1, kPathBic = 2, kL2Anchored = 3 };

struct VoxelResult {
    double alpha, residual_l2, bic, rss_l2;
    uint32_t iterations;
    bool converged;
};

struct Problem {
    const double* A;
    int N, K;
    double L;
    uint32_t max_iter;
    double tol;
    bool nonneg;
    int strategy;
    double alpha, ratio, path_eps, slack;
    int path_n;
    const std::vector<double>* tik;  // Cholesky factor for the L2 reference
};

// src/solver/alpha.rs:207-318 (resolve / resolve_with_l2_ref), including the
// all-zero-signal short circuit (:178-189, :231-233) and fit_path's warm
// starts (src/solver/fista.rs:165-180).
VoxelResult resolve(const Problem& P, const double* s, double* coef) {
    const int N = P.N, K = P.K;
    VoxelResult out{0, 0, 0, 0, 0, true};
    auto zero = [&]() {
        std::fill(coef, coef + K, 0.0);
        return VoxelResult{0.0, 0.0, 0.0, 0.0, 0, true};
    };
    if (P.strategy == kFixed || P.strategy == kAlphaRatio) {
        double a = P.alpha;
        if (P.strategy == kAlphaRatio) {
            const double amax = alpha_max(P.A, N, K, s);
            if (amax == 0.0) return zero();
            a = amax * P.ratio;
        }
        Diag d = fista(P.A, N, K, s, a, P.max_iter, P.tol, P.nonneg, P.L, nullptr, coef);
        out.alpha = a;
        out.residual_l2 = d.residual_l2;
        out.iterations = d.iterations;
        out.converged = d.converged;
        out.bic = bic(N, d.residual_l2 * d.residual_l2, count_nonzero(coef, K));
        return out;
    }
    const double amax = alpha_max(P.A, N, K, s);
    if (amax == 0.0) return zero();
    std::vector<double> alphas(P.path_n);
    log_path(amax, P.path_n, P.path_eps, alphas.data());
    std::vector<double> coefs(size_t(P.path_n) * K);
    std::vector<Diag> diags(P.path_n);
    for (int i = 0; i < P.path_n; ++i) {
        const double* warm = (i == 0) ? nullptr : &coefs[size_t(i - 1) * K];
        diags[i] = fista(P.A, N, K, s, alphas[i], P.max_iter, P.tol, P.nonneg, P.L, warm, &coefs[size_t(i) * K]);
    }
    int idx = 0;
    if (P.strategy == kPathBic) {  // src/solver/alpha.rs:107-128
        double best = std::numeric_limits<double>::infinity();
        for (int i = 0; i < P.path_n; ++i) {
            const double rss = diags[i].residual_l2 * diags[i].residual_l2;
            const double b = bic(N, rss, count_nonzero(&coefs[size_t(i) * K], K));
            if (b < best) {
                best = b;
                idx = i;
            }
        }
    } else {  // src/solver/alpha.rs:140-165
        std::vector<double> cl2(K);
        const Diag d2 = tikhonov_fit(P.A, N, K, *P.tik, s, cl2.data());
        const double rss_l2 = d2.residual_l2 * d2.residual_l2;
        out.rss_l2 = rss_l2;
        const double thr = (1.0 + P.slack) * rss_l2;
        idx = -1;
        for (int i = 0; i < P.path_n; ++i) {
            if (diags[i].residual_l2 * diags[i].residual_l2 <= thr) {
                idx = i;
                break;
            }
        }
        if (idx < 0) {
            idx = 0;
            double best = std::numeric_limits<double>::infinity();
            for (int i = 0; i < P.path_n; ++i) {
                const double rss = diags[i].residual_l2 * diags[i].residual_l2;
                if (rss < best) {
                    best = rss;
                    idx = i;
                }
            }
        }
    }
    std::copy(&coefs[size_t(idx) * K], &coefs[size_t(idx) * K] + K, coef);
    const Diag& d = diags[idx];
    out.alpha = alphas[idx];
    out.residual_l2 = d.residual_l2;
    out.iterations = d.iterations;
    out.converged = d.converged;
    out.bic = bic(N, d.residual_l2 * d.residual_l2, count_nonzero(coef, K));
    return out;
}


// ------------------------------------------------------ post-fit (config 3) ----

// src/math.rs:130-145 — Gauss 2F1 by its series (terminating for a = -n).
double hyp2f1(double a, double b, double c, double z) {
    double sum = 1.0, term = 1.0;
    for (int k = 0; k < 200; ++k) {
        const double kf = double(k);
        term *= (a + kf) * (b + kf) / ((c + kf) * (kf + 1.0)) * z;
        sum += term;
        if (term == 0.0) break;
        if (std::fabs(term) < 1e-15 * std::fmax(std::fabs(sum), 1.0)) break;
    }
    return sum;
}

// src/math.rs:150-152
double binomial_float(double n, unsigned k) { return gamma_fn(n + 1.0) / (factorial(k) * gamma_fn(n - double(k) + 1.0)); }

// src/odf.rs:35-50 — C(n,l) G(n,l) F(n,l).
double odf_radial_factor(unsigned n, unsigned ell, double zeta) {
    const double nf = double(n), lf = double(ell);
    const double sign = ((int(n) - int(ell) / 2) % 2 == 0) ? 1.0 : -1.0;
    const double c_nl = sign * std::sqrt(2.0 * factorial(n - ell) / (std::pow(zeta, 1.5) * gamma_fn(nf + 1.5)));
    const double g_nl = (gamma_fn(lf / 2.0 + 1.5) * gamma_fn(nf + 1.5)) / (gamma_fn(lf + 1.5) * factorial(n - ell)) *
                        std::pow(0.5, -lf / 2.0 - 1.5);
    const double f_nl = hyp2f1(-nf + lf, lf / 2.0 + 1.5, lf + 1.5, 2.0);
    return c_nl * g_nl * f_nl;
}

// src/odf.rs:58-77 — Tournier index and sign of BrainSuite slot k of block ell.
void brainsuite_to_tournier(unsigned ell, int k, int* row, double* sign) {
    const int base = int(ell * (ell + 1) / 2), kl = int(ell);
    if (k < kl) {
        *row = base + (kl - k);
        *sign = 1.0;
    } else if (k == kl) {
        *row = base;
        *sign = 1.0;
    } else {
        const int j = k - kl;
        *row = base - j;
        *sign = (j % 2 == 1) ? 1.0 : -1.0;
    }
}

// src/odf.rs:79-102 — row-major (lmax+1)(lmax+2)/2 x K.
void tournier_matrix(unsigned radial_order, double zeta, unsigned lmax, double* out) {
    const int rows = int((lmax + 1) * (lmax + 2) / 2);
    const int K = int(shore_indices(radial_order).size());
    std::fill(out, out + size_t(rows) * K, 0.0);
    int col = 0;
    for (unsigned n = 0; n <= radial_order; ++n)
        for (unsigned ell = 0; ell <= n; ell += 2) {
            const int bs = int(2 * ell + 1);
            if (ell <= lmax) {
                const double radial = odf_radial_factor(n, ell, zeta);
                for (int k = 0; k < bs; ++k) {
                    int row;
                    double sign;
                    brainsuite_to_tournier(ell, k, &row, &sign);
                    out[size_t(row) * K + col + k] = radial * sign;
                }
            }
            col += bs;
        }
}

// src/scalars/microstructure.rs:41-71 — (n, ell) blocks and the (n,0,0) columns.
struct NlBlock {
    unsigned n, ell;
    int start, len;
};
std::vector<NlBlock> nl_blocks(unsigned radial_order) {
    std::vector<NlBlock> b;
    int start = 0;
    for (unsigned n = 0; n <= radial_order; ++n)
        for (unsigned ell = 0; ell <= n; ell += 2) {
            b.push_back({n, ell, start, int(2 * ell + 1)});
            start += int(2 * ell + 1);
        }
    return b;
}
int n00_index(const std::vector<NlBlock>& b, unsigned n) {
    for (const auto& k : b)
        if (k.n == n && k.ell == 0) return k.start;
    return -1;
}

// src/scalars/microstructure.rs:79-87
double predict_e0(const double* c, unsigned R, double zeta, const std::vector<NlBlock>& b) {
    double e0 = 0.0;
    const double inv = 1.0 / std::sqrt(4.0 * kPi);
    for (unsigned n = 0; n <= R; ++n) {
        const double r0 = kappa(zeta, n, 0) * gen_laguerre(n, 0.5, 0.0);
        e0 += c[n00_index(b, n)] * r0 * inv;
    }
    return e0;
}

const double kNaN = std::numeric_limits<double>::quiet_NaN();

// src/scalars/microstructure.rs:96-113
double rtop(const double* c, unsigned R, double zeta, const std::vector<NlBlock>& b) {
    const double e0 = predict_e0(c, R, zeta, b);
    if (!std::isfinite(e0) || e0 == 0.0) return kNaN;
    double acc = 0.0;
    for (unsigned n = 0; n <= R; ++n) {
        const double cn = c[n00_index(b, n)] / e0;
        const double sign = (n % 2 == 0) ? 1.0 : -1.0;
        const double w = std::sqrt(16.0 * kPi * std::pow(zeta, 1.5) * gamma_fn(double(n) + 1.5) / factorial(n));
        acc += cn * sign * w;
    }
    return std::fmax(acc, 0.0);  // f64::max: a NaN operand yields the other
}

// src/scalars/microstructure.rs:117-135
double msd(const double* c, unsigned R, double zeta, const std::vector<NlBlock>& b) {
    const double e0 = predict_e0(c, R, zeta, b);
    if (!std::isfinite(e0) || e0 == 0.0) return kNaN;
    double acc = 0.0;
    for (unsigned n = 0; n <= R; ++n) {
        const double cn = c[n00_index(b, n)] / e0;
        const double sign = (n % 2 == 0) ? 1.0 : -1.0;
        const double w = std::sqrt(9.0 * gamma_fn(double(n) + 1.5) /
                                   (8.0 * std::pow(kPi, 6) * std::pow(zeta, 3.5) * factorial(n)));
        acc += cn * sign * w * hyp2f1(-double(n), 2.5, 1.5, 2.0);
    }
    return std::fmax(acc, 0.0);  // f64::max: a NaN operand yields the other
}

// src/scalars/microstructure.rs:150-177
double qiv(const double* c, unsigned R, double zeta, const std::vector<NlBlock>& b) {
    const double e0 = predict_e0(c, R, zeta, b);
    if (!std::isfinite(e0) || e0 == 0.0) return kNaN;
    const double pre = 6.0 * std::sqrt(2.0 * kPi) * std::pow(zeta, 2.5);
    double integral = 0.0;
    for (unsigned n = 0; n <= R; ++n) {
        const double cn = c[n00_index(b, n)] / e0;
        integral += cn * kappa(zeta, n, 0) * gamma_fn(double(n) + 1.5) / factorial(n) *
                    hyp2f1(-double(n), 2.5, 1.5, 2.0);
    }
    const double q2 = pre * integral;
    if (!std::isfinite(q2) || q2 <= 0.0) return kNaN;
    return 1.0 / q2;
}

// src/scalars/microstructure.rs:181-197
double ng(const double* c, int K, unsigned R, double zeta, const std::vector<NlBlock>& b) {
    const double e0 = predict_e0(c, R, zeta, b);
    if (!std::isfinite(e0) || e0 == 0.0) return kNaN;
    double total = 0.0;
    for (int i = 0; i < K; ++i) {
        const double cn = c[i] / e0;
        total += cn * cn;
    }
    if (total == 0.0) return 0.0;
    const double c0 = c[0] / e0;
    return std::sqrt(std::fmax(1.0 - (c0 * c0) / total, 0.0));
}

// src/scalars/microstructure.rs:278-284
double alpha_factor(unsigned n, unsigned ell, double zeta) {
    const double sign = ((ell / 2) % 2 == 0) ? 1.0 : -1.0;
    const double num = 2.0 * kPi * std::pow(zeta, 1.5) * gamma_fn(double(n) + 1.5);
    const double den = std::pow(2.0, int(ell)) * factorial(n - ell);
    return sign * std::sqrt(num / den);
}

// src/scalars/microstructure.rs:291-306
double rtap_kappa(unsigned j, unsigned ell) {
    const double jf = double(j), lf = double(ell);
    const double kap = std::pow(-1.0, int(j) - 1) * std::pow(2.0, -(lf + 3.0) / 2.0) / kPi;
    double sum = 0.0;
    for (unsigned k = 0; k < j; ++k) {
        const double kf = double(k);
        const double sign = (k % 2 == 0) ? 1.0 : -1.0;
        sum += sign * binomial_float(jf + lf - 0.5, j - k - 1) * gamma_fn((lf + 1.0) / 2.0 + kf) /
               (factorial(k) * std::pow(0.5, (lf + 1.0) / 2.0 + kf));
    }
    return 2.0 * kap * sum;
}

// src/scalars/microstructure.rs:309-322
double rtpp_kappa(unsigned j, unsigned ell) {
    const double jf = double(j), lf = double(ell);
    const double pre = std::pow(-0.5, lf / 2.0) / std::sqrt(kPi);
    double sum = 0.0;
    for (unsigned k = 0; k < j; ++k) {
        const double kf = double(k);
        const double sign = (k % 2 == 0) ? 1.0 : -1.0;
        sum += sign * binomial_float(jf + lf - 0.5, j - k - 1) * gamma_fn(lf / 2.0 + kf + 0.5) /
               (factorial(k) * std::pow(0.5, lf / 2.0 + 0.5 + kf));
    }
    return pre * sum;
}

// src/scalars/microstructure.rs:219-270 — RTAP (axis = true) or RTPP.
double directional(const double* c, unsigned R, double zeta, const std::vector<NlBlock>& b, const double* dir,
                   bool axis) {
    const double norm = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    if (!std::isfinite(norm) || norm == 0.0) return kNaN;
    const double e0 = predict_e0(c, R, zeta, b);
    if (!std::isfinite(e0) || e0 == 0.0) return kNaN;
    const double inv = 1.0 / norm;
    double r, theta, phi;
    cart2sphere(dir[0] * inv, dir[1] * inv, dir[2] * inv, &r, &theta, &phi);
    double acc = 0.0;
    std::vector<double> sh(2 * R + 1);
    for (const auto& k : b) {
        const unsigned j = k.n - k.ell + 1;
        const double kern = axis ? rtap_kappa(j, k.ell) : rtpp_kappa(j, k.ell);
        if (kern == 0.0) continue;
        const double scale = kern / alpha_factor(k.n, k.ell, zeta);
        brainsuite_sh_block(k.ell, theta, phi, sh.data());
        double inner = 0.0;
        for (int t = 0; t < k.len; ++t) inner += (c[k.start + t] / e0) * sh[t];
        acc += scale * inner;
    }
    const double mu = 1.0 / (2.0 * kPi * std::sqrt(zeta));
    const double pref = axis ? 1.0 / (mu * mu) : 1.0 / mu;
    return std::fmax(pref * acc, 0.0);
}

// src/io/microstructure.rs:303-313 — index round((n-1) p) of the sorted finite values.
bool percentile_value(const std::vector<float>& xs, double pct, float* out) {
    std::vector<float> f;
    for (float v : xs)
        if (std::isfinite(v)) f.push_back(v);
    if (f.empty()) return false;
    std::sort(f.begin(), f.end());
    const double p = std::min(std::max(pct, 0.0), 100.0) / 100.0;
    const size_t i = std::min(size_t(std::llround(p * double(f.size() - 1))), f.size() - 1);
    *out = f[i];
    return true;
}

// src/io/microstructure.rs:317-326
int nan_above(std::vector<float>& xs, float thr) {
    int n = 0;
    for (float& v : xs)
        if (std::isfinite(v) && v > thr) {
            v = std::numeric_limits<float>::quiet_NaN();
            ++n;
        }
    return n;
}

}  // namespace

// ======================================================== C entry points ====
extern "C" {

int orc_n_coeffs(unsigned radial_order) { return int(shore_indices(radial_order).size()); }

void orc_shore_indices(unsigned radial_order, int32_t* out_nlm) {
    const auto idx = shore_indices(radial_order);
    for (size_t i = 0; i < idx.size(); ++i) {
        out_nlm[3 * i] = int32_t(idx[i].n);
        out_nlm[3 * i + 1] = int32_t(idx[i].ell);
        out_nlm[3 * i + 2] = idx[i].m;
    }
}

double orc_factorial(unsigned n) { return factorial(n); }
double orc_gamma(double x) { return gamma_fn(x); }
double orc_assoc_legendre(unsigned l, unsigned m, double x) { return assoc_legendre(l, m, x); }
double orc_gen_laguerre(unsigned n, double a, double x) { return gen_laguerre(n, a, x); }
double orc_kappa(double zeta, unsigned n, unsigned ell) { return kappa(zeta, n, ell); }
void orc_brainsuite_sh_block(unsigned ell, double theta, double phi, double* out) {
    brainsuite_sh_block(ell, theta, phi, out);
}

// src/qspace.rs:213-227 — TORTOISE (Delta, delta) heuristic.
void orc_estimate_deltas_tortoise(const double* bvals, int n, double gmax, double* big, double* small) {
    double mb = 0.0;
    for (int i = 0; i < n; ++i) mb = std::max(mb, bvals[i]);
    if (mb <= 0.0) {
        const double sm = 1.0;
        *big = 3.0 * sm * 1e-3;
        *small = sm * 1e-3;
        return;
    }
    const double gyro = 267.51532e6;
    const double temp = mb / gyro / gyro / gmax / gmax / 2.0 * 1e6;
    const double small_ms = std::pow(temp, 1.0 / 3.0) * 1000.0;
    *big = small_ms * 3.0 * 1e-3;
    *small = small_ms * 1e-3;
}

void orc_qvals(const double* bvals, int n, double big_delta, double small_delta, double b0_thr, double* q) {
    qvals(bvals, n, big_delta, small_delta, b0_thr, q);
}

// src/qspace.rs:171-197 + src/fit.rs:615-617 — mean b-values of the non-zero
// shells, ascending. Returns the shell count; writes at most max_out values.
int orc_dwi_shells(const double* bvals, int n, double b0_thr, double tol, double* out, int max_out) {
    std::vector<double> b;
    std::vector<int> cnt;
    for (int i = 0; i < n; ++i) {
        if (bvals[i] <= b0_thr) continue;
        bool placed = false;
        for (size_t s = 0; s < b.size(); ++s) {
            if (std::fabs(b[s] - bvals[i]) <= tol) {
                const double c = double(cnt[s]);
                b[s] = (b[s] * c + bvals[i]) / (c + 1.0);
                cnt[s] += 1;
                placed = true;
                break;
            }
        }
        if (!placed) {
            b.push_back(bvals[i]);
            cnt.push_back(1);
        }
    }
    std::stable_sort(b.begin(), b.end());
    for (int i = 0; i < int(b.size()) && i < max_out; ++i) out[i] = b[i];
    return int(b.size());
}

// Design matrix from raw (bvals, bvecs, deltas): row-major N x K.
void orc_design_matrix(unsigned radial_order, double zeta, const double* bvals, const double* bvecs, int n,
                       double big_delta, double small_delta, double b0_thr, double* A) {
    std::vector<double> q(n);
    qvals(bvals, n, big_delta, small_delta, b0_thr, q.data());
    design_matrix(radial_order, zeta, q.data(), bvecs, n, A);
}

// src/basis/shore.rs:180-195
void orc_regularization(unsigned radial_order, double* dn, double* dl) {
    const auto idx = shore_indices(radial_order);
    for (size_t i = 0; i < idx.size(); ++i) {
        const double nn = double(idx[i].n * (idx[i].n + 1));
        const double ll = double(idx[i].ell * (idx[i].ell + 1));
        dn[i] = nn * nn;
        dl[i] = ll * ll;
    }
}

double orc_lipschitz(const double* A, int N, int K) { return lipschitz(A, N, K); }
double orc_sym_max_eig(const double* S, int K) { return sym_max_eig(std::vector<double>(S, S + size_t(K) * K), K); }

// One FISTA fit (src/solver/fista.rs:70-143). L <= 0 computes the Lipschitz
// constant as FistaSolver::new does. diag_out = {iterations, residual_l2,
// converged, alpha}.
void orc_fista(const double* A, int N, int K, const double* s, double alpha, uint32_t max_iter, double tol,
               int non_negative, double L, const double* warm, double* coef, double* diag_out) {
    if (!(L > 0.0)) L = lipschitz(A, N, K);
    Diag d = fista(A, N, K, s, alpha, max_iter, tol, non_negative != 0, L, warm, coef);
    diag_out[0] = d.iterations;
    diag_out[1] = d.residual_l2;
    diag_out[2] = d.converged ? 1.0 : 0.0;
    diag_out[3] = d.alpha;
}

// Tikhonov fit of V signals (row-major V x N) -> coef V x K, residual_l2 V.
// Returns 0, or -1 when the normal matrix is not positive definite.
int orc_tikhonov(const double* A, int N, int K, const double* dn, const double* dl, double lam_n, double lam_l,
                 const double* S, int V, double* C, double* resid) {
    std::vector<double> Lf;
    if (!tikhonov_factor(A, N, K, dn, dl, lam_n, lam_l, Lf)) return -1;
    for (int v = 0; v < V; ++v) {
        Diag d = tikhonov_fit(A, N, K, Lf, S + size_t(v) * N, C + size_t(v) * K);
        if (resid) resid[v] = d.residual_l2;
    }
    return 0;
}

double orc_alpha_max(const double* A, int N, int K, const double* s) { return alpha_max(A, N, K, s); }
void orc_log_path(double amax, int n, double eps, double* out) { log_path(amax, n, eps, out); }
double orc_bic(int n_grads, double rss, int k) { return bic(n_grads, rss, k); }

// Whole-volume fit mirroring src/fit.rs:74-175 (L2), :219-362 (L1 alpha
// strategy) and :372-527 (L2-anchored), with voxels collected in x->y->z
// order (src/voxel_loop.rs:28-42) and a per-task thread pool in place of
// rayon. A is row-major N x K (built by the caller).
//
// reg: 1 = L1, 2 = L2. strategy: 0 fixed, 1 alpha-ratio, 2 path-bic,
// 3 l2-anchored. Output maps may be NULL. Returns 0, -1 if the Tikhonov
// normal matrix is not PD.
int orc_fit_volume(const float* data, const uint8_t* mask, int nx, int ny, int nz, int N, const double* A, int K,
                   int reg, int strategy, double alpha, double alpha_ratio, int path_n, double path_eps,
                   double slack, uint32_t max_iter, double tol, int non_negative, const double* dn,
                   const double* dl, double lam_n, double lam_l, int n_threads, float* coef_out, float* r2_out,
                   float* resid_out, uint32_t* iters_out, uint8_t* kind_out, float* alpha_out, float* bic_out,
                   float* rss_l2_out, double* alpha_dist_out) {
    std::vector<int64_t> work;
    for (int64_t x = 0; x < nx; ++x)
        for (int64_t y = 0; y < ny; ++y)
            for (int64_t z = 0; z < nz; ++z) {
                const int64_t vi = (x * ny + y) * nz + z;
                if (!mask || mask[vi]) work.push_back(vi);
            }
    const int64_t nvox = int64_t(nx) * ny * nz;
    std::fill(coef_out, coef_out + nvox * K, 0.0f);
    if (r2_out) std::fill(r2_out, r2_out + nvox, 0.0f);
    if (resid_out) std::fill(resid_out, resid_out + nvox, 0.0f);
    if (iters_out) std::fill(iters_out, iters_out + nvox, 0u);
    if (kind_out) std::fill(kind_out, kind_out + nvox, uint8_t(0));
    if (alpha_out) std::fill(alpha_out, alpha_out + nvox, 0.0f);
    if (bic_out) std::fill(bic_out, bic_out + nvox, 0.0f);
    if (rss_l2_out) std::fill(rss_l2_out, rss_l2_out + nvox, 0.0f);

    std::vector<double> tik;
    const bool need_tik = (reg == 2) || (reg == 1 && strategy == kL2Anchored);
    if (need_tik && !tikhonov_factor(A, N, K, dn, dl, lam_n, lam_l, tik)) return -1;
    Problem P{A, N, K, 0.0, max_iter, tol, non_negative != 0, strategy, alpha, alpha_ratio, path_eps, slack, path_n,
              &tik};
    if (reg == 1) P.L = lipschitz(A, N, K);

    std::vector<double> alphas_seen(work.size(), 0.0);
    std::atomic<size_t> next{0};
    auto worker = [&]() {
        std::vector<double> s(N), c(K);
        for (;;) {
            const size_t w = next.fetch_add(1);
            if (w >= work.size()) break;
            const int64_t vi = work[w];
            const float* src = data + vi * N;
            double mean = 0.0;
            for (int i = 0; i < N; ++i) {
                s[i] = double(src[i]);
                mean += s[i];
            }
            mean /= double(N);
            double resid_l2, a = 0.0, b = 0.0, rss_l2 = 0.0;
            uint32_t its;
            if (reg == 2) {
                Diag d = tikhonov_fit(A, N, K, tik, s.data(), c.data());
                resid_l2 = d.residual_l2;
                its = d.iterations;
            } else {
                if (strategy == kL2Anchored) {
                    std::vector<double> cl2(K);
                    Diag d2 = tikhonov_fit(A, N, K, tik, s.data(), cl2.data());
                    rss_l2 = d2.residual_l2 * d2.residual_l2;
                }
                VoxelResult r = resolve(P, s.data(), c.data());
                resid_l2 = r.residual_l2;
                its = r.iterations;
                a = r.alpha;
                b = r.bic;
            }
            double ss_tot = 0.0;
            for (int i = 0; i < N; ++i) ss_tot += (s[i] - mean) * (s[i] - mean);
            const double ss_res = resid_l2 * resid_l2;
            const double r2 = ss_tot > 0.0 ? 1.0 - ss_res / ss_tot : 0.0;
            for (int k = 0; k < K; ++k) coef_out[vi * K + k] = float(c[k]);
            if (r2_out) r2_out[vi] = float(r2);
            if (resid_out) resid_out[vi] = float(resid_l2);
            if (iters_out) iters_out[vi] = its;
            if (kind_out) kind_out[vi] = uint8_t(reg == 2 ? 2 : 1);
            if (reg == 1) {
                if (alpha_out) alpha_out[vi] = float(a);
                if (bic_out) bic_out[vi] = float(b);
                if (rss_l2_out && strategy == kL2Anchored) rss_l2_out[vi] = float(rss_l2);
                alphas_seen[w] = a;
            }
        }
    };
    if (n_threads <= 0) n_threads = int(std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int t = 1; t < n_threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();

    // src/fit.rs:334-352 — alpha percentiles at index round((len-1) q).
    if (alpha_dist_out && reg == 1 && strategy != kFixed) {
        std::vector<double> srt = alphas_seen;
        std::sort(srt.begin(), srt.end());
        if (srt.empty()) {
            alpha_dist_out[0] = alpha_dist_out[1] = alpha_dist_out[2] = 0.0;
        } else {
            auto pct = [&](double q) {
                size_t i = size_t(std::llround(double(srt.size() - 1) * q));
                return srt[std::min(i, srt.size() - 1)];
            };
            alpha_dist_out[0] = pct(0.5);
            alpha_dist_out[1] = pct(0.10);
            alpha_dist_out[2] = pct(0.90);
        }
    }
    return 0;
}

// src/graddev.rs:160-185 — effective table for one deviation matrix T
// (row-major), with the frame rotation R (row-major) and voxel-frame bvecs.
void orc_effective_table(const double* T, const double* R, const double* bvals, const double* bvecs_vox,
                         const double* bvecs_nominal, int n, double b0_thr, double* b_out, double* dir_out) {
    for (int i = 0; i < n; ++i) {
        const double* g = bvecs_vox + 3 * i;
        const double gn = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        if (!(bvals[i] > b0_thr) || !(gn > 1e-12)) {
            b_out[i] = bvals[i];
            for (int k = 0; k < 3; ++k) dir_out[3 * i + k] = bvecs_nominal[3 * i + k];
            continue;
        }
        double u[3] = {g[0] / gn, g[1] / gn, g[2] / gn};
        double q[3];
        for (int r = 0; r < 3; ++r) q[r] = T[0 * 3 + r] * u[0] + T[1 * 3 + r] * u[1] + T[2 * 3 + r] * u[2];
        double d[3];
        for (int r = 0; r < 3; ++r) d[r] = R[r * 3 + 0] * q[0] + R[r * 3 + 1] * q[1] + R[r * 3 + 2] * q[2];
        const double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        b_out[i] = bvals[i] * (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        for (int k = 0; k < 3; ++k) dir_out[3 * i + k] = d[k] / dn;
    }
}

// Whole-volume L1 fit with a gradient-deviation field: the per-voxel loop of
// src/dti.rs:393-407 (the reference's pattern for graddev fits) applied to
// the SHORE fit. A voxel whose T is the identity to IDENTITY_TOL
// (src/graddev.rs:35-36, :89-93) is fitted with the nominal design; any other
// voxel gets the effective table (src/graddev.rs:160-185), q-values from that
// table (src/qspace.rs:130-149), its own design (src/basis/shore.rs:122-178),
// its own Lipschitz constant (src/solver/fista.rs:38-62) and then the fixed /
// alpha-ratio resolve (src/solver/alpha.rs:207-233), as the volume driver
// does (src/fit.rs:219-362). field: (X,Y,Z,9) f32 row-major T. strategy:
// 0 fixed, 1 alpha-ratio. lip_out (optional, f64 per voxel) receives the
// Lipschitz constant each fitted voxel used.
int orc_fit_graddev(const float* data, const uint8_t* mask, int nx, int ny, int nz, int N, const double* bvals,
                    const double* bvecs_vox, const double* bvecs_nominal, const double* R, double big_delta,
                    double small_delta, double b0_thr, unsigned radial_order, double zeta, const float* field,
                    int strategy, double alpha, double alpha_ratio, uint32_t max_iter, double tol, int non_negative,
                    int n_threads, float* coef_out, float* r2_out, float* resid_out, uint32_t* iters_out,
                    uint8_t* kind_out, float* alpha_out, float* bic_out, double* lip_out) {
    if (strategy != kFixed && strategy != kAlphaRatio) return -2;
    const int K = int(shore_indices(radial_order).size());
    std::vector<int64_t> work;
    for (int64_t x = 0; x < nx; ++x)
        for (int64_t y = 0; y < ny; ++y)
            for (int64_t z = 0; z < nz; ++z) {
                const int64_t vi = (x * ny + y) * nz + z;
                if (!mask || mask[vi]) work.push_back(vi);
            }
    const int64_t nvox = int64_t(nx) * ny * nz;
    std::fill(coef_out, coef_out + nvox * K, 0.0f);
    if (r2_out) std::fill(r2_out, r2_out + nvox, 0.0f);
    if (resid_out) std::fill(resid_out, resid_out + nvox, 0.0f);
    if (iters_out) std::fill(iters_out, iters_out + nvox, 0u);
    if (kind_out) std::fill(kind_out, kind_out + nvox, uint8_t(0));
    if (alpha_out) std::fill(alpha_out, alpha_out + nvox, 0.0f);
    if (bic_out) std::fill(bic_out, bic_out + nvox, 0.0f);
    if (lip_out) std::fill(lip_out, lip_out + nvox, 0.0);

    std::vector<double> q0(N), A0(size_t(N) * K);
    qvals(bvals, N, big_delta, small_delta, b0_thr, q0.data());
    design_matrix(radial_order, zeta, q0.data(), bvecs_nominal, N, A0.data());
    const double L0 = lipschitz(A0.data(), N, K);

    std::atomic<size_t> next{0};
    auto worker = [&]() {
        std::vector<double> s(N), c(K), beff(N), deff(size_t(N) * 3), q(N), Av(size_t(N) * K);
        for (;;) {
            const size_t w = next.fetch_add(1);
            if (w >= work.size()) break;
            const int64_t vi = work[w];
            double T[9];
            double dev = 0.0;
            for (int k = 0; k < 9; ++k) {
                T[k] = double(field[vi * 9 + k]);
                dev = std::max(dev, std::fabs(T[k] - ((k % 4 == 0) ? 1.0 : 0.0)));
            }
            const double* A = A0.data();
            double L = L0;
            if (!(dev < 1e-6)) {
                orc_effective_table(T, R, bvals, bvecs_vox, bvecs_nominal, N, b0_thr, beff.data(), deff.data());
                qvals(beff.data(), N, big_delta, small_delta, b0_thr, q.data());
                design_matrix(radial_order, zeta, q.data(), deff.data(), N, Av.data());
                A = Av.data();
                L = lipschitz(A, N, K);
            }
            const float* src = data + vi * N;
            double mean = 0.0;
            for (int i = 0; i < N; ++i) {
                s[i] = double(src[i]);
                mean += s[i];
            }
            mean /= double(N);
            Problem P{A, N, K, L, max_iter, tol, non_negative != 0, strategy, alpha, alpha_ratio, 0.0, 0.0, 0,
                      nullptr};
            const VoxelResult r = resolve(P, s.data(), c.data());
            double ss_tot = 0.0;
            for (int i = 0; i < N; ++i) ss_tot += (s[i] - mean) * (s[i] - mean);
            const double ss_res = r.residual_l2 * r.residual_l2;
            const double r2 = ss_tot > 0.0 ? 1.0 - ss_res / ss_tot : 0.0;
            for (int k = 0; k < K; ++k) coef_out[vi * K + k] = float(c[k]);
            if (r2_out) r2_out[vi] = float(r2);
            if (resid_out) resid_out[vi] = float(r.residual_l2);
            if (iters_out) iters_out[vi] = r.iterations;
            if (kind_out) kind_out[vi] = 1;
            if (alpha_out) alpha_out[vi] = float(r.alpha);
            if (bic_out) bic_out[vi] = float(r.bic);
            if (lip_out) lip_out[vi] = L;
        }
    };
    if (n_threads <= 0) n_threads = int(std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int t = 1; t < n_threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    return 0;
}

// ---------------------------------------------------- post-fit entry points ----
double orc_hyp2f1(double a, double b, double c, double z) { return hyp2f1(a, b, c, z); }
double orc_binomial_float(double n, unsigned k) { return binomial_float(n, k); }
double orc_odf_radial_factor(unsigned n, unsigned ell, double zeta) { return odf_radial_factor(n, ell, zeta); }
void orc_brainsuite_to_tournier(unsigned ell, int k, int* row, double* sign) { brainsuite_to_tournier(ell, k, row, sign); }
void orc_tournier_matrix(unsigned radial_order, double zeta, unsigned lmax, double* out) {
    tournier_matrix(radial_order, zeta, lmax, out);
}
double orc_alpha_factor(unsigned n, unsigned ell, double zeta) { return alpha_factor(n, ell, zeta); }

// Raw per-voxel scalars (src/scalars/microstructure.rs:79-270) of one FP64
// coefficient vector: out = rtop, rtap, rtpp, msd, qiv, ng. dir may be NULL
// (RTAP/RTPP NaN, as for a missing peak).
void orc_microstructure_raw(unsigned radial_order, double zeta, const double* c, const double* dir, double* out) {
    const auto b = nl_blocks(radial_order);
    const int K = int(shore_indices(radial_order).size());
    const double zero[3] = {0.0, 0.0, 0.0};
    const double* d = dir ? dir : zero;
    out[0] = rtop(c, radial_order, zeta, b);
    out[1] = directional(c, radial_order, zeta, b, d, true);
    out[2] = directional(c, radial_order, zeta, b, d, false);
    out[3] = msd(c, radial_order, zeta, b);
    out[4] = qiv(c, radial_order, zeta, b);
    out[5] = ng(c, K, radial_order, zeta, b);
}

// src/io/microstructure.rs:160-300 (the values it returns): per masked voxel
// (coefs V x K f32, dirs V x 3 f32 or NULL; a zero or non-finite direction is
// a missing peak, python/src/models.rs:318-326), the six scalars as f32 times
// the unit factor (units 0 = um, 1 = mm), then, when outlier_factor > 0, each
// scalar's finite values above outlier_factor x p99 set to NaN.
// out: 6 x V (rtop, rtap, rtpp, msd, qiv, ng).
void orc_compute_microstructure(unsigned radial_order, double zeta, const float* coefs, const float* dirs, int64_t V,
                                int units, float outlier_factor, float* out) {
    const int K = int(shore_indices(radial_order).size());
    const auto b = nl_blocks(radial_order);
    const float su[6] = {1e-9f, 1e-6f, 1e-3f, 1e6f, 1e15f, 1.0f};
    std::vector<std::vector<float>> vals(6, std::vector<float>(size_t(V)));
    std::vector<double> c(K);
    for (int64_t v = 0; v < V; ++v) {
        for (int k = 0; k < K; ++k) c[k] = double(coefs[size_t(v) * K + k]);
        double d[3] = {0.0, 0.0, 0.0};
        if (dirs) {
            const float* p = dirs + 3 * v;
            const bool any = p[0] != 0.0f || p[1] != 0.0f || p[2] != 0.0f;
            const bool fin = std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]);
            if (any && fin)
                for (int t = 0; t < 3; ++t) d[t] = double(p[t]);
        }
        double raw[6];
        raw[0] = rtop(c.data(), radial_order, zeta, b);
        raw[1] = directional(c.data(), radial_order, zeta, b, d, true);
        raw[2] = directional(c.data(), radial_order, zeta, b, d, false);
        raw[3] = msd(c.data(), radial_order, zeta, b);
        raw[4] = qiv(c.data(), radial_order, zeta, b);
        raw[5] = ng(c.data(), K, radial_order, zeta, b);
        for (int q = 0; q < 6; ++q) vals[q][size_t(v)] = float(raw[q]) * (units == 0 ? su[q] : 1.0f);
    }
    if (outlier_factor > 0.0f)
        for (int q = 0; q < 6; ++q) {
            float p99;
            if (percentile_value(vals[q], 99.0, &p99)) nan_above(vals[q], outlier_factor * p99);
        }
    for (int q = 0; q < 6; ++q) std::copy(vals[q].begin(), vals[q].end(), out + size_t(q) * V);
}

// src/synth.rs:32-84 — per voxel s = A c (FP64) as f32; an all-zero
// coefficient vector gives zeros. A is the N' x K target design (row-major).
void orc_synthesize(const double* A, int Np, int K, const float* coefs, int64_t V, float* out) {
    for (int64_t v = 0; v < V; ++v) {
        const float* c = coefs + size_t(v) * K;
        bool any = false;
        for (int k = 0; k < K; ++k) any = any || (c[k] != 0.0f);
        for (int i = 0; i < Np; ++i) {
            if (!any) {
                out[size_t(v) * Np + i] = 0.0f;
                continue;
            }
            double acc = 0.0;
            for (int k = 0; k < K; ++k) acc += A[size_t(i) * K + k] * double(c[k]);
            out[size_t(v) * Np + i] = float(acc);
        }
    }
}

// python/src/models.rs:228-247 — per voxel t = M c (FP64) as f32.
void orc_project(const double* M, int rows, int K, const float* coefs, int64_t V, float* out) {
    for (int64_t v = 0; v < V; ++v)
        for (int i = 0; i < rows; ++i) {
            double acc = 0.0;
            for (int k = 0; k < K; ++k) acc += M[size_t(i) * K + k] * double(coefs[size_t(v) * K + k]);
            out[size_t(v) * rows + i] = float(acc);
        }
}

// ------------------------------------------------------ differentiation ----
// No counterpart in the reference. For the lasso objective of
// src/solver/fista.rs (1/2N |s - A c|^2 + alpha |c|_1), the optimality
// conditions on the active set S (c_S != 0, z = sign(c_S)) are
// G_SS c_S = A_S^T s - N alpha z with G = A^T A. The textbook consequence:
// w = G_SS^-1 g_S, dL/ds = A_S w, dL/dalpha = -N z^T w for an upstream
// gradient g. FP64 Cholesky. Returns the number of signals whose G_SS is not
// positive definite (their outputs are NaN).
namespace {
bool cholesky_solve(std::vector<double>& M, int m, std::vector<double>& x) {
    for (int j = 0; j < m; ++j) {
        double d = M[size_t(j) * m + j];
        for (int k = 0; k < j; ++k) d -= M[size_t(j) * m + k] * M[size_t(j) * m + k];
        if (!(d > 0.0)) return false;
        const double ljj = std::sqrt(d);
        M[size_t(j) * m + j] = ljj;
        for (int i = j + 1; i < m; ++i) {
            double v = M[size_t(i) * m + j];
            for (int k = 0; k < j; ++k) v -= M[size_t(i) * m + k] * M[size_t(j) * m + k];
            M[size_t(i) * m + j] = v / ljj;
        }
    }
    for (int j = 0; j < m; ++j) {
        double v = x[j];
        for (int k = 0; k < j; ++k) v -= M[size_t(j) * m + k] * x[k];
        x[j] = v / M[size_t(j) * m + j];
    }
    for (int j = m - 1; j >= 0; --j) {
        double v = x[j];
        for (int k = j + 1; k < m; ++k) v -= M[size_t(k) * m + j] * x[k];
        x[j] = v / M[size_t(j) * m + j];
    }
    return true;
}
}  // namespace

int orc_lasso_adjoint(const double* A, int N, int K, const double* coef, const double* grad, int V, double* grad_s,
                      double* grad_alpha) {
    const std::vector<double> G = gram(A, N, K);
    int failed = 0;
    for (int v = 0; v < V; ++v) {
        const double* c = coef + size_t(v) * K;
        std::vector<int> S;
        for (int k = 0; k < K; ++k)
            if (c[k] != 0.0) S.push_back(k);
        const int m = int(S.size());
        std::vector<double> M(size_t(m) * m), w(m);
        for (int i = 0; i < m; ++i) {
            w[i] = grad[size_t(v) * K + S[i]];
            for (int j = 0; j < m; ++j) M[size_t(i) * m + j] = G[size_t(S[i]) * K + S[j]];
        }
        if (m > 0 && !cholesky_solve(M, m, w)) {
            ++failed;
            for (int n = 0; n < N; ++n) grad_s[size_t(v) * N + n] = std::nan("");
            grad_alpha[v] = std::nan("");
            continue;
        }
        double zw = 0.0;
        for (int i = 0; i < m; ++i) zw += (c[S[i]] > 0.0 ? w[i] : -w[i]);
        grad_alpha[v] = -double(N) * zw;
        for (int n = 0; n < N; ++n) {
            double acc = 0.0;
            for (int i = 0; i < m; ++i) acc += A[size_t(n) * K + S[i]] * w[i];
            grad_s[size_t(v) * N + n] = acc;
        }
    }
    return failed;
}

// The exact lasso minimiser on the active set and signs of c_in: c_S solves
// G_SS c_S = A_S^T s - N alpha z. Returns 0 when it is verified optimal
// (signs kept, |A_k^T (s - A c)| / N <= alpha (1 + 1e-9) off S), 1 on a sign
// change, 2 on a violated inactive condition, 3 if G_SS is not PD.
int orc_lasso_polish(const double* A, int N, int K, const double* s, double alpha, const double* c_in,
                     double* c_out) {
    const std::vector<double> G = gram(A, N, K);
    std::vector<int> S;
    for (int k = 0; k < K; ++k)
        if (c_in[k] != 0.0) S.push_back(k);
    const int m = int(S.size());
    std::vector<double> M(size_t(m) * m), x(m);
    for (int i = 0; i < m; ++i) {
        double ats = 0.0;
        for (int n = 0; n < N; ++n) ats += A[size_t(n) * K + S[i]] * s[n];
        x[i] = ats - double(N) * alpha * (c_in[S[i]] > 0.0 ? 1.0 : -1.0);
        for (int j = 0; j < m; ++j) M[size_t(i) * m + j] = G[size_t(S[i]) * K + S[j]];
    }
    if (m > 0 && !cholesky_solve(M, m, x)) return 3;
    std::fill(c_out, c_out + K, 0.0);
    for (int i = 0; i < m; ++i) c_out[S[i]] = x[i];
    for (int i = 0; i < m; ++i)
        if ((x[i] > 0.0) != (c_in[S[i]] > 0.0)) return 1;
    std::vector<double> r(N);
    for (int n = 0; n < N; ++n) {
        double acc = 0.0;
        for (int k = 0; k < K; ++k) acc += A[size_t(n) * K + k] * c_out[k];
        r[n] = s[n] - acc;
    }
    for (int k = 0; k < K; ++k) {
        if (c_out[k] != 0.0) continue;
        double g = 0.0;
        for (int n = 0; n < N; ++n) g += A[size_t(n) * K + k] * r[n];
        if (std::fabs(g) / double(N) > alpha * (1.0 + 1e-9)) return 2;
    }
    return 0;
}

}  // extern "C"
