// Post-fit maps from SHORE coefficients: the SHORE -> Tournier SH projection
// (src/odf.rs:22-102, python/src/models.rs:203-248), signal synthesis for a
// new gradient table (src/synth.rs:26-84) and the microstructure scalars
// RTOP / RTAP / RTPP / MSD / QIV / NG (src/scalars/microstructure.rs:79-322,
// src/io/microstructure.rs:160-326).
//
// The linear maps (projection, synthesis) are one FP64 matrix applied to each
// voxel's coefficient vector: 64 voxels per workgroup, their coefficients
// staged in LDS, one voxel per lane, the matrix read with wave-uniform loads.
// The scalars are per-voxel FP64 closed forms over precomputed per-mode
// weights; the direction-dependent RTAP / RTPP evaluate the BrainSuite SH
// blocks at each voxel's peak direction.
#include <cmath>
#include <vector>

#include "common.h"

namespace shorefit {

namespace {

constexpr double kPiP = 3.14159265358979323846264338327950288;

// ------------------------------------------------------- host FP64 helpers ----
double fact_h(unsigned n) {
    double f = 1.0;
    for (unsigned i = 2; i <= n; ++i) f *= double(i);
    return f;
}
double gamma_h(double x) { return std::tgamma(x); }

// src/math.rs:130-145
double hyp2f1_h(double a, double b, double c, double z) {
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
double binom_h(double n, unsigned k) { return gamma_h(n + 1.0) / (fact_h(k) * gamma_h(n - double(k) + 1.0)); }

// src/basis/shore.rs:56-61
double kappa_h(double zeta, unsigned n, unsigned ell) {
    return std::sqrt(2.0 * fact_h(n - ell) / (std::pow(zeta, 1.5) * gamma_h(double(n) + 1.5)));
}

// L_n^a(0) = binomial(n + a, n) by the three-term recurrence at x = 0
// (src/math.rs:108-125).
double laguerre0_h(unsigned n, double a) {
    double lm1 = 1.0;
    if (n == 0) return lm1;
    double l = 1.0 + a;
    for (unsigned k = 1; k < n; ++k) {
        const double lp1 = ((2.0 * k + 1.0 + a) * l - (k + a) * lm1) / (k + 1.0);
        lm1 = l;
        l = lp1;
    }
    return l;
}

// src/scalars/microstructure.rs:291-322
double rtap_kappa_h(unsigned j, unsigned ell) {
    const double jf = double(j), lf = double(ell);
    const double kap = std::pow(-1.0, int(j) - 1) * std::pow(2.0, -(lf + 3.0) / 2.0) / kPiP;
    double sum = 0.0;
    for (unsigned k = 0; k < j; ++k) {
        const double kf = double(k);
        sum += ((k % 2 == 0) ? 1.0 : -1.0) * binom_h(jf + lf - 0.5, j - k - 1) * gamma_h((lf + 1.0) / 2.0 + kf) /
               (fact_h(k) * std::pow(0.5, (lf + 1.0) / 2.0 + kf));
    }
    return 2.0 * kap * sum;
}
double rtpp_kappa_h(unsigned j, unsigned ell) {
    const double jf = double(j), lf = double(ell);
    const double pre = std::pow(-0.5, lf / 2.0) / std::sqrt(kPiP);
    double sum = 0.0;
    for (unsigned k = 0; k < j; ++k) {
        const double kf = double(k);
        sum += ((k % 2 == 0) ? 1.0 : -1.0) * binom_h(jf + lf - 0.5, j - k - 1) * gamma_h(lf / 2.0 + kf + 0.5) /
               (fact_h(k) * std::pow(0.5, lf / 2.0 + 0.5 + kf));
    }
    return pre * sum;
}
// src/scalars/microstructure.rs:278-284
double alpha_factor_h(unsigned n, unsigned ell, double zeta) {
    const double sign = ((ell / 2) % 2 == 0) ? 1.0 : -1.0;
    return sign * std::sqrt(2.0 * kPiP * std::pow(zeta, 1.5) * gamma_h(double(n) + 1.5) /
                            (std::pow(2.0, int(ell)) * fact_h(n - ell)));
}

// ------------------------------------------------------------- kernels ----

// out[v][r] = sum_k M[r][k] c_v[k] (FP64) for the voxels of the list (or all
// voxels when vox is null). With skip_zero, an all-zero coefficient vector
// gives zeros (src/synth.rs:57-61). Workgroup = 8 waves x 64 lanes x VPL
// voxels per lane, their coefficients staged in LDS (coefficient-major);
// wave w takes row blocks of RB rows and reads M transposed and padded
// (MT[k][r], Rp >= R rows) with wave-uniform scalar loads, one RB-row
// stretch per k, each element feeding VPL FMAs per lane. A workgroup whose
// voxels are all zero (outside the mask) skips the products.
template <int RB, int VPL>
__global__ void __launch_bounds__(512) apply_rows_kernel(const float* __restrict__ coef, const int64_t* __restrict__ vox,
                                                         int64_t V, int K, const double* __restrict__ MT, int Rp,
                                                         int R, int skip_zero, float* __restrict__ out) {
    constexpr int NV = 64 * VPL;
    extern __shared__ float sc[];  // [K][NV + 1]
    __shared__ int nz[NV];
    __shared__ int any_nz;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t v0 = int64_t(blockIdx.x) * NV;
    for (int t = threadIdx.x; t < NV; t += 512) nz[t] = 0;
    if (threadIdx.x == 0) any_nz = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < NV * K; t += 512) {
        const int vj = t / K, k = t % K;
        const int64_t v = v0 + vj;
        float x = 0.f;
        if (v < V) x = coef[(vox ? vox[v] : v) * int64_t(K) + k];
        sc[k * (NV + 1) + vj] = x;
        if (x != 0.f) {
            nz[vj] = 1;
            any_nz = 1;
        }
    }
    __syncthreads();
    const bool skip = skip_zero && any_nz == 0;
    for (int r0 = wave * RB; r0 < R; r0 += 8 * RB) {
        double acc[VPL][RB];
#pragma unroll
        for (int i = 0; i < VPL; ++i)
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) acc[i][rb] = 0.0;
        if (!skip) {
            for (int k = 0; k < K; ++k) {
                double ck[VPL];
#pragma unroll
                for (int i = 0; i < VPL; ++i) ck[i] = double(sc[k * (NV + 1) + lane + 64 * i]);
                const double* mrow = MT + size_t(k) * Rp + r0;
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
                    const double m = mrow[rb];
#pragma unroll
                    for (int i = 0; i < VPL; ++i) acc[i][rb] = fma(m, ck[i], acc[i][rb]);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < VPL; ++i) {
            const int64_t v = v0 + lane + 64 * i;
            if (v >= V) continue;
            const int64_t idx = vox ? vox[v] : v;
            const bool zero = skip_zero && nz[lane + 64 * i] == 0;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
                if (r0 + rb < R) out[idx * R + r0 + rb] = zero ? 0.f : float(acc[i][rb]);
        }
    }
}

// MT[k][r] = M[r][k] for r < R, zero for R <= r < Rp.
__global__ void transpose_pad_kernel(const double* __restrict__ M, int R, int K, int Rp, double* __restrict__ MT) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(K) * Rp) return;
    const int k = int(t / Rp), r = int(t % Rp);
    MT[t] = r < R ? M[size_t(r) * K + k] : 0.0;
}

// Associated Legendre P_l^m(x) with the Condon-Shortley phase (src/math.rs:70-101).
__device__ double legendre_pf(int l, int m, double x) {
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

__device__ double fact_d(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= double(i);
    return f;
}

// BrainSuite real-SH slot of block l at (theta, phi) (src/basis/shore.rs:76-115).
__device__ double bs_sh(int l, int slot, double theta, double phi) {
    if (l == 0) return 1.0 / sqrt(4.0 * kPiP);
    const int k = slot < l ? l - slot : slot - l;
    const double norm = sqrt((2.0 * l + 1.0) / (4.0 * kPiP) * fact_d(l - k) / fact_d(l + k));
    const double p = norm * legendre_pf(l, k, cos(phi));
    if (slot < l) return sqrt(2.0) * p * cos(k * theta);
    if (slot == l) return p;
    return sqrt(2.0) * ((k % 2 == 1) ? 1.0 : -1.0) * p * sin(k * theta);
}

// Raw scalars per masked voxel, as f32 times the unit factors:
// raw[q * V + v], q = rtop, rtap, rtpp, msd, qiv, ng.
__global__ void micro_kernel(const float* __restrict__ coef, const int64_t* __restrict__ vox, int64_t V, int K,
                             MicroTables T, const float* __restrict__ dirs, float* __restrict__ raw) {
    const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int64_t idx = vox ? vox[v] : v;
    const float* c = coef + idx * int64_t(K);
    const double nan = __builtin_nan("");
    double e0 = 0.0;
    for (int n = 0; n <= T.R; ++n) e0 += double(c[T.n00[n]]) * T.e0w[n];
    double out[6];
    if (!isfinite(e0) || e0 == 0.0) {
        for (int q = 0; q < 6; ++q) out[q] = nan;
    } else {
        double a_rtop = 0.0, a_msd = 0.0, a_qiv = 0.0;
        for (int n = 0; n <= T.R; ++n) {
            const double cn = double(c[T.n00[n]]) / e0;
            a_rtop += cn * T.rtopw[n];
            a_msd += cn * T.msdw[n];
            a_qiv += cn * T.qivw[n];
        }
        out[0] = fmax(a_rtop, 0.0);
        out[3] = fmax(a_msd, 0.0);
        const double q2 = T.qiv_pre * a_qiv;
        out[4] = (!isfinite(q2) || q2 <= 0.0) ? nan : 1.0 / q2;
        double total = 0.0;
        for (int k = 0; k < K; ++k) {
            const double cn = double(c[k]) / e0;
            total += cn * cn;
        }
        if (total == 0.0) {
            out[5] = 0.0;
        } else {
            const double c0 = double(c[0]) / e0;
            out[5] = sqrt(fmax(1.0 - (c0 * c0) / total, 0.0));
        }
        // RTAP / RTPP at the voxel's peak direction; a zero or non-finite
        // direction is a missing peak (python/src/models.rs:318-326).
        bool have = false;
        double d[3] = {0.0, 0.0, 0.0};
        if (dirs) {
            const float* p = dirs + idx * 3;
            const bool any = p[0] != 0.f || p[1] != 0.f || p[2] != 0.f;
            const bool fin = isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]);
            if (any && fin) {
                have = true;
                for (int t = 0; t < 3; ++t) d[t] = double(p[t]);
            }
        }
        const double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (!have || !isfinite(nrm) || nrm == 0.0) {
            out[1] = out[2] = nan;
        } else {
            const double inv = 1.0 / nrm;
            const double x = d[0] * inv, y = d[1] * inv, z = d[2] * inv;
            const double r = sqrt(x * x + y * y + z * z);
            const double theta = atan2(y, x);
            const double phi = r > 0.0 ? acos(fmin(1.0, fmax(-1.0, z / r))) : 0.0;
            double ap = 0.0, pp = 0.0;
            for (int b = 0; b < T.nblocks; ++b) {
                const int l = T.blk_l[b], start = T.blk_start[b];
                const double sa = T.scale_ap[b], sp = T.scale_pp[b];
                double inner = 0.0;
                for (int s = 0; s < 2 * l + 1; ++s) inner += (double(c[start + s]) / e0) * bs_sh(l, s, theta, phi);
                if (sa != 0.0) ap += sa * inner;
                if (sp != 0.0) pp += sp * inner;
            }
            out[1] = fmax(T.pref_ap * ap, 0.0);
            out[2] = fmax(T.pref_pp * pp, 0.0);
        }
    }
    for (int q = 0; q < 6; ++q) raw[q * V + v] = float(out[q]) * T.unit[q];
}

// maps[q][vox[v]] = raw[q][v], NaN where finite and above thr[q].
__global__ void micro_scatter_kernel(const float* __restrict__ raw, const int64_t* __restrict__ vox, int64_t V,
                                     const MicroThresholds* __restrict__ th, float* m0, float* m1, float* m2,
                                     float* m3, float* m4, float* m5) {
    const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int64_t idx = vox ? vox[v] : v;
    float* maps[6] = {m0, m1, m2, m3, m4, m5};
    for (int q = 0; q < 6; ++q) {
        float x = raw[q * V + v];
        if (th->active[q] && isfinite(x) && x > th->thr[q]) x = __builtin_nanf("");
        if (maps[q]) maps[q][idx] = x;
    }
}

// Order-preserving map of a float's bits to an unsigned key, and back.
__device__ __forceinline__ uint32_t float_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k ^ 0x80000000u) : ~k);
}

// Fit-failure thresholds (src/io/microstructure.rs:215-268, 303-326): for
// each map q, factor x the value of rank round((n - 1) 0.99) among the n
// finite values in ascending order, found by a four-pass radix selection over
// 8-bit digits of the order-preserving keys (the same element a full sort, or
// nth_element, puts at that rank). No finite value: inactive.
// The selection runs over kSelSlices workgroups per map (one workgroup per
// map, all passes in one launch, kept six CUs busy: 0.39 ms for the six maps
// of the whole-brain volume). SelState holds each map's count, rank, digit prefix and the
// histogram of the current pass in device memory; `sel_hist_kernel` adds a
// slice's histogram into it, `sel_pick_kernel` takes the digit and clears it.
struct SelState {
    unsigned long long n;
    long long rank;
    uint32_t prefix, mask;
    unsigned hist[256];
};
constexpr int kSelSlices = 32;  // workgroups per map

__global__ void sel_init_kernel(SelState* __restrict__ st) {
    SelState& S = st[blockIdx.x];
    for (int b = threadIdx.x; b < 256; b += blockDim.x) S.hist[b] = 0u;
    if (threadIdx.x == 0) {
        S.n = 0ull;
        S.rank = 0;
        S.prefix = 0u;
        S.mask = 0u;
    }
}

__global__ void __launch_bounds__(256) sel_count_kernel(const float* __restrict__ raw, int64_t V,
                                                        SelState* __restrict__ st) {
    const int q = blockIdx.y;
    const float* x = raw + q * V;
    __shared__ unsigned long long cnt;
    if (threadIdx.x == 0) cnt = 0ull;
    __syncthreads();
    unsigned long long c = 0;
    for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += int64_t(gridDim.x) * blockDim.x)
        c += isfinite(x[v]) ? 1 : 0;
    atomicAdd(&cnt, c);
    __syncthreads();
    if (threadIdx.x == 0 && cnt) atomicAdd(&st[q].n, cnt);
}

__global__ void __launch_bounds__(256) sel_hist_kernel(const float* __restrict__ raw, int64_t V, int shift,
                                                       SelState* __restrict__ st) {
    const int q = blockIdx.y;
    const float* x = raw + q * V;
    __shared__ unsigned hist[256];
    for (int b = threadIdx.x; b < 256; b += blockDim.x) hist[b] = 0u;
    __syncthreads();
    const uint32_t prefix = st[q].prefix, mask = st[q].mask;
    for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += int64_t(gridDim.x) * blockDim.x) {
        const float f = x[v];
        if (!isfinite(f)) continue;
        const uint32_t k = float_key(f);
        if ((k & mask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < 256; b += blockDim.x)
        if (hist[b]) atomicAdd(&st[q].hist[b], hist[b]);
}

// One wave per map: lane l holds bins 4l..4l+3; the digit is the first bin
// whose running count passes the rank (capped at 255), found with a prefix
// sum across the lanes instead of a serial scan of 256 bins.
__global__ void __launch_bounds__(64) sel_pick_kernel(int shift, float factor, SelState* __restrict__ st,
                                                      MicroThresholds* __restrict__ th) {
    const int q = blockIdx.x, lane = threadIdx.x;
    SelState& S = st[q];
    const long long n = (long long)S.n;
    if (n == 0) {
        if (lane == 0) th->active[q] = 0;
        return;
    }
    const long long r0 = shift == 24 ? std::min((long long)llround(0.99 * double(n - 1)), n - 1) : S.rank;
    unsigned hb[4];
    long long sum = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        hb[i] = S.hist[4 * lane + i];
        sum += hb[i];
    }
    long long incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const long long t = __shfl_up(incl, off);
        if (lane >= off) incl += t;
    }
    const unsigned long long past = __ballot(incl > r0);
    const int sel = past ? __builtin_ctzll(past) : 63;
#pragma unroll
    for (int i = 0; i < 4; ++i) S.hist[4 * lane + i] = 0u;  // cleared for the next pass
    if (lane != sel) return;
    long long r = r0 - (incl - sum);
    int d = 4 * lane;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (d == 255 || r < (long long)hb[i]) break;
        r -= hb[i];
        ++d;
    }
    S.rank = r;
    S.prefix |= uint32_t(d) << shift;
    S.mask |= 255u << shift;
    if (shift == 0) {
        th->active[q] = 1;
        th->thr[q] = factor * key_float(S.prefix);
    }
}

}  // namespace

int n_tournier(unsigned lmax) { return int((lmax + 1) * (lmax + 2) / 2); }

// src/odf.rs:35-102 — row-major n_tournier(lmax) x K.
void tournier_matrix_host(unsigned R, double zeta, unsigned lmax, double* out) {
    const int rows = n_tournier(lmax), K = n_coeffs(R);
    std::fill(out, out + size_t(rows) * K, 0.0);
    int col = 0;
    for (unsigned n = 0; n <= R; ++n)
        for (unsigned l = 0; l <= n; l += 2) {
            const int bs = int(2 * l + 1);
            if (l <= lmax) {
                const double nf = double(n), lf = double(l);
                const double sign = ((int(n) - int(l) / 2) % 2 == 0) ? 1.0 : -1.0;
                const double c_nl = sign * std::sqrt(2.0 * fact_h(n - l) / (std::pow(zeta, 1.5) * gamma_h(nf + 1.5)));
                const double g_nl = (gamma_h(lf / 2.0 + 1.5) * gamma_h(nf + 1.5)) / (gamma_h(lf + 1.5) * fact_h(n - l)) *
                                    std::pow(0.5, -lf / 2.0 - 1.5);
                const double radial = c_nl * g_nl * hyp2f1_h(-nf + lf, lf / 2.0 + 1.5, lf + 1.5, 2.0);
                const int base = int(l * (l + 1) / 2), kl = int(l);
                for (int k = 0; k < bs; ++k) {
                    int row;
                    double s = 1.0;
                    if (k < kl) {
                        row = base + (kl - k);
                    } else if (k == kl) {
                        row = base;
                    } else {
                        const int j = k - kl;
                        row = base - j;
                        s = (j % 2 == 1) ? 1.0 : -1.0;
                    }
                    out[size_t(row) * K + col + k] = radial * s;
                }
            }
            col += bs;
        }
}

// Per-mode weights of src/scalars/microstructure.rs:79-322.
void micro_tables_host(unsigned R, double zeta, int units, MicroTables& T) {
    T = MicroTables{};
    T.R = int(R);
    int start = 0, b = 0;
    for (unsigned n = 0; n <= R; ++n)
        for (unsigned l = 0; l <= n; l += 2) {
            if (l == 0) {
                T.n00[n] = start;
                const double sign = (n % 2 == 0) ? 1.0 : -1.0;
                T.e0w[n] = kappa_h(zeta, n, 0) * laguerre0_h(n, 0.5) / std::sqrt(4.0 * kPiP);
                T.rtopw[n] = sign * std::sqrt(16.0 * kPiP * std::pow(zeta, 1.5) * gamma_h(double(n) + 1.5) / fact_h(n));
                T.msdw[n] = sign *
                            std::sqrt(9.0 * gamma_h(double(n) + 1.5) /
                                      (8.0 * std::pow(kPiP, 6) * std::pow(zeta, 3.5) * fact_h(n))) *
                            hyp2f1_h(-double(n), 2.5, 1.5, 2.0);
                T.qivw[n] = kappa_h(zeta, n, 0) * gamma_h(double(n) + 1.5) / fact_h(n) *
                            hyp2f1_h(-double(n), 2.5, 1.5, 2.0);
            }
            const unsigned j = n - l + 1;
            const double alpha = alpha_factor_h(n, l, zeta);
            T.blk_l[b] = int(l);
            T.blk_start[b] = start;
            T.scale_ap[b] = rtap_kappa_h(j, l) / alpha;
            T.scale_pp[b] = rtpp_kappa_h(j, l) / alpha;
            ++b;
            start += int(2 * l + 1);
        }
    T.nblocks = b;
    T.qiv_pre = 6.0 * std::sqrt(2.0 * kPiP) * std::pow(zeta, 2.5);
    const double mu = 1.0 / (2.0 * kPiP * std::sqrt(zeta));
    T.pref_ap = 1.0 / (mu * mu);
    T.pref_pp = 1.0 / mu;
    const float um[6] = {1e-9f, 1e-6f, 1e-3f, 1e6f, 1e15f, 1.0f};  // src/io/microstructure.rs:168-178
    for (int q = 0; q < 6; ++q) T.unit[q] = units == 0 ? um[q] : 1.0f;
}

void launch_apply_rows(const float* coef, const int64_t* vox, int64_t V, int K, const double* M, int R, bool skip_zero,
                       float* out, hipStream_t s) {
    if (V <= 0) return;
    constexpr int VPL = 2;
    const bool small = R <= 64;  // 8-row blocks keep more waves busy on short matrices
    const int RB = small ? 8 : 16;
    const int Rp = (R + RB - 1) / RB * RB;
    DevBuf<double> MT(size_t(K) * Rp);
    hipLaunchKernelGGL(transpose_pad_kernel, dim3(unsigned((int64_t(K) * Rp + 255) / 256)), dim3(256), 0, s, M, R, K,
                       Rp, MT.p);
    SF_HIP(hipGetLastError());
    const size_t lds = size_t(K) * (64 * VPL + 1) * sizeof(float);
    auto kern = small ? apply_rows_kernel<8, VPL> : apply_rows_kernel<16, VPL>;
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(lds));
    hipLaunchKernelGGL(kern, dim3(unsigned((V + 64 * VPL - 1) / (64 * VPL))), dim3(512), lds, s, coef, vox, V, K, MT.p,
                       Rp, R, int(skip_zero), out);
    SF_HIP(hipGetLastError());
    sync_if_unscoped(s);  // MT is released at scope exit
}

void launch_micro(const float* coef, const int64_t* vox, int64_t V, int K, const MicroTables& T, const float* dirs,
                  float* raw, hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(micro_kernel, dim3(unsigned((V + 127) / 128)), dim3(128), 0, s, coef, vox, V, K, T, dirs, raw);
    SF_HIP(hipGetLastError());
}

void launch_micro_thresholds(const float* raw, int64_t V, float factor, MicroThresholds* d_th, hipStream_t s) {
    if (V <= 0) return;
    DevBuf<SelState> st(6);
    hipLaunchKernelGGL(sel_init_kernel, dim3(6), dim3(256), 0, s, st.p);
    SF_HIP(hipGetLastError());
    hipLaunchKernelGGL(sel_count_kernel, dim3(kSelSlices, 6), dim3(256), 0, s, raw, V, st.p);
    SF_HIP(hipGetLastError());
    for (int shift = 24; shift >= 0; shift -= 8) {
        hipLaunchKernelGGL(sel_hist_kernel, dim3(kSelSlices, 6), dim3(256), 0, s, raw, V, shift, st.p);
        SF_HIP(hipGetLastError());
        hipLaunchKernelGGL(sel_pick_kernel, dim3(6), dim3(64), 0, s, shift, factor, st.p, d_th);
        SF_HIP(hipGetLastError());
    }
    sync_if_unscoped(s);  // st is released at scope exit
    SF_HIP(hipGetLastError());
}

void launch_micro_scatter(const float* raw, const int64_t* vox, int64_t V, const MicroThresholds* d_th,
                          float* const* maps, hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(micro_scatter_kernel, dim3(unsigned((V + 255) / 256)), dim3(256), 0, s, raw, vox, V, d_th,
                       maps[0], maps[1], maps[2], maps[3], maps[4], maps[5]);
    SF_HIP(hipGetLastError());
}

}  // namespace shorefit
