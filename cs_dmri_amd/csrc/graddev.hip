// Per-voxel SHORE fits under a gradient-deviation field.
//
// A deviated voxel v has its own gradient table (src/graddev.rs:160-185)
//     q = T_v^T g_vox / |g_vox|,  b_eff = b |q|^2,  dir = normalize(R q)
// and therefore its own design A_v (N x K) and Lipschitz constant. The
// reference pattern for such fits is a per-voxel design build followed by the
// ordinary solver (src/dti.rs:393-407); here A_v is never stored anywhere but
// LDS, and the fit runs in Gram form on G_v = A_v^T A_v:
//
//   gd_gram_kernel      one workgroup per voxel. Effective table and per-row
//                       geometry (one thread per row), then RC design rows at
//                       a time, each built by one thread from compile-time
//                       unrolled Laguerre / Legendre recurrences into an FP64
//                       row block in LDS; b_v += A^T s (FP64) and the upper
//                       16x16 tiles of G_v += A^T A on the FP64 MFMA. G_v is
//                       written as FP32 (Kp x Kp, both triangles), b_v, |s|^2
//                       and the sum of squares about the mean as FP64, to a
//                       workspace.
//   gd_fista_kernel     persistent; a workgroup claims voxels from a counter.
//                       Lane (row group rb, column group cb) keeps the RB x CB
//                       block of G_v in registers. Lanczos (kLanczos steps,
//                       FP64 dot products) gives sigma_max(A_v)^2, then FISTA
//                       (src/solver/fista.rs:70-143) in Gram form
//                           v = y + b/(L N) - (G y)/(L N)
//                       with one workgroup barrier per iteration; the
//                       residual |s - A_v x|^2 = s's - 2 x'b + x'G x in FP64.
//   gd_finish_kernel    R^2, residual, BIC and the maps per voxel.
//
// FP64 is used where the shared-design path uses it: design, Gram matrix,
// right-hand side, Lipschitz estimate, residual. The iteration itself is FP32.
#include <type_traits>

#include "common.h"
#include "special.h"

namespace shorefit {

// ------------------------------------------------------------ host tables ----

namespace {
double fact_h(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= double(i);
    return f;
}
double gamma_n32_h(int n) {
    double g = std::sqrt(kPiD) * 0.5;
    for (int i = 1; i <= n; ++i) g *= (double(i) + 0.5);
    return g;
}
}  // namespace

GdBasis gd_basis_host(uint32_t R, double zeta) {
    GdBasis b{};
    b.R = int(R);
    b.zeta = zeta;
    int col = 0, p = 0;
    for (int n = 0; n <= int(R); ++n)
        for (int l = 0; l <= n; l += 2) {
            b.pair_n[p] = uint8_t(n);
            b.pair_l[p] = uint8_t(l);
            b.pair_col0[p] = int16_t(col);
            // src/basis/shore.rs:56-61
            b.kappa[p] = std::sqrt(2.0 * fact_h(n - l) / (std::pow(zeta, 1.5) * gamma_n32_h(n)));
            for (int slot = 0; slot <= 2 * l; ++slot) {
                b.col_pair[col] = uint8_t(p);
                b.col_ang[col] = uint8_t(l * (l - 1) / 2 + slot);
                ++col;
            }
            ++p;
        }
    b.K = col;
    b.P = p;
    int a = 0;
    for (int l = 0; l <= int(R); l += 2)
        for (int slot = 0; slot <= 2 * l; ++slot, ++a) {
            // src/basis/shore.rs:76-115 (BrainSuite block)
            b.ang_l[a] = uint8_t(l);
            if (l == 0) {
                b.ang_kind[a] = 0;
                b.ang_k[a] = 0;
                b.ang_c[a] = 1.0 / std::sqrt(4.0 * kPiD);
                continue;
            }
            const int k = slot < l ? l - slot : slot - l;
            const double norm = std::sqrt((2.0 * l + 1.0) / (4.0 * kPiD) * fact_h(l - k) / fact_h(l + k));
            b.ang_k[a] = uint8_t(k);
            if (slot < l) {
                b.ang_kind[a] = 1;
                b.ang_c[a] = std::sqrt(2.0) * norm;
            } else if (slot == l) {
                b.ang_kind[a] = 2;
                b.ang_c[a] = norm;
            } else {
                b.ang_kind[a] = 3;
                b.ang_c[a] = std::sqrt(2.0) * ((k % 2 == 1) ? 1.0 : -1.0) * norm;
            }
        }
    b.Q = a;
    return b;
}

namespace {

typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int kLanczos = 24;  // Lanczos steps for sigma_max^2 (~1e-8 relative)

constexpr int n_pairs_c(int R) {
    int p = 0;
    for (int n = 0; n <= R; ++n) p += n / 2 + 1;
    return p;
}
constexpr int n_coeffs_c(int R) {
    int k = 0;
    for (int n = 0; n <= R; ++n)
        for (int l = 0; l <= n; l += 2) k += 2 * l + 1;
    return k;
}
constexpr int n_ang_c(int R) { return (R + 1) * (R + 2) / 2; }  // sum over even l <= R of 2l+1
struct RowGeom {
    double u, e, cp, ct, st;
};

// Effective table row i of deviation T (row-major, g_eff = T^T g): b and
// direction (src/graddev.rs:160-185), q-value (src/qspace.rs:130-149), and
// the spherical coordinates of the design (src/basis/shore.rs:129-138,
// src/math.rs:157-166): u = q^2 / zeta, e = exp(-u / 2), cos(phi), cos and
// sin of theta.
__device__ inline RowGeom gd_row(const GdTable& t, const double (&T)[9], int i, double zeta) {
    const double b = t.bvals[i];
    const double g0 = t.gvox[3 * i], g1 = t.gvox[3 * i + 1], g2 = t.gvox[3 * i + 2];
    const double gn = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
    double be, d0, d1, d2;
    if (!(b > t.b0_thr) || !(gn > 1e-12)) {
        be = b;
        d0 = t.gnom[3 * i];
        d1 = t.gnom[3 * i + 1];
        d2 = t.gnom[3 * i + 2];
    } else {
        const double u0 = g0 / gn, u1 = g1 / gn, u2 = g2 / gn;
        const double q0 = T[0] * u0 + T[3] * u1 + T[6] * u2;
        const double q1 = T[1] * u0 + T[4] * u1 + T[7] * u2;
        const double q2 = T[2] * u0 + T[5] * u1 + T[8] * u2;
        const double e0 = t.R[0] * q0 + t.R[1] * q1 + t.R[2] * q2;
        const double e1 = t.R[3] * q0 + t.R[4] * q1 + t.R[5] * q2;
        const double e2 = t.R[6] * q0 + t.R[7] * q1 + t.R[8] * q2;
        const double dn = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
        be = b * (q0 * q0 + q1 * q1 + q2 * q2);
        d0 = e0 / dn;
        d1 = e1 / dn;
        d2 = e2 / dn;
    }
    const double qv = (be <= t.b0_thr) ? 0.0 : sqrt(be / t.qdenom);
    const double x = d0 * qv, y = d1 * qv, z = d2 * qv;
    const double r = sqrt(x * x + y * y + z * z);
    double th = atan2(y, x);
    if (isnan(th)) th = 0.0;
    const double ph = r > 0.0 ? acos(fmin(1.0, fmax(-1.0, z / r))) : 0.0;
    RowGeom g;
    g.u = r * r / zeta;
    g.e = exp(-g.u * 0.5);
    g.cp = cos(ph);
    sincos(th, &g.st, &g.ct);
    return g;
}

// Compile-time loop: f(std::integral_constant<int, i>) for i = B .. E-1.
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Basis index set (src/basis/shore.rs:29-44): n = 0..R, even l <= n, 2l + 1
// slots (m = -l..l) per (n, l). Pair (n, l) is entry pair_c(n, l) of the
// radial factors, its first column col0_c(n, l); slot s of even l is entry
// ang_c(l, s) of the angular factors.
constexpr int pair_c(int n, int l) {
    int p = 0;
    for (int j = 0; j < n; ++j) p += j / 2 + 1;
    return p + l / 2;
}
constexpr int col0_c(int n, int l) {
    int c = 0;
    for (int j = 0; j < n; ++j)
        for (int q = 0; q <= j; q += 2) c += 2 * q + 1;
    for (int q = 0; q < l; q += 2) c += 2 * q + 1;
    return c;
}
constexpr int ang_c(int l, int slot) { return l * (l - 1) / 2 + slot; }

// Tile t of the upper triangle (I <= J) of a KT x KT tile grid in row-major
// order: its block row and column; (0, 0) for t past the last tile.
constexpr int tile_row(int KT, int t) {
    int I = 0;
    while (I < KT && t >= KT - I) t -= KT - I++;
    return I < KT ? I : 0;
}
constexpr int tile_col(int KT, int t) {
    int I = 0;
    while (I < KT && t >= KT - I) t -= KT - I++;
    return I < KT ? I + t : 0;
}

// One design row of A_v from its geometry, written to arow[0 .. K-1]
// (src/basis/shore.rs:122-178): column (n, l, slot) is
//     kappa(n, l) L_{n-l}^{(l+1/2)}(u) u^(l/2) e^(-u/2) * c_l^slot P_l^k(cos phi) trig(k theta)
// with k = |l - slot| and trig = cos for slot < l, sin for slot > l
// (src/basis/shore.rs:56-115). The loops are unrolled at compile time, so the
// factors live in registers and every column index is a constant:
//   radial  for each even l, the Laguerre recurrence in n - l (the recurrence
//           of laguerre_d) over n = l..R;
//   angular for k = 0..R, cos / sin (k theta) and P_k^k by carrying the
//           k-1 values one step, then the upward Legendre recurrence in l (the
//           recurrence of legendre_d); each even l >= max(k, 2) (and l = 0 for
//           k = 0) emits its columns for every n >= l.
// Divisions by the recurrence index are multiplications by its reciprocal.
// kap[p] = kappa of pair p and angc[a] = normalisation of angular entry a are
// read from LDS (uniform addresses): held in scalar registers, the 70-128
// constants would not fit.
template <int R>
__device__ __forceinline__ void build_row(const double* kap, const double* angc, const RowGeom& g,
                                          double* __restrict__ arow) {
    constexpr int P = n_pairs_c(R);
    double rad[P];
    static_for<0, R / 2 + 1>([&](auto lh) {
        constexpr int l = 2 * decltype(lh)::value;
        constexpr double a = l + 0.5;
        double ep = g.e;
#pragma unroll
        for (int j = 0; j < l / 2; ++j) ep *= g.u;
        double lp = 0.0, lc = 1.0;
        static_for<l, R + 1>([&](auto nn) {
            constexpr int n = decltype(nn)::value, k = n - l;
            if constexpr (k == 1) {
                lp = 1.0;
                lc = 1.0 + a - g.u;
            } else if constexpr (k >= 2) {
                constexpr double c1 = 2.0 * (k - 1) + 1.0 + a, c2 = (k - 1) + a, rk = 1.0 / k;
                const double nw = ((c1 - g.u) * lc - c2 * lp) * rk;
                lp = lc;
                lc = nw;
            }
            rad[pair_c(n, l)] = kap[pair_c(n, l)] * lc * ep;
        });
        __builtin_amdgcn_sched_barrier(0);
    });

    // the columns of angular slot `slot` of even l, value av
    auto emit = [&](auto lc_, auto slot_, double av) {
        constexpr int l = decltype(lc_)::value, slot = decltype(slot_)::value;
        static_for<l, R + 1>([&](auto nn) {
            constexpr int n = decltype(nn)::value;
            arow[col0_c(n, l) + slot] = rad[pair_c(n, l)] * av;
        });
    };
    const double x = g.cp;
    const double sq = sqrt(fmax((1.0 - x) * (1.0 + x), 0.0));
    double ck = 1.0, sk = 0.0, pmm = 1.0;
    static_for<0, R + 1>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        if constexpr (k > 0) {
            const double cn = ck * g.ct - sk * g.st;
            sk = sk * g.ct + ck * g.st;
            ck = cn;
            pmm *= -(2.0 * k - 1.0) * sq;
        }
        // P_l^k for l = k..R; emits even l
        auto at = [&](auto ll, double pv) {
            constexpr int l = decltype(ll)::value;
            if constexpr (l == 0) {
                emit(std::integral_constant<int, 0>{}, std::integral_constant<int, 0>{}, angc[0]);
            } else if constexpr (l % 2 == 0) {
                if constexpr (k == 0) {
                    emit(ll, std::integral_constant<int, l>{}, angc[ang_c(l, l)] * pv);
                } else {
                    emit(ll, std::integral_constant<int, l - k>{}, angc[ang_c(l, l - k)] * pv * ck);
                    emit(ll, std::integral_constant<int, l + k>{}, angc[ang_c(l, l + k)] * pv * sk);
                }
            }
        };
        at(std::integral_constant<int, k>{}, pmm);
        if constexpr (k + 1 <= R) {
            double p0 = pmm, p1 = x * (2.0 * k + 1.0) * pmm;
            at(std::integral_constant<int, k + 1>{}, p1);
            static_for<k + 2, R + 1>([&](auto ll) {
                constexpr int l = decltype(ll)::value;
                constexpr double c1 = 2.0 * l - 1.0, c2 = l + k - 1.0, rl = 1.0 / (l - k);
                const double pl = (c1 * x * p1 - c2 * p0) * rl;
                p0 = p1;
                p1 = pl;
                at(ll, pl);
            });
        }
        // keeps the scheduler from hoisting every constant of the row at
        // once (which would not fit in registers)
        __builtin_amdgcn_sched_barrier(0);
    });
}

__device__ inline void load_T(const float* field, int64_t v, double (&T)[9]) {
#pragma unroll
    for (int k = 0; k < 9; ++k) T[k] = double(field[v * 9 + k]);
}

// ------------------------------------------------------------ reductions ----
// Sum over the 16 lanes of a row (DPP): every lane ends with the same value.
__device__ __forceinline__ float row16_sum(float x) {
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));  // quad [1,0,3,2]
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x4E, 0xF, 0xF, false));  // quad [2,3,0,1]
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x141, 0xF, 0xF, false)); // half mirror
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x140, 0xF, 0xF, false)); // row mirror
    return x;
}

// Sum over the lanes of each group of CG = 8 or 16 lanes (DPP).
template <int CG>
__device__ __forceinline__ float group_sum(float x) {
    static_assert(CG == 8 || CG == 16, "lane group");
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));  // quad [1,0,3,2]
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x4E, 0xF, 0xF, false));  // quad [2,3,0,1]
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x141, 0xF, 0xF, false)); // half mirror
    if constexpr (CG == 16)
        x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x140, 0xF, 0xF, false));  // row mirror
    return x;
}

// Lanes per column group of the FISTA kernel: 8 (three DPP steps per row
// sum instead of four, twice the columns per lane) where the rows divide
// evenly and a lane keeps at most 20 columns, else 16. Measured on 250,176
// deviated voxels with 258 gradients: orders 6 and 8 are 18% / 14% faster
// with 8, order 10 (32 columns per lane) 6% slower.
constexpr int fista_cg(int KP, int W) {
    return (KP % (8 * W) == 0 && KP / (8 * W) <= 8 && KP % 8 == 0 && KP / 8 <= 20) ? 8 : 16;
}

// Wave sum of a value that is uniform within each row of 16 lanes.
__device__ __forceinline__ float rows4_sum(float x) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 0)) +
           __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 16)) +
           __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 32)) +
           __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 48));
}

__device__ __forceinline__ double wave_sum_d(double x) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) x += __shfl_xor(x, m);
    return x;
}
__device__ __forceinline__ double wave_max_d(double x) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) x = fmax(x, __shfl_xor(x, m));
    return x;
}

// ------------------------------------------------------------ Gram kernel ----
template <int R, int KP, int RC, int WAVES>
__global__ void __launch_bounds__(64 * WAVES, 2)
    gd_gram_kernel(const GdBasis B, const GdTable tab, const float* __restrict__ data, int64_t sig_stride,
                   const float* __restrict__ field, const int64_t* __restrict__ vox, int64_t n,
                   float* __restrict__ Gws, double* __restrict__ bws, double* __restrict__ sstat) {
    constexpr int K = n_coeffs_c(R), P = n_pairs_c(R), Q = n_ang_c(R);
    constexpr int KT = KP / 16, NT = KT * (KT + 1) / 2, TPW = (NT + WAVES - 1) / WAVES;
    constexpr int LDA = KP + 2;
    constexpr int NTH = 64 * WAVES;
    static_assert(KP % 16 == 0 && KP >= K && RC % 4 == 0 && RC <= 64 && KP <= NTH && Q <= NTH, "tile shape");
    const int64_t slot = blockIdx.x;
    if (slot >= n) return;
    const int64_t v = vox[slot];
    const int N = tab.N;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);

    extern __shared__ double smem[];
    double* s_kap = smem;        // P
    double* s_angc = s_kap + P;  // Q
    double* s_sig = s_angc + Q;  // N
    double* s_geo = s_sig + N;   // 5 x N: u, e, cos phi, cos theta, sin theta
    double* s_A = s_geo + 5 * N; // RC x LDA

    double T[9];
    load_T(field, v, T);
    // one thread per row: the effective table row and its geometry
    for (int i = tid; i < N; i += NTH) {
        const RowGeom g = gd_row(tab, T, i, B.zeta);
        s_geo[i] = g.u;
        s_geo[N + i] = g.e;
        s_geo[2 * N + i] = g.cp;
        s_geo[3 * N + i] = g.ct;
        s_geo[4 * N + i] = g.st;
        s_sig[i] = double(data[v * sig_stride + i]);
    }
    if (tid < P) s_kap[tid] = B.kappa[tid];
    if (tid < Q) s_angc[tid] = B.ang_c[tid];
    // the padding columns K..KP-1 stay zero
    if constexpr (KP > K) {
        for (int it = tid; it < RC * (KP - K); it += NTH) {
            const int r = it / (KP - K);
            s_A[r * LDA + K + (it - r * (KP - K))] = 0.0;
        }
    }
    __syncthreads();
    if (w == 0) {
        // |s|^2 and the total sum of squares about the mean (FP64), for the
        // residual by the Gram identity and R^2 (src/fit.rs:120-128)
        double sum = 0.0, sq = 0.0;
        for (int i = lane; i < N; i += 64) {
            sum += s_sig[i];
            sq += s_sig[i] * s_sig[i];
        }
        sum = wave_sum_d(sum);
        sq = wave_sum_d(sq);
        const double mean = sum / double(N);
        double sst = 0.0;
        for (int i = lane; i < N; i += 64) sst += (s_sig[i] - mean) * (s_sig[i] - mean);
        sst = wave_sum_d(sst);
        if (lane == 0) {
            sstat[2 * slot] = sq;
            sstat[2 * slot + 1] = sst;
        }
    }

    // tiles of this wave: t = w + WAVES j in row-major order of the upper
    // triangle (I <= J); the slots past the last tile (at most one per wave
    // for the shapes used) repeat tile (0, 0) and are not written, so the
    // MFMA loop has no branches
    int tI[TPW], tJ[TPW];
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
        tI[j] = tile_row(KT, w + WAVES * j);
        tJ[j] = tile_col(KT, w + WAVES * j);
    }
    doublex4 acc[TPW];
#pragma unroll
    for (int j = 0; j < TPW; ++j) acc[j] = doublex4{0.0, 0.0, 0.0, 0.0};
    double bacc = 0.0;

    for (int i0 = 0; i0 < N; i0 += RC) {
        __syncthreads();  // previous chunk's A reads done
        const int rmax = min(RC, N - i0);
        // one thread per row of the chunk, in one wave (rows past the end
        // are zero); the wave that builds rotates from chunk to chunk
        if (w == (i0 / RC) % WAVES && lane < RC) {
            double* arow = s_A + lane * LDA;
            if (lane < rmax) {
                const int i = i0 + lane;
                const RowGeom g{s_geo[i], s_geo[N + i], s_geo[2 * N + i], s_geo[3 * N + i], s_geo[4 * N + i]};
                build_row<R>(s_kap, s_angc, g, arow);
            } else {
                for (int c = 0; c < K; ++c) arow[c] = 0.0;
            }
        }
        __syncthreads();
        if (tid < KP) {
#pragma unroll 8
            for (int r = 0; r < rmax; ++r) bacc += s_A[r * LDA + tid] * s_sig[i0 + r];
        }
        // four rows per step (rows past rmax up to the next multiple of 4
        // are zero)
        const int ksteps = (rmax + 3) / 4;
        // the tile list is a compile-time constant per wave, so every LDS
        // read has an immediate offset and reads shared by two tiles of the
        // wave are issued once
        const double* rowp = s_A + (lane >> 4) * LDA + (lane & 15);
        static_for<0, WAVES>([&](auto wc) {
            constexpr int ww = decltype(wc)::value;
            if (w != ww) return;
            const double* rp = rowp;
            for (int kk = 0; kk < ksteps; ++kk, rp += 4 * LDA) {
                double a[TPW], b[TPW];
                static_for<0, TPW>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    a[j] = rp[16 * tile_row(KT, ww + WAVES * j)];
                    b[j] = rp[16 * tile_col(KT, ww + WAVES * j)];
                });
#pragma unroll
                for (int j = 0; j < TPW; ++j)
                    acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[j], b[j], acc[j], 0, 0, 0);
            }
        });
    }

    // D layout of the 16x16x4 f64 MFMA (measured on gfx950; it differs from
    // the f32 variant's): lane l, register r holds row 4 r + l / 16, column
    // l % 16 of the tile.
    float* Gv = Gws + size_t(slot) * KP * KP;
    const int rq = lane >> 4, cl = lane & 15;
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
        if (w + WAVES * j >= NT) continue;
        const int I = tI[j], J = tJ[j];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float gv = float(acc[j][r]);
            Gv[size_t(16 * I + 4 * r + rq) * KP + 16 * J + cl] = gv;
            if (I != J) Gv[size_t(16 * J + cl) * KP + 16 * I + 4 * r + rq] = gv;
        }
    }
    if (tid < KP) bws[size_t(slot) * KP + tid] = bacc;
}

// ------------------------------------------------------------ FISTA kernel ----

template <int KP, int W>
__global__ void __launch_bounds__(64 * W, (KP >= 256 ? 2 : 3)) gd_fista_kernel(const GdFistaArgs a) {
    constexpr int CG = fista_cg(KP, W), GW = 64 / CG;  // column groups, row groups per wave
    constexpr int RB = KP / (GW * W), CB = KP / CG;
    static_assert(RB * GW * W == KP && CB * CG == KP && RB <= CG, "block shape");
    __shared__ float s_vec[2][KP];
    __shared__ double s_red[2][W][3];
    __shared__ double s_al[kLanczos], s_be[kLanczos];
    __shared__ int64_t s_slot;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int rb = GW * w + lane / CG, cb = lane % CG;
    const int r0 = rb * RB, c0 = cb * CB;
    const bool owner = cb < RB;
    const int j = r0 + cb;  // the row this lane owns (owner lanes)
    const int N = a.N, K = a.K;
    int rp = 0;             // reduction buffer parity

    // Block sums of FP64 values (one barrier each).
    auto block_sum2 = [&](double x, double y, double& sx, double& sy) {
        x = wave_sum_d(x);
        y = wave_sum_d(y);
        if (lane == 0) {
            s_red[rp][w][0] = x;
            s_red[rp][w][1] = y;
        }
        __syncthreads();
        sx = 0.0;
        sy = 0.0;
#pragma unroll
        for (int q = 0; q < W; ++q) {
            sx += s_red[rp][q][0];
            sy += s_red[rp][q][1];
        }
        rp ^= 1;
    };
    auto block_sum3 = [&](double x, double y, double z, double& sx, double& sy, double& sz) {
        x = wave_sum_d(x);
        y = wave_sum_d(y);
        z = wave_sum_d(z);
        if (lane == 0) {
            s_red[rp][w][0] = x;
            s_red[rp][w][1] = y;
            s_red[rp][w][2] = z;
        }
        __syncthreads();
        sx = 0.0;
        sy = 0.0;
        sz = 0.0;
#pragma unroll
        for (int q = 0; q < W; ++q) {
            sx += s_red[rp][q][0];
            sy += s_red[rp][q][1];
            sz += s_red[rp][q][2];
        }
        rp ^= 1;
    };
    // (G v)_j for owner lanes, v = s_vec[buf]
    auto matvec = [&](const float (&g)[RB][CB], int buf) -> float {
        float vv[CB];
#pragma unroll
        for (int c = 0; c < CB; ++c) vv[c] = s_vec[buf][c0 + c];
        float accr[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            float s = 0.f;
#pragma unroll
            for (int c = 0; c < CB; ++c) s = fmaf(g[r][c], vv[c], s);
            accr[r] = group_sum<CG>(s);
        }
        float mine = accr[0];
#pragma unroll
        for (int r = 1; r < RB; ++r) mine = (cb == r) ? accr[r] : mine;
        return mine;
    };

    for (;;) {
        if (tid == 0) s_slot = int64_t(atomicAdd(a.counter, 1u));
        __syncthreads();
        const int64_t slot = s_slot;
        if (slot >= a.n) break;

        float g[RB][CB];
        {
            const float* Gp = a.G + size_t(slot) * KP * KP + size_t(r0) * KP + c0;
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int c = 0; c < CB; ++c) g[r][c] = Gp[size_t(r) * KP + c];
        }
        const double bj = (owner && j < K) ? a.b[size_t(slot) * KP + j] : 0.0;

        // --- Lanczos on G_v for sigma_max(A_v)^2 ------------------------------
        // deterministic start vector with every component non-zero
        const double q0 = (owner && j < K) ? 1.0 + 0.5 * double((j * 37) % 11) / 11.0 : 0.0;
        double qn2, amax, dummy;
        block_sum2(q0 * q0, 0.0, qn2, dummy);
        {
            const double mx = wave_max_d(fabs(bj));
            if (lane == 0) s_red[rp][w][0] = mx;
            __syncthreads();
            amax = 0.0;
#pragma unroll
            for (int q = 0; q < W; ++q) amax = fmax(amax, s_red[rp][q][0]);
            rp ^= 1;
        }
        double qv = q0 / sqrt(qn2);
        double qprev = 0.0, beta_prev = 0.0;
        int m = 0;
        for (; m < kLanczos && m < K; ++m) {
            if (owner) s_vec[m & 1][j] = float(qv);
            __syncthreads();
            const double gq = double(matvec(g, m & 1));
            double wv = owner ? gq - beta_prev * qprev : 0.0;
            double al;
            block_sum2(wv * qv, 0.0, al, dummy);
            wv -= al * qv;
            double bb;
            block_sum2(wv * wv, 0.0, bb, dummy);
            const double be = sqrt(bb);
            if (tid == 0) {
                s_al[m] = al;
                s_be[m] = be;
            }
            qprev = qv;
            beta_prev = be;
            if (!(be > 1e-30 * fabs(al))) {
                ++m;
                break;
            }
            qv = wv / be;
        }
        __syncthreads();
        // largest eigenvalue of the m x m tridiagonal: Sturm counts at 64
        // points per round (one per lane), 7 rounds of 65x narrowing
        double lam;
        {
            double lo = 0.0, hi = 0.0;
            for (int i = 0; i < m; ++i) {
                const double off = (i > 0 ? fabs(s_be[i - 1]) : 0.0) + (i + 1 < m ? fabs(s_be[i]) : 0.0);
                lo = i == 0 ? s_al[i] - off : fmin(lo, s_al[i] - off);
                hi = i == 0 ? s_al[i] + off : fmax(hi, s_al[i] + off);
            }
            for (int round = 0; round < 7; ++round) {
                const double x = lo + (hi - lo) * double(lane + 1) / 65.0;
                int cnt = 0;
                double d = 1.0;
                for (int i = 0; i < m; ++i) {
                    const double b2 = i > 0 ? s_be[i - 1] * s_be[i - 1] : 0.0;
                    d = s_al[i] - x - (i > 0 ? b2 / d : 0.0);
                    if (d == 0.0) d = -1e-300;
                    cnt += d < 0.0;
                }
                const uint64_t above = __ballot(cnt == m);  // lanes with x > lambda_max
                const int f = above ? __builtin_ctzll(above) : 64;
                const double nlo = f == 0 ? lo : lo + (hi - lo) * double(f) / 65.0;
                const double nhi = f == 64 ? hi : lo + (hi - lo) * double(f + 1) / 65.0;
                lo = nlo;
                hi = nhi;
            }
            lam = 0.5 * (lo + hi);
        }
        // src/solver/fista.rs:51,58: L = max(sigma_max^2 / N, eps)
        const double L = fmax(fmax(lam, 0.0) / double(N), 2.220446049250313e-16);
        const double step = 1.0 / L;
        double alpha_v = a.alpha;
        bool zero_voxel = false;
        if (a.ratio > 0.0) {  // src/solver/alpha.rs:47-53, :178-189
            const double amx = amax / double(N);
            alpha_v = amx * a.ratio;
            zero_voxel = amx == 0.0;
        }
        const float c = float(step / double(N));
        const float bs = float(bj * (step / double(N)));
        const float thr = float(step * alpha_v);

        // --- FISTA ----------------------------------------------------------
        float x = 0.f, y = 0.f;
        uint32_t its = 0;
        bool conv = false;
        int buf = 0;
        if (zero_voxel) {
            conv = true;
        } else {
            if (owner) s_vec[0][j] = 0.f;
            __syncthreads();
            for (uint32_t it = 0; it < a.max_iter; ++it) {
                const float gy = matvec(g, buf);
                float d2 = 0.f, n2 = 0.f;
                if (owner) {
                    const float vj = fmaf(-c, gy, y + bs);
                    float xn = vj > thr ? vj - thr : (vj < -thr ? vj + thr : 0.f);
                    if (a.nonneg) xn = fmaxf(xn, 0.f);
                    const float dx = xn - x;
                    const float yn = fmaf(a.mom[it], dx, xn);
                    d2 = dx * dx;
                    n2 = xn * xn;
                    x = xn;
                    y = yn;
                    s_vec[buf ^ 1][j] = yn;
                }
                d2 = rows4_sum(row16_sum(d2));
                n2 = rows4_sum(row16_sum(n2));
                if (lane == 0) {
                    s_red[rp][w][0] = d2;
                    s_red[rp][w][1] = n2;
                }
                __syncthreads();
                double D2 = 0.0, N2 = 0.0;
#pragma unroll
                for (int q = 0; q < W; ++q) {
                    D2 += s_red[rp][q][0];
                    N2 += s_red[rp][q][1];
                }
                rp ^= 1;
                buf ^= 1;
                its = it + 1;
                // src/solver/fista.rs:113-127: |dx| / max(|x+|, 1e-12) < tol,
                // compared in squares
                if (D2 < a.tol * a.tol * fmax(N2, 1e-24)) {
                    conv = true;
                    break;
                }
            }
        }
        const int64_t out = a.vox[slot];
        if (owner && j < K) a.coef[out * K + j] = x;
        // residual by the Gram identity in FP64 with the FP32 coefficients:
        // |s - A x|^2 = s's - 2 x'(A^T s) + x'(A^T A)x
        if (owner) s_vec[buf ^ 1][j] = x;
        __syncthreads();
        double xgx = 0.0;
        {
            float xc[CB];
#pragma unroll
            for (int c = 0; c < CB; ++c) xc[c] = s_vec[buf ^ 1][c0 + c];
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                float acc = 0.f;
#pragma unroll
                for (int c = 0; c < CB; ++c) acc = fmaf(g[r][c], xc[c], acc);
                xgx = fma(double(s_vec[buf ^ 1][r0 + r]), double(acc), xgx);
            }
        }
        double XGX, XB, NZ;
        block_sum3(xgx, owner ? double(x) * bj : 0.0, (owner && j < K && x != 0.f) ? 1.0 : 0.0, XGX, XB, NZ);
        if (tid == 0) {
            GdSlotOut so;
            so.rss = zero_voxel ? 0.0 : fmax(a.sstat[2 * slot] - 2.0 * XB + XGX, 0.0);
            so.alpha = float(alpha_v);
            so.its = its;
            so.nnz = uint16_t(NZ);
            so.flags = uint16_t((conv ? 1 : 0) | (zero_voxel ? 2 : 0));
            a.slot_out[slot] = so;
            if (a.lip) a.lip[slot] = L;
        }
        __syncthreads();  // s_slot / s_vec reuse
    }
}

// identity flags: 1 when max |T - I| < IDENTITY_TOL (src/graddev.rs:35-36, :89-93)
__global__ void gd_identity_kernel(const float* __restrict__ field, const int64_t* __restrict__ vox, int64_t n,
                                   uint8_t* __restrict__ ident, uint8_t* __restrict__ devi) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t v = vox ? vox[i] : i;
    double dev = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) dev = fmax(dev, fabs(double(field[v * 9 + k]) - ((k % 4 == 0) ? 1.0 : 0.0)));
    const bool id = dev < 1e-6;
    ident[i] = id;
    devi[i] = !id;
}

// Maps of the deviated voxels from the FISTA kernel's per-slot results
// (src/fit.rs:120-128 R^2, src/solver/alpha.rs:99-103 BIC, :178-189 the
// all-zero voxel), one thread per slot.
__global__ void gd_finish_kernel(const GdSlotOut* __restrict__ so, const double* __restrict__ sstat,
                                 const int64_t* __restrict__ vox, int64_t n, int N, GdMaps m) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GdSlotOut o = so[i];
    const int64_t v = vox[i];
    const bool zero_voxel = (o.flags & 2) != 0;
    const double sst = sstat[2 * i + 1];
    const double resid_l2 = sqrt(o.rss);
    if (m.resid) m.resid[v] = float(resid_l2);
    if (m.r2) m.r2[v] = float(sst > 0.0 ? 1.0 - o.rss / sst : 0.0);
    if (m.iters) m.iters[v] = o.its;
    if (m.kind) m.kind[v] = 1;
    if (m.alpha_map) m.alpha_map[v] = o.alpha;
    if (m.bic) {
        const double nn = double(N);
        const double rc = fmax(resid_l2 * resid_l2, 2.2250738585072014e-308);
        m.bic[v] = zero_voxel ? 0.f : float(nn * log(rc / nn) + log(nn) * double(o.nnz));
    }
}

__global__ void gather_i64_kernel(const int64_t* __restrict__ src, const int64_t* __restrict__ pos, int64_t n,
                                  int64_t* __restrict__ dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src ? src[pos[i]] : pos[i];
}

template <int R, int KP, int RC, int WAVES>
size_t gram_lds(int N) {
    return sizeof(double) * (size_t(n_pairs_c(R) + n_ang_c(R)) + size_t(6) * N + size_t(RC) * (KP + 2));
}

template <int R, int KP, int RC, int WAVES>
void launch_gram_t(const GdChunk& c, hipStream_t s) {
    const size_t lds = gram_lds<R, KP, RC, WAVES>(c.tab.N);
    if (lds > 160 * 1024) throw Error{SHOREFIT_E_UNSUPPORTED, "gradient-deviation fit: too many gradients for LDS"};
    auto kern = gd_gram_kernel<R, KP, RC, WAVES>;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                              int(lds));
    hipLaunchKernelGGL(kern, dim3(unsigned(c.n)), dim3(64 * WAVES), lds, s, c.B, c.tab, c.data, c.sig_stride,
                       c.field, c.vox, c.n, c.G, c.b, c.sstat);
    SF_HIP(hipGetLastError());
}

template <int KP, int W>
void launch_fista_t(const GdFistaArgs& a, int max_blocks, hipStream_t s) {
    auto kern = gd_fista_kernel<KP, W>;
    int occ = 0;
    SF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 64 * W, 0));
    int dev = 0, cus = 0;
    SF_HIP(hipGetDevice(&dev));
    SF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int64_t blocks = int64_t(std::max(occ, 1)) * cus;
    if (max_blocks > 0) blocks = std::min<int64_t>(blocks, max_blocks);
    blocks = std::min<int64_t>(blocks, std::max<int64_t>(a.n, 1));
    hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(64 * W), 0, s, a);
    SF_HIP(hipGetLastError());
}

}  // namespace

// Variants: (order, Kp, gram rows per chunk, gram waves, FISTA waves).
static const GdVariant kVariants[] = {
    // each order's variants are contiguous; the first (the untuned default)
    // is the one the tuner picks on 250k-voxel volumes with 258 gradients
    {2, 16, 32, 1, 1},   {4, 32, 32, 1, 1},   {4, 32, 32, 1, 2},   {6, 80, 48, 4, 2},
    {6, 80, 32, 4, 2},   {6, 80, 32, 4, 4},   {6, 80, 16, 4, 4},   {6, 80, 64, 4, 2},
    {8, 160, 48, 4, 4},  {8, 160, 32, 4, 4},  {8, 160, 32, 4, 5},  {8, 160, 32, 4, 8},
    {8, 160, 16, 4, 5},  {8, 160, 64, 4, 4},  {10, 256, 64, 8, 8}, {10, 256, 16, 8, 8},
    {10, 256, 32, 8, 8},
};

bool gd_gram_fits(const GdVariant& v, int N) {
    const int R = v.order;
    const size_t lds = sizeof(double) * (size_t(n_pairs_c(R) + n_ang_c(R)) + size_t(6) * N + size_t(v.rc) * (v.kp + 2));
    return lds <= 160 * 1024;
}

int gd_variants(uint32_t order, const GdVariant** out) {
    int first = -1, count = 0;
    for (int i = 0; i < int(sizeof(kVariants) / sizeof(kVariants[0])); ++i)
        if (kVariants[i].order == int(order)) {
            if (first < 0) first = i;
            ++count;
        }
    *out = first >= 0 ? &kVariants[first] : nullptr;
    return count;
}

void gd_launch_gram(const GdVariant& v, const GdChunk& c, hipStream_t s) {
    if (c.n == 0) return;
#define SF_GRAM(O, KP, RC, WV) \
    if (v.order == O && v.kp == KP && v.rc == RC && v.gram_waves == WV) return launch_gram_t<O, KP, RC, WV>(c, s);
    SF_GRAM(2, 16, 32, 1)
    SF_GRAM(4, 32, 32, 1)
    SF_GRAM(6, 80, 32, 4)
    SF_GRAM(6, 80, 16, 4)
    SF_GRAM(8, 160, 32, 4)
    SF_GRAM(8, 160, 16, 4)
    SF_GRAM(8, 160, 64, 4)
    SF_GRAM(8, 160, 48, 4)
    SF_GRAM(6, 80, 48, 4)
    SF_GRAM(6, 80, 64, 4)
    SF_GRAM(10, 256, 16, 8)
    SF_GRAM(10, 256, 32, 8)
    SF_GRAM(10, 256, 64, 8)
#undef SF_GRAM
    throw Error{SHOREFIT_E_UNSUPPORTED, "no Gram kernel for this variant"};
}

void gd_launch_fista(const GdVariant& v, const GdFistaArgs& a, int max_blocks, hipStream_t s) {
    if (a.n == 0) return;
#define SF_FI(KP, W) \
    if (v.kp == KP && v.fista_waves == W) return launch_fista_t<KP, W>(a, max_blocks, s);
    SF_FI(16, 1)
    SF_FI(32, 1)
    SF_FI(32, 2)
    SF_FI(80, 2)
    SF_FI(80, 4)
    SF_FI(160, 4)
    SF_FI(160, 5)
    SF_FI(160, 8)
    SF_FI(256, 8)
#undef SF_FI
    throw Error{SHOREFIT_E_UNSUPPORTED, "no FISTA kernel for this variant"};
}

void gd_launch_identity(const float* field, const int64_t* vox, int64_t n, uint8_t* ident, uint8_t* devi,
                        hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(gd_identity_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, field, vox, n, ident,
                       devi);
    SF_HIP(hipGetLastError());
}

void gd_launch_finish(const GdSlotOut* so, const double* sstat, const int64_t* vox, int64_t n, int N,
                      const GdMaps& m, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(gd_finish_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, so, sstat, vox, n, N, m);
    SF_HIP(hipGetLastError());
}

void launch_gather_i64(const int64_t* src, const int64_t* pos, int64_t n, int64_t* dst, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(gather_i64_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, src, pos, n, dst);
    SF_HIP(hipGetLastError());
}

}  // namespace shorefit
