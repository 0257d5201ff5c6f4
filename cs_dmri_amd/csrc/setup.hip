// FP64 set-up kernels: SHORE design matrix, Gram matrix, Lipschitz constant
// (power iteration), Cholesky factor, and the packing of FP32 / FP16 / BF16
// MFMA operands. Everything here runs once per fit (or once per design).
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "common.h"
#include "special.h"

namespace shorefit {

int n_coeffs(uint32_t radial_order) {
    // src/basis/shore.rs:29-44: for n in 0..=R, even l <= n, 2l+1 slots.
    int k = 0;
    for (uint32_t n = 0; n <= radial_order; ++n)
        for (uint32_t l = 0; l <= n; l += 2) k += int(2 * l + 1);
    return k;
}

namespace {

constexpr double kPi = kPiD;

// One entry of the SHORE design matrix. Column col maps to (n, l, slot)
// in the order of src/basis/shore.rs:29-44; the angular factor follows the
// BrainSuite real-SH block of src/basis/shore.rs:76-115 and the radial
// factor src/basis/shore.rs:56-70.
__global__ void design_matrix_kernel(const double* __restrict__ q, const double* __restrict__ bvecs, int N, int R,
                                     double zeta, int K, double* __restrict__ A) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(N) * K) return;
    const int i = int(t / K), col = int(t % K);
    int n = 0, l = 0, slot = 0, c = 0;
    for (int nn = 0; nn <= R; ++nn) {
        bool found = false;
        for (int ll = 0; ll <= nn; ll += 2) {
            if (col < c + 2 * ll + 1) {
                n = nn;
                l = ll;
                slot = col - c;
                found = true;
                break;
            }
            c += 2 * ll + 1;
        }
        if (found) break;
    }
    const double qq = q[i];
    const double x = bvecs[3 * i] * qq, y = bvecs[3 * i + 1] * qq, z = bvecs[3 * i + 2] * qq;
    const double r = sqrt(x * x + y * y + z * z);
    double theta = atan2(y, x);
    if (isnan(theta)) theta = 0.0;
    const double phi = r > 0.0 ? acos(fmin(1.0, fmax(-1.0, z / r))) : 0.0;

    const double u = r * r / zeta;
    const double kap = sqrt(2.0 * factorial_d(n - l) / (pow(zeta, 1.5) * gamma_n_three_halves(n)));
    const double radial = kap * laguerre_d(n - l, l + 0.5, u) * exp(-0.5 * u) * pow(u, 0.5 * l);

    double ang;
    if (l == 0) {
        ang = 1.0 / sqrt(4.0 * kPi);
    } else {
        const int k = slot < l ? l - slot : slot - l;  // |order| of the slot
        const double norm = sqrt((2.0 * l + 1.0) / (4.0 * kPi) * factorial_d(l - k) / factorial_d(l + k));
        const double p = norm * legendre_d(l, k, cos(phi));
        if (slot < l)
            ang = sqrt(2.0) * p * cos(k * theta);
        else if (slot == l)
            ang = p;
        else
            ang = sqrt(2.0) * ((k % 2 == 1) ? 1.0 : -1.0) * p * sin(k * theta);
    }
    A[t] = radial * ang;
}

__global__ void gram_kernel(const double* __restrict__ A, int N, int K, double* __restrict__ G) {
    const int p = blockIdx.y, q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= K) return;
    double acc = 0.0;
    // unrolled so that several loads are in flight (the loop is latency-bound)
#pragma unroll 8
    for (int i = 0; i < N; ++i) acc += A[size_t(i) * K + p] * A[size_t(i) * K + q];
    G[size_t(p) * K + q] = acc;
}


// S_out = S S / trace(S)^2 for a symmetric PSD K x K matrix S (one thread per
// entry). With r = l_max / trace in [1/K, 1], the output's largest
// eigenvalue is r^2 >= 1/K^2 and its own ratio is >= r, so repeated squaring
// stays in range. Every block sums the trace in the same order (and the
// butterfly leaves the same sum on every lane), so the output is symmetric
// bit for bit. Launched with one wave (64 threads) per block.
__global__ void __launch_bounds__(64) square_scaled_kernel(const double* __restrict__ S, int K,
                                                           double* __restrict__ out) {
    const int p = blockIdx.y, q = blockIdx.x * blockDim.x + threadIdx.x;
    double tr = 0.0;
    for (int i = threadIdx.x; i < K; i += 64) tr += S[size_t(i) * K + i];
    for (int o = 32; o > 0; o >>= 1) tr += __shfl_xor(tr, o);  // the same sum on every lane
    const double f = tr > 0.0 ? 1.0 / tr : 0.0;
    if (q >= K) return;
    double acc = 0.0;
#pragma unroll 8
    for (int i = 0; i < K; ++i) acc += S[size_t(p) * K + i] * S[size_t(i) * K + q];
    out[size_t(p) * K + q] = (acc * f) * f;
}

__device__ double block_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < int(blockDim.x >> 6); ++i) s += red[i];
    return s;
}

// Largest eigenvalue of the symmetric PSD Gram matrix G. The dominant
// eigenvector is found by power iteration on P = G^64 (scaled), whose
// eigenvalue ratio is (l2/l1)^64: SHORE Gram matrices have l2/l1 ~ 0.86-0.94,
// so P's ratio is <= 0.02 and a handful of sweeps pin the vector to ~1e-13.
// The eigenvalue is then the Rayleigh quotient v^T G v / v^T v with G itself,
// polished by power sweeps on G until it stops moving (its error is quadratic
// in the eigenvector error, so it reaches the FP64 floor; it never exceeds
// l1). This replaces the reference's full SVD (src/solver/fista.rs:45-51).
constexpr int kMvSplit = 7;  // threads per row of a matrix-vector product (1024 threads: 146 rows a pass)

__global__ void __launch_bounds__(1024) power_iteration_kernel(const double* __restrict__ P,
                                                               const double* __restrict__ G, int K,
                                                               double* __restrict__ out) {
    extern __shared__ double sm[];
    double* v = sm;
    double* w = sm + K;
    double* parts = sm + 2 * K;  // [kMvSplit][rows a pass]
    __shared__ double red[16];
    const int t = threadIdx.x, nrow = int(blockDim.x) / kMvSplit;
    const int chunk = (K + kMvSplit - 1) / kMvSplit;
    // w = M v for symmetric M (row p read as column p): thread (part, p)
    // sums the products of columns part * chunk .. over its row, consecutive
    // threads on consecutive rows; the kMvSplit partial sums of a row are
    // then added in part order
    auto matvec = [&](const double* __restrict__ M) {
        for (int p0 = 0; p0 < K; p0 += nrow) {
            const int part = t / nrow, p = p0 + t % nrow;
            if (part < kMvSplit && p < K) {
                const int q0 = part * chunk, q1 = min(K, q0 + chunk);
                double acc = 0.0;
#pragma unroll 8
                for (int q = q0; q < q1; ++q) acc += M[size_t(q) * K + p] * v[q];
                parts[part * nrow + (p - p0)] = acc;
            }
            __syncthreads();
            if (t < nrow && p0 + t < K) {
                double acc = 0.0;
                for (int i = 0; i < kMvSplit; ++i) acc += parts[i * nrow + t];
                w[p0 + t] = acc;
            }
            __syncthreads();
        }
    };
    for (int i = t; i < K; i += blockDim.x) v[i] = 1.0 + 1e-3 * i;
    __syncthreads();
    // phase 1: eigenvector of P
    for (int it = 0; it < 400; ++it) {
        matvec(P);
        double nrm = 0.0, diff = 0.0;
        for (int p = t; p < K; p += blockDim.x) nrm += w[p] * w[p];
        const double ww = block_sum(nrm, red);
        if (!(ww > 0.0)) break;
        const double inv = 1.0 / sqrt(ww);
        for (int p = t; p < K; p += blockDim.x) diff = fmax(diff, fabs(w[p] * inv - v[p]));
        __syncthreads();
        for (int p = t; p < K; p += blockDim.x) v[p] = w[p] * inv;
        for (int o = 32; o > 0; o >>= 1) diff = fmax(diff, __shfl_xor(diff, o));
        __syncthreads();
        if ((t & 63) == 0) red[t >> 6] = diff;
        __syncthreads();
        double dmax = 0.0;
        for (int i = 0; i < int(blockDim.x >> 6); ++i) dmax = fmax(dmax, red[i]);
        __syncthreads();
        if (dmax < 1e-13) break;
    }
    // phase 2: Rayleigh quotient with G, polished by power sweeps on G
    double lam = 0.0;
    int stable = 0;
    for (int it = 0; it < 200 && stable < 2; ++it) {
        matvec(G);
        double part = 0.0, nrm = 0.0, vpart = 0.0;
        for (int p = t; p < K; p += blockDim.x) {
            part += w[p] * v[p];
            nrm += w[p] * w[p];
            vpart += v[p] * v[p];
        }
        const double vv = block_sum(vpart, red);
        const double num = block_sum(part, red);
        const double ww = block_sum(nrm, red);
        const double lam_new = vv > 0.0 ? num / vv : 0.0;
        stable = (fabs(lam_new - lam) <= 1e-15 * fabs(lam_new)) ? stable + 1 : 0;
        lam = fmax(lam, lam_new);
        const double inv = ww > 0.0 ? 1.0 / sqrt(ww) : 0.0;
        __syncthreads();
        for (int p = t; p < K; p += blockDim.x) v[p] = w[p] * inv;
        __syncthreads();
        if (!(ww > 0.0)) break;
    }
    if (t == 0) out[0] = lam;
}

// Right-looking Cholesky in one workgroup (K <= a few hundred).
__global__ void __launch_bounds__(256) cholesky_kernel(double* __restrict__ G, int K, int* __restrict__ status) {
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < K; ++j) {
        __syncthreads();
        const double d = G[size_t(j) * K + j];
        if (!(d > 0.0)) {
            if (threadIdx.x == 0) bad = j + 1;
            __syncthreads();
            break;
        }
        const double s = sqrt(d);
        __syncthreads();
        for (int i = j + 1 + threadIdx.x; i < K; i += blockDim.x) G[size_t(i) * K + j] /= s;
        __syncthreads();
        if (threadIdx.x == 0) G[size_t(j) * K + j] = s;
        // trailing update of the lower triangle
        const int m = K - j - 1;
        for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
            const int a = j + 1 + t / m, b = j + 1 + t % m;
            if (b <= a) G[size_t(a) * K + b] -= G[size_t(a) * K + j] * G[size_t(b) * K + j];
        }
        __syncthreads();
    }
    __syncthreads();
    for (int t = threadIdx.x; t < K * K; t += blockDim.x) {
        const int a = t / K, b = t % K;
        if (b > a) G[t] = 0.0;
    }
    if (threadIdx.x == 0) status[0] = bad;
}

// The same right-looking factorization (the same operations in the same
// order, the column scaling by the pivot's reciprocal aside) on the packed
// lower triangle held in LDS
// (L[a][b] at a (a + 1) / 2 + b): the trailing updates are LDS traffic
// instead of L2 round trips. Used when K (K + 1) / 2 doubles fit in LDS.
// Blocked right-looking Cholesky on the packed lower triangle in LDS: panels
// of kPanel columns. Wave 0 factors a panel on its own, in registers (no
// workgroup barrier; a column's entries are scaled by the pivot's
// reciprocal), then every wave subtracts the panel's columns from the
// trailing triangle, one column after another, so each entry receives the
// updates of the column-by-column algorithm in the same order, with two
// workgroup barriers per panel instead of three per column.
constexpr int kPanel = 8;
constexpr int kPanelRows = 3;  // rows per lane in the panel: K <= 192

// Value of lane `src` (wave-uniform) of a double.
__device__ __forceinline__ double lane_double(double x, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), src);
    return __hiloint2double(hi, lo);
}

__global__ void __launch_bounds__(1024) cholesky_lds_kernel(double* __restrict__ G, int K, int* __restrict__ status) {
    extern __shared__ double Lp[];
    __shared__ int bad;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    auto at = [&](int a, int b) -> double& { return Lp[a * (a + 1) / 2 + b]; };
    if (tid == 0) bad = 0;
    // one flat pass over G, so each thread's loads are independent
    for (int e = tid; e < K * K; e += nt) {
        const int a = e / K, b = e - a * K;
        if (b <= a) at(a, b) = G[e];
    }
    __syncthreads();
    for (int j0 = 0; j0 < K; j0 += kPanel) {
        const int j1 = min(j0 + kPanel, K);
        if (wave == 0) {
            // Panel columns j0..j1-1 in registers: lane t holds rows
            // j0 + t + 64 q (q < kPanelRows); the pivot and the entries
            // L[k][j] come from the lane that holds row k.
            const int jb = j1 - j0;
            double pv[kPanelRows][kPanel];
#pragma unroll
            for (int q = 0; q < kPanelRows; ++q) {
                const int i = j0 + lane + 64 * q;
#pragma unroll
                for (int c = 0; c < kPanel; ++c) pv[q][c] = (c < jb && i < K && j0 + c <= i) ? at(i, j0 + c) : 0.0;
            }
            bool fail = false;
#pragma unroll
            for (int c = 0; c < kPanel; ++c) {
                if (c >= jb || fail) continue;
                const int j = j0 + c;
                const double d = lane_double(pv[0][c], c);
                if (!(d > 0.0)) {
                    if (lane == 0) bad = j + 1;
                    fail = true;
                    continue;
                }
                // Branch-free: one reciprocal of the pivot per column and
                // selects instead of per-row branches (exec-masked FP64
                // divisions had made each column ~1 us of latency).
                const double sj = sqrt(d);
                const double rj = 1.0 / sj;
#pragma unroll
                for (int q = 0; q < kPanelRows; ++q) {
                    const int i = j0 + lane + 64 * q;
                    const double sc = pv[q][c] * rj;
                    pv[q][c] = i > j ? sc : (i == j ? sj : pv[q][c]);
                }
                // the panel's later columns, rows k..K-1 (rows past K hold zeros)
#pragma unroll
                for (int k = c + 1; k < kPanel; ++k) {
                    if (k >= jb) continue;
                    const double lkj = lane_double(pv[0][c], k);
#pragma unroll
                    for (int q = 0; q < kPanelRows; ++q) {
                        const int i = j0 + lane + 64 * q;
                        const double u = fma(-pv[q][c], lkj, pv[q][k]);
                        pv[q][k] = i >= j0 + k ? u : pv[q][k];
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < kPanelRows; ++q) {
                const int i = j0 + lane + 64 * q;
#pragma unroll
                for (int c = 0; c < kPanel; ++c)
                    if (c < jb && i < K && j0 + c <= i) at(i, j0 + c) = pv[q][c];
            }
        }
        __syncthreads();
        if (bad) break;
        // trailing triangle: rows a >= j1 (wave w takes a = j1 + w + nw i),
        // columns j1..a (lanes), minus the panel's columns in order
        for (int a = j1 + wave; a < K; a += nw) {
            for (int b = j1 + lane; b <= a; b += 64) {
                double v = at(a, b);
                for (int j = j0; j < j1; ++j) v -= at(a, j) * at(b, j);
                at(a, b) = v;
            }
        }
        __syncthreads();
    }
    __syncthreads();
    for (int e = tid; e < K * K; e += nt) {
        const int a = e / K, b = e - a * K;
        G[e] = b > a ? 0.0 : Lp[a * (a + 1) / 2 + b];
    }
    if (tid == 0) status[0] = bad;
}

__global__ void add_diagonals_kernel(const double* __restrict__ G, int K, const double* __restrict__ dn, double lam_n,
                                     const double* __restrict__ dl, double lam_l, double* __restrict__ out) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(K) * K) return;
    double g = G[t];
    const int i = int(t / K);
    if (t == int64_t(i) * K + i) {
        if (dn) g += lam_n * dn[i];
        if (dl) g += lam_l * dl[i];
    }
    out[t] = g;
}

// Fragment-major operand of __builtin_amdgcn_mfma_f32_16x16x4f32: for
// 16x16 block (rb, cb) and lane l, four floats X[16rb + l%16][16cb + 4(l/16) + r].
__global__ void pack_rowblock_kernel(const double* __restrict__ src, int rows, int cols, int ld, int transpose,
                                     double scale, int rows_p, int cols_p, float* __restrict__ frag) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int CB = cols_p / 16;
    const int64_t total = int64_t(rows_p) * cols_p;
    if (t >= total) return;
    const int r = int(t & 3), l = int((t >> 2) & 63);
    const int64_t blk = t >> 8;
    const int rb = int(blk / CB), cb = int(blk % CB);
    const int i = 16 * rb + (l & 15), j = 16 * cb + 4 * (l >> 4) + r;
    double v = 0.0;
    if (i < rows && j < cols) v = scale * (transpose ? src[size_t(j) * ld + i] : src[size_t(i) * ld + j]);
    frag[t] = float(v);
}

// -c G in the same fragment order, zero outside K x K (the FISTA kernels add
// the identity term of M = I - c G outside the MFMA chain).
__global__ void pack_M_kernel(const double* __restrict__ G, int K, int Kp, double c, float* __restrict__ frag) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int MB = Kp / 16;
    if (t >= int64_t(Kp) * Kp) return;
    const int r = int(t & 3), l = int((t >> 2) & 63);
    const int64_t blk = t >> 8;
    const int rb = int(blk / MB), cb = int(blk % MB);
    const int i = 16 * rb + (l & 15), j = 16 * cb + 4 * (l >> 4) + r;
    double v = 0.0;
    if (i < K && j < K) v = -c * G[size_t(i) * K + j];
    frag[t] = float(v);
}

// -c G as the A operand of the 16x16x32 FP16/BF16 MFMA: for output
// block mb, k-chunk kc (32 rows of Y, i.e. 16-row blocks 2kc and 2kc+1) and
// lane l = 16h + i, eight values M[16mb + i][p(v)], v = 0..7, where
// p(v) = 32kc + 16(v/4) + 4h + v%4 is the row of Y that lane (h, j) of the
// accumulator layout holds in register v%4 of block 2kc + v/4.
// part 0: the FP16 (BF16) rounding of -cG; part 1 (FP16 only): the rounding
// remainder -cG - fp16(-cG), scaled by 2^12 so it stays in FP16's normal range
// (the exact refresh); part 2: the remainder -cG - lowp(-cG) unscaled, in FP16
// (subnormal where |cG| < 2^-2) or BF16 (the increments).
__global__ void pack_M_lowp_kernel(const double* __restrict__ G, int K, int Kp, double c, int bf16, int part,
                                   uint16_t* __restrict__ frag) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int KC = Kp / 32;
    if (t >= int64_t(Kp) * Kp) return;
    const int v = int(t & 7), l = int((t >> 3) & 63);
    const int64_t blk = t >> 9;
    const int mb = int(blk / KC), kc = int(blk % KC);
    const int h = l >> 4, i = 16 * mb + (l & 15);
    const int p = 32 * kc + 16 * (v >> 2) + 4 * h + (v & 3);
    double x = 0.0;
    if (i < K && p < K) x = -c * G[size_t(i) * K + p];
    if (bf16) {
        __hip_bfloat16 b = __float2bfloat16(float(x));
        if (part == 2) b = __float2bfloat16(float(x - double(__bfloat162float(b))));
        frag[t] = *reinterpret_cast<uint16_t*>(&b);
    } else {
        __half hv = __float2half(float(x));
        if (part == 1) hv = __float2half(float((x - double(__half2float(hv))) * 4096.0));
        if (part == 2) hv = __float2half(float(x - double(__half2float(hv))));
        frag[t] = *reinterpret_cast<uint16_t*>(&hv);
    }
}

// Operands of the split-FP16 Tikhonov apply (tikhonov_h16_kernel). Every row
// of P (K x N) and of A (N x K) is scaled by a power of two, 2^(15 - E) with
// max_j |x_ij| = m 2^E, m in [0.5, 1) (1 for a zero row), so that it lies in
// [2^14, 2^15) in magnitude; inv[] receives the inverse scales, the Kp rows
// of P first, then the rows of A padded to Np2.
__global__ void __launch_bounds__(64) tik16_scales_kernel(const double* __restrict__ P, const double* __restrict__ A,
                                                          int K, int N, int Kp, float* __restrict__ inv) {
    const int t = blockIdx.x, lane = threadIdx.x;  // one wave per row
    const double* row = nullptr;
    int len = 0;
    if (t < Kp) {
        if (t < K && P) row = P + size_t(t) * N, len = N;  // P == nullptr: A's chunks only
    } else if (t - Kp < N) {
        row = A + size_t(t - Kp) * K, len = K;
    }
    double m = 0.0;
    for (int j = lane; j < len; j += 64) m = fmax(m, fabs(row[j]));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if (lane != 0) return;
    int e = 0;
    (void)frexp(m, &e);
    // the inverse scale stays a normal FP32 number (rows below ~2^-111 keep
    // the smallest scale instead of flushing to zero)
    inv[t] = m > 0.0 ? float(ldexp(1.0, std::min(std::max(e - 15, -126), 127))) : 1.f;
}

// The fragments, chunk by chunk: chunks g < G1e hold, for coefficient block
// mb, slots 2 mb (hi) and 2 mb + 1 (lo) of P's rows 16 mb + i and columns
// p(v) = 16 (2g + v/4) + 4h + v%4 for lane 16h + i and half v (the samples
// lane (h, j) keeps in registers v%4 of sample blocks 2g and 2g + 1); chunks
// G1e + g2 hold, for sample block 2 g2 + b (b = 0, 1) and coefficient pair
// kc, slots 2 (b MB2 + kc) (hi) and + 1 (lo) of A's rows 16 (2 g2 + b) + i
// and columns 16 (2 kc + v/4) + 4h + v%4 (the coefficients lane (h, j) holds
// in registers v%4 of coefficient blocks 2 kc and 2 kc + 1). hi = fp16(x')
// and lo = fp16(x' - hi) of the scaled value x', zero outside the matrices
// (two chunks of zeros follow the G1e + G2: 16 (2 g2 + b) >= N there).
__global__ void pack_tik16_kernel(const double* __restrict__ P, const double* __restrict__ A, int K, int N, int MB,
                                  int G1e, int G2, const float* __restrict__ inv, uint16_t* __restrict__ frag) {
    const int MB2 = (MB + 1) / 2, CF = 4 * MB2;
    const int64_t total = int64_t(G1e + G2 + 2) * CF * 64 * 8;  // the last two chunks: zeros
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int v = int(t & 7), l = int((t >> 3) & 63);
    const int64_t cs = t >> 9;
    const int slot = int(cs % CF), g = int(cs / CF);
    const int h = l >> 4, i = l & 15, hl = slot & 1;
    double x = 0.0;
    if (g < G1e) {
        const int mb = slot >> 1, row = 16 * mb + i, p = 16 * (2 * g + (v >> 2)) + 4 * h + (v & 3);
        if (mb < MB && row < K && p < N) x = P[size_t(row) * N + p] / double(inv[row]);
    } else {
        const int q = slot >> 1, b = q / MB2, kc = q % MB2;
        const int n = 16 * (2 * (g - G1e) + b) + i, k = 16 * (2 * kc + (v >> 2)) + 4 * h + (v & 3);
        if (n < N && k < K) x = A[size_t(n) * K + k] / double(inv[16 * MB + n]);
    }
    __half hv = __float2half(float(x));
    if (hl) hv = __float2half(float(x - double(__half2float(hv))));
    frag[t] = *reinterpret_cast<uint16_t*>(&hv);
}

// P = L^-T L^-1 A^T, 16 columns of P per workgroup (column n solves for row n
// of A; P is K x N row-major). The 16 lanes of a column
// (all in one wave) keep its vector in registers, lane u the entries
// u + 16 r; a substitution step divides the solved entry on its lane, hands it
// to the other 15 with a 16-lane shuffle and lets each lane update its own
// entries (L y = a: the entries below; L^T p = y: the entries above), so no
// step waits on a workgroup barrier.
constexpr int kSolveRows = 12;  // entries per lane: K <= 192

template <int CB>
__global__ void __launch_bounds__(CB * 16) chol_solve_lds_kernel(const double* __restrict__ Lf, int K,
                                                                 const double* __restrict__ A, int N,
                                                                 double* __restrict__ P) {
    extern __shared__ double Lp[];  // K (K + 1) / 2
    const int tid = threadIdx.x, nt = CB * 16;
    const int c = tid >> 4, u = tid & 15;
    const int n = blockIdx.x * CB + c;
    for (int e = tid; e < K * K; e += nt) {  // flat, so each thread's loads are independent
        const int a = e / K, b = e - a * K;
        if (b <= a) Lp[a * (a + 1) / 2 + b] = Lf[e];
    }
    double yv[kSolveRows];
#pragma unroll
    for (int r = 0; r < kSolveRows; ++r) {
        const int k = u + 16 * r;
        yv[r] = (n < N && k < K) ? A[size_t(n) * K + k] : 0.0;
    }
    __syncthreads();
    // L y = a
#pragma unroll
    for (int r = 0; r < kSolveRows; ++r) {
        for (int uu = 0; uu < 16; ++uu) {
            const int i = 16 * r + uu;
            if (i >= K) break;
            const double yi = __shfl(yv[r], uu, 16) / Lp[i * (i + 1) / 2 + i];
            if (u == uu) yv[r] = yi;
#pragma unroll
            for (int rr = r; rr < kSolveRows; ++rr) {
                const int k = u + 16 * rr;
                if (k > i && k < K) yv[rr] -= Lp[k * (k + 1) / 2 + i] * yi;
            }
        }
    }
    // L^T p = y
#pragma unroll
    for (int r = kSolveRows - 1; r >= 0; --r) {
        for (int uu = 15; uu >= 0; --uu) {
            const int i = 16 * r + uu;
            if (i >= K) continue;
            const double pi = __shfl(yv[r], uu, 16) / Lp[i * (i + 1) / 2 + i];
            if (u == uu) yv[r] = pi;
#pragma unroll
            for (int rr = 0; rr <= r; ++rr) {
                const int k = u + 16 * rr;
                if (k < i) yv[rr] -= Lp[i * (i + 1) / 2 + k] * pi;
            }
        }
    }
    if (n < N)
#pragma unroll
        for (int r = 0; r < kSolveRows; ++r) {
            const int k = u + 16 * r;
            if (k < K) P[size_t(k) * N + n] = yv[r];
        }
}

__global__ void chol_solve_columns_kernel(const double* __restrict__ Lf, int K, const double* __restrict__ A, int N,
                                          double* __restrict__ P) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    for (int i = 0; i < K; ++i) {
        double v = A[size_t(n) * K + i];
        for (int k = 0; k < i; ++k) v -= Lf[size_t(i) * K + k] * P[size_t(k) * N + n];
        P[size_t(i) * N + n] = v / Lf[size_t(i) * K + i];
    }
    for (int i = K - 1; i >= 0; --i) {
        double v = P[size_t(i) * N + n];
        for (int k = i + 1; k < K; ++k) v -= Lf[size_t(k) * K + i] * P[size_t(k) * N + n];
        P[size_t(i) * N + n] = v / Lf[size_t(i) * K + i];
    }
}

__global__ void pad_cols_kernel(const double* __restrict__ src, int rows, int cols, int cols_p,
                                double* __restrict__ dst) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(rows) * cols_p) return;
    const int i = int(t / cols_p), j = int(t % cols_p);
    dst[t] = j < cols ? src[size_t(i) * cols + j] : 0.0;
}

inline unsigned blocks_for(int64_t n, int b) { return unsigned((n + b - 1) / b); }

// ------------------------------------------------------ mask compaction ----
// Masked voxels in linear (x -> y -> z) order (src/voxel_loop.rs:28-42),
// on the device: per-tile counts, one scan, then an ordered write per tile.
constexpr int kMaskTile = 4096;  // voxels per workgroup: 256 threads x 16 rounds

__global__ void __launch_bounds__(256) mask_count_kernel(const uint8_t* __restrict__ mask, int64_t n,
                                                         uint32_t* __restrict__ counts) {
    const int64_t base = int64_t(blockIdx.x) * kMaskTile;
    int c = 0;
    for (int i = 0; i < kMaskTile / 256; ++i) {
        const int64_t v = base + i * 256 + threadIdx.x;
        c += (v < n && mask[v] != 0) ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    __shared__ int ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = uint32_t(ws[0] + ws[1] + ws[2] + ws[3]);
}

// Exclusive scan of the tile counts in one workgroup; the total goes to *total.
__global__ void __launch_bounds__(1024) mask_scan_kernel(const uint32_t* __restrict__ counts, int64_t nb,
                                                         int64_t* __restrict__ offsets, int64_t* __restrict__ total) {
    __shared__ int64_t part[1024];
    const int64_t per = (nb + 1023) / 1024;
    const int64_t b0 = int64_t(threadIdx.x) * per, b1 = b0 + per < nb ? b0 + per : nb;
    int64_t sum = 0;
    for (int64_t b = b0; b < b1; ++b) sum += counts[b];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = 0;
        for (int t = 0; t < 1024; ++t) {
            const int64_t x = part[t];
            part[t] = run;
            run += x;
        }
        *total = run;
    }
    __syncthreads();
    int64_t run = part[threadIdx.x];
    for (int64_t b = b0; b < b1; ++b) {
        offsets[b] = run;
        run += counts[b];
    }
}

__global__ void __launch_bounds__(256) mask_write_kernel(const uint8_t* __restrict__ mask, int64_t n,
                                                         const int64_t* __restrict__ offsets,
                                                         int64_t* __restrict__ vox) {
    __shared__ int wsum[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = int64_t(blockIdx.x) * kMaskTile;
    int64_t off = offsets[blockIdx.x];
    for (int i = 0; i < kMaskTile / 256; ++i) {
        const int64_t v = base + i * 256 + threadIdx.x;
        const bool f = v < n && mask[v] != 0;
        const unsigned long long bal = __ballot(f);
        const int pre = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = __popcll(bal);
        __syncthreads();
        int wpre = 0;
        for (int w = 0; w < wave; ++w) wpre += wsum[w];
        const int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (f) vox[off + wpre + pre] = v;
        off += tot;
        __syncthreads();
    }
}

__global__ void fit_scalars_kernel(const double* __restrict__ lam, int N, double* __restrict__ out) {
    const double L = fmax(fmax(*lam, 0.0) / double(N), 2.220446049250313e-16);
    out[0] = L;
    out[1] = (1.0 / L) / double(N);
}

__global__ void gather_f32_kernel(const float* __restrict__ src, const int64_t* __restrict__ idx, int64_t n,
                                  float* __restrict__ dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx ? idx[i] : i];
}

}  // namespace

int64_t compact_mask(const uint8_t* d_mask, int64_t n, DevBuf<int64_t>& vox, hipStream_t s) {
    if (n <= 0) return 0;
    const int64_t nb = (n + kMaskTile - 1) / kMaskTile;
    DevBuf<uint32_t> counts(nb);
    DevBuf<int64_t> offsets(nb), total(1);
    vox.alloc(n);  // upper bound: no round trip before the write pass
    hipLaunchKernelGGL(mask_count_kernel, dim3(unsigned(nb)), dim3(256), 0, s, d_mask, n, counts.p);
    SF_HIP(hipGetLastError());
    hipLaunchKernelGGL(mask_scan_kernel, dim3(1), dim3(1024), 0, s, counts.p, nb, offsets.p, total.p);
    SF_HIP(hipGetLastError());
    hipLaunchKernelGGL(mask_write_kernel, dim3(unsigned(nb)), dim3(256), 0, s, d_mask, n, offsets.p, vox.p);
    SF_HIP(hipGetLastError());
    int64_t V = 0;
    SF_HIP(hipMemcpyAsync(&V, total.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SF_HIP(hipStreamSynchronize(s));
    return V;
}

void launch_gather_f32(const float* d_src, const int64_t* d_idx, int64_t n, float* d_dst, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(gather_f32_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, d_src, d_idx, n, d_dst);
    SF_HIP(hipGetLastError());
}

void launch_design_matrix(const double* d_q, const double* d_bvecs, int N, uint32_t R, double zeta, double* d_A,
                          hipStream_t s) {
    const int K = n_coeffs(R);
    const int64_t total = int64_t(N) * K;
    hipLaunchKernelGGL(design_matrix_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, d_q, d_bvecs, N,
                       int(R), zeta, K, d_A);
    SF_HIP(hipGetLastError());
}

void launch_gram(const double* d_A, int N, int K, double* d_G, hipStream_t s) {
    hipLaunchKernelGGL(gram_kernel, dim3(blocks_for(K, 64), K), dim3(64), 0, s, d_A, N, K, d_G);
    SF_HIP(hipGetLastError());
}

void launch_power_iteration(const double* d_G, int K, double* d_lambda, hipStream_t s) {
    // P = G^64 (scaled) by six squarings (ping-pong buffers)
    DevBuf<double> p0(size_t(K) * K), p1(size_t(K) * K);
    const double* src = d_G;
    double* bufs[2] = {p0.p, p1.p};
    for (int i = 0; i < 6; ++i) {
        double* dst = bufs[i & 1];
        hipLaunchKernelGGL(square_scaled_kernel, dim3(blocks_for(K, 64), K), dim3(64), 0, s, src, K, dst);
        src = dst;
    }
    hipLaunchKernelGGL(power_iteration_kernel, dim3(1), dim3(1024),
                       size_t(2 * K + kMvSplit * (1024 / kMvSplit)) * sizeof(double), s, src, d_G, K, d_lambda);
    SF_HIP(hipGetLastError());
    sync_if_unscoped(s);  // p0 / p1 are released at scope exit
}

void launch_fit_scalars(const double* d_lam, int N, double* d_out, hipStream_t s) {
    hipLaunchKernelGGL(fit_scalars_kernel, dim3(1), dim3(1), 0, s, d_lam, N, d_out);
    SF_HIP(hipGetLastError());
}

void sync_if_unscoped(hipStream_t s) {
    if (!stream_scope_active()) SF_HIP(hipStreamSynchronize(s));
}

void launch_add_diagonals(const double* d_G, int K, const double* d_dn, double lam_n, const double* d_dl,
                          double lam_l, double* d_out, hipStream_t s) {
    hipLaunchKernelGGL(add_diagonals_kernel, dim3(blocks_for(int64_t(K) * K, 256)), dim3(256), 0, s, d_G, K, d_dn,
                       lam_n, d_dl, lam_l, d_out);
    SF_HIP(hipGetLastError());
}

constexpr size_t kCholLds = 150 * 1024;  // LDS the packed-triangle kernels may take

void launch_cholesky(double* d_G, int K, int* d_status, hipStream_t s) {
    const size_t lds = size_t(K) * (K + 1) / 2 * sizeof(double);
    if (lds <= kCholLds && K <= kPanelRows * 64) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(cholesky_lds_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        hipLaunchKernelGGL(cholesky_lds_kernel, dim3(1), dim3(1024), lds, s, d_G, K, d_status);
    } else {
        hipLaunchKernelGGL(cholesky_kernel, dim3(1), dim3(256), 0, s, d_G, K, d_status);
    }
    SF_HIP(hipGetLastError());
}

void launch_pack_M(const double* d_G, int K, int Kp, double c, float* d_frag, hipStream_t s) {
    const int64_t total = int64_t(Kp) * Kp;
    hipLaunchKernelGGL(pack_M_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, d_G, K, Kp, c, d_frag);
    SF_HIP(hipGetLastError());
}

void launch_pack_tik16(const double* d_P, const double* d_A, int K, int N, int MB, int G1e, int G2, float* d_inv,
                       uint16_t* d_frag, hipStream_t s) {
    const int Kp = 16 * MB, Np2 = 32 * G2;
    hipLaunchKernelGGL(tik16_scales_kernel, dim3(unsigned(Kp + Np2)), dim3(64), 0, s, d_P, d_A, K, N, Kp, d_inv);
    SF_HIP(hipGetLastError());
    const int64_t total = int64_t(G1e + G2 + 2) * (4 * ((MB + 1) / 2)) * 64 * 8;
    hipLaunchKernelGGL(pack_tik16_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, d_P, d_A, K, N, MB, G1e, G2,
                       d_inv, d_frag);
    SF_HIP(hipGetLastError());
}

void launch_pack_rowblock(const double* d_src, int rows, int cols, int src_ld, bool transpose, double scale,
                          int rows_p, int cols_p, float* d_frag, hipStream_t s) {
    const int64_t total = int64_t(rows_p) * cols_p;
    hipLaunchKernelGGL(pack_rowblock_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, d_src, rows, cols,
                       src_ld, int(transpose), scale, rows_p, cols_p, d_frag);
    SF_HIP(hipGetLastError());
}

void launch_pack_M_lowp(const double* d_G, int K, int Kp, double c, bool bf16, uint16_t* d_frag, hipStream_t s,
                        int part) {
    const int64_t total = int64_t(Kp) * Kp;
    hipLaunchKernelGGL(pack_M_lowp_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, d_G, K, Kp, c,
                       int(bf16), part, d_frag);
    SF_HIP(hipGetLastError());
}

void launch_chol_solve_columns(const double* d_L, int K, const double* d_A, int N, double* d_P, hipStream_t s) {
    constexpr int CB = 16;  // columns per workgroup
    const size_t lds = size_t(K) * (K + 1) / 2 * sizeof(double);
    if (lds <= kCholLds && K <= 16 * kSolveRows) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(chol_solve_lds_kernel<CB>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        hipLaunchKernelGGL(chol_solve_lds_kernel<CB>, dim3(unsigned((N + CB - 1) / CB)), dim3(CB * 16), lds, s, d_L,
                           K, d_A, N, d_P);
    } else {
        hipLaunchKernelGGL(chol_solve_columns_kernel, dim3(blocks_for(N, 64)), dim3(64), 0, s, d_L, K, d_A, N, d_P);
    }
    SF_HIP(hipGetLastError());
}

void launch_pad_cols(const double* d_src, int rows, int cols, int cols_p, double* d_dst, hipStream_t s) {
    hipLaunchKernelGGL(pad_cols_kernel, dim3(blocks_for(int64_t(rows) * cols_p, 256)), dim3(256), 0, s, d_src, rows,
                       cols, cols_p, d_dst);
    SF_HIP(hipGetLastError());
}

}  // namespace shorefit
