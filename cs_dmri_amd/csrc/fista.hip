// Batched FISTA-L1 for the SHORE fit on gfx950 (CDNA4).
//
// The reference loop (src/solver/fista.rs:89-128) is, per voxel,
//     r = A y - s;  g = A^T r / N;  x+ = soft(y - g/L, alpha/L);
//     t+ = (1 + sqrt(1 + 4 t^2)) / 2;  y+ = x+ + ((t - 1)/t+)(x+ - x);
//     stop when |x+ - x| / max(|x+|, 1e-12) < tol.
// A is shared by every voxel, so y - g/L = M y + b with
//     M = I - A^T A / (N L)   (K x K, once per fit)
//     b = A^T s / (N L)       (K per voxel, once per fit)
// and one iteration over a tile of voxels is the single contraction M * Y
// plus an element-wise epilogue. The momentum sequence t_k depends only on
// the iteration number, so it is shared by every voxel of a tile.
//
// Layout: a wave owns 16 voxels for the whole solve. The 16x16x4 f32 MFMA
// accumulator layout puts coefficient row 16mb + 4h + r of voxel j in
// register r of block mb on lane 16h + j; the same registers are the B
// operand of the next iteration's MFMAs (k-chunk (kb, r) = rows
// {16kb + 4h + r : h}), so Y never leaves registers. M is staged once per
// workgroup in LDS in "fragment-major" order (one float4 per lane per 16x16
// block, read with conflict-free ds_read_b128).
#include "common.h"

namespace shorefit {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float lane_sum4(float v) {
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}
__device__ __forceinline__ float lane_max4(float v) {
    v = fmaxf(v, __shfl_xor(v, 16));
    v = fmaxf(v, __shfl_xor(v, 32));
    return v;
}

struct FistaArgs {
    const float* data;
    const int64_t* vox;
    int64_t n_vox;
    int64_t sig_stride;
    int N, K, NC;  // NC = Np / 16
    const floatx4* Mfrag;
    const floatx4* ATfrag;
    const floatx4* Afrag;
    float thr, step, ratio;
    double alpha;
    uint32_t max_iter;
    float tol;
    int nonneg;
    float* coef;
    float* resid;
    float* r2;
    uint32_t* iters;
    uint8_t* converged;
    uint8_t* kind;
    float* alpha_map;
    float* bic;
};

// Prologue: b = c A^T s for the wave's 16 voxels (c = 1/(N L) folded into
// ATfrag). Signal chunk c of 16 samples: lane (h, j) loads s_j[16c + 4h + r].
template <int MB>
__device__ __forceinline__ void load_rhs(const FistaArgs& a, const float* sig, bool valid, int h, int lane,
                                         floatx4 (&b)[MB], float& sum) {
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) b[mb] = floatx4{0.f, 0.f, 0.f, 0.f};
    sum = 0.f;
    for (int c = 0; c < a.NC; ++c) {
        float s4[4];
        const int n0 = 16 * c + 4 * h;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int n = n0 + r;
            s4[r] = (valid && n < a.N) ? sig[n] : 0.f;
            sum += s4[r];
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const floatx4 af = a.ATfrag[(size_t(mb) * a.NC + c) * 64 + lane];
#pragma unroll
            for (int r = 0; r < 4; ++r) b[mb] = mfma4(af[r], s4[r], b[mb]);
        }
    }
}

// Epilogue: residual |s - A x|^2 and total sum of squares about the mean.
template <int MB>
__device__ __forceinline__ void residual(const FistaArgs& a, const float* sig, bool valid, int h, int lane,
                                         const floatx4 (&x)[MB], float mean, float& rss, float& sst) {
    rss = 0.f;
    sst = 0.f;
    for (int nb = 0; nb < a.NC; ++nb) {
        floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < MB; ++kb) {
            const floatx4 af = a.Afrag[(size_t(nb) * MB + kb) * 64 + lane];
#pragma unroll
            for (int r = 0; r < 4; ++r) acc = mfma4(af[r], x[kb][r], acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int n = 16 * nb + 4 * h + r;
            if (valid && n < a.N) {
                const float s = sig[n];
                const float e = s - acc[r];
                rss = fmaf(e, e, rss);
                const float dm = s - mean;
                sst = fmaf(dm, dm, sst);
            }
        }
    }
    rss = lane_sum4(rss);
    sst = lane_sum4(sst);
}

template <int MB>
__device__ __forceinline__ void write_voxel(const FistaArgs& a, int64_t out_idx, int h, const floatx4 (&x)[MB],
                                            float rss, float sst, uint32_t its, bool conv, double alpha_v,
                                            bool zero_voxel) {
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k = 16 * mb + 4 * h + r;
            if (k < a.K) a.coef[out_idx * a.K + k] = x[mb][r];
        }
    int nnz = 0;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) nnz += (x[mb][r] != 0.f);
    nnz += __shfl_xor(nnz, 16);
    nnz += __shfl_xor(nnz, 32);
    if (h != 0) return;
    // src/fit.rs:120-128 (R^2), src/solver/alpha.rs:99-103 (BIC),
    // src/solver/alpha.rs:178-189 (all-zero voxel of a per-voxel strategy).
    const double ss_res = zero_voxel ? 0.0 : double(rss);
    const double resid_l2 = sqrt(ss_res);
    const double r2 = sst > 0.f ? 1.0 - ss_res / double(sst) : 0.0;
    if (a.resid) a.resid[out_idx] = float(resid_l2);
    if (a.r2) a.r2[out_idx] = float(r2);
    if (a.iters) a.iters[out_idx] = its;
    if (a.converged) a.converged[out_idx] = conv ? 1 : 0;
    if (a.kind) a.kind[out_idx] = 1;
    if (a.alpha_map) a.alpha_map[out_idx] = float(alpha_v);
    if (a.bic) {
        const double n = double(a.N);
        const double rc = fmax(resid_l2 * resid_l2, 2.2250738585072014e-308);
        a.bic[out_idx] = zero_voxel ? 0.f : float(n * log(rc / n) + log(n) * double(nnz));
    }
}

// Plain FP32 form: V = M Y + B recomputed exactly on FP32 MFMA every
// iteration. WAVES waves per workgroup, 16 voxels per wave.
template <int MB, int WAVES, bool M_IN_LDS>
__global__ void __launch_bounds__(64 * WAVES) fista_f32_kernel(const FistaArgs a) {
    extern __shared__ floatx4 Ms[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int h = lane >> 4, j = lane & 15;
    if (M_IN_LDS) {
        for (int i = threadIdx.x; i < MB * MB * 64; i += 64 * WAVES) Ms[i] = a.Mfrag[i];
        __syncthreads();
    }
    auto load_m = [&](int idx) -> floatx4 {
        if constexpr (M_IN_LDS) return Ms[idx];
        else return a.Mfrag[idx];
    };

    const int64_t v = (int64_t(blockIdx.x) * WAVES + wave) * 16 + j;
    const bool valid = v < a.n_vox;
    const int64_t out_idx = valid ? (a.vox ? a.vox[v] : v) : 0;
    const float* sig = a.data + out_idx * a.sig_stride;

    floatx4 b[MB];
    float sum;
    load_rhs<MB>(a, sig, valid, h, lane, b, sum);
    const float mean = lane_sum4(sum) / float(a.N);

    // alpha-ratio strategy: alpha_v = ratio * |A^T s|_inf / N, i.e.
    // thr_v = ratio * max|b| (src/solver/alpha.rs:229-235).
    float thr = a.thr;
    double alpha_v = a.alpha;
    bool zero_voxel = false;
    if (a.ratio > 0.f) {
        float m = 0.f;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) m = fmaxf(m, fabsf(b[mb][r]));
        m = lane_max4(m);
        thr = a.ratio * m;
        alpha_v = double(m) / double(a.step) * double(a.ratio);
        zero_voxel = (m == 0.f);
    }

    floatx4 x[MB], y[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) x[mb] = y[mb] = floatx4{0.f, 0.f, 0.f, 0.f};
    bool active = valid && !zero_voxel;
    bool conv = zero_voxel;
    uint32_t its = 0;
    double t = 1.0;
    const float tol = a.tol;

    for (uint32_t it = 0; it < a.max_iter; ++it) {
        if (__ballot(active) == 0ull) break;
        // One accumulation chain per 16-row block; back-to-back MFMAs that
        // accumulate into the same registers issue without extra waits.
        floatx4 acc[MB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            // keep the scheduler from hoisting every block's M loads at once
            asm volatile("" ::: "memory");
            floatx4 c = b[mb];
#pragma unroll
            for (int kb = 0; kb < MB; ++kb) {
                const floatx4 mf = load_m((mb * MB + kb) * 64 + lane);
#pragma unroll
                for (int r = 0; r < 4; ++r) c = mfma4(mf[r], y[kb][r], c);
            }
            acc[mb] = c;
        }
        const double tn = 0.5 * (1.0 + sqrt(1.0 + 4.0 * t * t));
        const float mom = float((t - 1.0) / tn);
        t = tn;
        float dd = 0.f, xx = 0.f;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float val = acc[mb][r];
                float xn = val - fminf(fmaxf(val, -thr), thr);  // strict soft threshold
                if (a.nonneg) xn = fmaxf(xn, 0.f);
                const float d = xn - x[mb][r];
                dd = fmaf(d, d, dd);
                xx = fmaf(xn, xn, xx);
                if (active) {
                    y[mb][r] = fmaf(mom, d, xn);
                    x[mb][r] = xn;
                }
            }
        dd = lane_sum4(dd);
        xx = lane_sum4(xx);
        if (active) {
            its = it + 1;
            const float rel = sqrtf(dd) / fmaxf(sqrtf(xx), 1e-12f);
            if (rel < tol) {
                conv = true;
                active = false;
            }
        }
    }

    float rss, sst;
    residual<MB>(a, sig, valid, h, lane, x, mean, rss, sst);
    if (valid) write_voxel<MB>(a, out_idx, h, x, rss, sst, its, conv, zero_voxel ? 0.0 : alpha_v, zero_voxel);
}

// Tikhonov L2 (src/solver/tikhonov.rs:59-76): c = P s with
// P = (A^T A + lambda_n D_n + lambda_l D_l)^-1 A^T factored in FP64 at set-up
// and applied here with the prologue contraction (P in place of c A^T),
// followed by the same residual epilogue.
template <int MB>
__global__ void __launch_bounds__(256) tikhonov_kernel(const FistaArgs a, float* rss_l2) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int h = lane >> 4, j = lane & 15;
    const int64_t v = (int64_t(blockIdx.x) * 4 + wave) * 16 + j;
    const bool valid = v < a.n_vox;
    const int64_t out_idx = valid ? (a.vox ? a.vox[v] : v) : 0;
    const float* sig = a.data + out_idx * a.sig_stride;
    floatx4 c[MB];
    float sum;
    load_rhs<MB>(a, sig, valid, h, lane, c, sum);
    const float mean = lane_sum4(sum) / float(a.N);
    float rss, sst;
    residual<MB>(a, sig, valid, h, lane, c, mean, rss, sst);
    if (!valid) return;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k = 16 * mb + 4 * h + r;
            if (k < a.K) a.coef[out_idx * a.K + k] = c[mb][r];
        }
    if (h != 0) return;
    const double ss_res = double(rss);
    if (a.resid) a.resid[out_idx] = float(sqrt(ss_res));
    if (a.r2) a.r2[out_idx] = float(sst > 0.f ? 1.0 - ss_res / double(sst) : 0.0);
    if (a.iters) a.iters[out_idx] = 1;
    if (a.converged) a.converged[out_idx] = 1;
    if (a.kind) a.kind[out_idx] = 2;
    if (rss_l2) rss_l2[out_idx] = float(ss_res);
}

template <int MB>
void launch_tik(const FistaArgs& args, int64_t n_vox, float* rss_l2, hipStream_t s) {
    const int64_t tiles = (n_vox + 15) / 16;
    hipLaunchKernelGGL(tikhonov_kernel<MB>, dim3(unsigned((tiles + 3) / 4)), dim3(256), 0, s, args, rss_l2);
    SF_HIP(hipGetLastError());
}

template <int MB>
void launch_f32(const FistaArgs& args, int64_t n_vox, hipStream_t s) {
    constexpr int WAVES = (MB >= 16) ? 4 : 8;
    constexpr bool IN_LDS = (MB * MB * 64 * 16) <= 112 * 1024;
    const int64_t tiles = (n_vox + 15) / 16;
    const unsigned grid = unsigned((tiles + WAVES - 1) / WAVES);
    const size_t lds = IN_LDS ? size_t(MB) * MB * 64 * 16 : 0;
    auto kern = fista_f32_kernel<MB, WAVES, IN_LDS>;
    if (lds > 64 * 1024)  // gfx950 has 160 KB of LDS per CU; a refusal surfaces at launch
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(lds));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * WAVES), lds, s, args);
    SF_HIP(hipGetLastError());
}

}  // namespace

void launch_fista(const FistaLaunch& L, hipStream_t s) {
    if (L.n_vox <= 0) return;
    FistaArgs a;
    a.data = L.data;
    a.vox = L.vox;
    a.n_vox = L.n_vox;
    a.sig_stride = L.sig_stride;
    a.N = L.N;
    a.K = L.K;
    a.NC = L.Np / 16;
    a.Mfrag = reinterpret_cast<const floatx4*>(L.Mfrag);
    a.ATfrag = reinterpret_cast<const floatx4*>(L.ATfrag);
    a.Afrag = reinterpret_cast<const floatx4*>(L.Afrag);
    a.thr = L.thr;
    a.step = L.step;
    a.ratio = L.ratio;
    a.alpha = L.alpha;
    a.max_iter = L.max_iter;
    a.tol = L.tol;
    a.nonneg = L.nonneg;
    a.coef = L.coef;
    a.resid = L.resid;
    a.r2 = L.r2;
    a.iters = L.iters;
    a.converged = L.converged;
    a.kind = L.kind;
    a.alpha_map = L.alpha_map;
    a.bic = L.bic;
    switch (L.Kp / 16) {
        case 1: launch_f32<1>(a, L.n_vox, s); break;
        case 2: launch_f32<2>(a, L.n_vox, s); break;
        case 3: launch_f32<3>(a, L.n_vox, s); break;
        case 4: launch_f32<4>(a, L.n_vox, s); break;
        case 5: launch_f32<5>(a, L.n_vox, s); break;
        case 6: launch_f32<6>(a, L.n_vox, s); break;
        case 8: launch_f32<8>(a, L.n_vox, s); break;
        case 10: launch_f32<10>(a, L.n_vox, s); break;
        case 12: launch_f32<12>(a, L.n_vox, s); break;
        case 16: launch_f32<16>(a, L.n_vox, s); break;
        default:
            throw Error{SHOREFIT_E_UNSUPPORTED,
                        "FISTA kernel: padded coefficient count " + std::to_string(L.Kp) + " is not instantiated"};
    }
}

void launch_tikhonov(const TikhonovLaunch& L, hipStream_t s) {
    if (L.n_vox <= 0) return;
    FistaArgs a{};
    a.data = L.data;
    a.vox = L.vox;
    a.n_vox = L.n_vox;
    a.sig_stride = L.sig_stride;
    a.N = L.N;
    a.K = L.K;
    a.NC = L.Np / 16;
    a.ATfrag = reinterpret_cast<const floatx4*>(L.Pfrag);
    a.Afrag = reinterpret_cast<const floatx4*>(L.Afrag);
    a.coef = L.coef;
    a.resid = L.resid;
    a.r2 = L.r2;
    a.iters = L.iters;
    a.kind = L.kind;
    switch (L.Kp / 16) {
        case 1: launch_tik<1>(a, L.n_vox, L.rss_l2, s); break;
        case 2: launch_tik<2>(a, L.n_vox, L.rss_l2, s); break;
        case 3: launch_tik<3>(a, L.n_vox, L.rss_l2, s); break;
        case 4: launch_tik<4>(a, L.n_vox, L.rss_l2, s); break;
        case 5: launch_tik<5>(a, L.n_vox, L.rss_l2, s); break;
        case 6: launch_tik<6>(a, L.n_vox, L.rss_l2, s); break;
        case 8: launch_tik<8>(a, L.n_vox, L.rss_l2, s); break;
        case 10: launch_tik<10>(a, L.n_vox, L.rss_l2, s); break;
        case 12: launch_tik<12>(a, L.n_vox, L.rss_l2, s); break;
        case 16: launch_tik<16>(a, L.n_vox, L.rss_l2, s); break;
        default:
            throw Error{SHOREFIT_E_UNSUPPORTED,
                        "Tikhonov kernel: padded coefficient count " + std::to_string(L.Kp) + " is not instantiated"};
    }
}

}  // namespace shorefit
