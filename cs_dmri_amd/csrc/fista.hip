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
#include <type_traits>

#include "common.h"

namespace shorefit {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Combine lanes l, l^16, l^32, l^48 (the four 4-row groups of one voxel
// column in the MFMA D layout) with the gfx950 lane swaps: permlane32_swap of
// a register with itself leaves lanes 0-31 of the pair in the first result
// and lanes 32-63 in the second, so their sum holds x[l] + x[l^32] in every
// lane; permlane16_swap does the same for rows of 16. VALU only, no LDS.
__device__ __forceinline__ float xor32_sum(float v) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float xor16_sum(float v) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float lane_sum4(float v) { return xor32_sum(xor16_sum(v)); }
__device__ __forceinline__ float lane_max4(float v) {
    auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = fmaxf(__uint_as_float(r[0]), __uint_as_float(r[1]));
    r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return fmaxf(__uint_as_float(r[0]), __uint_as_float(r[1]));
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
    const float* B;       // [n_vox][Kp] right-hand sides b = c A^T s (FP64-accumulated)
    const float* mean;    // [n_vox] signal means (FP64-accumulated)
    int Kp;
    float thr, step, ratio;
    double alpha;
    double step64;             // 1 / L in FP64 (per-voxel thresholds)
    const float* alpha_vox;    // optional per-voxel alpha, work order (FP32 kernel, fixed alpha)
    const float* warm;         // optional warm start, indexed like coef (FP32 kernel)
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
    // alpha path (path_n > 0): voxel v solves at alpha_i = alpha_max f_i,
    // i = 0 .. path_n-1, each solve warm-started from the previous one
    // (src/solver/fista.rs:165-180); point i of voxel v goes to
    // Xpath[(v path_n + i) Kp ..], itspath / convpath[v path_n + i].
    int path_n;
    const float* path_f;       // [path_n] f_i = eps^(i/(path_n-1))
    const double* path_f64;    // same in FP64 (reported alphas)
    const float* mom;          // momentum table (t_k - 1)/t_{k+1}, k < max_iter
    float* Xpath;
    uint32_t* itspath;
    uint8_t* convpath;         // bit0 converged, bit1 all-zero voxel (alpha_max = 0)
    int path_rule;             // 0: BIC argmin, 1: L2-anchored
    double slack;
    const float* rss_l2_ref;   // L2-anchored: Tikhonov RSS, indexed like the outputs
    unsigned int* progress;    // host-pinned finished-voxel count per wave (may be null)
    int progress_slots;
};

// Publishes a wave's count of finished voxels to the host-pinned progress
// array the host polls while the solve runs: a system-scope (vector) store,
// so the value reaches host memory without waiting for the kernel's end.
__device__ __forceinline__ void publish_progress(unsigned int* progress, int slots, int waves, unsigned int done) {
    if (!progress) return;
    const int slot = int(blockIdx.x) * waves + int(threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0 && slot < slots)
        __hip_atomic_store(progress + slot, done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

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

// Coefficients of one voxel (lane (h, j) holds rows 16 mb + 4h + r), stored
// when `store`, and its non-zero count summed over the voxel's four lanes
// (every lane takes part in the shuffles).
template <int MB>
__device__ __forceinline__ int write_coefs(const FistaArgs& a, int64_t out_idx, int h, const floatx4 (&x)[MB],
                                           bool store = true) {
    if (store)
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
    return nnz;
}

// The per-voxel maps, written by lane h == 0.
// src/fit.rs:120-128 (R^2), src/solver/alpha.rs:99-103 (BIC),
// src/solver/alpha.rs:178-189 (all-zero voxel of a per-voxel strategy).
__device__ __forceinline__ void write_maps(const FistaArgs& a, int64_t out_idx, int h, int nnz, float rss, float sst,
                                           uint32_t its, bool conv, double alpha_v, bool zero_voxel) {
    if (h != 0) return;
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

template <int MB>
__device__ __forceinline__ void write_voxel(const FistaArgs& a, int64_t out_idx, int h, const floatx4 (&x)[MB],
                                            float rss, float sst, uint32_t its, bool conv, double alpha_v,
                                            bool zero_voxel) {
    const int nnz = write_coefs<MB>(a, out_idx, h, x);
    write_maps(a, out_idx, h, nnz, rss, sst, its, conv, alpha_v, zero_voxel);
}

// Right-hand sides b_v = c A^T s_v (c = 1/(N L)) and signal means, with
// FP64 accumulation: b enters every FISTA iteration, so its rounding is a
// perturbation of the problem itself; accumulating the N-term dot products
// in FP32 measurably moved the FP32 trajectories away from the FP64 ones.
// FP64 FMAs on the vector ALUs: on gfx950 they outrun the FP64 MFMA (65 vs
// 49 TFLOP/s measured), and an LDS-broadcast variant of this kernel ran 1.6x
// slower than these scalar loads.
// Workgroup = 8 waves x 64 lanes x VPL voxels per lane (signal chunks staged
// in LDS); wave w owns the contiguous coefficient range [w KQ, (w+1) KQ) of
// the zero-padded design Ap (N x Kp, FP64), which it reads with wave-uniform
// (scalar) loads. Each loaded element feeds VPL FMAs per lane. The next
// signal chunk is loaded into registers (PER independent loads per thread,
// each thread keeping its voxel and sample slot from chunk to chunk) while
// the current chunk is summed, so the global-load latency is paid once per
// workgroup instead of once per staged element (0.69 -> 0.55 ms on 250,176
// voxels at N = 258, Kp = 160).
template <int KQ, int VPL>  // KQ = Kp / 8 coefficients per wave
__global__ void __launch_bounds__(512) rhs_kernel(const float* __restrict__ data, const int64_t* __restrict__ vox,
                                                  int64_t n_vox, int64_t sig_stride, int N, int Kp,
                                                  const double* __restrict__ Ap, const double* __restrict__ cp,
                                                  float* __restrict__ B, float* __restrict__ mean) {
    static_assert(KQ % 2 == 0, "coefficients are stored in pairs");
    constexpr int CH = 32;        // samples per LDS chunk
    constexpr int NV = 64 * VPL;  // voxels per workgroup
    constexpr int NT = 512;
    constexpr int PER = (NV * CH + NT - 1) / NT;  // staged elements per thread and chunk
    __shared__ float sm[NV][CH + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int k0 = __builtin_amdgcn_readfirstlane(tid >> 6) * KQ;
    const int64_t v0 = int64_t(blockIdx.x) * NV;
    // staging slot i of this thread: element e = tid + i NT of every chunk,
    // i.e. voxel e / CH and sample offset e % CH
    const float* src[PER];
    int snn[PER], svj[PER];
    bool sok[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = tid + i * NT;
        svj[i] = e / CH;
        snn[i] = e % CH;
        const int64_t v = v0 + svj[i];
        sok[i] = e < NV * CH && v < n_vox;
        src[i] = sok[i] ? data + (vox ? vox[v] : v) * sig_stride + snn[i] : data;
    }
    float pre[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) pre[i] = (sok[i] && snn[i] < N) ? src[i][0] : 0.f;
    double acc[VPL][KQ];
    double sum[VPL];
#pragma unroll
    for (int i = 0; i < VPL; ++i) {
        sum[i] = 0.0;
#pragma unroll
        for (int q = 0; q < KQ; ++q) acc[i][q] = 0.0;
    }
    for (int n0 = 0; n0 < N; n0 += CH) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PER; ++i)
            if (tid + i * NT < NV * CH) sm[svj[i]][snn[i]] = pre[i];
        __syncthreads();
        const int n1 = n0 + CH;
        if (n1 < N) {  // the next chunk, in flight while this one is summed
#pragma unroll
            for (int i = 0; i < PER; ++i) pre[i] = (sok[i] && n1 + snn[i] < N) ? src[i][n1] : 0.f;
        }
        const int nmax = min(CH, N - n0);
        for (int nn = 0; nn < nmax; ++nn) {
            double sv[VPL];
#pragma unroll
            for (int i = 0; i < VPL; ++i) {
                sv[i] = double(sm[lane + 64 * i][nn]);
                sum[i] += sv[i];
            }
            const double* arow = Ap + size_t(n0 + nn) * Kp + k0;
#pragma unroll
            for (int q = 0; q < KQ; ++q) {
                const double aq = arow[q];
#pragma unroll
                for (int i = 0; i < VPL; ++i) acc[i][q] = fma(aq, sv[i], acc[i][q]);
            }
        }
    }
    const double c = *cp;  // step / N, computed on the device from the Lipschitz constant
#pragma unroll
    for (int i = 0; i < VPL; ++i) {
        const int64_t v = v0 + lane + 64 * i;
        if (v >= n_vox) continue;
#pragma unroll
        for (int q = 0; q < KQ; q += 2)
            *reinterpret_cast<float2*>(B + v * Kp + k0 + q) = make_float2(float(c * acc[i][q]), float(c * acc[i][q + 1]));
        if (k0 == 0) mean[v] = float(sum[i] / double(N));
    }
}

// Plain FP32 form: V = M Y + B recomputed exactly on FP32 MFMA every
// iteration. WAVES waves per workgroup, 16 voxels per wave.
template <int MB, int WAVES, bool M_IN_LDS, bool PATH>
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
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
        b[mb] = valid ? *reinterpret_cast<const floatx4*>(a.B + v * a.Kp + 16 * mb + 4 * h)
                      : floatx4{0.f, 0.f, 0.f, 0.f};
    const float mean = valid ? a.mean[v] : 0.f;

    // alpha-ratio strategy: alpha_v = ratio * |A^T s|_inf / N, i.e.
    // thr_v = ratio * max|b| (src/solver/alpha.rs:229-235).
    // Path strategies: thr_i = step alpha_max f_i = max|b| f_i.
    float thr = a.thr, bmax = 0.f;
    double alpha_v = a.alpha;
    bool zero_voxel = false;
    if (PATH || a.ratio > 0.f) {
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) bmax = fmaxf(bmax, fabsf(b[mb][r]));
        bmax = lane_max4(bmax);
        zero_voxel = (bmax == 0.f);
        if constexpr (PATH) {
            thr = bmax * a.path_f[0];
        } else {
            thr = a.ratio * bmax;
            alpha_v = double(bmax) / double(a.step) * double(a.ratio);
        }
    }
    if (!PATH && a.alpha_vox) {
        // per-voxel alpha (fixed strategy): thr_v = alpha_v / L
        const float av = valid ? a.alpha_vox[v] : 0.f;
        thr = float(a.step64 * double(av));
        alpha_v = double(av);
    }
    if (PATH && valid && zero_voxel && h == 0)
        for (int i = 0; i < a.path_n; ++i) {
            a.itspath[v * a.path_n + i] = 0;
            a.convpath[v * a.path_n + i] = 3;
        }

    floatx4 x[MB], y[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) x[mb] = y[mb] = floatx4{0.f, 0.f, 0.f, 0.f};
    if (!PATH && a.warm && valid) {
        // warm start x0 = y0 = warm, t = 1 (src/solver/fista.rs:76-80)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 16 * mb + 4 * h + r;
                x[mb][r] = y[mb][r] = k < a.K ? a.warm[out_idx * a.K + k] : 0.f;
            }
    }
    bool active = valid && !zero_voxel;
    bool conv = zero_voxel;
    uint32_t its = 0;
    double t = 1.0;
    const float tol = a.tol;
    int pi = 0;  // path point of this slot

    for (uint32_t it = 0; PATH || it < a.max_iter; ++it) {
        if (__ballot(active) == 0ull) break;
        // An opaque zero offset keeps the (loop-invariant) M fragment loads
        // inside the loop instead of hoisted into registers for all blocks.
        int moff = 0;
        asm volatile("" : "+s"(moff));
        // acc = b - c G y on MFMA, one accumulation chain per 16-row block
        // (back-to-back MFMAs into the same registers issue without extra
        // waits); v = y + acc adds the identity term outside the chain so the
        // FP32 partial sums stay small.
        floatx4 acc[MB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            floatx4 c = b[mb];
#pragma unroll
            for (int kb = 0; kb < MB; ++kb) {
                const floatx4 mf = load_m(moff + (mb * MB + kb) * 64 + lane);
#pragma unroll
                for (int r = 0; r < 4; ++r) c = mfma4(mf[r], y[kb][r], c);
            }
            acc[mb] = c;
        }
        float mom;
        if constexpr (PATH) {
            // per-slot iteration counts differ along the path: the momentum
            // comes from the table (the same FP64 recurrence, rounded)
            mom = a.mom[active ? its : 0];
        } else {
            const double tn = 0.5 * (1.0 + sqrt(1.0 + 4.0 * t * t));
            mom = float((t - 1.0) / tn);
            t = tn;
        }
        float dd = 0.f, xx = 0.f;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float val = y[mb][r] + acc[mb][r];
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
        if constexpr (PATH) {
            bool next = false;
            if (active) {
                ++its;
                const float rel = sqrtf(dd) / fmaxf(sqrtf(xx), 1e-12f);
                const bool c = rel < tol;
                if (c || its >= a.max_iter) {
                    const int64_t pt = v * a.path_n + pi;
#pragma unroll
                    for (int mb = 0; mb < MB; ++mb)
                        *reinterpret_cast<floatx4*>(a.Xpath + pt * a.Kp + 16 * mb + 4 * h) = x[mb];
                    if (h == 0) {
                        a.itspath[pt] = its;
                        a.convpath[pt] = c ? 1 : 0;
                    }
                    if (pi + 1 < a.path_n) {
                        // warm start of the next point: x0 = y0 = x, t = 1
                        ++pi;
                        thr = bmax * a.path_f[pi];
                        its = 0;
                        next = true;
                    } else {
                        active = false;
                    }
                }
            }
            if (__ballot(next) != 0ull) {
#pragma unroll
                for (int mb = 0; mb < MB; ++mb)
                    if (next) y[mb] = x[mb];
            }
        } else if (active) {
            its = it + 1;
            const float rel = sqrtf(dd) / fmaxf(sqrtf(xx), 1e-12f);
            if (rel < tol) {
                conv = true;
                active = false;
            }
        }
    }
    if constexpr (PATH) return;  // the selection pass writes the outputs

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
    if (a.coef)
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

// Tikhonov with T tiles of 16 voxels per wave: every fragment of P (and of A
// in the residual) is loaded once per wave and used for the T tiles, which
// divides the fragment traffic from L2 by T (P and A are 174 KB each at order
// 8, re-read by every wave).
template <int MB, int T>
__global__ void __launch_bounds__(256) tikhonov_multi_kernel(const FistaArgs a, float* rss_l2) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int h = lane >> 4, j = lane & 15;
    int64_t v[T], out_idx[T];
    bool valid[T];
    const float* sig[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        v[t] = ((int64_t(blockIdx.x) * 4 + wave) * T + t) * 16 + j;
        valid[t] = v[t] < a.n_vox;
        out_idx[t] = valid[t] ? (a.vox ? a.vox[v[t]] : v[t]) : 0;
        sig[t] = a.data + out_idx[t] * a.sig_stride;
    }
    // c = P s (the prologue contraction with P in place of c A^T)
    floatx4 c[T][MB];
    float sum[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        sum[t] = 0.f;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) c[t][mb] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
    for (int ch = 0; ch < a.NC; ++ch) {
        float s4[T][4];
        const int n0 = 16 * ch + 4 * h;
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int n = n0 + r;
                s4[t][r] = (valid[t] && n < a.N) ? sig[t][n] : 0.f;
                sum[t] += s4[t][r];
            }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const floatx4 af = a.ATfrag[(size_t(mb) * a.NC + ch) * 64 + lane];
#pragma unroll
            for (int t = 0; t < T; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) c[t][mb] = mfma4(af[r], s4[t][r], c[t][mb]);
        }
    }
    float mean[T], rss[T], sst[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        mean[t] = lane_sum4(sum[t]) / float(a.N);
        rss[t] = 0.f;
        sst[t] = 0.f;
    }
    // residual |s - A c|^2 and the sum of squares about the mean
    for (int nb = 0; nb < a.NC; ++nb) {
        floatx4 acc[T];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < MB; ++kb) {
            const floatx4 af = a.Afrag[(size_t(nb) * MB + kb) * 64 + lane];
#pragma unroll
            for (int t = 0; t < T; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[t] = mfma4(af[r], c[t][kb][r], acc[t]);
        }
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int n = 16 * nb + 4 * h + r;
                if (valid[t] && n < a.N) {
                    const float sv = sig[t][n];
                    const float e = sv - acc[t][r];
                    rss[t] = fmaf(e, e, rss[t]);
                    const float dm = sv - mean[t];
                    sst[t] = fmaf(dm, dm, sst[t]);
                }
            }
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
        rss[t] = lane_sum4(rss[t]);
        sst[t] = lane_sum4(sst[t]);
        if (!valid[t]) continue;
        if (a.coef)
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int k = 16 * mb + 4 * h + r;
                    if (k < a.K) a.coef[out_idx[t] * a.K + k] = c[t][mb][r];
                }
        if (h != 0) continue;
        const double ss_res = double(rss[t]);
        if (a.resid) a.resid[out_idx[t]] = float(sqrt(ss_res));
        if (a.r2) a.r2[out_idx[t]] = float(sst[t] > 0.f ? 1.0 - ss_res / double(sst[t]) : 0.0);
        if (a.iters) a.iters[out_idx[t]] = 1;
        if (a.converged) a.converged[out_idx[t]] = 1;
        if (a.kind) a.kind[out_idx[t]] = 2;
        if (rss_l2) rss_l2[out_idx[t]] = float(ss_res);
    }
}

// ---------------------------------------------------------------------------
// Apply kernels with the fixed operand staged in LDS.
//
// The operand sequence is cut into chunks of MB 16x16x4 fragments: chunk g <
// nc1 holds the fragments (mb, g) of ATfrag (the c = P s pass: one 16-sample
// chunk of P, every coefficient block; chunks from nc then hold zeros), chunk
// nc1 + nb the fragments (nb, kb) of Afrag (the residual pass: one 16-row
// block of A, every coefficient block). A workgroup stages each chunk once in
// LDS for all its waves, double buffered, and requests each chunk from L2 two
// chunks before it is stored (two register sets), so the load latency is
// covered by two chunks of products instead of every wave fetching every
// fragment from L2 with the latency of each load exposed.
// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima), gfx9 encoding.
constexpr int kWaitVmcnt0 = 0x0F70;

template <int MB, int NT>
struct ChunkStage {
    static constexpr int NF = MB * 64;  // floatx4 per chunk
    static constexpr int PER = (NF + NT - 1) / NT;
    floatx4 pre[2][PER];
    // Branch-free: every lane issues its PER loads from in-bounds addresses
    // (indices clamped, the zero chunk read from chunk 0 and zeroed after),
    // so no exec-masked branch sits between a load and the products that
    // follow it; such branches made the compiler wait for every outstanding
    // load before the next chunk's products.
    template <int SET>
    __device__ __forceinline__ void load(const FistaArgs& a, int g, int nc1, int nc, int tid) {
        const bool second = g >= nc1;
        const bool zero = !second && g >= nc;
        const floatx4* base = second ? a.Afrag + size_t(g - nc1) * MB * 64 : a.ATfrag + size_t(zero ? 0 : g) * 64;
        const size_t smb = second ? 64 : size_t(a.NC) * 64;  // stride between coefficient blocks
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = min(tid + i * NT, NF - 1);
            const floatx4 f = base[size_t(e >> 6) * smb + (e & 63)];
            pre[SET][i] = zero ? floatx4{0.f, 0.f, 0.f, 0.f} : f;
        }
    }
    template <int SET>
    __device__ __forceinline__ void store(floatx4* buf, int tid) const {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = tid + i * NT;
            if (e < NF) buf[e] = pre[SET][i];
        }
    }
};

// Tikhonov L2 (src/solver/tikhonov.rs:59-76) with the signals read once:
// c = P s accumulates over the 16-sample chunks of each wave's 16 voxels,
// and the chunks stay in registers (4 samples per lane and chunk, in the
// layout the residual pass needs: rows 16 nb + 4h + r of voxel j on lane
// 16h + j), so |s - A c|^2 and the sum of squares about the mean use them
// without a second pass over the signals. NCM bounds the chunk count
// (N <= 16 NCM); W waves of 16 voxels per workgroup. The c = P s pass runs
// over an even number of chunks (a zero chunk after P's last when NC is
// odd), so that chunk parities, and with them the LDS buffer and register
// set of every step, are compile-time constants.
template <int MB, int NCM, int W>
__global__ void __launch_bounds__(64 * W) tikhonov_lds_kernel(const FistaArgs a, float* rss_l2) {
    static_assert(NCM % 2 == 0, "the first pass runs over pairs of chunks");
    __shared__ floatx4 fb[2][MB * 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 4, j = lane & 15;
    const int64_t v = (int64_t(blockIdx.x) * W + wave) * 16 + j;
    const bool valid = v < a.n_vox;
    const int64_t out_idx = valid ? (a.vox ? a.vox[v] : v) : 0;
    const float* sig = a.data + out_idx * a.sig_stride;
    const int NC = a.NC;             // <= NCM, checked by the host
    const int NC1 = (NC + 1) & ~1;   // first-pass chunks (even)
    const int total = NC1 + NC;
    ChunkStage<MB, 64 * W> st;
    st.template load<0>(a, 0, NC1, NC, tid);
    st.template load<1>(a, 1, NC1, NC, tid);  // total >= 3
    // every signal load of the tile is issued here, ahead of the products
    // (loops fully unrolled, so sreg stays in registers)
    float sreg[NCM][4];
    // Unguarded: chunks before the last are inside the signal (NC =
    // ceil(N / 16)) and load from one base with immediate offsets; in the
    // last chunk, rows past N read sample N - 1; they meet zero columns of
    // P in the first pass and are left out of the sums. A per-lane guard (a
    // select or an exec-masked load) would make the compiler wait for each
    // load at once, serializing all of them.
    const float* sb = sig + 4 * h;
#pragma unroll
    for (int ch = 0; ch < NCM; ++ch) {
        if (ch + 1 < NC) {  // uniform
#pragma unroll
            for (int r = 0; r < 4; ++r) sreg[ch][r] = sb[16 * ch + r];
        } else if (ch + 1 == NC) {
#pragma unroll
            for (int r = 0; r < 4; ++r) sreg[ch][r] = sig[min(16 * ch + 4 * h + r, a.N - 1)];
        } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) sreg[ch][r] = 0.f;
        }
    }
    st.template store<0>(fb[0], tid);
    __syncthreads();
    // one step of the chunk pipeline: request chunk g + 2 into the register
    // set chunk g came from, multiply chunk g, store chunk g + 1
    auto step_begin = [&](int g, auto par) {
        constexpr int P = decltype(par)::value;
        if (g + 2 < total) st.template load<P>(a, g + 2, NC1, NC, tid);
    };
    auto step_end = [&](int g, auto par) {
        constexpr int P = decltype(par)::value;
        if (g + 1 < total) st.template store<1 - P>(fb[1 - P], tid);
        __syncthreads();
    };
    floatx4 c[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) c[mb] = floatx4{0.f, 0.f, 0.f, 0.f};
    float sum = 0.f;
#pragma unroll
    for (int ch = 0; ch < NCM; ++ch) {
        if (ch < NC1) {
            if (ch & 1) step_begin(ch, std::integral_constant<int, 1>{});
            else step_begin(ch, std::integral_constant<int, 0>{});
#pragma unroll
            for (int r = 0; r < 4; ++r) sum += (16 * ch + 4 * h + r < a.N) ? sreg[ch][r] : 0.f;
            const floatx4* f = fb[ch & 1];
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                const floatx4 af = f[mb * 64 + lane];
#pragma unroll
                for (int r = 0; r < 4; ++r) c[mb] = mfma4(af[r], sreg[ch][r], c[mb]);
            }
            if (ch & 1) step_end(ch, std::integral_constant<int, 1>{});
            else step_end(ch, std::integral_constant<int, 0>{});
        }
    }
    const float mean = lane_sum4(sum) / float(a.N);
    float rss = 0.f, sst = 0.f;
#pragma unroll
    for (int nb = 0; nb < NCM; ++nb) {
        if (nb < NC) {
            const int g = NC1 + nb;  // same parity as nb
            if (nb & 1) step_begin(g, std::integral_constant<int, 1>{});
            else step_begin(g, std::integral_constant<int, 0>{});
            const floatx4* f = fb[nb & 1];
            // four independent accumulation chains (one per k-slot r)
            floatx4 ac[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) ac[r] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < MB; ++kb) {
                const floatx4 af = f[kb * 64 + lane];
#pragma unroll
                for (int r = 0; r < 4; ++r) ac[r] = mfma4(af[r], c[kb][r], ac[r]);
            }
            const floatx4 acc = (ac[0] + ac[1]) + (ac[2] + ac[3]);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int n = 16 * nb + 4 * h + r;
                if (valid && n < a.N) {
                    const float e = sreg[nb][r] - acc[r];
                    rss = fmaf(e, e, rss);
                    const float dm = sreg[nb][r] - mean;
                    sst = fmaf(dm, dm, sst);
                }
            }
            if (nb & 1) step_end(g, std::integral_constant<int, 1>{});
            else step_end(g, std::integral_constant<int, 0>{});
        }
    }
    rss = lane_sum4(rss);
    sst = lane_sum4(sst);
    if (!valid) return;
    if (a.coef)
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

// Epilogue of the persistent solvers (residual, R^2, BIC, scatter of the
// iterates they wrote in work order) with A staged in LDS by chunks of 16
// rows; each wave takes 16 voxels and reads their signals once, each
// chunk's signals requested one chunk ahead. The chunk loop is unrolled by
// two so that buffers and register sets are compile-time constants.
template <int MB, int W>
__global__ void __launch_bounds__(64 * W) epilogue_lds_kernel(const FistaArgs a, const float* X, const uint32_t* its,
                                                              const uint8_t* status, const float* alpha_v) {
    __shared__ floatx4 fb[2][MB * 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 4, j = lane & 15;
    const int64_t v = (int64_t(blockIdx.x) * W + wave) * 16 + j;
    const bool valid = v < a.n_vox;
    const int64_t out_idx = valid ? (a.vox ? a.vox[v] : v) : 0;
    const float* sig = a.data + out_idx * a.sig_stride;
    const int NC = a.NC;
    ChunkStage<MB, 64 * W> st;
    st.template load<0>(a, 0, 0, 0, tid);
    if (NC > 1) st.template load<1>(a, 1, 0, 0, tid);
    floatx4 x[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
        x[mb] = valid ? *reinterpret_cast<const floatx4*>(X + v * a.Kp + 16 * mb + 4 * h)
                      : floatx4{0.f, 0.f, 0.f, 0.f};
    const float mean = valid ? a.mean[v] : 0.f;
    float rss = 0.f, sst = 0.f;
    // Rows past N (and chunks past the last) read sample N - 1 and are left
    // out of the sums below, so the loads need no guard: an exec-masked load
    // or a select right after it would make the compiler wait for the load
    // at once instead of where the value is used.
    auto sload = [&](int nb, float (&s4)[4]) {
#pragma unroll
        for (int r = 0; r < 4; ++r) s4[r] = sig[min(16 * nb + 4 * h + r, a.N - 1)];
    };
    auto chunk = [&](int nb, const float (&s4)[4], auto par) {
        constexpr int P = decltype(par)::value;
        if (nb + 2 < NC) st.template load<P>(a, nb + 2, 0, 0, tid);
        const floatx4* f = fb[P];
        floatx4 ac[4];  // four independent accumulation chains (one per k-slot r)
#pragma unroll
        for (int r = 0; r < 4; ++r) ac[r] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < MB; ++kb) {
            const floatx4 af = f[kb * 64 + lane];
#pragma unroll
            for (int r = 0; r < 4; ++r) ac[r] = mfma4(af[r], x[kb][r], ac[r]);
        }
        const floatx4 acc = (ac[0] + ac[1]) + (ac[2] + ac[3]);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int n = 16 * nb + 4 * h + r;
            if (valid && n < a.N) {
                const float e = s4[r] - acc[r];
                rss = fmaf(e, e, rss);
                const float dm = s4[r] - mean;
                sst = fmaf(dm, dm, sst);
            }
        }
        if (nb + 1 < NC) st.template store<1 - P>(fb[1 - P], tid);
        __syncthreads();
    };
    // x, the mean and the first chunk are complete before the loop (one
    // wait here; left to itself the compiler keeps a full wait for the
    // loads of x inside every step)
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    float sa[4], sb[4];
    sload(0, sa);
    st.template store<0>(fb[0], tid);
    __syncthreads();
    for (int nb = 0; nb < NC; nb += 2) {
        sload(nb + 1, sb);
        chunk(nb, sa, std::integral_constant<int, 0>{});
        if (nb + 1 < NC) {
            sload(nb + 2, sa);
            chunk(nb + 1, sb, std::integral_constant<int, 1>{});
        }
    }
    rss = lane_sum4(rss);
    sst = lane_sum4(sst);
    if (!valid) return;
    const uint8_t st_v = status[v];
    write_voxel<MB>(a, out_idx, h, x, rss, sst, its[v], st_v & 1, double(alpha_v[v]), (st_v & 2) != 0);
}

// Chunks of the LDS Tikhonov kernel's signal registers: N <= 16 * 18 = 288.
constexpr int kTikLdsChunks = 18;

template <int MB>
void launch_rhs(const FistaLaunch& L, hipStream_t s) {
    constexpr int VPL = 2;
    const unsigned grid = unsigned((L.n_vox + 64 * VPL - 1) / (64 * VPL));
    hipLaunchKernelGGL((rhs_kernel<2 * MB, VPL>), dim3(grid), dim3(512), 0, s, L.data, L.vox, L.n_vox, L.sig_stride,
                       L.N, L.Kp, L.A64p, L.c_dev, L.Bbuf, L.meanbuf);
    SF_HIP(hipGetLastError());
}

// Persistent FP32 form (fixed alpha, scalar or per voxel; warm starts): the
// iteration of fista_f32_kernel, but each of a wave's 16 voxel slots takes a
// new voxel from a global counter as soon as its voxel stops, so a wave does
// not idle until the slowest of 16 voxels has converged. The momentum of a
// slot comes from the table of the FP64 recurrence (the values the plain
// kernel computes), and every voxel's arithmetic is the plain kernel's, so
// the coefficients and iteration counts are the same bit for bit.
// Each voxel's outputs are written as it stops: the coefficients and
// diagnostics directly, or (when residual maps are wanted) the iterate and
// its status in work order for epilogue_lds_kernel, as the mixed kernel
// does.
template <int MB, int WAVES, bool M_IN_LDS, bool PATH>
__global__ void __launch_bounds__(64 * WAVES) fista_f32_persist_kernel(const FistaArgs a, unsigned int* counter,
                                                                       float* X, uint32_t* its_out,
                                                                       uint8_t* status_out, float* alpha_out) {
    extern __shared__ floatx4 Ms[];
    const int lane = threadIdx.x & 63;
    const int h = lane >> 4, j = lane & 15;
    if (M_IN_LDS) {
        for (int i = threadIdx.x; i < MB * MB * 64; i += 64 * WAVES) Ms[i] = a.Mfrag[i];
        __syncthreads();
    }
    auto load_m = [&](int idx) -> floatx4 {
        if constexpr (M_IN_LDS) return Ms[idx];
        else return a.Mfrag[idx];
    };

    floatx4 b[MB], x[MB], y[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) b[mb] = x[mb] = y[mb] = floatx4{0.f, 0.f, 0.f, 0.f};
    int64_t v = 0, out_idx = 0;
    uint32_t its = 0;
    float thr = a.thr;
    double alpha_v = a.alpha;
    bool active = false, zero_voxel = false;
    bool need = true;  // this slot (all four lanes of column j agree) wants a voxel
    float bmax = 0.f;
    unsigned int wdone_lane = 0;  // voxels this slot has finished (on each of its four lanes)
    int pi = 0;        // alpha-path point of the slot (PATH)

    // outputs of the slot's voxel (src/solver/fista.rs:113-143 diagnostics):
    // with X, the iterate, iterations, status (bit 0 converged, bit 1 zero
    // voxel) and alpha in work order, for epilogue_lds_kernel; without,
    // the coefficients and diagnostics directly
    auto finish = [&](bool conv) {
        if (X) {
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) *reinterpret_cast<floatx4*>(X + v * a.Kp + 16 * mb + 4 * h) = x[mb];
            if (h == 0) {
                its_out[v] = its;
                status_out[v] = uint8_t((conv ? 1 : 0) | (zero_voxel ? 2 : 0));
                alpha_out[v] = zero_voxel ? 0.f : float(alpha_v);
            }
            return;
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 16 * mb + 4 * h + r;
                if (k < a.K) a.coef[out_idx * a.K + k] = x[mb][r];
            }
        if (h == 0) {
            if (a.iters) a.iters[out_idx] = its;
            if (a.converged) a.converged[out_idx] = conv ? 1 : 0;
            if (a.kind) a.kind[out_idx] = 1;
            if (a.alpha_map) a.alpha_map[out_idx] = zero_voxel ? 0.f : float(alpha_v);
        }
    };

    for (;;) {
        if (__ballot(need) != 0ull) {
            if (a.progress) {
                unsigned int tot = h == 0 ? wdone_lane : 0u;
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) tot += __shfl_xor(tot, o);
                publish_progress(a.progress, a.progress_slots, WAVES, tot);
            }
            // slots without a voxel take the next one (lane j of the column
            // draws it, the other three read it from lane j)
            unsigned int nv = 0;
            if (need && h == 0) nv = atomicAdd(counter, 1u);
            nv = __shfl(nv, j);
            bool load = false;
            if (need) {
                v = int64_t(nv);
                if (v >= a.n_vox) {
                    need = false;  // no work left for this slot
                } else {
                    load = true;
                    out_idx = a.vox ? a.vox[v] : v;
                }
            }
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
                if (load) b[mb] = *reinterpret_cast<const floatx4*>(a.B + v * a.Kp + 16 * mb + 4 * h);
            // alpha-ratio strategy: thr_v = ratio * max|b| (src/solver/alpha.rs:229-235);
            // path strategies: thr_i = step alpha_max f_i = max|b| f_i
            float bm = 0.f;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int r = 0; r < 4; ++r) bm = fmaxf(bm, fabsf(b[mb][r]));
            bm = lane_max4(bm);
            if (load) {
                bmax = bm;
                thr = a.thr;
                alpha_v = a.alpha;
                zero_voxel = false;
                pi = 0;
                if constexpr (PATH) {
                    zero_voxel = (bmax == 0.f);
                    thr = bmax * a.path_f[0];
                } else if (a.ratio > 0.f) {
                    zero_voxel = (bmax == 0.f);
                    thr = a.ratio * bmax;
                    alpha_v = double(bmax) / double(a.step) * double(a.ratio);
                }
                if (!PATH && a.alpha_vox) {  // per-voxel alpha (fixed strategy): thr_v = alpha_v / L
                    const float av = a.alpha_vox[v];
                    thr = float(a.step64 * double(av));
                    alpha_v = double(av);
                }
#pragma unroll
                for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        // warm start x0 = y0 = warm, t = 1 (src/solver/fista.rs:76-80)
                        const int k = 16 * mb + 4 * h + r;
                        x[mb][r] = (a.warm && k < a.K) ? a.warm[out_idx * a.K + k] : 0.f;
                        y[mb][r] = x[mb][r];
                    }
                its = 0;
                if (PATH && zero_voxel) {
                    ++wdone_lane;
                    // every path point of an all-zero voxel: 0 iterations, status 3
                    if (h == 0)
                        for (int i = 0; i < a.path_n; ++i) {
                            a.itspath[v * a.path_n + i] = 0;
                            a.convpath[v * a.path_n + i] = 3;
                        }
                } else if (!PATH && (zero_voxel || a.max_iter == 0)) {
                    ++wdone_lane;
                    finish(zero_voxel);  // nothing to iterate: take another voxel next round
                } else {
                    active = true;
                    need = false;
                }
            }
        }
        if (__ballot(active) == 0ull) {
            if (__ballot(need) == 0ull) break;
            continue;
        }
        int moff = 0;
        asm volatile("" : "+s"(moff));
        floatx4 acc[MB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            floatx4 c = b[mb];
#pragma unroll
            for (int kb = 0; kb < MB; ++kb) {
                const floatx4 mf = load_m(moff + (mb * MB + kb) * 64 + lane);
#pragma unroll
                for (int r = 0; r < 4; ++r) c = mfma4(mf[r], y[kb][r], c);
            }
            acc[mb] = c;
        }
        const float mom = a.mom[active ? its : 0];
        float dd = 0.f, xx = 0.f;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float val = y[mb][r] + acc[mb][r];
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
        bool next = false;
        if (active) {
            ++its;
            const float rel = sqrtf(dd) / fmaxf(sqrtf(xx), 1e-12f);
            const bool conv = rel < a.tol;
            if (conv || its >= a.max_iter) {
                if constexpr (PATH) {
                    const int64_t pt = v * a.path_n + pi;
#pragma unroll
                    for (int mb = 0; mb < MB; ++mb)
                        *reinterpret_cast<floatx4*>(a.Xpath + pt * a.Kp + 16 * mb + 4 * h) = x[mb];
                    if (h == 0) {
                        a.itspath[pt] = its;
                        a.convpath[pt] = conv ? 1 : 0;
                    }
                    if (pi + 1 < a.path_n) {
                        // warm start of the next point: x0 = y0 = x, t = 1
                        ++pi;
                        thr = bmax * a.path_f[pi];
                        its = 0;
                        next = true;
                    } else {
                        ++wdone_lane;
                        active = false;
                        need = true;
                    }
                } else {
                    ++wdone_lane;
                    finish(conv);
                    active = false;
                    need = true;
                }
            }
        }
        if constexpr (PATH) {
            if (__ballot(next) != 0ull) {
#pragma unroll
                for (int mb = 0; mb < MB; ++mb)
                    if (next) y[mb] = x[mb];
            }
        }
    }
}

template <int MB>
void launch_f32(const FistaArgs& args, int64_t n_vox, hipStream_t s) {
    constexpr int WAVES = (MB >= 16) ? 4 : 8;
    constexpr bool IN_LDS = (MB * MB * 64 * 16) <= 112 * 1024;
    const int64_t tiles = (n_vox + 15) / 16;
    const unsigned grid = unsigned((tiles + WAVES - 1) / WAVES);
    const size_t lds = IN_LDS ? size_t(MB) * MB * 64 * 16 : 0;
    auto kern = args.path_n > 0 ? fista_f32_kernel<MB, WAVES, IN_LDS, true> : fista_f32_kernel<MB, WAVES, IN_LDS, false>;
    if (lds > 64 * 1024)  // gfx950 has 160 KB of LDS per CU; a refusal surfaces at launch
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(lds));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * WAVES), lds, s, args);
    SF_HIP(hipGetLastError());
}


// ---------------------------------------------------------------------------
// Mixed-precision persistent form.
//
// Per voxel the kernel keeps x, y and acc = b - c G y, all scaled by a
// per-voxel power of two 2^e; the pre-threshold value is y + acc. Between
// exact refreshes acc advances by -c G dy on the 16x16x32 FP16 (or BF16)
// MFMA, accumulating straight into acc in FP32 with the FP16 copy of -cG
// staged in LDS. Only this -cG dy term sees the FP16 rounding of dy, and that
// rounding is relative to |dy|, which shrinks as the iterate converges (acc,
// not the larger y + acc, carries the FP32 accumulation, which keeps the
// rounding noise under the stopping test small). Every `refresh` iterations
// acc is recomputed exactly as b 2^e - c G y, which bounds the drift: with
// FP16 increments from hi/lo FP16 splits of -cG and y on three FP16 MFMA
// passes (~2^-22 relative, FP32-grade); with BF16 increments on the FP32
// MFMA (FP32 -cG read from L2). The scale is then re-chosen: 2^e puts the
// bound |dy| <= (1 + m_k)|d_k| + m_{k-1}|d_{k-1}| near 2^10, so the FP16 copy
// of dy has headroom; a slot whose bound would pass 2^15 forces an exact
// refresh before it is converted. Scaling by powers of two is exact, and the
// stopping ratio |d|/|x+| does not depend on it.
//
// Persistent scheduling with active-voxel compaction: a wave owns 16 slots;
// when a slot's voxel converges or reaches max_iter its solution is written
// out and the slot is refilled with the next voxel from a global counter, so
// MFMA tiles stay full while iteration counts vary from voxel to voxel. Empty
// slots hold zeros and need no predication. Residual, R^2 and BIC are
// computed afterwards by a batched epilogue.
// ---------------------------------------------------------------------------
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float2v __attribute__((ext_vector_type(2)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));

struct MixedArgs {
    const int64_t* vox;
    int64_t n_vox;
    int K, Kp;
    const float* B;         // [n_vox][Kp]
    const floatx4* M32;     // -cG, 16x16x4 f32 fragments (MB*MB*64)
    const uint4* M16;       // -cG, 16x16x32 f16/bf16 fragments (MB*KC*64)
    const uint4* M16lo;     // FP16 remainder of -cG, x 2^12 (split refresh)
    const uint4* Mlo_inc;   // remainder -cG - lowp(-cG), unscaled (split increments)
    const float* mom;       // momentum (t_k - 1)/t_{k+1}, k < max_iter
    float thr, step, ratio, alpha_fixed;
    uint32_t max_iter;
    float tol;
    int refresh;
    unsigned int* counter;  // work counter, zero at launch
    float* X;               // [n_vox][Kp] solutions, work order
    uint32_t* its;          // [n_vox]
    uint8_t* status;        // [n_vox] bit0 converged, bit1 all-zero voxel (alpha-ratio)
    float* alpha_v;         // [n_vox] alpha used
    // alpha path (PATH kernels): as in FistaArgs
    int path_n;
    const float* path_f;
    float* Xpath;
    uint32_t* itspath;
    uint8_t* convpath;
    unsigned int* progress;  // host-pinned finished-voxel count per wave (may be null)
    int progress_slots;
};

template <bool BF16>
__device__ __forceinline__ floatx4 mfma16(uint4 a, uint4 b, floatx4 c) {
    if constexpr (BF16)
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                       c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8, a), __builtin_bit_cast(half8, b), c,
                                                      0, 0, 0);
}

template <bool BF16>
__device__ __forceinline__ float2v unpack2(uint32_t u) {
    if constexpr (BF16)
        return __builtin_convertvector(__builtin_bit_cast(bf16x2v, u), float2v);
    else
        return __builtin_convertvector(__builtin_bit_cast(half2v, u), float2v);
}

template <bool BF16>
__device__ __forceinline__ uint32_t pack2(float lo, float hi) {
    if constexpr (BF16)
        return __builtin_bit_cast(uint32_t, __builtin_convertvector(float2v{lo, hi}, bf16x2v));
    else
        return __builtin_bit_cast(uint32_t, __builtin_convertvector(float2v{lo, hi}, half2v));
}

// 2^(target - E) where bound = m 2^E, m in [0.5, 1); 1 for bound == 0.
__device__ __forceinline__ float pow2_to(float bound, int target) {
    if (!(bound > 0.f) || !isfinite(bound)) return 1.f;
    int e = 0;
    (void)frexpf(bound, &e);
    return ldexpf(1.f, target - e);
}

// ---------------------------------------------------------------------------
// Tikhonov L2 apply on the FP16 MFMA (src/solver/tikhonov.rs:59-76), c = P s
// and |s - A c|^2 with the signals read once.
//
// Each operand x is carried as two FP16 numbers: x' = x 2^e scaled by a power
// of two into [2^14, 2^15) in magnitude (per row of P and of A, per voxel for
// s and for c), hi = fp16(x') and lo = fp16(x' - hi), and each product as
// hi*hi + hi*lo + lo*hi on one FP32 accumulation chain: the representation
// keeps ~22 bits and the dropped lo*lo term is 2^-22 of the product, so the
// error is of the order of FP32 arithmetic's (where lo is subnormal, below
// ~2^-3 in scaled units, it keeps 2^-25 of the scaled unit in absolute terms).
// Three 16x16x32 FP16 MFMAs replace eight 16x16x4 FP32 ones per 32 products,
// about 4.4x fewer matrix-core cycles.
//
// Pass 1 runs over chunks of two 16-sample blocks: the signals, scaled and
// split, are the B operand (lane (h, j) holds voxel j's samples 16 nb + 4h +
// r of blocks nb = 2g, 2g + 1, the layout the residual needs), P's fragments
// the A operand. Pass 2 runs over the same chunks: A's rows of the two
// sample blocks against c, scaled and split, as the B operand (c in the
// accumulator layout of pass 1: coefficients 16 mb + 4h + r, blocks 2 kc and
// 2 kc + 1 per k-chunk). The chunks are staged in LDS for the W waves of a
// workgroup as in the FP32 kernel (double buffered, requested from L2 two
// chunks ahead). Chunk counts are those of the largest signal (NCM sample
// blocks), zeros past the actual one, and pass 1 runs over an even number of
// them, so that chunk parities are compile-time constants.
struct Tik16Args {
    const float* data;
    const int64_t* vox;
    int64_t n_vox;
    int64_t sig_stride;
    int N, K, NC;
    const floatx4* H;    // G1 + G2 + 2 chunks (G2 = 9 for 18 sample blocks, G1 = 10) of 4 ceil(MB / 2) x 64 fragments
    const float* inv;    // inverse row scales: 16 MB rows of P, then the rows of A
    float* coef;
    float* resid;
    float* r2;
    uint32_t* iters;
    uint8_t* kind;
    float* rss_l2;
};

// (fragments as floatx4: a native vector type, which the compiler keeps in
// registers where HIP's uint4 struct made it stage the set in scratch)
// s_buffer resource word 3 for raw (untyped) buffer loads on gfx9 parts
constexpr int kBufferWord3 = 0x00020000;

template <int NF, int NT>
struct FlatStage {
    static constexpr int PER = (NF + NT - 1) / NT;
    floatx4 pre[2][PER];
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t off[PER];  // byte offsets of this thread's fragments in a chunk
    // Buffer loads: the chunk offset goes in the scalar offset and the lane
    // offsets stay fixed, so a chunk request costs no vector address
    // arithmetic (a global load takes a 64-bit address per lane).
    __device__ __forceinline__ void init(const floatx4* base, int n_chunks, int tid) {
        rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<floatx4*>(base), 0, n_chunks * NF * 16, kBufferWord3);
#pragma unroll
        for (int i = 0; i < PER; ++i) off[i] = uint32_t(min(tid + i * NT, NF - 1)) * 16u;
    }
    template <int SET>
    __device__ __forceinline__ void load(int chunk) {
#pragma unroll
        for (int i = 0; i < PER; ++i)
            pre[SET][i] = __builtin_bit_cast(floatx4, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off[i],
                                                                                            chunk * NF * 16, 0));
    }
    template <int SET>
    __device__ __forceinline__ void store(floatx4* buf, int tid) const {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = tid + i * NT;
            if (NF % NT == 0 || e < NF) buf[e] = pre[SET][i];
        }
    }
};

// eight scaled values as the hi and lo halves of an FP16 B operand
__device__ __forceinline__ void split8(const float (&t)[8], uint4& hi, uint4& lo) {
    uint32_t hv[4], lv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        hv[q] = pack2<false>(t[2 * q], t[2 * q + 1]);
        const float2v b = unpack2<false>(hv[q]);
        lv[q] = pack2<false>(t[2 * q] - b[0], t[2 * q + 1] - b[1]);
    }
    hi = uint4{hv[0], hv[1], hv[2], hv[3]};
    lo = uint4{lv[0], lv[1], lv[2], lv[3]};
}

template <int MB, int NCM, int W>
__global__ void __launch_bounds__(64 * W, 2) tikhonov_h16_kernel(const Tik16Args a) {  // two waves a SIMD
    static_assert(NCM % 2 == 0, "chunks of two sample blocks");
    constexpr int MB2 = (MB + 1) / 2;       // coefficient k-chunks of pass 2
    constexpr int NF = 4 * MB2 * 64;        // fragments per chunk (pass 1 uses 2 MB x 64)
    constexpr int GM = NCM / 2;             // pass-2 chunks
    constexpr int G1M = (GM + 1) & ~1;      // pass-1 chunks (even; zeros past the signal)
    constexpr int NR = 16 * MB + 32 * GM;   // row scales
    constexpr int NT = 64 * W;
    __shared__ floatx4 fb[2][NF];
    __shared__ float rs[NR];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 4, j = lane & 15;
    const int64_t v = (int64_t(blockIdx.x) * W + wave) * 16 + j;
    const bool valid = v < a.n_vox;
    const int64_t out_idx = valid ? (a.vox ? a.vox[v] : v) : 0;
    const float* sig = a.data + out_idx * a.sig_stride;
    const int NC = a.NC;                    // <= NCM, checked by the host
    // Fixed chunk counts and unconditional staging loads (the packer adds
    // two zero chunks at the end): a load issued under a branch makes the
    // compiler wait for every outstanding load where the paths join, which
    // would expose the L2 latency of the chunk requested two steps ahead.
    FlatStage<NF, NT> st;
    st.init(a.H, G1M + GM + 2, tid);
    st.template load<0>(0);
    st.template load<1>(1);
#pragma unroll
    for (int i = 0; i < (NR + NT - 1) / NT; ++i) {
        const int e = tid + i * NT;
        if (NR % NT == 0 || e < NR) rs[e] = a.inv[e];
    }
    // every signal load is issued here (unguarded: rows past N in the last
    // block read sample N - 1, meet zero columns of P and are left out of
    // the sums)
    float sreg[NCM][4];
    const float* sb = sig + 4 * h;
#pragma unroll
    for (int ch = 0; ch < NCM; ++ch) {
        if (ch + 1 < NC) {
#pragma unroll
            for (int r = 0; r < 4; ++r) sreg[ch][r] = sb[16 * ch + r];
        } else if (ch + 1 == NC) {
#pragma unroll
            for (int r = 0; r < 4; ++r) sreg[ch][r] = sig[min(16 * ch + 4 * h + r, a.N - 1)];
        } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) sreg[ch][r] = 0.f;
        }
    }
    st.template store<0>(fb[0], tid);
    // rows past N become zeros once here (every load has arrived: the scale
    // needs them all), so that no sum below needs a per-row mask
    float smax = 0.f, sum = 0.f;
#pragma unroll
    for (int ch = 0; ch < NCM; ++ch)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (ch + 1 >= NC) sreg[ch][r] = (16 * ch + 4 * h + r < a.N) ? sreg[ch][r] : 0.f;
            smax = fmaxf(smax, fabsf(sreg[ch][r]));
            sum += sreg[ch][r];
        }
    const float ss = fminf(pow2_to(lane_max4(smax), 15), 0x1p100f);  // finite for any signal
    __syncthreads();
    // one step of the chunk pipeline: request chunk g + 2 into the register
    // set chunk g came from, multiply chunk g, store chunk g + 1
    // (scheduling fences keep the requests at the start of the step and the
    // LDS stores at its end: left to itself the compiler sinks the requests
    // towards their stores, one step before the data is needed)
    auto step_begin = [&](int g, auto par) __attribute__((always_inline)) {
        constexpr int P = decltype(par)::value;
        st.template load<P>(g + 2);
        __builtin_amdgcn_sched_barrier(0);
    };
    auto step_end = [&](auto par) __attribute__((always_inline)) {
        constexpr int P = decltype(par)::value;
        __builtin_amdgcn_sched_barrier(0);
        st.template store<1 - P>(fb[1 - P], tid);
        __syncthreads();
    };
    floatx4 c[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) c[mb] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int g = 0; g < G1M; ++g) {
        if (g & 1) step_begin(g, std::integral_constant<int, 1>{});
        else step_begin(g, std::integral_constant<int, 0>{});
        if (g < GM) {  // compile-time: the last pass-1 chunk may be padding
            float t[8];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                t[r] = sreg[2 * g < NCM ? 2 * g : 0][r] * ss;
                t[4 + r] = sreg[2 * g + 1 < NCM ? 2 * g + 1 : 0][r] * ss;
            }
            uint4 bh, bl;
            split8(t, bh, bl);
            const floatx4* f = fb[g & 1];
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                const uint4 fh = __builtin_bit_cast(uint4, f[(2 * mb) * 64 + lane]);
                const uint4 fl = __builtin_bit_cast(uint4, f[(2 * mb + 1) * 64 + lane]);
                c[mb] = mfma16<false>(fh, bh, c[mb]);
                c[mb] = mfma16<false>(fh, bl, c[mb]);
                c[mb] = mfma16<false>(fl, bh, c[mb]);
            }
            // two blocks of fragments in flight from LDS ahead of the MFMAs
            __builtin_amdgcn_sched_group_barrier(0x100, MB < 2 ? 2 * MB : 4, 0);
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                __builtin_amdgcn_sched_group_barrier(0x008, 3, 0);
                if (mb + 2 < MB) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
            }
        }
        if (g & 1) step_end(std::integral_constant<int, 1>{});
        else step_end(std::integral_constant<int, 0>{});
    }
    // c = c' / (row scale x ss), written out; then scaled and split as the
    // B operand of pass 2
    const float iss = 1.f / ss;
    float cm = 0.f;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            c[mb][r] *= rs[16 * mb + 4 * h + r] * iss;
            cm = fmaxf(cm, fabsf(c[mb][r]));
        }
    if (valid && a.coef)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 16 * mb + 4 * h + r;
                if (k < a.K) a.coef[out_idx * a.K + k] = c[mb][r];
            }
    const float cs = fminf(pow2_to(lane_max4(cm), 15), 0x1p100f), ics = 1.f / cs;
    uint4 chh[MB2], cll[MB2];
#pragma unroll
    for (int kc = 0; kc < MB2; ++kc) {
        float t[8];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            t[r] = c[2 * kc][r] * cs;
            t[4 + r] = 2 * kc + 1 < MB ? c[2 * kc + 1 < MB ? 2 * kc + 1 : 0][r] * cs : 0.f;
        }
        split8(t, chh[kc], cll[kc]);
    }
    const float mean = lane_sum4(sum) / float(a.N);
    float lastm[4];  // rows of the last sample block inside the signal
#pragma unroll
    for (int r = 0; r < 4; ++r) lastm[r] = 16 * (NC - 1) + 4 * h + r < a.N ? 1.f : 0.f;
    float rss = 0.f, sst = 0.f;
    const float* ars = rs + 16 * MB;
#pragma unroll
    for (int g2 = 0; g2 < GM; ++g2) {
        const int g = G1M + g2;  // same parity as g2
        if (g2 & 1) step_begin(g, std::integral_constant<int, 1>{});
        else step_begin(g, std::integral_constant<int, 0>{});
        const floatx4* f = fb[g2 & 1];
        floatx4 ac[2];  // sample blocks 2 g2 and 2 g2 + 1, two chains
#pragma unroll
        for (int b = 0; b < 2; ++b) ac[b] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < MB2; ++kc)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const uint4 fh = __builtin_bit_cast(uint4, f[(2 * (b * MB2 + kc)) * 64 + lane]);
                const uint4 fl = __builtin_bit_cast(uint4, f[(2 * (b * MB2 + kc) + 1) * 64 + lane]);
                ac[b] = mfma16<false>(fh, chh[kc], ac[b]);
                ac[b] = mfma16<false>(fh, cll[kc], ac[b]);
                ac[b] = mfma16<false>(fl, chh[kc], ac[b]);
            }
        __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
#pragma unroll
        for (int kc = 0; kc < MB2; ++kc) {
            __builtin_amdgcn_sched_group_barrier(0x008, 6, 0);
            if (kc + 1 < MB2) __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
        }
        // rows past N: zero signal against zero rows of A, so e = 0; blocks
        // past the signal are skipped and the last one masks x - mean
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int nb = 2 * g2 + b;
            if (nb < NC) {  // uniform
                float dm[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float x = sreg[nb][r];
                    const float e = fmaf(-ac[b][r], ars[16 * nb + 4 * h + r] * ics, x);
                    rss = fmaf(e, e, rss);
                    dm[r] = x - mean;
                }
                if (nb + 1 == NC)  // uniform
#pragma unroll
                    for (int r = 0; r < 4; ++r) dm[r] *= lastm[r];
#pragma unroll
                for (int r = 0; r < 4; ++r) sst = fmaf(dm[r], dm[r], sst);
            }
        }
        if (g2 & 1) step_end(std::integral_constant<int, 1>{});
        else step_end(std::integral_constant<int, 0>{});
    }
    rss = lane_sum4(rss);
    sst = lane_sum4(sst);
    if (!valid || h != 0) return;
    const double ss_res = double(rss);
    if (a.resid) a.resid[out_idx] = float(sqrt(ss_res));
    if (a.r2) a.r2[out_idx] = float(sst > 0.f ? 1.0 - ss_res / double(sst) : 0.0);
    if (a.iters) a.iters[out_idx] = 1;
    if (a.kind) a.kind[out_idx] = 2;
    if (a.rss_l2) a.rss_l2[out_idx] = float(ss_res);
}

// Epilogue of the persistent solvers with the residual pass of the split-FP16
// Tikhonov kernel: |s - A x|^2 on the FP16 MFMA (hi/lo splits of A's scaled
// rows and of the voxel's scaled x, hi*hi + hi*lo + lo*hi on FP32 chains),
// A staged in LDS by chunks of two 16-sample blocks, the signals in
// registers; then R^2, BIC and the scatter of x (src/fit.rs:120-128,
// src/solver/alpha.rs:99-103).
template <int MB, int NCM, int W>
__global__ void __launch_bounds__(64 * W, 2) epilogue_h16_kernel(const FistaArgs a, const floatx4* H,
                                                                 const float* inv, const float* X,
                                                                 const uint32_t* its, const uint8_t* status,
                                                                 const float* alpha_v) {
    static_assert(NCM % 2 == 0, "chunks of two sample blocks");
    constexpr int MB2 = (MB + 1) / 2;
    constexpr int NF = 4 * MB2 * 64;
    constexpr int GM = NCM / 2;
    constexpr int NR = 16 * MB + 32 * GM;
    constexpr int NT = 64 * W;
    __shared__ floatx4 fb[2][NF];
    __shared__ float rs[NR];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 4, j = lane & 15;
    const int64_t v = (int64_t(blockIdx.x) * W + wave) * 16 + j;
    const bool valid = v < a.n_vox;
    const int64_t out_idx = valid ? (a.vox ? a.vox[v] : v) : 0;
    const float* sig = a.data + out_idx * a.sig_stride;
    const int NC = a.NC;  // <= NCM, checked by the host
    FlatStage<NF, NT> st;
    st.init(H, GM + 2, tid);
    st.template load<0>(0);
    st.template load<1>(1);
#pragma unroll
    for (int i = 0; i < (NR + NT - 1) / NT; ++i) {
        const int e = tid + i * NT;
        if (NR % NT == 0 || e < NR) rs[e] = inv[e];
    }
    floatx4 x[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
        x[mb] = valid ? *reinterpret_cast<const floatx4*>(X + v * a.Kp + 16 * mb + 4 * h)
                      : floatx4{0.f, 0.f, 0.f, 0.f};
    const float mean = valid ? a.mean[v] : 0.f;
    // every signal load is issued here, before the products (as in the
    // Tikhonov kernel; requesting each chunk's samples one chunk ahead left
    // the steps waiting on HBM latency). Rows past N read sample N - 1: they
    // meet zero rows of A and are masked out of the sums.
    float sreg[NCM][4];
    {
        const float* sbp = sig + 4 * h;
#pragma unroll
        for (int ch = 0; ch < NCM; ++ch) {
            if (ch + 1 < NC) {
#pragma unroll
                for (int r = 0; r < 4; ++r) sreg[ch][r] = sbp[16 * ch + r];
            } else if (ch + 1 == NC) {
#pragma unroll
                for (int r = 0; r < 4; ++r) sreg[ch][r] = sig[min(16 * ch + 4 * h + r, a.N - 1)];
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) sreg[ch][r] = 0.f;
            }
        }
    }
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    float xm = 0.f;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) xm = fmaxf(xm, fabsf(x[mb][r]));
    const float cs = fminf(pow2_to(lane_max4(xm), 15), 0x1p100f), ics = 1.f / cs;
    uint4 chh[MB2], cll[MB2];
#pragma unroll
    for (int kc = 0; kc < MB2; ++kc) {
        float t[8];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            t[r] = x[2 * kc][r] * cs;
            t[4 + r] = 2 * kc + 1 < MB ? x[2 * kc + 1 < MB ? 2 * kc + 1 : 0][r] * cs : 0.f;
        }
        split8(t, chh[kc], cll[kc]);
    }
    // the coefficients go out here, so that x is dead during the residual pass
    const int nnz = write_coefs<MB>(a, out_idx, h, x, valid);
    float lastm[4];  // rows of the last sample block inside the signal
#pragma unroll
    for (int r = 0; r < 4; ++r) lastm[r] = 16 * (NC - 1) + 4 * h + r < a.N ? 1.f : 0.f;
    st.template store<0>(fb[0], tid);
    __syncthreads();
    const float* ars = rs + 16 * MB;
    float rss = 0.f, sst = 0.f;
    auto chunk = [&](int g, auto par) __attribute__((always_inline)) {
        constexpr int P = decltype(par)::value;
        st.template load<P>(g + 2);
        __builtin_amdgcn_sched_barrier(0);
        const floatx4* f = fb[P];
        floatx4 ac[2];
#pragma unroll
        for (int b = 0; b < 2; ++b) ac[b] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < MB2; ++kc)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const uint4 fh = __builtin_bit_cast(uint4, f[(2 * (b * MB2 + kc)) * 64 + lane]);
                const uint4 fl = __builtin_bit_cast(uint4, f[(2 * (b * MB2 + kc) + 1) * 64 + lane]);
                ac[b] = mfma16<false>(fh, chh[kc], ac[b]);
                ac[b] = mfma16<false>(fh, cll[kc], ac[b]);
                ac[b] = mfma16<false>(fl, chh[kc], ac[b]);
            }
        __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
#pragma unroll
        for (int kc = 0; kc < MB2; ++kc) {
            __builtin_amdgcn_sched_group_barrier(0x008, 6, 0);
            if (kc + 1 < MB2) __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int nb = 2 * g + b;
            if (nb < NC) {  // uniform
                float e[4], dm[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    e[r] = fmaf(-ac[b][r], ars[16 * nb + 4 * h + r] * ics, sreg[nb][r]);
                    dm[r] = sreg[nb][r] - mean;
                }
                if (nb + 1 == NC)  // uniform
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        e[r] *= lastm[r];
                        dm[r] *= lastm[r];
                    }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    rss = fmaf(e[r], e[r], rss);
                    sst = fmaf(dm[r], dm[r], sst);
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        st.template store<1 - P>(fb[1 - P], tid);
        __syncthreads();
    };
#pragma unroll
    for (int g = 0; g < GM; g += 2) {
        chunk(g, std::integral_constant<int, 0>{});
        if (g + 1 < GM) chunk(g + 1, std::integral_constant<int, 1>{});
    }
    rss = lane_sum4(rss);
    sst = lane_sum4(sst);
    if (!valid) return;
    const uint8_t st_v = status[v];
    write_maps(a, out_idx, h, nnz, rss, sst, its[v], st_v & 1, double(alpha_v[v]), (st_v & 2) != 0);
}

template <int MB>
void launch_epilogue(const FistaArgs& args, const float* X, const uint32_t* its, const uint8_t* status,
                     const float* alpha_v, const uint16_t* AH16, const float* Ainv16, hipStream_t s) {
    const int64_t tiles = (args.n_vox + 15) / 16;
    if constexpr (MB <= kTik16MaxMB) {
        if (AH16 && args.NC <= kTik16MaxBlocks) {
            constexpr int W = 4;
            hipLaunchKernelGGL((epilogue_h16_kernel<MB, kTik16MaxBlocks, W>), dim3(unsigned((tiles + W - 1) / W)),
                               dim3(64 * W), 0, s, args, reinterpret_cast<const floatx4*>(AH16), Ainv16, X, its,
                               status, alpha_v);
            SF_HIP(hipGetLastError());
            return;
        }
    }
    constexpr int W = 4;
    hipLaunchKernelGGL((epilogue_lds_kernel<MB, W>), dim3(unsigned((tiles + W - 1) / W)), dim3(64 * W), 0, s, args, X,
                       its, status, alpha_v);
    SF_HIP(hipGetLastError());
}

template <int MB>
void launch_tik(const FistaArgs& args, const TikhonovLaunch& L, hipStream_t s) {
    const int64_t n_vox = L.n_vox;
    const int64_t tiles = (n_vox + 15) / 16;
    if constexpr (MB <= kTik16MaxMB) {
        if (L.H16 && args.NC <= kTik16MaxBlocks) {
            constexpr int W = 4;
            Tik16Args t{};
            t.data = args.data;
            t.vox = args.vox;
            t.n_vox = n_vox;
            t.sig_stride = args.sig_stride;
            t.N = args.N;
            t.K = args.K;
            t.NC = args.NC;
            t.H = reinterpret_cast<const floatx4*>(L.H16);
            t.inv = L.inv16;
            t.coef = args.coef;
            t.resid = args.resid;
            t.r2 = args.r2;
            t.iters = args.iters;
            t.kind = args.kind;
            t.rss_l2 = L.rss_l2;
            hipLaunchKernelGGL((tikhonov_h16_kernel<MB, kTik16MaxBlocks, W>), dim3(unsigned((tiles + W - 1) / W)),
                               dim3(64 * W), 0, s, t);
            SF_HIP(hipGetLastError());
            return;
        }
    }
    float* rss_l2 = L.rss_l2;
    if (args.NC <= kTikLdsChunks) {
        constexpr int W = 4;
        hipLaunchKernelGGL((tikhonov_lds_kernel<MB, kTikLdsChunks, W>), dim3(unsigned((tiles + W - 1) / W)),
                           dim3(64 * W), 0, s, args, rss_l2);
        SF_HIP(hipGetLastError());
        return;
    }
    if constexpr (MB >= 8 && MB <= 10) {
        constexpr int T = 2;
        hipLaunchKernelGGL((tikhonov_multi_kernel<MB, T>), dim3(unsigned((tiles + 4 * T - 1) / (4 * T))), dim3(256),
                           0, s, args, rss_l2);
        SF_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(tikhonov_kernel<MB>, dim3(unsigned((tiles + 3) / 4)), dim3(256), 0, s, args, rss_l2);
    SF_HIP(hipGetLastError());
}

// SPLITINC: the increments apply both halves of the split operator,
// lowp(-cG) and the unscaled remainder -cG - lowp(-cG) (FP16 subnormal where
// |cG| < 2^-2; the FP16 MFMA takes subnormal inputs as they are), chained on
// one accumulator, so the operator acting on dy carries ~2^-22 (FP16) /
// 2^-17 (BF16) relative error instead of the single rounding's 2^-12 / 2^-9. That error would otherwise shift the
// small eigenvalues of c G, the slow modes that decide the iteration at which
// |dx| / |x| crosses tol, and so move the stopping iteration away from the
// FP64 fit's. Both halves live in LDS (2 x 51 KB at order 8); each slot's b
// is read from global memory when the refresh needs it. Without SPLITINC
// (order 10, whose 128 KB FP16 matrix fills LDS alone) only lowp(-cG) is
// applied and b stays in registers.
// Each increment's MFMA chain starts from zero and is added to acc with one
// rounding, so the chain's roundings scale with the increment, not with acc.
// SPLIT: refresh with FP16 hi/lo splits (three FP16 MFMA passes, ~2^-22)
// instead of the FP32 MFMA (FP16 increments only). Measured on the whole-brain
// fit, the FP32 refresh moves fewer stopping iterations off the FP64 fit's
// (92.8% equal against 89.5% at refresh 40) for 6% more solve time, so the
// split refresh is kept only where the increments are not split.
// KL: the coefficient count when it is known at compile time (0: every padded
// row is treated as live). Rows from KL to the padded count stay exactly zero
// (zero rows of -cG and of b), so registers that hold only such rows are
// skipped: at K = 145 in 160 rows, registers 1-3 of the last block; at K = 72
// in 96, the whole last block, MFMAs included.
template <int MB, int KL>
struct LiveRows {
    static constexpr int MBL = KL > 0 ? (KL + 15) / 16 : MB;                 // blocks with a real row
    static constexpr int RL = KL > 0 ? (KL - 16 * (MBL - 1) < 4 ? KL - 16 * (MBL - 1) : 4) : 4;
    static constexpr bool reg(int mb, int r) { return mb < MBL - 1 || (mb == MBL - 1 && r < RL); }
};

template <int MB, int WAVES, bool BF16, bool SPLITINC, bool NONNEG, bool SPLIT, bool PATH, int KL = 0>
__global__ void __launch_bounds__(64 * WAVES) fista_mixed_kernel(const MixedArgs a) {
    constexpr int KC = MB / 2;
    using LV = LiveRows<MB, KL>;
    constexpr int NF = MB * KC * 64;  // uint4 fragments of one operator half
    extern __shared__ uint4 smem[];
    uint4* Ml = smem;  // lowp(-cG), then (SPLITINC) the remainder
    const int lane = threadIdx.x & 63;
    const int h = lane >> 4, j = lane & 15;
    for (int i = threadIdx.x; i < NF; i += 64 * WAVES) {
        Ml[i] = a.M16[i];
        if constexpr (SPLITINC) Ml[NF + i] = a.Mlo_inc[i];
    }
    __syncthreads();

    floatx4 x[MB], y[MB], acc[MB], breg[MB];
    floatx4 x2[MB];  // the other register set of x in the inner loop
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) x2[mb] = floatx4{0.f, 0.f, 0.f, 0.f};  // its dead rows stay zero
    int64_t vid = -1;
    uint32_t it = 0;
    float thr_u = a.thr, thr = a.thr, sc = 1.f, mprev = 0.f, dprev = 0.f, mom = 0.f;
    float xref = 0.f, drift = 0.f;  // |x| at its last evaluation, sum of |dx| since
    bool alive = false, zero = false;
    int pi = 0;        // path point of this slot (PATH)
    float bmx = 0.f;   // max|b| of this slot's voxel (PATH)

    // Claims new voxels for the slots flagged in `need` (one atomic per
    // wave) and loads them: x = y = 0, v = b 2^e. An all-zero voxel of the
    // alpha-ratio strategy (max|b| = 0) runs like any other: it converges at
    // its first iteration with x = 0, and is reported the way
    // src/solver/alpha.rs:178-189 does.
    auto refill = [&](bool need) {
        const unsigned long long bal = __ballot(need && h == 0);
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(a.counter, unsigned(__popcll(bal)));
        base = __shfl(base, 0);
        if (need) {
            vid = int64_t(base) + __popcll(bal & ((1ull << j) - 1ull));
            alive = vid < a.n_vox;
            it = 0;
            mprev = dprev = mom = 0.f;  // mom = (t_0 - 1) / t_1
            xref = drift = 0.f;
        }
        float m = 0.f, n2 = 0.f;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            floatx4 bq = floatx4{0.f, 0.f, 0.f, 0.f};
            if (need && alive) bq = *reinterpret_cast<const floatx4*>(a.B + vid * a.Kp + 16 * mb + 4 * h);
            if constexpr (!SPLITINC)
                if (need) breg[mb] = bq;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                m = fmaxf(m, fabsf(bq[r]));
                n2 = fmaf(bq[r], bq[r], n2);
            }
            if (need) {
                x[mb] = y[mb] = floatx4{0.f, 0.f, 0.f, 0.f};
                acc[mb] = bq;
            }
        }
        m = lane_max4(m);  // alpha-ratio: thr_v = ratio * max|b| (src/solver/alpha.rs:229-235)
        n2 = lane_sum4(n2);
        if (need) {
            if constexpr (PATH) {
                // path point i: thr = step alpha_max f_i = max|b| f_i
                pi = 0;
                bmx = m;
                thr_u = m * a.path_f[0];
                zero = alive && m == 0.f;
            } else {
                thr_u = a.ratio > 0.f ? a.ratio * m : a.thr;
                zero = alive && a.ratio > 0.f && m == 0.f;
            }
            // the first steps move x by about |b|
            sc = pow2_to(sqrtf(n2), 10);
            thr = thr_u * sc;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) acc[mb] *= sc;
        }
    };
    refill(true);

    // |x| of every slot's voxel (FP32 sum of squares in a fixed order).
    // The opaque start value keeps the compiler from merging the two uses
    // and hoisting the sum onto the every-iteration path.
    auto xnorm = [&](const floatx4(&xv)[MB]) {
        float q = 0.f;
        asm volatile("" : "+v"(q));
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (LV::reg(mb, r)) q = fmaf(xv[mb][r], xv[mb][r], q);
        return sqrtf(lane_sum4(q));
    };

    // Writes out a finished slot (or, PATH, its last path point).
    auto write_out = [&](const floatx4(&xv)[MB], bool cv) {
        const float isc = 1.f / sc;  // exact: sc is a power of two
        if constexpr (PATH) {
            const int64_t pt = vid * a.path_n + pi;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
                *reinterpret_cast<floatx4*>(a.Xpath + pt * a.Kp + 16 * mb + 4 * h) = xv[mb] * isc;
            if (h == 0) {
                a.itspath[pt] = zero ? 0u : it;
                a.convpath[pt] = uint8_t((cv ? 1 : 0) | (zero ? 2 : 0));
            }
        } else {
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
                *reinterpret_cast<floatx4*>(a.X + vid * a.Kp + 16 * mb + 4 * h) = xv[mb] * isc;
            if (h == 0) {
                a.its[vid] = zero ? 0u : it;
                a.status[vid] = uint8_t((cv ? 1 : 0) | (zero ? 2 : 0));
                a.alpha_v[vid] = zero ? 0.f : (a.ratio > 0.f ? thr_u / a.step : a.alpha_fixed);
            }
        }
    };

    uint32_t wit = 0;
    unsigned int wdone = 0;  // voxels this wave has finished
    const float tol = a.tol;
    // Outer loop: write out finished slots and refill them. Inner loop:
    // iterate until some slot finishes. Keeping the refill (which rewrites
    // x, y, acc) out of the inner loop keeps register copies off the
    // per-iteration path.
    for (;;) {
        if (__ballot(alive) == 0ull) break;
        bool done = false, conv = false, wrote = false;
        // One FISTA iteration from xi to xo (y and acc in place). The inner
        // loop runs two per trip and alternates the two register sets of x,
        // so x_new, which is live together with x until d = x_new - x, is
        // never copied over x. A slot that finishes in the first half is
        // written out there and sits out the second half (its lanes compute
        // one discarded iteration); the trip then ends and the slots refill.
        auto step = [&](floatx4(&xi)[MB], floatx4(&xo)[MB], auto first) __attribute__((always_inline)) -> bool {
            // Wave priority: the element-wise phase runs at the base level,
            // the stopping test and the increment (MFMA) phase one level up,
            // so that when both waves of a SIMD are ready the one feeding the
            // matrix cores issues first and the other fills in VALU work
            // (3% on the whole-brain fit).
            __builtin_amdgcn_s_setprio(0);
            // An opaque zero offset keeps the (loop-invariant) fragment loads
            // inside the loop instead of hoisted into registers for all blocks.
            int moff = 0;
            asm volatile("" : "+s"(moff));
            float dd = 0.f;
            // dy = y_new - y_old goes to FP16/BF16 here, while y_old is still
            // in its registers; the opaque asm keeps the conversion from
            // being sunk into the increment branch, which would keep both
            // copies of y alive across the branch.
            uint32_t dyh[KC][4];
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                float t[8];
#pragma unroll
                for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int mb = 2 * kc + s2;
                        if (!LV::reg(mb, r)) {  // padding rows: x, y, acc and dy are zero
                            t[4 * s2 + r] = 0.f;
                            continue;
                        }
                        const float val = y[mb][r] + acc[mb][r];
                        float xn = val - __builtin_amdgcn_fmed3f(val, -thr, thr);  // strict soft threshold
                        if constexpr (NONNEG) xn = fmaxf(xn, 0.f);
                        const float d = xn - xi[mb][r];
                        const float yn = fmaf(mom, d, xn);
                        dd = fmaf(d, d, dd);
                        t[4 * s2 + r] = yn - y[mb][r];
                        xo[mb][r] = xn;
                        y[mb][r] = yn;
                    }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    dyh[kc][q] = pack2<BF16>(t[2 * q], t[2 * q + 1]);
                    asm volatile("" : "+v"(dyh[kc][q]));
                }
            }
            dd = lane_sum4(dd);
            const float nd = __builtin_amdgcn_sqrtf(dd);  // V_SQRT_F32, 1 ulp
            // Stopping rule nd / max(|x_new|, 1e-12) < tol without summing
            // |x_new|^2 every iteration: |x_new| lies within xref -+ drift
            // (triangle inequality over the steps since xref was computed),
            // widened by 4e-5 for the rounding of the FP32 norms. Only when
            // that interval straddles the threshold is |x_new| summed, in the
            // same order as always, so the decision is the one the exact test
            // gives.
            drift += nd;
            const float fl = 1e-12f * sc;
            const float lo = fmaxf(fmaf(xref, 1.f - 4e-5f, -drift * (1.f + 4e-5f)), fl);
            const float hi = fmaxf((xref + drift) * (1.f + 4e-5f), fl);
            bool cv = nd < tol * lo * (1.f - 1e-6f);
            const bool amb = !cv && !(nd > tol * hi * (1.f + 1e-6f));
            if (__ballot(alive && amb) != 0ull) {
                const float xs = xnorm(xo);
                if (amb) cv = nd / fmaxf(xs, fl) < tol;
                xref = xs;
                drift = 0.f;
            }
            if (alive) {
                ++it;
                conv = cv;
                done = conv || it >= a.max_iter;
            }
            bool next = false;  // PATH: this slot moves on to its next path point
            if constexpr (PATH) {
                next = done && pi + 1 < a.path_n;
                if (next) done = false;
                if (__ballot(next) != 0ull) {
                    if (next) {
                        const float isc = 1.f / sc;
                        const int64_t pt = vid * a.path_n + pi;
#pragma unroll
                        for (int mb = 0; mb < MB; ++mb)
                            *reinterpret_cast<floatx4*>(a.Xpath + pt * a.Kp + 16 * mb + 4 * h) = xo[mb] * isc;
                        if (h == 0) {
                            a.itspath[pt] = zero ? 0u : it;
                            a.convpath[pt] = uint8_t((conv ? 1 : 0) | (zero ? 2 : 0));
                        }
                        ++pi;
                        thr_u = bmx * a.path_f[pi];
                        thr = thr_u * sc;
                        it = 0;
                    }
                    // Warm start x0 = y0 = x, t = 1 (src/solver/fista.rs:76-80,
                    // 165-180): the increment takes y from y_k to x instead of
                    // to the extrapolated point, so dy loses y - x.
#pragma unroll
                    for (int kc = 0; kc < KC; ++kc)
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int mb = 2 * kc + (q >> 1), r0 = 2 * (q & 1);
                            float2v t2 = unpack2<BF16>(dyh[kc][q]);
                            t2[0] -= y[mb][r0] - xo[mb][r0];
                            t2[1] -= y[mb][r0 + 1] - xo[mb][r0 + 1];
                            if (next) dyh[kc][q] = pack2<BF16>(t2[0], t2[1]);
                        }
#pragma unroll
                    for (int mb = 0; mb < MB; ++mb)
                        if (next) y[mb] = xo[mb];
                }
            }
            __builtin_amdgcn_s_setprio(2);
            const float bound = (1.f + mom) * nd + mprev * dprev;
            mprev = next ? 0.f : mom;
            dprev = nd;
            // next iteration's momentum, fetched while the MFMA phase runs
            // (it < max_iter for every slot that is still running)
            const float mom_next = a.mom[it < a.max_iter ? it : 0];
            ++wit;
            const bool overflow = __ballot(alive && !(bound <= 32768.f)) != 0ull;
            if (overflow || wit % uint32_t(a.refresh) == 0) {
                // exact refresh: a new scale 2^e for the next window (capped
                // so |x| stays < 2^40), then acc = b 2^e - c G y.
                // Matrix fragments from global memory (L2): a uniform base plus
                // a 32-bit per-lane offset (one VALU add per load).
                const uint32_t loff = (uint32_t(moff) + uint32_t(lane)) * 16u;
                const char* lo16 = reinterpret_cast<const char*>(a.M16lo);
                if constexpr (SPLITINC) {
                    // b of the slot's voxel into acc (dead until rewritten
                    // below); the loads complete behind the VALU work
#pragma unroll
                    for (int mb = 0; mb < MB; ++mb)
                        acc[mb] = alive ? *reinterpret_cast<const floatx4*>(a.B + vid * a.Kp + 16 * mb + 4 * h)
                                        : floatx4{0.f, 0.f, 0.f, 0.f};
                }
                constexpr int LA = MB < 2 ? MB : 2;  // blocks of remainder fragments in flight
                uint4 glo[LA + 1][KC];
                if constexpr (SPLIT) {
                    // the first blocks' remainder fragments are requested before
                    // the VALU work below, which covers their L2 latency
#pragma unroll
                    for (int i = 0; i < LA * KC; ++i)
                        glo[i / KC][i % KC] = *reinterpret_cast<const uint4*>(lo16 + (loff + uint32_t(i) * 1024u));
                    __builtin_amdgcn_sched_barrier(0);
                }
                const float xs = xnorm(xo);
                const float f = fminf(pow2_to(bound, 10), pow2_to(xs, 40));
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) {
                    xo[mb] *= f;
                    y[mb] *= f;
                }
                sc *= f;
                thr = thr_u * sc;
                dprev *= f;
                xref = xs * f;
                drift = 0.f;
                if constexpr (SPLIT) {
                    // -cG y from FP16 hi/lo splits of both operands:
                    // hi*hi + hi*lo on one chain, lo*hi (lo x 2^12) on another;
                    // the dropped lo*lo term is ~2^-22 of |cG||y|, like FP32.
                    float ym = 0.f;
#pragma unroll
                    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                        for (int r = 0; r < 4; ++r) ym = fmaxf(ym, fabsf(y[mb][r]));
                    ym = lane_max4(ym);
                    const float sy = pow2_to(ym, 14), isy = 1.f / sy;
                    uint4 yh[KC], yl[KC];
#pragma unroll
                    for (int kc = 0; kc < KC; ++kc) {
                        float t[8], e[8];
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            t[q] = y[2 * kc][q] * sy;
                            t[4 + q] = y[2 * kc + 1][q] * sy;
                        }
#pragma unroll
                        for (int q = 0; q < 8; ++q) e[q] = t[q] - float(_Float16(t[q]));
                        yh[kc] = uint4{pack2<false>(t[0], t[1]), pack2<false>(t[2], t[3]), pack2<false>(t[4], t[5]),
                                       pack2<false>(t[6], t[7])};
                        yl[kc] = uint4{pack2<false>(e[0], e[1]), pack2<false>(e[2], e[3]), pack2<false>(e[4], e[5]),
                                       pack2<false>(e[6], e[7])};
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int mb = 0; mb < MB; ++mb) {
                        // later blocks' remainder fragments load while this
                        // block's MFMAs run
                        if (mb + LA < MB) {
#pragma unroll
                            for (int kc = 0; kc < KC; ++kc)
                                glo[(mb + LA) % (LA + 1)][kc] = *reinterpret_cast<const uint4*>(
                                    lo16 + (loff + uint32_t((mb + LA) * KC + kc) * 1024u));
                        }
                        floatx4 c1 = floatx4{0.f, 0.f, 0.f, 0.f}, c2 = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                        for (int kc = 0; kc < KC; ++kc) {
                            const uint4 mh = Ml[moff + (mb * KC + kc) * 64 + lane];
                            c1 = mfma16<false>(mh, yh[kc], c1);
                            c1 = mfma16<false>(mh, yl[kc], c1);
                            c2 = mfma16<false>(glo[mb % (LA + 1)][kc], yh[kc], c2);
                        }
                        const floatx4 bq = SPLITINC ? acc[mb] : breg[mb];
                        acc[mb] = bq * sc + (c1 + c2 * (1.f / 4096.f)) * isy;
                        // one output block at a time: bounds the registers in flight
                        __builtin_amdgcn_sched_barrier(0);
                    }
                } else {
#pragma unroll
                    for (int mb = 0; mb < LV::MBL; ++mb) {  // padding blocks keep acc = 0
                        floatx4 c = SPLITINC ? acc[mb] : breg[mb];
                        c *= sc;
#pragma unroll
                        for (int kb = 0; kb < LV::MBL; ++kb) {
                            const floatx4 mf = *reinterpret_cast<const floatx4*>(
                                reinterpret_cast<const char*>(a.M32) + (loff + uint32_t(mb * MB + kb) * 1024u));
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                if (LV::reg(kb, r)) c = mfma4(mf[r], y[kb][r], c);
                        }
                        acc[mb] = c;
                    }
                }
            } else if constexpr (SPLITINC) {
                // acc += (lowp(-cG) + remainder) lowp(dy), both halves on one
                // chain. The schedule keeps the next output block's 2 KC
                // fragments in flight from LDS while one block's MFMAs run.
                uint4 fr[2 * MB * KC];
#pragma unroll
                for (int i = 0; i < LV::MBL * KC; ++i) {
                    fr[2 * i] = Ml[moff + i * 64 + lane];
                    fr[2 * i + 1] = Ml[moff + NF + i * 64 + lane];
                }
#pragma unroll
                for (int mb = 0; mb < LV::MBL; ++mb) {
                    // the chain starts from zero, so its roundings are relative
                    // to the increment, not to acc; one rounding adds it
                    floatx4 c = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int kc = 0; kc < KC; ++kc) {
                        const uint4 dv = uint4{dyh[kc][0], dyh[kc][1], dyh[kc][2], dyh[kc][3]};
                        c = mfma16<BF16>(fr[2 * (mb * KC + kc)], dv, c);
                        c = mfma16<BF16>(fr[2 * (mb * KC + kc) + 1], dv, c);
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (LV::reg(mb, r)) acc[mb][r] += c[r];
                }
                __builtin_amdgcn_sched_group_barrier(0x100, 2 * KC, 0);  // DS reads
#pragma unroll
                for (int mb = 0; mb < LV::MBL; ++mb) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 2 * KC, 0);  // MFMA
                    if (mb + 1 < LV::MBL) __builtin_amdgcn_sched_group_barrier(0x100, 2 * KC, 0);
                }
            } else {
                // acc += (-cG) fp16(dy). The schedule keeps the fragments of
                // the next two output blocks in flight from LDS while one
                // block's MFMAs run (left to itself the scheduler keeps two
                // fragments in flight and the MFMAs wait on LDS latency).
                uint4 fr[MB * KC];
#pragma unroll
                for (int i = 0; i < MB * KC; ++i) fr[i] = Ml[moff + i * 64 + lane];
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) {
                    floatx4 c = acc[mb];
#pragma unroll
                    for (int kc = 0; kc < KC; ++kc)
                        c = mfma16<BF16>(fr[mb * KC + kc], uint4{dyh[kc][0], dyh[kc][1], dyh[kc][2], dyh[kc][3]}, c);
                    acc[mb] = c;
                }
                constexpr int AHEAD = MB < 2 ? MB : 2;
                __builtin_amdgcn_sched_group_barrier(0x100, AHEAD * KC, 0);  // DS reads
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) {
                    __builtin_amdgcn_sched_group_barrier(0x008, KC, 0);  // MFMA
                    if (mb + AHEAD < MB) __builtin_amdgcn_sched_group_barrier(0x100, KC, 0);
                }
            }
            mom = mom_next;
            if constexpr (decltype(first)::value) {
                if (__ballot(done) != 0ull && done) {
                    write_out(xo, conv);
                    wrote = true;
                    alive = false;  // until the refill: no iteration counts, no ballots
                }
                return false;
            } else {
                return __ballot(done) != 0ull;
            }
        };
        // The loop exits only after a second half, so no path out of it
        // leaves x in the other register set (which would need a copy).
        for (;;) {
            (void)step(x, x2, std::true_type{});
            if (step(x2, x, std::false_type{})) break;
        }
        if (done && !wrote) write_out(x, conv);
        wdone += unsigned(__popcll(__ballot(done && h == 0)));
        publish_progress(a.progress, a.progress_slots, WAVES, wdone);
        refill(done);
    }
}

// Selection pass of the alpha-path strategies (src/solver/alpha.rs:107-165,
// 244-263, 285-316): per voxel, the residual sum of squares and non-zero count
// of every path point, then the BIC argmin (ties to the larger alpha) or the
// L2-anchored rule (largest alpha with RSS <= (1 + slack) RSS_L2, else the
// RSS argmin), and the chosen point's outputs.
template <int MB>
__global__ void __launch_bounds__(256) path_select_kernel(const FistaArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int h = lane >> 4, j = lane & 15;
    const int64_t v = (int64_t(blockIdx.x) * 4 + wave) * 16 + j;
    const bool valid = v < a.n_vox;
    const int64_t out_idx = valid ? (a.vox ? a.vox[v] : v) : 0;
    const float* sig = a.data + out_idx * a.sig_stride;
    const int P = a.path_n;
    const float mean = valid ? a.mean[v] : 0.f;
    const bool zero = valid && (a.convpath[v * P] & 2) != 0;
    const bool live = valid && !zero;
    const double n = double(a.N), ln_n = log(n);
    const double thr_l2 = (a.path_rule == 1 && valid) ? (1.0 + a.slack) * double(a.rss_l2_ref[out_idx]) : 0.0;
    double best = INFINITY, best_rss = INFINITY;
    int idx = 0, idx_min = 0;
    float rss_idx = 0.f, rss_min = 0.f, sst = 0.f;
    bool found = false;
    for (int i = 0; i < P; ++i) {
        floatx4 x[MB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
            x[mb] = live ? *reinterpret_cast<const floatx4*>(a.Xpath + (v * P + i) * a.Kp + 16 * mb + 4 * h)
                         : floatx4{0.f, 0.f, 0.f, 0.f};
        float rss;
        residual<MB>(a, sig, valid, h, lane, x, mean, rss, sst);
        int nnz = 0;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) nnz += (x[mb][r] != 0.f);
        nnz += __shfl_xor(nnz, 16);
        nnz += __shfl_xor(nnz, 32);
        const double rs = double(rss);
        if (a.path_rule == 0) {
            const double b = n * log(fmax(rs, 2.2250738585072014e-308) / n) + ln_n * double(nnz);
            if (b < best) {
                best = b;
                idx = i;
                rss_idx = rss;
            }
        } else {
            if (!found && rs <= thr_l2) {
                found = true;
                idx = i;
                rss_idx = rss;
            }
            if (rs < best_rss) {
                best_rss = rs;
                idx_min = i;
                rss_min = rss;
            }
        }
    }
    if (a.path_rule == 1 && !found) {
        idx = idx_min;
        rss_idx = rss_min;
    }
    if (!valid) return;
    floatx4 x[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
        x[mb] = live ? *reinterpret_cast<const floatx4*>(a.Xpath + (v * P + idx) * a.Kp + 16 * mb + 4 * h)
                     : floatx4{0.f, 0.f, 0.f, 0.f};
    // alpha_i = alpha_max f_i with alpha_max = max|b| / step
    float bm = 0.f;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const floatx4 bq = *reinterpret_cast<const floatx4*>(a.B + v * a.Kp + 16 * mb + 4 * h);
#pragma unroll
        for (int r = 0; r < 4; ++r) bm = fmaxf(bm, fabsf(bq[r]));
    }
    bm = fmaxf(bm, __shfl_xor(bm, 16));
    bm = fmaxf(bm, __shfl_xor(bm, 32));
    const double alpha_i = double(bm) / double(a.step) * a.path_f64[idx];
    const int64_t pt = v * P + idx;
    write_voxel<MB>(a, out_idx, h, x, rss_idx, sst, zero ? 0u : a.itspath[pt], (a.convpath[pt] & 1) != 0,
                    zero ? 0.0 : alpha_i, zero);
}

template <int MB>
void launch_path_select(const FistaArgs& a, hipStream_t s) {
    const int64_t tiles = (a.n_vox + 15) / 16;
    hipLaunchKernelGGL(path_select_kernel<MB>, dim3(unsigned((tiles + 3) / 4)), dim3(256), 0, s, a);
    SF_HIP(hipGetLastError());
}

template <int MB>
void launch_f32_persist(const FistaArgs& args, const FistaLaunch& L, hipStream_t s) {
    unsigned int* counter = L.counter;
    constexpr int WAVES = (MB >= 16) ? 4 : 8;
    constexpr bool IN_LDS = (MB * MB * 64 * 16) <= 112 * 1024;
    const size_t lds = IN_LDS ? size_t(MB) * MB * 64 * 16 : 0;
    auto kern = L.path_n > 0 ? fista_f32_persist_kernel<MB, WAVES, IN_LDS, true>
                             : fista_f32_persist_kernel<MB, WAVES, IN_LDS, false>;
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(lds));
    int dev = 0, cus = 0, per_cu = 0;
    SF_HIP(hipGetDevice(&dev));
    SF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    SF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), 64 * WAVES,
                                                        lds));
    if (per_cu < 1) per_cu = 1;
    const int64_t need = (args.n_vox + 16 * WAVES - 1) / (16 * WAVES);
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(need, int64_t(cus) * per_cu));
    if (!L.counter_zeroed) SF_HIP(hipMemsetAsync(counter, 0, sizeof(unsigned int), s));
    hipLaunchKernelGGL(kern, dim3(unsigned(grid)), dim3(64 * WAVES), lds, s, args, counter, L.Xbuf, L.itsbuf,
                       L.statusbuf, L.alphabuf);
    SF_HIP(hipGetLastError());
    if (L.event_mid) SF_HIP(hipEventRecord(static_cast<hipEvent_t>(L.event_mid), s));
    if (L.path_n > 0) {
        launch_path_select<MB>(args, s);  // the selection pass writes the outputs
        return;
    }
    if (!L.Xbuf) return;
    launch_epilogue<MB>(args, L.Xbuf, L.itsbuf, L.statusbuf, L.alphabuf, L.AH16, L.Ainv16, s);
}

template <int MB, bool BF16>
void launch_mixed_mb(const FistaLaunch& L, const FistaArgs& ea, hipStream_t s) {
    constexpr bool SPLITINC = 16 * MB <= kMixedSplitMaxKp;
    constexpr int WAVES = SPLITINC ? 8 : 4;
    constexpr int KC = MB / 2;
    MixedArgs m{};
    m.vox = L.vox;
    m.n_vox = L.n_vox;
    m.K = L.K;
    m.Kp = L.Kp;
    m.B = L.Bbuf;
    m.M32 = reinterpret_cast<const floatx4*>(L.Mfrag);
    m.M16 = reinterpret_cast<const uint4*>(L.Mlow);
    m.M16lo = reinterpret_cast<const uint4*>(L.Mlow2);
    m.Mlo_inc = reinterpret_cast<const uint4*>(L.Mlow3);
    if (SPLITINC && !m.Mlo_inc) throw Error{SHOREFIT_E_DEVICE, "mixed-precision FISTA: split operator not packed"};
    m.mom = L.mom;
    m.thr = L.thr;
    m.step = L.step;
    m.ratio = L.ratio;
    m.alpha_fixed = float(L.alpha);
    m.max_iter = L.max_iter;
    m.tol = L.tol;
    m.refresh = L.refresh < 1 ? 1 : L.refresh;
    m.counter = L.counter;
    m.X = L.Xbuf;
    m.its = L.itsbuf;
    m.status = L.statusbuf;
    m.alpha_v = L.alphabuf;
    m.path_n = L.path_n;
    m.path_f = L.path_f;
    m.Xpath = L.Xpath;
    m.itspath = L.itspath;
    m.convpath = L.convpath;
    m.progress = L.progress;
    m.progress_slots = L.progress_slots;
    const bool path = L.path_n > 0;
    const size_t lds = size_t(MB) * KC * 64 * 16 * (SPLITINC ? 2 : 1);
    // refresh on the FP32 MFMA, except FP16 without split increments (order
    // 10), which refreshes from FP16 hi/lo splits
    constexpr bool SPLIT = !BF16 && !SPLITINC;
    // the coefficient counts of radial orders 6 and 8 are compiled in
    constexpr int KL = MB == 10 ? 145 : (MB == 6 ? 72 : 0);
    const bool kl = KL > 0 && L.K == KL;
    auto pick = [&](auto nonneg, auto pth) {
        constexpr bool NN = decltype(nonneg)::value, PT = decltype(pth)::value;
        return kl ? fista_mixed_kernel<MB, WAVES, BF16, SPLITINC, NN, SPLIT, PT, KL>
                  : fista_mixed_kernel<MB, WAVES, BF16, SPLITINC, NN, SPLIT, PT, 0>;
    };
    auto kern = path ? (L.nonneg ? pick(std::true_type{}, std::true_type{}) : pick(std::false_type{}, std::true_type{}))
                     : (L.nonneg ? pick(std::true_type{}, std::false_type{})
                                 : pick(std::false_type{}, std::false_type{}));
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                              int(lds));
    int dev = 0, cus = 0, per_cu = 0;
    SF_HIP(hipGetDevice(&dev));
    SF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    SF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), 64 * WAVES,
                                                        lds));
    if (per_cu < 1) per_cu = 1;
    const int64_t need = (L.n_vox + 16 * WAVES - 1) / (16 * WAVES);
    const int64_t grid = std::min<int64_t>(need, int64_t(cus) * per_cu);
    SF_HIP(hipMemsetAsync(L.counter, 0, sizeof(unsigned int), s));
    hipLaunchKernelGGL(kern, dim3(unsigned(grid)), dim3(64 * WAVES), lds, s, m);
    SF_HIP(hipGetLastError());
    if (L.event_mid) SF_HIP(hipEventRecord(static_cast<hipEvent_t>(L.event_mid), s));
    if (path) {
        launch_path_select<MB>(ea, s);
        return;
    }
    launch_epilogue<MB>(ea, L.Xbuf, L.itsbuf, L.statusbuf, L.alphabuf, L.AH16, L.Ainv16, s);
}

}  // namespace

void launch_fista_rhs(const FistaLaunch& L, hipStream_t s) {
    if (L.n_vox <= 0) return;
    switch (L.Kp / 16) {
#define SF_RHS(MB) \
    case MB: launch_rhs<MB>(L, s); break;
        SF_RHS(1) SF_RHS(2) SF_RHS(3) SF_RHS(4) SF_RHS(5) SF_RHS(6) SF_RHS(8) SF_RHS(10) SF_RHS(12) SF_RHS(16)
#undef SF_RHS
        default:
            throw Error{SHOREFIT_E_UNSUPPORTED,
                        "right-hand-side kernel: padded coefficient count " + std::to_string(L.Kp) + " is not instantiated"};
    }
}

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
    a.step64 = L.step64;
    a.alpha_vox = L.alpha_vox;
    a.warm = L.warm;
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
    a.B = L.Bbuf;
    a.mean = L.meanbuf;
    a.Kp = L.Kp;
    a.path_n = L.path_n;
    a.path_f = L.path_f;
    a.path_f64 = L.path_f64;
    a.mom = L.mom;
    a.Xpath = L.Xpath;
    a.itspath = L.itspath;
    a.convpath = L.convpath;
    a.path_rule = L.path_rule;
    a.slack = L.slack;
    a.rss_l2_ref = L.rss_l2_ref;
    a.progress = L.progress;
    a.progress_slots = L.progress_slots;
    if (L.precision != SHOREFIT_PREC_FP32) {
        const bool bf = L.precision == SHOREFIT_PREC_BF16;
        switch (L.Kp / 16) {
#define SF_MIX(MB)                                                    \
    case MB:                                                          \
        if (bf)                                                       \
            launch_mixed_mb<MB, true>(L, a, s);                       \
        else                                                          \
            launch_mixed_mb<MB, false>(L, a, s);                      \
        break;
            SF_MIX(2) SF_MIX(4) SF_MIX(6) SF_MIX(8) SF_MIX(10) SF_MIX(16)
#undef SF_MIX
            default:
                throw Error{SHOREFIT_E_UNSUPPORTED, "mixed-precision FISTA: padded coefficient count " +
                                                        std::to_string(L.Kp) + " is not instantiated"};
        }
        return;
    }
    // persistent FP32 form: a work counter and a momentum table supplied,
    // work-order buffers when residual maps are wanted without an alpha path,
    // and a voxel count the 32-bit work counter can hand out
    const bool maps = L.resid || L.r2 || L.bic;
    const bool persist = L.persist && L.counter && L.mom && (L.path_n > 0 || L.Xbuf || !maps) &&
                         L.n_vox < (int64_t(1) << 31);
    switch (L.Kp / 16) {
#define SF_F32(MB)                                                  \
    case MB:                                                        \
        if (persist) {                                              \
            launch_f32_persist<MB>(a, L, s);                        \
            break;                                                  \
        }                                                           \
        launch_f32<MB>(a, L.n_vox, s);                              \
        if (L.event_mid) SF_HIP(hipEventRecord(static_cast<hipEvent_t>(L.event_mid), s)); \
        if (L.path_n > 0) launch_path_select<MB>(a, s);             \
        break;
        SF_F32(1) SF_F32(2) SF_F32(3) SF_F32(4) SF_F32(5) SF_F32(6) SF_F32(8) SF_F32(10) SF_F32(12) SF_F32(16)
#undef SF_F32
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
        case 1: launch_tik<1>(a, L, s); break;
        case 2: launch_tik<2>(a, L, s); break;
        case 3: launch_tik<3>(a, L, s); break;
        case 4: launch_tik<4>(a, L, s); break;
        case 5: launch_tik<5>(a, L, s); break;
        case 6: launch_tik<6>(a, L, s); break;
        case 8: launch_tik<8>(a, L, s); break;
        case 10: launch_tik<10>(a, L, s); break;
        case 12: launch_tik<12>(a, L, s); break;
        case 16: launch_tik<16>(a, L, s); break;
        default:
            throw Error{SHOREFIT_E_UNSUPPORTED,
                        "Tikhonov kernel: padded coefficient count " + std::to_string(L.Kp) + " is not instantiated"};
    }
}

}  // namespace shorefit
