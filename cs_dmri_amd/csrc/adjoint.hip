// Adjoint of the lasso solution by implicit differentiation on the active set.
//
// The FISTA fixed point (src/solver/fista.rs:70-143) minimises
//     (1 / 2N) |s - A c|^2 + alpha |c|_1 .
// With S = {k : c_k != 0} and z = sign(c_S), its optimality conditions on S
// read G_SS c_S = A_S^T s - N alpha z (G = A^T A), and c is 0 off S. While S
// and z stay fixed, c is an affine function of (s, alpha):
//     dc_S / ds = G_SS^-1 A_S^T ,   dc_S / dalpha = -N G_SS^-1 z .
// For an upstream gradient g = dL/dc the vector-Jacobian products are
//     w = G_SS^-1 g_S ,   dL/ds = A_S w ,   dL/dalpha = -N z^T w .
//
// One wave per signal: the active set by ballots, G_SS gathered from the FP64
// Gram matrix into a packed lower triangle in LDS (rows, sized by the batch's
// largest active set), an in-place right-looking Cholesky (FP64) in which the
// lanes own rows, the two triangular solves with the right-hand side in
// registers, then A_S w with A read from L2.
#include <type_traits>

#include "common.h"

namespace shorefit {
namespace {

__device__ __forceinline__ int tri(int i) { return i * (i + 1) / 2; }

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) x += __shfl_xor(x, m);
    return x;
}

// Value of lane `src` (wave-uniform) of a double.
__device__ __forceinline__ double lane_value(double x, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), src);
    return __hiloint2double(hi, lo);
}

constexpr int kThreads = 64;
constexpr int kRowsPerLane = 3;  // LDS kernel: active sets up to 192
constexpr int kRegRows = 64;     // register kernel: active sets up to 64

// Compile-time loop: f(std::integral_constant<int, i>) for i = B .. E-1.
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Active set S of signal c in coefficient order into s_idx[0 .. m-1] (the
// rest of s_idx is 0); returns m. One wave; ends with a barrier.
__device__ int active_set(const float* c, int K, int lane, int* s_idx) {
    for (int i = lane; i < 256; i += kThreads) s_idx[i] = 0;
    __syncthreads();
    int m = 0;
    for (int k0 = 0; k0 < K; k0 += 64) {
        const bool act = k0 + lane < K && c[k0 + lane] != 0.f;
        const uint64_t bal = __ballot(act);
        if (act) s_idx[m + __popcll(bal & ((uint64_t(1) << lane) - 1))] = k0 + lane;
        m += __popcll(bal);
    }
    __syncthreads();
    return m;
}

// The same value for every output of signal v: 0 with status 0 (empty
// active set: c = 0 is locally constant) or NaN with status 1 (G_SS not
// positive definite).
__device__ void uniform_outputs(int64_t v, int N, int lane, bool nan, float* gsig, float* galpha,
                                uint8_t* status) {
    const float val = nan ? __int_as_float(0x7fc00000) : 0.f;
    if (gsig)
        for (int n = lane; n < N; n += kThreads) gsig[v * N + n] = val;
    if (lane == 0) {
        if (galpha) galpha[v] = val;
        if (status) status[v] = nan ? 1 : 0;
    }
}

// Signals with at most 64 active coefficients: lane i keeps row i of G_SS,
// then of L, in registers (128 VGPRs); the column loops are unrolled at
// compile time, so every register index is a constant. Column j: the pivot
// is read from lane j, every lane forms sqrt and 1 / L_jj (uniform), the
// lanes below scale their entry, the column goes through LDS once (64
// doubles) and each lane updates its row. Entries right of a lane's diagonal
// receive the same updates and are never read. L y = g broadcasts one
// solved entry per step; L^T w = y sums one column per step over the lanes.
// Signals with larger active sets are left to lasso_adjoint_kernel.
__global__ void __launch_bounds__(kThreads)
    lasso_adjoint_reg_kernel(const double* __restrict__ G, const double* __restrict__ A, int N, int K,
                             const float* __restrict__ coef, const float* __restrict__ gcoef, int64_t V,
                             float* __restrict__ gsig, float* __restrict__ galpha, uint8_t* __restrict__ status) {
    __shared__ int s_idx[256];
    __shared__ double s_col[kRegRows];
    const int64_t v = blockIdx.x;
    if (v >= V) return;
    const int lane = threadIdx.x;
    const float* c = coef + v * K;
    const int m = active_set(c, K, lane, s_idx);
    if (m > kRegRows) return;
    if (m == 0) {
        uniform_outputs(v, N, lane, false, gsig, galpha, status);
        return;
    }

    const bool own = lane < m;
    const double* grow = G + size_t(s_idx[own ? lane : 0]) * K;
    double Lr[kRegRows];
    static_for<0, kRegRows>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        Lr[k] = (own && k <= lane) ? grow[s_idx[k]] : 0.0;
    });

    double invd = 0.0;  // 1 / L_ii of this lane's row
    bool fail = false;
    static_for<0, kRegRows>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (j >= m || fail) return;
        const double piv = lane_value(Lr[j], j);
        if (!(piv > 0.0)) {
            fail = true;
            return;
        }
        const double ljj = sqrt(piv), inv = 1.0 / ljj;
        const double lij = Lr[j] * inv;
        if (lane == j) {
            Lr[j] = ljj;
            invd = inv;
        } else if (lane > j) {
            Lr[j] = lij;
        }
        s_col[lane] = lane > j ? lij : 0.0;
        __syncthreads();
        static_for<j + 1, kRegRows>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            Lr[k] = fma(-lij, s_col[k], Lr[k]);
        });
        __syncthreads();
    });
    if (fail) {
        uniform_outputs(v, N, lane, true, gsig, galpha, status);
        return;
    }

    // L y = g, then L^T w = y; lane i keeps entry i
    double x = own ? double(gcoef[v * K + s_idx[lane]]) : 0.0;
    static_for<0, kRegRows>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (j >= m) return;
        const double yj = lane_value(x, j) * lane_value(invd, j);
        x = lane > j ? fma(-Lr[j], yj, x) : (lane == j ? yj : x);
    });
    static_for<0, kRegRows>([&](auto jr) {
        constexpr int j = kRegRows - 1 - decltype(jr)::value;
        if (j >= m) return;
        const double sum = wave_sum((lane > j && own) ? Lr[j] * x : 0.0);
        const double wj = (lane_value(x, j) - sum) * lane_value(invd, j);
        if (lane == j) x = wj;
    });

    // dL/dalpha = -N z^T w ; dL/ds = A_S w
    const double zw = wave_sum(own ? (c[s_idx[lane]] > 0.f ? x : -x) : 0.0);
    s_col[lane] = own ? x : 0.0;
    __syncthreads();
    if (gsig)
        for (int n = lane; n < N; n += kThreads) {
            const double* arow = A + size_t(n) * K;
            double acc = 0.0;
            for (int i = 0; i < m; ++i) acc = fma(arow[s_idx[i]], s_col[i], acc);
            gsig[v * N + n] = float(acc);
        }
    if (lane == 0) {
        if (galpha) galpha[v] = float(-double(N) * zw);
        if (status) status[v] = 0;
    }
}

// The signals of `list` (those with more than 64 active coefficients). One
// wave per signal: the workgroup barriers below synchronise
// a single wave, so they cost next to nothing, and the LDS per signal is
// sized by the batch's largest active set, so several signals share a CU.

__global__ void __launch_bounds__(kThreads)
    lasso_adjoint_kernel(const double* __restrict__ G, const double* __restrict__ A, int N, int K,
                         const float* __restrict__ coef, const float* __restrict__ gcoef,
                         const int64_t* __restrict__ list, int64_t n_list, float* __restrict__ gsig,
                         float* __restrict__ galpha, uint8_t* __restrict__ status) {
    extern __shared__ double sm[];
    __shared__ int s_idx[256];
    if (int64_t(blockIdx.x) >= n_list) return;
    const int64_t v = list[blockIdx.x];
    const int lane = threadIdx.x;
    const float* c = coef + v * K;
    const int m = active_set(c, K, lane, s_idx);
    if (m == 0) {
        uniform_outputs(v, N, lane, false, gsig, galpha, status);
        return;
    }
    double* L = sm;               // packed lower triangle of G_SS by rows, tri(m)
    double* s_col = L + tri(m);   // the column being eliminated, then w
    double* s_inv = s_col + m;    // 1 / L_jj
    for (int i = 0; i < m; ++i) {
        const double* grow = G + size_t(s_idx[i]) * K;
        for (int k = lane; k <= i; k += kThreads) L[tri(i) + k] = grow[s_idx[k]];
    }

    // Cholesky G_SS = L L^T, column by column; every lane forms the pivot
    // (a uniform value), lane t owns rows j + 1 + t + 64 q below it
    bool fail = false;
    for (int j = 0; j < m; ++j) {
        __syncthreads();  // the previous column's updates are visible
        const double d = L[tri(j) + j];
        if (!(d > 0.0)) {
            fail = true;
            break;
        }
        const double ljj = sqrt(d), inv = 1.0 / ljj;
        for (int i = j + 1 + lane; i < m; i += kThreads) {
            const double lij = L[tri(i) + j] * inv;
            L[tri(i) + j] = lij;
            s_col[i] = lij;
        }
        if (lane == 0) {
            L[tri(j) + j] = ljj;
            s_inv[j] = inv;
        }
        __syncthreads();
        for (int i = j + 1 + lane; i < m; i += kThreads) {
            const double lij = s_col[i];
            double* row = L + tri(i);
            int k = j + 1;
            // eight at a time, every load issued before the stores (the
            // column and the row share LDS, so loads after a store would
            // wait for them)
            for (; k + 7 <= i; k += 8) {
                double cv[8], rv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    cv[u] = s_col[k + u];
                    rv[u] = row[k + u];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) row[k + u] = fma(-lij, cv[u], rv[u]);
            }
            for (; k <= i; ++k) row[k] = fma(-lij, s_col[k], row[k]);
        }
    }
    if (fail) {
        uniform_outputs(v, N, lane, true, gsig, galpha, status);
        return;
    }
    __syncthreads();

    // L y = g, then L^T w = y; lane t keeps entries t + 64 q in registers and
    // each step broadcasts one solved entry
    double xr[kRowsPerLane];
#pragma unroll
    for (int q = 0; q < kRowsPerLane; ++q) {
        const int i = lane + kThreads * q;
        xr[q] = i < m ? double(gcoef[v * K + s_idx[i]]) : 0.0;
    }
    // entry j of the vector (read from each register, then chosen, so the
    // registers are never indexed at run time)
    auto entry = [&](int j) {
        const int q = j / kThreads, src = j % kThreads;
        const double e0 = lane_value(xr[0], src), e1 = lane_value(xr[1], src), e2 = lane_value(xr[2], src);
        return q == 0 ? e0 : (q == 1 ? e1 : e2);
    };
    for (int j = 0; j < m; ++j) {
        const double yj = entry(j) * s_inv[j];
#pragma unroll
        for (int q = 0; q < kRowsPerLane; ++q) {
            const int i = lane + kThreads * q;
            if (i > j && i < m) xr[q] = fma(-L[tri(i) + j], yj, xr[q]);
            else if (i == j) xr[q] = yj;
        }
    }
    for (int j = m - 1; j >= 0; --j) {
        const double wj = entry(j) * s_inv[j];
#pragma unroll
        for (int q = 0; q < kRowsPerLane; ++q) {
            const int i = lane + kThreads * q;
            if (i < j) xr[q] = fma(-L[tri(j) + i], wj, xr[q]);
            else if (i == j) xr[q] = wj;
        }
    }

    // dL/dalpha = -N z^T w
    double zw = 0.0;
#pragma unroll
    for (int q = 0; q < kRowsPerLane; ++q) {
        const int i = lane + kThreads * q;
        if (i < m) {
            zw += c[s_idx[i]] > 0.f ? xr[q] : -xr[q];
            s_col[i] = xr[q];
        }
    }
    zw = wave_sum(zw);
    __syncthreads();
    // dL/ds = A_S w
    if (gsig)
        for (int n = lane; n < N; n += kThreads) {
            const double* arow = A + size_t(n) * K;
            double acc = 0.0;
            for (int i = 0; i < m; ++i) acc = fma(arow[s_idx[i]], s_col[i], acc);
            gsig[v * N + n] = float(acc);
        }
    if (lane == 0) {
        if (galpha) galpha[v] = float(-double(N) * zw);
        if (status) status[v] = 0;
    }
}

// Largest active-set size of the batch (atomic max over the signals), and
// the list of signals with more than kRegRows active coefficients (order not
// kept; each signal's outputs are its own).
__global__ void active_scan_kernel(const float* __restrict__ coef, int K, int64_t V, int* __restrict__ m_max,
                                   unsigned long long* __restrict__ n_list, int64_t* __restrict__ list) {
    const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int m = 0;
    if (v < V)
        for (int k = 0; k < K; ++k) m += coef[v * K + k] != 0.f;
    if (v < V && m > kRegRows) list[atomicAdd(n_list, 1ull)] = v;
    m = max(m, __shfl_xor(m, 1));
    m = max(m, __shfl_xor(m, 2));
    m = max(m, __shfl_xor(m, 4));
    m = max(m, __shfl_xor(m, 8));
    m = max(m, __shfl_xor(m, 16));
    m = max(m, __shfl_xor(m, 32));
    if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(m_max, m);
}

}  // namespace

// packed triangle, column buffer, 1 / L_jj
size_t lasso_adjoint_lds(int m) { return sizeof(double) * (size_t(m) * (m + 1) / 2 + 2 * size_t(m)); }

void launch_lasso_adjoint(const double* G, const double* A, int N, int K, const float* coef, const float* gcoef,
                          int64_t V, float* gsig, float* galpha, uint8_t* status, hipStream_t s) {
    if (V <= 0) return;
    if (K > 256) throw Error{SHOREFIT_E_UNSUPPORTED, "lasso adjoint: at most 256 coefficients"};
    struct Scan {
        unsigned long long n_list;
        int m_max;
    };
    DevBuf<Scan> d_scan(1);
    DevBuf<int64_t> d_list{size_t(V)};
    SF_HIP(hipMemsetAsync(d_scan.p, 0, sizeof(Scan), s));
    hipLaunchKernelGGL(active_scan_kernel, dim3(unsigned((V + 255) / 256)), dim3(256), 0, s, coef, K, V,
                       &d_scan.p->m_max, &d_scan.p->n_list, d_list.p);
    SF_HIP(hipGetLastError());
    Scan h{};
    SF_HIP(hipMemcpyAsync(&h, d_scan.p, sizeof(Scan), hipMemcpyDeviceToHost, s));
    SF_HIP(hipStreamSynchronize(s));
    const int m_max = h.m_max;
    const size_t lds = lasso_adjoint_lds(std::max(m_max, 1));
    if (m_max > kRowsPerLane * kThreads || lds > 150 * 1024)
        throw Error{SHOREFIT_E_UNSUPPORTED, "lasso adjoint: an active set of " + std::to_string(m_max) +
                                                " coefficients exceeds the kernel's limit (at most 192)"};
    // active sets up to 64 (and empty ones) in registers, the listed larger
    // ones in LDS
    hipLaunchKernelGGL(lasso_adjoint_reg_kernel, dim3(unsigned(V)), dim3(kThreads), 0, s, G, A, N, K, coef, gcoef,
                       V, gsig, galpha, status);
    SF_HIP(hipGetLastError());
    const int64_t n_list = int64_t(h.n_list);
    if (n_list == 0) return;
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(lasso_adjoint_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    hipLaunchKernelGGL(lasso_adjoint_kernel, dim3(unsigned(n_list)), dim3(kThreads), lds, s, G, A, N, K, coef, gcoef,
                       d_list.p, n_list, gsig, galpha, status);
    SF_HIP(hipGetLastError());
}

}  // namespace shorefit
