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

// One wave per signal: the workgroup barriers below synchronise a single
// wave, so they cost next to nothing, and the LDS per signal is sized by the
// batch's largest active set, so several signals share a CU.
constexpr int kThreads = 64;
constexpr int kRowsPerLane = 3;  // active sets up to 192 (LDS allows 190)

__global__ void __launch_bounds__(kThreads)
    lasso_adjoint_kernel(const double* __restrict__ G, const double* __restrict__ A, int N, int K,
                         const float* __restrict__ coef, const float* __restrict__ gcoef, int64_t V,
                         float* __restrict__ gsig, float* __restrict__ galpha, uint8_t* __restrict__ status) {
    extern __shared__ double sm[];
    __shared__ int s_idx[256];
    const int64_t v = blockIdx.x;
    if (v >= V) return;
    const int lane = threadIdx.x;
    const float* c = coef + v * K;

    // active set S in coefficient order
    int m = 0;
    for (int k0 = 0; k0 < K; k0 += 64) {
        const bool act = k0 + lane < K && c[k0 + lane] != 0.f;
        const uint64_t bal = __ballot(act);
        if (act) s_idx[m + __popcll(bal & ((uint64_t(1) << lane) - 1))] = k0 + lane;
        m += __popcll(bal);
    }
    __syncthreads();

    if (m == 0) {  // c = 0 is locally constant: zero gradients
        if (gsig)
            for (int n = lane; n < N; n += kThreads) gsig[v * N + n] = 0.f;
        if (lane == 0) {
            if (galpha) galpha[v] = 0.f;
            if (status) status[v] = 0;
        }
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
        const float nan = __int_as_float(0x7fc00000);
        if (gsig)
            for (int n = lane; n < N; n += kThreads) gsig[v * N + n] = nan;
        if (lane == 0) {
            if (galpha) galpha[v] = nan;
            if (status) status[v] = 1;
        }
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

// Largest active-set size of the batch (atomic max over the signals).
__global__ void active_max_kernel(const float* __restrict__ coef, int K, int64_t V, int* __restrict__ out) {
    const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int m = 0;
    if (v < V)
        for (int k = 0; k < K; ++k) m += coef[v * K + k] != 0.f;
    m = max(m, __shfl_xor(m, 1));
    m = max(m, __shfl_xor(m, 2));
    m = max(m, __shfl_xor(m, 4));
    m = max(m, __shfl_xor(m, 8));
    m = max(m, __shfl_xor(m, 16));
    m = max(m, __shfl_xor(m, 32));
    if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(out, m);
}

}  // namespace

// packed triangle, column buffer, 1 / L_jj
size_t lasso_adjoint_lds(int m) { return sizeof(double) * (size_t(m) * (m + 1) / 2 + 2 * size_t(m)); }

void launch_lasso_adjoint(const double* G, const double* A, int N, int K, const float* coef, const float* gcoef,
                          int64_t V, float* gsig, float* galpha, uint8_t* status, hipStream_t s) {
    if (V <= 0) return;
    if (K > 256) throw Error{SHOREFIT_E_UNSUPPORTED, "lasso adjoint: at most 256 coefficients"};
    DevBuf<int> d_m(1);
    SF_HIP(hipMemsetAsync(d_m.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(active_max_kernel, dim3(unsigned((V + 255) / 256)), dim3(256), 0, s, coef, K, V, d_m.p);
    SF_HIP(hipGetLastError());
    int m_max = 0;
    SF_HIP(hipMemcpyAsync(&m_max, d_m.p, sizeof(int), hipMemcpyDeviceToHost, s));
    SF_HIP(hipStreamSynchronize(s));
    const size_t lds = lasso_adjoint_lds(std::max(m_max, 1));
    if (m_max > kRowsPerLane * kThreads || lds > 150 * 1024)
        throw Error{SHOREFIT_E_UNSUPPORTED, "lasso adjoint: an active set of " + std::to_string(m_max) +
                                                " coefficients exceeds the kernel's limit (at most 192)"};
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(lasso_adjoint_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    hipLaunchKernelGGL(lasso_adjoint_kernel, dim3(unsigned(V)), dim3(kThreads), lds, s, G, A, N, K, coef, gcoef, V,
                       gsig, galpha, status);
    SF_HIP(hipGetLastError());
}

}  // namespace shorefit
