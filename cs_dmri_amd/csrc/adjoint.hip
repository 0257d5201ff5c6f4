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
// Gram matrix into a packed lower triangle in LDS (sized by the batch's
// largest active set), an in-place right-looking Cholesky (FP64), the two
// triangular solves, then A_S w with A read from L2.
#include "common.h"

namespace shorefit {
namespace {

__device__ __forceinline__ int tri(int i) { return i * (i + 1) / 2; }

// Row i and column j (j <= i) of entry e of a packed lower triangle.
__device__ __forceinline__ void tri_decode(int e, int& i, int& j) {
    int r = int((sqrtf(8.f * float(e) + 1.f) - 1.f) * 0.5f);
    if (tri(r + 1) <= e) ++r;
    if (tri(r) > e) --r;
    i = r;
    j = e - tri(r);
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) x += __shfl_xor(x, m);
    return x;
}

// One wave per signal: the workgroup barriers below synchronise a single
// wave, so they cost next to nothing, and the LDS per signal is sized by the
// batch's largest active set, so several signals share a CU.
constexpr int kThreads = 64;

__global__ void __launch_bounds__(kThreads)
    lasso_adjoint_kernel(const double* __restrict__ G, const double* __restrict__ A, int N, int K,
                         const float* __restrict__ coef, const float* __restrict__ gcoef, int64_t V,
                         float* __restrict__ gsig, float* __restrict__ galpha, uint8_t* __restrict__ status) {
    extern __shared__ double sm[];
    __shared__ int s_idx[256];
    __shared__ int s_fail;
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
    if (lane == 0) s_fail = 0;
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
    double* L = sm;                // packed lower triangle of G_SS, m (m + 1) / 2
    double* x = sm + tri(m);       // right-hand side, then the solution w
    const int nt = tri(m);
    for (int e = lane; e < nt; e += kThreads) {
        int i, j;
        tri_decode(e, i, j);
        L[e] = G[size_t(s_idx[i]) * K + s_idx[j]];
    }
    for (int i = lane; i < m; i += kThreads) x[i] = double(gcoef[v * K + s_idx[i]]);
    __syncthreads();

    // Cholesky G_SS = L L^T, column by column
    for (int j = 0; j < m; ++j) {
        const int jj = tri(j) + j;
        if (lane == 0) {
            const double d = L[jj];
            if (!(d > 0.0)) s_fail = 1;
            L[jj] = sqrt(d);
        }
        __syncthreads();
        if (s_fail) break;
        const double ljj = L[jj];
        for (int i = j + 1 + lane; i < m; i += kThreads) L[tri(i) + j] /= ljj;
        __syncthreads();
        const int n = m - j - 1, cnt = tri(n);
        for (int e = lane; e < cnt; e += kThreads) {
            int ii, kk;
            tri_decode(e, ii, kk);
            const int i = j + 1 + ii, k = j + 1 + kk;
            L[tri(i) + k] -= L[tri(i) + j] * L[tri(k) + j];
        }
        __syncthreads();
    }
    if (s_fail) {
        const float nan = __int_as_float(0x7fc00000);
        if (gsig)
            for (int n = lane; n < N; n += kThreads) gsig[v * N + n] = nan;
        if (lane == 0) {
            if (galpha) galpha[v] = nan;
            if (status) status[v] = 1;
        }
        return;
    }
    // L y = g, then L^T w = y
    for (int j = 0; j < m; ++j) {
        if (lane == 0) x[j] /= L[tri(j) + j];
        __syncthreads();
        const double xj = x[j];
        for (int i = j + 1 + lane; i < m; i += kThreads) x[i] -= L[tri(i) + j] * xj;
        __syncthreads();
    }
    for (int j = m - 1; j >= 0; --j) {
        if (lane == 0) x[j] /= L[tri(j) + j];
        __syncthreads();
        const double xj = x[j];
        for (int i = lane; i < j; i += kThreads) x[i] -= L[tri(j) + i] * xj;
        __syncthreads();
    }

    // dL/dalpha = -N z^T w
    double zw = 0.0;
    for (int i = lane; i < m; i += kThreads) zw += (c[s_idx[i]] > 0.f ? x[i] : -x[i]);
    zw = wave_sum(zw);
    // dL/ds = A_S w
    if (gsig)
        for (int n = lane; n < N; n += kThreads) {
            const double* arow = A + size_t(n) * K;
            double acc = 0.0;
            for (int i = 0; i < m; ++i) acc = fma(arow[s_idx[i]], x[i], acc);
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

size_t lasso_adjoint_lds(int m) { return sizeof(double) * (size_t(m) * (m + 1) / 2 + m); }

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
    if (lds > 150 * 1024)
        throw Error{SHOREFIT_E_UNSUPPORTED, "lasso adjoint: an active set of " + std::to_string(m_max) +
                                                " coefficients exceeds LDS (at most 190)"};
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(lasso_adjoint_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    hipLaunchKernelGGL(lasso_adjoint_kernel, dim3(unsigned(V)), dim3(kThreads), lds, s, G, A, N, K, coef, gcoef, V,
                       gsig, galpha, status);
    SF_HIP(hipGetLastError());
}

}  // namespace shorefit
