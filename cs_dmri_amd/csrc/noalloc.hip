// Device scratch for the allocation-free lasso forward
// (shorefit_lasso_forward_noalloc): a file-scope device array that every call
// on a GPU shares, so the call itself allocates no device memory.
//
// Layout of one call's scratch, each region 256-byte aligned:
//   [0, 256)            the work counter of the persistent kernel
//   momentum table      max_iter floats
//   right-hand sides    V x Kp floats
//   signal means        V floats
// Calls that overlap on one GPU use the same array and race; callers
// serialise them (one stream, or events).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "common.h"

namespace shorefit {
namespace {

constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

// Momentum (t_k - 1) / t_{k+1}, t_0 = 1 (src/solver/fista.rs:111-112) in FP64,
// rounded to FP32: the table the host builds for the allocating entry points,
// bit for bit (correctly rounded operations, no contraction into FMAs).
__global__ void momentum_table_kernel(float* __restrict__ mom, uint32_t n) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double t = 1.0;
    for (uint32_t k = 0; k < n; ++k) {
        const double tn = __dmul_rn(0.5, __dadd_rn(1.0, sqrt(__dadd_rn(1.0, __dmul_rn(4.0, __dmul_rn(t, t))))));
        mom[k] = float(__ddiv_rn(__dsub_rn(t, 1.0), tn));
        t = tn;
    }
}

// The persistent kernel's work counter starts at zero (a kernel, not a
// memset, so the call can be captured into a graph).
__global__ void clear_counter_kernel(unsigned int* __restrict__ counter) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
}

}  // namespace

void* noalloc_scratch() {
    int dev = 0;
    SF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        throw Error{SHOREFIT_E_DEVICE, "allocation-free entry point: device index out of range"};
    static std::mutex mu;
    static void* ptr[kMaxDevices] = {};
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        // each device holds its own copy of the array at its own address
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error{SHOREFIT_E_DEVICE,
                        std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        ptr[dev] = p;
    }
    return ptr[dev];
}

void launch_noalloc_prepare(float* mom, uint32_t n_mom, unsigned int* counter, hipStream_t s) {
    hipLaunchKernelGGL(clear_counter_kernel, dim3(1), dim3(64), 0, s, counter);
    SF_HIP(hipGetLastError());
    hipLaunchKernelGGL(momentum_table_kernel, dim3(1), dim3(64), 0, s, mom, n_mom);
    SF_HIP(hipGetLastError());
}

}  // namespace shorefit
