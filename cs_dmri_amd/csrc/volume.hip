// Volume layout conversion between the NIfTI storage order and the
// voxel-major order the fits use.
//
// A NIfTI file stores a 4D series volume after volume, each volume with x
// fastest: element (x, y, z, t) at ((t Z + z) Y + y) X + x. The fit wants
// each voxel's samples contiguous, (X, Y, Z, T) in C order: ((x Y + y) Z + z)
// T + t (src/io/dwi.rs:219-313). Both directions are a reversal of the four
// axes; on the device each (y, z) pair is a T x X transpose, done through LDS
// in 64 x 64 tiles so that the reads run along x and the writes along t (or
// the other way round), both coalesced. The reading direction also converts
// the stored type to f32 and applies scl_slope / scl_inter.
#include "common.h"

namespace shorefit {
namespace {

constexpr int kTile = 64;

// raw (file order, type T) -> out (voxel-major f32); value = f32(raw) when
// slope == 0, else f32(raw) * slope + inter in f32 (not fused).
template <typename T>
__global__ void __launch_bounds__(256) to_voxel_major_kernel(const T* __restrict__ raw, int64_t nx, int64_t ny,
                                                             int64_t nz, int64_t nt, float slope, float inter,
                                                             float* __restrict__ out) {
    __shared__ float tile[kTile][kTile + 1];
    const int64_t yz = blockIdx.x;  // y + Y z
    const int64_t x0 = int64_t(blockIdx.y) * kTile;
    const int64_t t0 = int64_t(blockIdx.z) * kTile;
    const int64_t y = yz % ny, z = yz / ny;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4 threads
    for (int r = ty; r < kTile; r += 4) {
        const int64_t t = t0 + r, x = x0 + tx;
        float v = 0.f;
        if (t < nt && x < nx) {
            v = float(raw[((t * nz + z) * ny + y) * nx + x]);
            if (slope != 0.f) {
#pragma clang fp contract(off)  // two roundings, as the reference's v * slope + inter
                v = v * slope + inter;
            }
        }
        tile[r][tx] = v;
    }
    __syncthreads();
    for (int r = ty; r < kTile; r += 4) {
        const int64_t x = x0 + r, t = t0 + tx;
        if (x < nx && t < nt) out[((x * ny + y) * nz + z) * nt + t] = tile[tx][r];
    }
}

// in (voxel-major f32) -> out (file order f32).
__global__ void __launch_bounds__(256) to_file_order_kernel(const float* __restrict__ in, int64_t nx, int64_t ny,
                                                            int64_t nz, int64_t nt, float* __restrict__ out) {
    __shared__ float tile[kTile][kTile + 1];
    const int64_t yz = blockIdx.x;
    const int64_t x0 = int64_t(blockIdx.y) * kTile;
    const int64_t t0 = int64_t(blockIdx.z) * kTile;
    const int64_t y = yz % ny, z = yz / ny;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < kTile; r += 4) {
        const int64_t x = x0 + r, t = t0 + tx;
        tile[r][tx] = (x < nx && t < nt) ? in[((x * ny + y) * nz + z) * nt + t] : 0.f;
    }
    __syncthreads();
    for (int r = ty; r < kTile; r += 4) {
        const int64_t t = t0 + r, x = x0 + tx;
        if (t < nt && x < nx) out[((t * nz + z) * ny + y) * nx + x] = tile[tx][r];
    }
}

dim3 layout_grid(int64_t nx, int64_t ny, int64_t nz, int64_t nt) {
    return dim3(unsigned(ny * nz), unsigned((nx + kTile - 1) / kTile), unsigned((nt + kTile - 1) / kTile));
}

}  // namespace

void launch_to_voxel_major(const void* raw, int datatype, int64_t nx, int64_t ny, int64_t nz, int64_t nt, float slope,
                           float inter, float* out, hipStream_t s) {
    if (nx * ny * nz * nt == 0) return;
    const dim3 grid = layout_grid(nx, ny, nz, nt);
#define SF_LAYOUT(T)                                                                                          \
    hipLaunchKernelGGL(to_voxel_major_kernel<T>, grid, dim3(256), 0, s, static_cast<const T*>(raw), nx, ny, nz, \
                       nt, slope, inter, out)
    switch (datatype) {  // NIfTI-1 datatype codes
        case 2: SF_LAYOUT(uint8_t); break;
        case 4: SF_LAYOUT(int16_t); break;
        case 8: SF_LAYOUT(int32_t); break;
        case 16: SF_LAYOUT(float); break;
        case 64: SF_LAYOUT(double); break;
        case 256: SF_LAYOUT(int8_t); break;
        case 512: SF_LAYOUT(uint16_t); break;
        case 768: SF_LAYOUT(uint32_t); break;
        default:
            throw Error{SHOREFIT_E_UNSUPPORTED, "NIfTI datatype " + std::to_string(datatype) +
                                                    " is not supported (uint8, int8, int16, uint16, int32, uint32, "
                                                    "float32, float64 are)"};
    }
#undef SF_LAYOUT
    SF_HIP(hipGetLastError());
}

void launch_to_file_order(const float* in, int64_t nx, int64_t ny, int64_t nz, int64_t nt, float* out,
                          hipStream_t s) {
    if (nx * ny * nz * nt == 0) return;
    hipLaunchKernelGGL(to_file_order_kernel, layout_grid(nx, ny, nz, nt), dim3(256), 0, s, in, nx, ny, nz, nt, out);
    SF_HIP(hipGetLastError());
}

}  // namespace shorefit
