// dev_common.hpp -- wave64 reduction helpers shared by the device translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ruv {

constexpr int BLK = 256;          // threads per workgroup of the vector / SpMV kernels
constexpr int WAVES = BLK / 64;   // wave64
constexpr int MAXG = 1024;        // cap on workgroups, hence on partials per reduction

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the workgroup; every thread returns the same value.  sh holds WAVES doubles.
__device__ __forceinline__ double block_sum_all(double v, double* sh) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < WAVES; ++i) s += sh[i];
    return s;
}

// Maximum over the workgroup; every thread returns the same value.  sh holds WAVES doubles.
__device__ __forceinline__ double block_max_all(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = sh[0];
#pragma unroll
    for (int i = 1; i < WAVES; ++i) m = fmax(m, sh[i]);
    return m;
}

// Sum of partials part[a, b) in a fixed order; identical in every workgroup.
__device__ __forceinline__ double sum_range(const double* __restrict__ part, int64_t a, int64_t b, double* sh) {
    double v = 0.0;
    for (int64_t i = a + threadIdx.x; i < b; i += BLK) v += part[i];
    return block_sum_all(v, sh);
}

__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int G, double* sh) {
    return sum_range(part, 0, G, sh);
}

}  // namespace ruv
