// amg_kernels.hpp -- device kernels of the AMG V(1,1) cycle (included by pcg.hip).
//
// Per level l with operator A, inverse diagonal d, damping w, prolongator P and restriction R = P^T, the
// cycle maps a residual r to a correction e:
//   t  = r - A (w d r)                 k_amg_pre        (pre-smooth from zero, then the residual)
//   rc = R t                           k_amg_spmv       (restrict)
//   ec = cycle(l + 1, rc)
//   e1 = w d r + P ec                  k_amg_prolong
//   e  = e1 + w d (r - A e1)           k_amg_post       (post-smooth)
// Equal pre- and post-smoothing with R = P^T and an exact coarsest solve make the cycle a fixed SPD
// operator.  The coarsest level applies its dense FP64 inverse (one wave per row) or, when aggregation
// stalled, its Jacobi diagonal (amg.rs:470-483).
#pragma once

#include <hip/hip_runtime.h>

#include "dev_common.hpp"

namespace ruv {

template <typename T, int R>
__global__ __launch_bounds__(BLK) void k_amg_pre(int n, const int* __restrict__ ptr, const int* __restrict__ cols,
                                                 const T* __restrict__ vals, const T* __restrict__ dinv, T w,
                                                 const T* __restrict__ r, T* __restrict__ t, const int32_t* done) {
    if (done && *done) return;
    for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
        T s[R];
#pragma unroll
        for (int c = 0; c < R; ++c) s[c] = T(0);
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
            const int j = cols[k];
            const T v = vals[k] * (w * dinv[j]);
#pragma unroll
            for (int c = 0; c < R; ++c) s[c] += v * r[(int64_t)j * R + c];
        }
#pragma unroll
        for (int c = 0; c < R; ++c) t[(int64_t)i * R + c] = r[(int64_t)i * R + c] - s[c];
    }
}

template <typename T, int R>
__global__ __launch_bounds__(BLK) void k_amg_spmv(int n, const int* __restrict__ ptr, const int* __restrict__ cols,
                                                  const T* __restrict__ vals, const T* __restrict__ x,
                                                  T* __restrict__ y, const int32_t* done) {
    if (done && *done) return;
    for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
        T s[R];
#pragma unroll
        for (int c = 0; c < R; ++c) s[c] = T(0);
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
            const T v = vals[k];
            const int64_t j = (int64_t)cols[k] * R;
#pragma unroll
            for (int c = 0; c < R; ++c) s[c] += v * x[j + c];
        }
#pragma unroll
        for (int c = 0; c < R; ++c) y[(int64_t)i * R + c] = s[c];
    }
}

template <typename T, int R>
__global__ __launch_bounds__(BLK) void k_amg_prolong(int n, const int* __restrict__ ptr, const int* __restrict__ cols,
                                                     const T* __restrict__ vals, const T* __restrict__ ec,
                                                     const T* __restrict__ dinv, T w, const T* __restrict__ r,
                                                     T* __restrict__ e1, const int32_t* done) {
    if (done && *done) return;
    for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
        const T dw = w * dinv[i];
        T s[R];
#pragma unroll
        for (int c = 0; c < R; ++c) s[c] = dw * r[(int64_t)i * R + c];
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
            const T v = vals[k];
            const int64_t j = (int64_t)cols[k] * R;
#pragma unroll
            for (int c = 0; c < R; ++c) s[c] += v * ec[j + c];
        }
#pragma unroll
        for (int c = 0; c < R; ++c) e1[(int64_t)i * R + c] = s[c];
    }
}

template <typename T, int R>
__global__ __launch_bounds__(BLK) void k_amg_post(int n, const int* __restrict__ ptr, const int* __restrict__ cols,
                                                  const T* __restrict__ vals, const T* __restrict__ dinv, T w,
                                                  const T* __restrict__ r, const T* __restrict__ e1,
                                                  T* __restrict__ e, const int32_t* done) {
    if (done && *done) return;
    for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
        T s[R];
#pragma unroll
        for (int c = 0; c < R; ++c) s[c] = T(0);
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
            const T v = vals[k];
            const int64_t j = (int64_t)cols[k] * R;
#pragma unroll
            for (int c = 0; c < R; ++c) s[c] += v * e1[j + c];
        }
        const T dw = w * dinv[i];
#pragma unroll
        for (int c = 0; c < R; ++c) {
            const int64_t k = (int64_t)i * R + c;
            e[k] = e1[k] + dw * (r[k] - s[c]);
        }
    }
}

// e = Ainv r with the dense FP64 inverse of the coarsest operator: one wave64 per row, lanes stride the
// columns (coalesced), FP64 accumulation.
template <typename T, int R>
__global__ __launch_bounds__(BLK) void k_amg_dense(int n, const double* __restrict__ inv, const T* __restrict__ r,
                                                   T* __restrict__ e, const int32_t* done) {
    if (done && *done) return;
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * WAVES + (threadIdx.x >> 6);
    if (row >= n) return;
    double s[R];
#pragma unroll
    for (int c = 0; c < R; ++c) s[c] = 0.0;
    for (int j = lane; j < n; j += 64) {
        const double a = inv[(int64_t)row * n + j];
#pragma unroll
        for (int c = 0; c < R; ++c) s[c] += a * (double)r[(int64_t)j * R + c];
    }
#pragma unroll
    for (int c = 0; c < R; ++c) {
        const double v = wave_sum(s[c]);
        if (lane == 0) e[(int64_t)row * R + c] = (T)v;
    }
}

template <typename T, int R>
__global__ __launch_bounds__(BLK) void k_amg_diag(int n, const T* __restrict__ dinv, const T* __restrict__ r,
                                                  T* __restrict__ e, const int32_t* done) {
    if (done && *done) return;
    for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK)
#pragma unroll
        for (int c = 0; c < R; ++c) e[(int64_t)i * R + c] = dinv[i] * r[(int64_t)i * R + c];
}

}  // namespace ruv
