// pcg_kernels.hpp -- HIP kernels of the device PCG (included by pcg.hip only).
//
// Vectors hold R interleaved right-hand sides per operator row (R = 2 for the LSCM (u, v) pair).  Every
// reduction is deterministic: a wave64 xor-butterfly, then the waves of the workgroup in index order, then
// the per-workgroup partials in index order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dev_common.hpp"

namespace ruv {

constexpr int TRI_BLK = 1024;     // single-workgroup triangular sweep
constexpr int GRAPH_ITERS = 8;    // PCG iterations per captured HIP graph

enum Done : int32_t { RUNNING = 0, CONVERGED = 1, BREAKDOWN = 2, MAXITER = 3 };

// Device-resident scalar state of one PCG run.
struct SolveState {
    double rz;        // r.z of the current residual
    double tol_abs;   // stop when ||r|| <= tol_abs
    double rr;        // last r.r
    double pap;       // last p.Ap
    uint64_t iters;   // iterations started
    uint64_t max_iter;
    int32_t done;
    int32_t pad;
};

template <typename T>
__global__ __launch_bounds__(BLK) void k_fill0(int64_t n, T* __restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) x[i] = T(0);
}

template <typename T>
__global__ __launch_bounds__(BLK) void k_fill(int64_t n, T v, T* __restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) x[i] = v;
}

template <typename T>
__global__ __launch_bounds__(BLK) void k_copy(int64_t n, const T* __restrict__ src, T* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) dst[i] = src[i];
}

__global__ __launch_bounds__(BLK) void k_to_f32(int64_t n, const double* __restrict__ src, float* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) dst[i] = (float)src[i];
}

// x64 += d32 (the FP64 solution update of the refinement loop); per-workgroup maxima of |d| and of the updated
// |x| into pd[blockIdx.x], px[blockIdx.x]
__global__ __launch_bounds__(BLK) void k_axpy_f32(int64_t n, const float* __restrict__ d, double* __restrict__ x,
                                                  double* __restrict__ pd, double* __restrict__ px) {
    __shared__ double sh[WAVES];
    double dmax = 0.0, xmax = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
        const double di = (double)d[i];
        const double v = x[i] + di;
        x[i] = v;
        dmax = fmax(dmax, fabs(di));
        xmax = fmax(xmax, fabs(v));
    }
    dmax = block_max_all(dmax, sh);
    xmax = block_max_all(xmax, sh);
    if (threadIdx.x == 0) {
        pd[blockIdx.x] = dmax;
        px[blockIdx.x] = xmax;
    }
}

// per-workgroup partials of a.b (FP64 accumulation)
template <typename T, int R>
__global__ __launch_bounds__(BLK) void k_dot(int n, const T* __restrict__ a, const T* __restrict__ b,
                                             double* __restrict__ part, const SolveState* st) {
    __shared__ double sh[WAVES];
    if (st && st->done) return;
    double acc = 0.0;
    const int64_t N = (int64_t)n * R;
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < N; i += (int64_t)gridDim.x * BLK)
        acc += (double)a[i] * (double)b[i];
    const double s = block_sum_all(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// State at x0 = 0, r0 = b (solver.rs:461-468): tol_abs = tol*||b|| (tol when ||b|| <= 1e-300).
__global__ __launch_bounds__(BLK) void k_init_state(const double* __restrict__ part_bb, int G, double tol,
                                                    uint64_t max_iter, SolveState* st) {
    __shared__ double sh[WAVES];
    const double bb = sum_partials(part_bb, G, sh);
    if (threadIdx.x == 0) {
        const double bnorm = sqrt(bb);
        st->tol_abs = (bnorm <= 1e-300) ? tol : tol * bnorm;
        st->rr = bb;
        st->rz = 0.0;
        st->pap = 0.0;
        st->iters = 0;
        st->max_iter = max_iter;
        st->done = (sqrt(bb) <= st->tol_abs) ? CONVERGED : RUNNING;
    }
}

template <typename T, int R>
__global__ __launch_bounds__(BLK) void k_jacobi(int n, const T* __restrict__ dinv, const T* __restrict__ r,
                                                T* __restrict__ z) {
    for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
        const T d = dinv[i];
#pragma unroll
        for (int c = 0; c < R; ++c) z[(int64_t)i * R + c] = d * r[(int64_t)i * R + c];
    }
}

// q = A p and partials of p.q.  Workgroup 0 also opens the iteration: rz <- sum(part_rz), iteration
// count, and the max_iter stop of `while !converged && iters < max_iter` (solver.rs:489-490).
template <typename T, int R>
__global__ __launch_bounds__(BLK) void k_spmv_pq(int n, const int* __restrict__ ptr, const int* __restrict__ cols,
                                                 const T* __restrict__ vals, const T* __restrict__ p,
                                                 T* __restrict__ q, double* __restrict__ part_pq,
                                                 const double* __restrict__ part_rz, int G, SolveState* st) {
    __shared__ double sh[WAVES];
    if (st->done) return;
    if (blockIdx.x == 0) {
        const double rz = sum_partials(part_rz, G, sh);
        if (threadIdx.x == 0) {
            if (st->iters >= st->max_iter) {
                st->done = MAXITER;
            } else {
                st->iters += 1;
                st->rz = rz;
            }
        }
    }
    double acc = 0.0;
    for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
        T s[R];
#pragma unroll
        for (int c = 0; c < R; ++c) s[c] = T(0);
        const int a = ptr[i], b = ptr[i + 1];
        for (int k = a; k < b; ++k) {
            const T v = vals[k];
            const int64_t col = (int64_t)cols[k] * R;
#pragma unroll
            for (int c = 0; c < R; ++c) s[c] += v * p[col + c];
        }
#pragma unroll
        for (int c = 0; c < R; ++c) {
            q[(int64_t)i * R + c] = s[c];
            acc += (double)p[(int64_t)i * R + c] * (double)s[c];
        }
    }
    const double bs = block_sum_all(acc, sh);
    if (threadIdx.x == 0) part_pq[blockIdx.x] = bs;
}

// alpha = rz / pAp (breakdown check, solver.rs:505-509); x += alpha p; r -= alpha q; z = M^-1 r for
// PC = 1 (Jacobi); partials of r.r and (PC 0/1) r.z.
template <typename T, int R, int PC>
__global__ __launch_bounds__(BLK) void k_update(int n, const T* __restrict__ p, const T* __restrict__ q,
                                                T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                const T* __restrict__ dinv, const double* __restrict__ part_pq,
                                                double* __restrict__ part_rr, double* __restrict__ part_rz, int G,
                                                SolveState* st) {
    __shared__ double sh[WAVES];
    if (st->done) return;
    const double pap = sum_partials(part_pq, G, sh);
    if (!(fabs(pap) > 1e-300) || !isfinite(pap)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->done = BREAKDOWN;
            st->pap = pap;
        }
        return;
    }
    const double alpha = st->rz / pap;
    const T a = (T)alpha;
    double arr = 0.0, arz = 0.0;
    for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
        const T d = (PC == 1) ? dinv[i] : T(1);
#pragma unroll
        for (int c = 0; c < R; ++c) {
            const int64_t k = (int64_t)i * R + c;
            const T xi = x[k] + a * p[k];
            const T ri = r[k] - a * q[k];
            x[k] = xi;
            r[k] = ri;
            arr += (double)ri * (double)ri;
            if (PC == 1) {
                const T zi = d * ri;
                z[k] = zi;
                arz += (double)ri * (double)zi;
            }
        }
    }
    const double srr = block_sum_all(arr, sh);
    const double srz = (PC == 1) ? block_sum_all(arz, sh) : srr;
    if (threadIdx.x == 0) {
        part_rr[blockIdx.x] = srr;
        if (PC != 2) part_rz[blockIdx.x] = srz;
        if (blockIdx.x == 0) st->pap = pap;
    }
}

// Convergence on the recursive residual (solver.rs:518-522), then beta = rz_new / rz and p = z + beta p.
template <typename T, int R, int PC>
__global__ __launch_bounds__(BLK) void k_pupdate(int n, const T* __restrict__ z, T* __restrict__ p,
                                                 const double* __restrict__ part_rr,
                                                 const double* __restrict__ part_rz, int G, SolveState* st) {
    __shared__ double sh[WAVES];
    if (st->done) return;
    const double rr = sum_partials(part_rr, G, sh);
    if (sqrt(rr) <= st->tol_abs) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->done = CONVERGED;
            st->rr = rr;
        }
        return;
    }
    const double rznew = (PC == 0) ? rr : sum_partials(part_rz, G, sh);
    const double beta = rznew / st->rz;
    const T bt = (T)beta;
    if (blockIdx.x == 0 && threadIdx.x == 0) st->rr = rr;
    const int64_t N = (int64_t)n * R;
    for (int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x; k < N; k += (int64_t)gridDim.x * BLK)
        p[k] = z[k] + bt * p[k];
}

// One triangular solve over a diagonal-last CSR factor, all levels in one single-workgroup launch: the
// rows of a level are independent; the workgroup barrier separates levels (the reference launched one
// kernel per level, solver.rs:359-403, kernels.cu:81-95).
template <typename T, int R>
__global__ __launch_bounds__(TRI_BLK) void k_tri_sweep(int nlev, const int* __restrict__ off,
                                                       const int* __restrict__ rows, const int* __restrict__ ptr,
                                                       const int* __restrict__ cols, const T* __restrict__ vals,
                                                       const T* __restrict__ in, T* out, const SolveState* st) {
    if (st && st->done) return;
    for (int l = 0; l < nlev; ++l) {
        const int a = off[l], b = off[l + 1];
        for (int t = a + (int)threadIdx.x; t < b; t += TRI_BLK) {
            const int i = rows[t];
            const int e = ptr[i + 1] - 1;  // diagonal last
            T s[R];
#pragma unroll
            for (int c = 0; c < R; ++c) s[c] = in[(int64_t)i * R + c];
            for (int k = ptr[i]; k < e; ++k) {
                const T v = vals[k];
                const int64_t col = (int64_t)cols[k] * R;
#pragma unroll
                for (int c = 0; c < R; ++c) s[c] -= v * out[col + c];
            }
            const T d = vals[e];
#pragma unroll
            for (int c = 0; c < R; ++c) out[(int64_t)i * R + c] = s[c] / d;
        }
        __syncthreads();
    }
}

// One wide level of a triangular sweep over many workgroups: rows rows[a .. b), the same arithmetic per row as
// k_tri_sweep (a level's rows depend only on earlier levels).
template <typename T, int R>
__global__ __launch_bounds__(BLK) void k_tri_level(int a, int b, const int* __restrict__ rows,
                                                   const int* __restrict__ ptr, const int* __restrict__ cols,
                                                   const T* __restrict__ vals, const T* __restrict__ in, T* out,
                                                   const SolveState* st) {
    if (st && st->done) return;
    const int t = a + (int)(blockIdx.x * BLK + threadIdx.x);
    if (t >= b) return;
    const int i = rows[t];
    const int e = ptr[i + 1] - 1;  // diagonal last
    T s[R];
#pragma unroll
    for (int c = 0; c < R; ++c) s[c] = in[(int64_t)i * R + c];
    for (int k = ptr[i]; k < e; ++k) {
        const T v = vals[k];
        const int64_t col = (int64_t)cols[k] * R;
#pragma unroll
        for (int c = 0; c < R; ++c) s[c] -= v * out[col + c];
    }
    const T d = vals[e];
#pragma unroll
    for (int c = 0; c < R; ++c) out[(int64_t)i * R + c] = s[c] / d;
}

// dst[i] = src[idx[i]] per row (permutation round trips, kernels.cu:98-102)
template <typename T, int R>
__global__ __launch_bounds__(BLK) void k_gather(int n, const int* __restrict__ idx, const T* __restrict__ src,
                                                T* __restrict__ dst, const SolveState* st) {
    if (st && st->done) return;
    for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
        const int64_t s = (int64_t)idx[i] * R;
#pragma unroll
        for (int c = 0; c < R; ++c) dst[(int64_t)i * R + c] = src[s + c];
    }
}

// r = b - A x in FP64 and partials of r.r (the refinement residual and the final true residual)
template <int R>
__global__ __launch_bounds__(BLK) void k_residual(int n, const int* __restrict__ ptr, const int* __restrict__ cols,
                                                  const double* __restrict__ vals, const double* __restrict__ x,
                                                  const double* __restrict__ b, double* __restrict__ r,
                                                  double* __restrict__ part_rr) {
    __shared__ double sh[WAVES];
    double acc = 0.0;
    for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
        double s[R];
#pragma unroll
        for (int c = 0; c < R; ++c) s[c] = 0.0;
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
            const double v = vals[k];
            const int64_t col = (int64_t)cols[k] * R;
#pragma unroll
            for (int c = 0; c < R; ++c) s[c] += v * x[col + c];
        }
#pragma unroll
        for (int c = 0; c < R; ++c) {
            const double rc = b[(int64_t)i * R + c] - s[c];
            r[(int64_t)i * R + c] = rc;
            acc += rc * rc;
        }
    }
    const double bs = block_sum_all(acc, sh);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = bs;
}

// ---------------------------------------------------------------- IC(0) factorisation on the device
// Row i of the IC(0) factor (factor_ic0, precond.rs:145-219) in FP64 from A's values: for each strict-lower
// entry (ascending column k) L(i,k) = (A(i,k) - sum_{j<k} L(i,j) L(k,j)) / L(k,k), the products subtracted in
// ascending j (a merge of row i's entries so far with row k's); then L(i,i) = sqrt(max(A(i,i) - sum L(i,k)^2,
// 1e-12 max(|A(i,i)|, 1e-30))).  Rows k < i are those of earlier levels.  The reference's operations in its order
// and without fused multiply-adds, so the factor is the host's to the bit.  A vanishing pivot or diagonal sets
// *fail (the host factorisation stops there; here the row goes on, and the caller discards the factor).
__device__ __forceinline__ void ic0_row(int i, const double* __restrict__ av, const int64_t* __restrict__ src,
                                        const int64_t* __restrict__ dsrc, const int* __restrict__ ptr,
                                        const int* __restrict__ cols, double* lv, int* fail) {
#pragma clang fp contract(off)
    const int start = ptr[i], e = ptr[i + 1] - 1;
    for (int q = start; q < e; ++q) {
        const int k = cols[q];
        double s = av[src[q]];
        int qk = ptr[k];
        const int qke = ptr[k + 1] - 1;  // row k's diagonal
        for (int qi = start; qi < q; ++qi) {
            const int j = cols[qi];
            while (qk < qke && cols[qk] < j) ++qk;
            if (qk < qke && cols[qk] == j) s -= lv[qi] * lv[qk];
        }
        const double lkk = lv[qke];
        if (fabs(lkk) < 1e-300) {
            atomicOr(fail, 1);
            lv[q] = 0.0;
            continue;
        }
        lv[q] = s / lkk;
    }
    const double aii = dsrc[i] >= 0 ? av[dsrc[i]] : 0.0;
    if (fabs(aii) < 1e-300) {
        atomicOr(fail, 1);
        lv[e] = 1.0;
        return;
    }
    double sum = aii;
    for (int q = start; q < e; ++q) sum -= lv[q] * lv[q];
    const double fl = 1e-12 * fmax(fabs(aii), 1e-30);
    if (sum < fl) sum = fl;
    lv[e] = sqrt(sum);
}

// a run of narrow levels in one workgroup (the levels' rows one per thread, a barrier between levels)
__global__ __launch_bounds__(TRI_BLK) void k_ic0_sweep(int nlev, const int* __restrict__ off,
                                                       const int* __restrict__ rows, const double* __restrict__ av,
                                                       const int64_t* __restrict__ src, const int64_t* __restrict__ dsrc,
                                                       const int* __restrict__ ptr, const int* __restrict__ cols,
                                                       double* lv, int* fail) {
    for (int l = 0; l < nlev; ++l) {
        for (int t = off[l] + (int)threadIdx.x; t < off[l + 1]; t += TRI_BLK)
            ic0_row(rows[t], av, src, dsrc, ptr, cols, lv, fail);
        __syncthreads();
    }
}

// one wide level over many workgroups: rows rows[a .. b)
__global__ __launch_bounds__(BLK) void k_ic0_level(int a, int b, const int* __restrict__ rows,
                                                   const double* __restrict__ av, const int64_t* __restrict__ src,
                                                   const int64_t* __restrict__ dsrc, const int* __restrict__ ptr,
                                                   const int* __restrict__ cols, double* lv, int* fail) {
    const int t = a + (int)(blockIdx.x * BLK + threadIdx.x);
    if (t < b) ic0_row(rows[t], av, src, dsrc, ptr, cols, lv, fail);
}

// the factor in the sweeps' precision: L (unless it is the FP64 workspace itself) and L^T (lt_src: each L^T entry's
// L entry)
template <typename T>
__global__ __launch_bounds__(BLK) void k_ic0_store(int64_t nl, const double* lv, const int32_t* __restrict__ lt_src,
                                                   T* l_out, T* __restrict__ lt_out) {
    const int64_t q = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (q >= nl) return;
    if ((const void*)l_out != (const void*)lv) l_out[q] = (T)lv[q];
    lt_out[q] = (T)lv[lt_src[q]];
}

}  // namespace ruv
