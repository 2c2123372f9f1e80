// batch.hip -- batched LSCM solves: thousands of independent charts, one set of launches per iteration.
//
// The reference solved charts one at a time from worker threads (crates/researchuv-unwrap/src/pipeline.rs:
// 171-194 -> lscm.rs:471-506 -> solver.rs:284-565), each solve paying its own uploads, host-synchronised dot
// products and launches.  Here every chart of a batch is one diagonal block of a single block-diagonal CSR
// (rows of L, vectors interleaved (u, v)); the rows are cut into tiles of at most TILE rows that never span
// two charts, and each workgroup owns one tile.  Each chart keeps its OWN CG scalars (alpha, beta, r.z,
// convergence): a tile's partial dot products are summed per chart by every workgroup of that chart, in a
// fixed order, so the recurrence of every chart is exactly an independent PCG.  Converged charts drop out:
// the host compacts the list of live tiles between chunks of iterations, so late iterations launch only the
// tiles of the charts still running, spread over the whole GPU.
//
// Mixed precision: the inner PCG runs in FP32 (matrix, Jacobi diagonal, Krylov vectors); the outer loop keeps
// x and r = b - A x in FP64 per chart and stops a chart when its true residual ||b - A x|| <= tol ||b||.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "dev_common.hpp"
#include "ruv_internal.hpp"

namespace ruv {

#define HIPCHK(call)                                                                                         \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess) return fail(RUV_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

constexpr int TILE = 512;  // rows per workgroup tile (two per thread)
enum : int32_t { B_RUNNING = 0, B_CONVERGED = 1, B_BREAKDOWN = 2, B_MAXITER = 3, B_SKIP = 4 };

struct BDev {
    const int64_t* ptr;
    const int* cols;
    const int* tile_chart;
    const int64_t* tile_r0;
    const int64_t* tile_r1;
    const int* chart_t0;  // C + 1
    double* c_rz;
    double* c_tol;
    double* c_bnorm;
    double* c_true;
    int* c_done;
    int* c_final;
    unsigned* c_iters;
    double* part_a;
    double* part_b;
    double* part_c;
    unsigned max_iter;
    int nchart;
};

template <typename T, int R>
__device__ __forceinline__ double row_spmv(const BDev& B, const T* __restrict__ vals, const T* __restrict__ p,
                                           T* __restrict__ q, int64_t i) {
    T s[R];
#pragma unroll
    for (int c = 0; c < R; ++c) s[c] = T(0);
    const int64_t a = B.ptr[i], b = B.ptr[i + 1];
    for (int64_t k = a; k < b; ++k) {
        const T v = vals[k];
        const int64_t col = (int64_t)B.cols[k] * R;
#pragma unroll
        for (int c = 0; c < R; ++c) s[c] += v * p[col + c];
    }
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < R; ++c) {
        q[i * R + c] = s[c];
        acc += (double)p[i * R + c] * (double)s[c];
    }
    return acc;
}

// q = A p over a tile, partial p.q; the chart's first tile opens the iteration (rz, count, max_iter stop).
template <typename T, int R>
__global__ __launch_bounds__(BLK) void kb_spmv(BDev B, const int* __restrict__ act, const T* __restrict__ vals,
                                               const T* __restrict__ p, T* __restrict__ q) {
    __shared__ double sh[WAVES];
    const int t = act[blockIdx.x];
    const int c = B.tile_chart[t];
    if (B.c_done[c] != B_RUNNING) return;
    const int t0 = B.chart_t0[c];
    if (t == t0) {
        const double rz = sum_range(B.part_c, t0, B.chart_t0[c + 1], sh);
        if (threadIdx.x == 0) {
            if (B.c_iters[c] >= B.max_iter) {
                B.c_done[c] = B_MAXITER;
            } else {
                B.c_iters[c] += 1;
                B.c_rz[c] = rz;
            }
        }
    }
    double acc = 0.0;
    const int64_t r1 = B.tile_r1[t];
    for (int64_t i = B.tile_r0[t] + threadIdx.x; i < r1; i += BLK) acc += row_spmv<T, R>(B, vals, p, q, i);
    const double s = block_sum_all(acc, sh);
    if (threadIdx.x == 0) B.part_a[t] = s;
}

// alpha = rz / pAp per chart; x += alpha p; r -= alpha q; z = D^-1 r (PC 1) ; partials r.r, r.z
template <typename T, int R, int PC>
__global__ __launch_bounds__(BLK) void kb_update(BDev B, const int* __restrict__ act, const T* __restrict__ dinv,
                                                 const T* __restrict__ p, const T* __restrict__ q,
                                                 T* __restrict__ x, T* __restrict__ r, T* __restrict__ z) {
    __shared__ double sh[WAVES];
    const int t = act[blockIdx.x];
    const int c = B.tile_chart[t];
    if (B.c_done[c] != B_RUNNING) return;
    const int t0 = B.chart_t0[c];
    const double pap = sum_range(B.part_a, t0, B.chart_t0[c + 1], sh);
    if (!(fabs(pap) > 1e-300) || !isfinite(pap)) {
        if (t == t0 && threadIdx.x == 0) B.c_done[c] = B_BREAKDOWN;
        return;
    }
    const T a = (T)(B.c_rz[c] / pap);
    double arr = 0.0, arz = 0.0;
    const int64_t r1 = B.tile_r1[t];
    for (int64_t i = B.tile_r0[t] + threadIdx.x; i < r1; i += BLK) {
        const T d = (PC == 1) ? dinv[i] : T(1);
#pragma unroll
        for (int cc = 0; cc < R; ++cc) {
            const int64_t k = i * R + cc;
            x[k] = x[k] + a * p[k];
            const T ri = r[k] - a * q[k];
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
        B.part_b[t] = srr;
        B.part_c[t] = srz;
    }
}

// per-chart convergence on the recursive residual, then p = z + beta p
template <typename T, int R, int PC>
__global__ __launch_bounds__(BLK) void kb_pupdate(BDev B, const int* __restrict__ act, const T* __restrict__ z,
                                                  T* __restrict__ p) {
    __shared__ double sh[WAVES];
    const int t = act[blockIdx.x];
    const int c = B.tile_chart[t];
    if (B.c_done[c] != B_RUNNING) return;
    const int t0 = B.chart_t0[c], t1 = B.chart_t0[c + 1];
    const double rr = sum_range(B.part_b, t0, t1, sh);
    if (sqrt(rr) <= B.c_tol[c]) {
        if (t == t0 && threadIdx.x == 0) B.c_done[c] = B_CONVERGED;
        return;
    }
    const double rznew = (PC == 0) ? rr : sum_range(B.part_c, t0, t1, sh);
    const T bt = (T)(rznew / B.c_rz[c]);
    const int64_t r1 = B.tile_r1[t];
    for (int64_t k = B.tile_r0[t] * R + threadIdx.x; k < r1 * R; k += BLK) p[k] = z[k] + bt * p[k];
}

// start of a solve: x64 = 0, r64 = b, partials of b.b
template <int R>
__global__ __launch_bounds__(BLK) void kb_start(BDev B, const double* __restrict__ b, double* __restrict__ x64,
                                                double* __restrict__ r64) {
    __shared__ double sh[WAVES];
    const int t = blockIdx.x;
    double acc = 0.0;
    const int64_t r1 = B.tile_r1[t];
    for (int64_t k = B.tile_r0[t] * R + threadIdx.x; k < r1 * R; k += BLK) {
        const double v = b[k];
        x64[k] = 0.0;
        r64[k] = v;
        acc += v * v;
    }
    const double s = block_sum_all(acc, sh);
    if (threadIdx.x == 0) B.part_b[t] = s;
}

// per chart: ||b||, not final
__global__ __launch_bounds__(BLK) void kb_bnorm(BDev B) {
    const int c = blockIdx.x * BLK + threadIdx.x;
    if (c >= B.nchart) return;
    double s = 0.0;
    for (int t = B.chart_t0[c]; t < B.chart_t0[c + 1]; ++t) s += B.part_b[t];
    B.c_bnorm[c] = sqrt(s);
    B.c_final[c] = 0;
    B.c_true[c] = 1.0;
}

// inner start for the charts not yet final: x = 0, r = T(r64), z = D^-1 r, p = z; partials r.r and r.z
template <typename T, int R, int PC>
__global__ __launch_bounds__(BLK) void kb_prep(BDev B, const int* __restrict__ act, const double* __restrict__ r64,
                                               const T* __restrict__ dinv, T* __restrict__ x, T* __restrict__ r,
                                               T* __restrict__ z, T* __restrict__ p) {
    __shared__ double sh[WAVES];
    const int t = act[blockIdx.x];
    const int c = B.tile_chart[t];
    double arr = 0.0, arz = 0.0;
    if (!B.c_final[c]) {
        const int64_t r1 = B.tile_r1[t];
        for (int64_t i = B.tile_r0[t] + threadIdx.x; i < r1; i += BLK) {
            const T d = (PC == 1) ? dinv[i] : T(1);
#pragma unroll
            for (int cc = 0; cc < R; ++cc) {
                const int64_t k = i * R + cc;
                const T ri = (T)r64[k];
                const T zi = d * ri;
                x[k] = T(0);
                r[k] = ri;
                if (PC == 1) z[k] = zi;
                p[k] = zi;
                arr += (double)ri * (double)ri;
                arz += (double)ri * (double)zi;
            }
        }
    }
    const double srr = block_sum_all(arr, sh);
    const double srz = block_sum_all(arz, sh);
    if (threadIdx.x == 0) {
        B.part_b[t] = srr;
        B.part_c[t] = srz;
    }
}

// per chart inner tolerance and state.  plain: tol * ||b|| on the recursive residual (solver.rs:465-468);
// mixed: inner_tol * ||r64|| but no tighter than a quarter of the outer target.
__global__ __launch_bounds__(BLK) void kb_init(BDev B, double tol, double inner_tol, int mixed) {
    const int c = blockIdx.x * BLK + threadIdx.x;
    if (c >= B.nchart) return;
    if (B.c_final[c]) {
        B.c_done[c] = B_SKIP;
        return;
    }
    double s = 0.0;
    for (int t = B.chart_t0[c]; t < B.chart_t0[c + 1]; ++t) s += B.part_b[t];
    const double rn = sqrt(s), bn = B.c_bnorm[c];
    double ta;
    if (mixed) ta = fmax(inner_tol * rn, 0.25 * ((bn <= 1e-300) ? tol : tol * bn));
    else ta = (bn <= 1e-300) ? tol : tol * bn;
    B.c_tol[c] = ta;
    B.c_iters[c] = 0;
    B.c_done[c] = (rn <= ta) ? B_CONVERGED : B_RUNNING;
}

// x64 += x (the inner correction) on the live charts
template <typename T, int R>
__global__ __launch_bounds__(BLK) void kb_axpy(BDev B, const int* __restrict__ act, const T* __restrict__ xin,
                                               double* __restrict__ x64) {
    const int t = act[blockIdx.x];
    if (B.c_final[B.tile_chart[t]]) return;
    const int64_t r1 = B.tile_r1[t] * R;
    for (int64_t k = B.tile_r0[t] * R + threadIdx.x; k < r1; k += BLK) x64[k] += (double)xin[k];
}

// r64 = b - A64 x64 in FP64 on the live charts, partials of r64.r64 (runs after kb_axpy: rows of a tile read
// x64 of rows other tiles update)
template <int R>
__global__ __launch_bounds__(BLK) void kb_residual(BDev B, const int* __restrict__ act,
                                                   const double* __restrict__ vals64, const double* __restrict__ b,
                                                   const double* __restrict__ x64, double* __restrict__ r64) {
    __shared__ double sh[WAVES];
    const int t = act[blockIdx.x];
    const int c = B.tile_chart[t];
    double acc = 0.0;
    if (!B.c_final[c]) {
        const int64_t r1 = B.tile_r1[t];
        for (int64_t i = B.tile_r0[t] + threadIdx.x; i < r1; i += BLK) {
            double s[R];
#pragma unroll
            for (int cc = 0; cc < R; ++cc) s[cc] = 0.0;
            for (int64_t k = B.ptr[i]; k < B.ptr[i + 1]; ++k) {
                const double v = vals64[k];
                const int64_t col = (int64_t)B.cols[k] * R;
#pragma unroll
                for (int cc = 0; cc < R; ++cc) s[cc] += v * x64[col + cc];
            }
#pragma unroll
            for (int cc = 0; cc < R; ++cc) {
                const double rc = b[i * R + cc] - s[cc];
                r64[i * R + cc] = rc;
                acc += rc * rc;
            }
        }
    }
    const double ss = block_sum_all(acc, sh);
    if (threadIdx.x == 0) B.part_b[t] = ss;
}

// per chart: true residual; final when ||b - A x|| <= tol ||b||
__global__ __launch_bounds__(BLK) void kb_outer(BDev B, double tol) {
    const int c = blockIdx.x * BLK + threadIdx.x;
    if (c >= B.nchart || B.c_final[c]) return;
    double s = 0.0;
    for (int t = B.chart_t0[c]; t < B.chart_t0[c + 1]; ++t) s += B.part_b[t];
    const double rn = sqrt(s), bn = B.c_bnorm[c];
    B.c_true[c] = (bn > 0.0) ? rn / bn : rn;
    if (rn <= ((bn <= 1e-300) ? tol : tol * bn)) B.c_final[c] = 1;
}

}  // namespace

struct BatchImpl {
    int R = 2, mode = 0, precond = 0, nchart = 0, ntile = 0;
    int64_t nrows = 0, nnz = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<void*> mem;
    BDev B{};
    double* vals64 = nullptr;
    float* vals32 = nullptr;
    double* dinv64 = nullptr;
    float* dinv32 = nullptr;
    double *x64 = nullptr, *r64 = nullptr;
    double *xd = nullptr, *rd = nullptr, *zd = nullptr, *pd = nullptr, *qd = nullptr;
    float *xf = nullptr, *rf = nullptr, *zf = nullptr, *pf = nullptr, *qf = nullptr;
    int* act = nullptr;
    int* h_act = nullptr;
    int* h_done = nullptr;
    int* h_final = nullptr;
    unsigned* h_iters = nullptr;
    double* h_true = nullptr;
    std::vector<int> chart_t0;
    std::vector<uint64_t> iters_total;
    double inner_tol = 1e-4;
    int max_rounds = 40;

    template <typename T>
    hipError_t alloc(T** p, size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) mem.push_back(q);
        *p = (T*)q;
        return e;
    }
    template <typename T>
    hipError_t upload(T** p, const T* h, size_t n) {
        hipError_t e = alloc(p, n);
        if (e == hipSuccess && n) e = hipMemcpyAsync(*p, h, n * sizeof(T), hipMemcpyHostToDevice, stream);
        return e;
    }
    ~BatchImpl() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : mem) (void)hipFree(p);
        for (void* p : {(void*)h_act, (void*)h_done, (void*)h_final, (void*)h_iters, (void*)h_true})
            if (p) (void)hipHostFree(p);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int batch_setup(BatchImpl& Q, int64_t nrows, const int64_t* ptr, const int32_t* cols, const double* vals,
                int nchart, const int64_t* crow, int R, int precond, int mode) {
    Q.R = R;
    Q.mode = mode;
    Q.precond = precond;
    Q.nchart = nchart;
    Q.nrows = nrows;
    Q.nnz = ptr[nrows];
    HIPCHK(hipStreamCreateWithFlags(&Q.stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&Q.ev0));
    HIPCHK(hipEventCreate(&Q.ev1));
    // tiles
    std::vector<int> tile_chart;
    std::vector<int64_t> tr0, tr1;
    Q.chart_t0.assign(nchart + 1, 0);
    for (int c = 0; c < nchart; ++c) {
        Q.chart_t0[c] = (int)tile_chart.size();
        for (int64_t r = crow[c]; r < crow[c + 1]; r += TILE) {
            tile_chart.push_back(c);
            tr0.push_back(r);
            tr1.push_back(std::min<int64_t>(r + TILE, crow[c + 1]));
        }
        if (crow[c + 1] == crow[c]) {  // an empty chart still owns one (empty) tile so its scalars exist
            tile_chart.push_back(c);
            tr0.push_back(crow[c]);
            tr1.push_back(crow[c]);
        }
    }
    Q.chart_t0[nchart] = (int)tile_chart.size();
    Q.ntile = (int)tile_chart.size();
    const int T = Q.ntile;
    const size_t N = (size_t)nrows * R;
    int64_t *d_ptr = nullptr, *d_r0 = nullptr, *d_r1 = nullptr;
    int *d_cols = nullptr, *d_tc = nullptr, *d_ct0 = nullptr;
    HIPCHK(Q.upload(&d_ptr, ptr, (size_t)nrows + 1));
    HIPCHK(Q.upload(&d_cols, cols, (size_t)Q.nnz));
    HIPCHK(Q.upload(&d_tc, tile_chart.data(), tile_chart.size()));
    HIPCHK(Q.upload(&d_r0, tr0.data(), tr0.size()));
    HIPCHK(Q.upload(&d_r1, tr1.data(), tr1.size()));
    HIPCHK(Q.upload(&d_ct0, Q.chart_t0.data(), Q.chart_t0.size()));
    HIPCHK(Q.upload(&Q.vals64, vals, (size_t)Q.nnz));
    std::vector<float> v32;
    if (mode == RUV_MODE_MIXED) {
        v32.assign(vals, vals + Q.nnz);
        HIPCHK(Q.upload(&Q.vals32, v32.data(), v32.size()));
    }
    std::vector<double> dinv;
    std::vector<float> dinv32;
    if (precond == RUV_PC_JACOBI) {
        dinv.assign(nrows, 0.0);
        for (int64_t i = 0; i < nrows; ++i) {
            double d = 0.0;
            for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k)
                if (cols[k] == i) d = vals[k];
            if (std::fabs(d) < 1e-300) return fail(RUV_ERR_PRECOND, "jacobi: zero diagonal at row " + std::to_string(i));
            dinv[i] = 1.0 / d;
        }
        HIPCHK(Q.upload(&Q.dinv64, dinv.data(), dinv.size()));
        if (mode == RUV_MODE_MIXED) {
            dinv32.assign(dinv.begin(), dinv.end());
            HIPCHK(Q.upload(&Q.dinv32, dinv32.data(), dinv32.size()));
        }
    }
    HIPCHK(Q.alloc(&Q.r64, N));
    if (mode == RUV_MODE_MIXED) {
        HIPCHK(Q.alloc(&Q.xf, N));
        HIPCHK(Q.alloc(&Q.rf, N));
        HIPCHK(Q.alloc(&Q.zf, N));
        HIPCHK(Q.alloc(&Q.pf, N));
        HIPCHK(Q.alloc(&Q.qf, N));
    } else {
        HIPCHK(Q.alloc(&Q.xd, N));
        HIPCHK(Q.alloc(&Q.rd, N));
        HIPCHK(Q.alloc(&Q.zd, N));
        HIPCHK(Q.alloc(&Q.pd, N));
        HIPCHK(Q.alloc(&Q.qd, N));
    }
    BDev& B = Q.B;
    B.ptr = d_ptr;
    B.cols = d_cols;
    B.tile_chart = d_tc;
    B.tile_r0 = d_r0;
    B.tile_r1 = d_r1;
    B.chart_t0 = d_ct0;
    B.nchart = nchart;
    HIPCHK(Q.alloc(&B.c_rz, nchart));
    HIPCHK(Q.alloc(&B.c_tol, nchart));
    HIPCHK(Q.alloc(&B.c_bnorm, nchart));
    HIPCHK(Q.alloc(&B.c_true, nchart));
    HIPCHK(Q.alloc(&B.c_done, nchart));
    HIPCHK(Q.alloc(&B.c_final, nchart));
    HIPCHK(Q.alloc(&B.c_iters, nchart));
    HIPCHK(Q.alloc(&B.part_a, T));
    HIPCHK(Q.alloc(&B.part_b, T));
    HIPCHK(Q.alloc(&B.part_c, T));
    HIPCHK(Q.alloc(&Q.act, T));
    HIPCHK(hipHostMalloc((void**)&Q.h_act, sizeof(int) * (size_t)T, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&Q.h_done, sizeof(int) * (size_t)std::max(nchart, 1), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&Q.h_final, sizeof(int) * (size_t)std::max(nchart, 1), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&Q.h_iters, sizeof(unsigned) * (size_t)std::max(nchart, 1), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&Q.h_true, sizeof(double) * (size_t)std::max(nchart, 1), hipHostMallocDefault));
    HIPCHK(hipStreamSynchronize(Q.stream));  // staging vectors above must outlive their copies
    return RUV_OK;
}

// upload the tiles of the charts whose flag[c] satisfies pred; returns the count
template <typename Pred>
int set_active(BatchImpl& Q, Pred pred, int* count) {
    int n = 0;
    for (int c = 0; c < Q.nchart; ++c)
        if (pred(c))
            for (int t = Q.chart_t0[c]; t < Q.chart_t0[c + 1]; ++t) Q.h_act[n++] = t;
    *count = n;
    if (n) HIPCHK(hipMemcpyAsync(Q.act, Q.h_act, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, Q.stream));
    return RUV_OK;
}

template <typename T, int R, int PC>
int inner_pcg(BatchImpl& Q, const T* vals, const T* dinv, T* x, T* r, T* z, T* p, T* q, double tol, uint64_t max_iter) {
    hipStream_t s = Q.stream;
    BDev& B = Q.B;
    B.max_iter = (unsigned)std::min<uint64_t>(max_iter, 0xffffffffu);
    int na = 0;
    int rc = set_active(Q, [&](int c) { return Q.h_final[c] == 0; }, &na);
    if (rc) return rc;
    if (na == 0) return RUV_OK;
    hipLaunchKernelGGL((kb_prep<T, R, PC>), dim3(na), dim3(BLK), 0, s, B, (const int*)Q.act, (const double*)Q.r64,
                       dinv, x, r, z, p);
    hipLaunchKernelGGL(kb_init, dim3((Q.nchart + BLK - 1) / BLK), dim3(BLK), 0, s, B, tol, Q.inner_tol,
                       Q.mode == RUV_MODE_MIXED ? 1 : 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(Q.h_done, B.c_done, sizeof(int) * Q.nchart, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int chunk = 8;
    const T* zz = (PC == 0) ? (const T*)r : (const T*)z;
    for (;;) {
        rc = set_active(Q, [&](int c) { return Q.h_done[c] == B_RUNNING; }, &na);
        if (rc) return rc;
        if (na == 0) break;
        for (int it = 0; it < chunk; ++it) {
            hipLaunchKernelGGL((kb_spmv<T, R>), dim3(na), dim3(BLK), 0, s, B, (const int*)Q.act, vals, (const T*)p, q);
            hipLaunchKernelGGL((kb_update<T, R, PC>), dim3(na), dim3(BLK), 0, s, B, (const int*)Q.act, dinv,
                               (const T*)p, (const T*)q, x, r, z);
            hipLaunchKernelGGL((kb_pupdate<T, R, PC>), dim3(na), dim3(BLK), 0, s, B, (const int*)Q.act, zz, p);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(Q.h_done, B.c_done, sizeof(int) * Q.nchart, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        chunk = std::min(chunk * 2, 64);
    }
    HIPCHK(hipMemcpyAsync(Q.h_iters, B.c_iters, sizeof(unsigned) * Q.nchart, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int c = 0; c < Q.nchart; ++c) {
        if (Q.h_final[c]) continue;
        Q.iters_total[c] += Q.h_iters[c];
        if (Q.h_done[c] == B_BREAKDOWN)
            return fail(RUV_ERR_BREAKDOWN, "batch: chart " + std::to_string(c) + " broke down");
        if (Q.h_done[c] == B_MAXITER)
            return fail(RUV_ERR_NOT_CONVERGED, "batch: chart " + std::to_string(c) + " hit max_iter");
    }
    return RUV_OK;
}

template <typename T, int R, int PC>
int batch_run(BatchImpl& Q, const double* d_b, double* d_x, uint64_t max_iter, double tol) {
    hipStream_t s = Q.stream;
    BDev& B = Q.B;
    const bool mixed = Q.mode == RUV_MODE_MIXED;
    const T* vals = mixed ? (const T*)(const void*)Q.vals32 : (const T*)(const void*)Q.vals64;
    const T* dinv = mixed ? (const T*)(const void*)Q.dinv32 : (const T*)(const void*)Q.dinv64;
    T* x = mixed ? (T*)(void*)Q.xf : (T*)(void*)Q.xd;
    T* r = mixed ? (T*)(void*)Q.rf : (T*)(void*)Q.rd;
    T* z = mixed ? (T*)(void*)Q.zf : (T*)(void*)Q.zd;
    T* p = mixed ? (T*)(void*)Q.pf : (T*)(void*)Q.pd;
    T* q = mixed ? (T*)(void*)Q.qf : (T*)(void*)Q.qd;
    Q.iters_total.assign(Q.nchart, 0);
    hipLaunchKernelGGL((kb_start<R>), dim3(Q.ntile), dim3(BLK), 0, s, B, d_b, d_x, Q.r64);
    hipLaunchKernelGGL(kb_bnorm, dim3((Q.nchart + BLK - 1) / BLK), dim3(BLK), 0, s, B);
    HIPCHK(hipGetLastError());
    std::fill(Q.h_final, Q.h_final + Q.nchart, 0);
    int na = 0;
    for (int round = 0;; ++round) {
        if (round >= Q.max_rounds) return fail(RUV_ERR_NOT_CONVERGED, "batch: refinement did not converge");
        int rc = inner_pcg<T, R, PC>(Q, vals, dinv, x, r, z, p, q, tol, max_iter);
        if (rc) return rc;
        rc = set_active(Q, [&](int c) { return Q.h_final[c] == 0; }, &na);
        if (rc) return rc;
        if (na == 0) break;
        // x64 += x_inner over the live charts, then the FP64 residual and the per-chart outer test
        hipLaunchKernelGGL((kb_axpy<T, R>), dim3(na), dim3(BLK), 0, s, B, (const int*)Q.act, (const T*)x, d_x);
        hipLaunchKernelGGL((kb_residual<R>), dim3(na), dim3(BLK), 0, s, B, (const int*)Q.act, (const double*)Q.vals64,
                           d_b, (const double*)d_x, Q.r64);
        hipLaunchKernelGGL(kb_outer, dim3((Q.nchart + BLK - 1) / BLK), dim3(BLK), 0, s, B, tol);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(Q.h_final, B.c_final, sizeof(int) * Q.nchart, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (!mixed) break;  // plain FP64: one PCG per chart, the true residual is only reported
        bool all = true;
        for (int c = 0; c < Q.nchart; ++c) all = all && Q.h_final[c];
        if (all) break;
    }
    HIPCHK(hipMemcpyAsync(Q.h_true, B.c_true, sizeof(double) * Q.nchart, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return RUV_OK;
}

}  // namespace

int batch_create(int64_t nrows, const int64_t* ptr, const int32_t* cols, const double* vals, int nchart,
                 const int64_t* crow, int R, int precond, int mode, BatchImpl** out) {
    if (R != 1 && R != 2) return fail(RUV_ERR_INVALID, "rhs_per_row must be 1 or 2");
    if (precond != RUV_PC_NONE && precond != RUV_PC_JACOBI)
        return fail(RUV_ERR_INVALID, "the batched solver supports the none and jacobi preconditioners");
    if (mode != RUV_MODE_FP64 && mode != RUV_MODE_MIXED) return fail(RUV_ERR_INVALID, "unknown mode");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(RUV_ERR_UNAVAILABLE, "no HIP device");
    std::unique_ptr<BatchImpl> Q(new BatchImpl());
    int rc = batch_setup(*Q, nrows, ptr, cols, vals, nchart, crow, R, precond, mode);
    if (rc) return rc;
    *out = Q.release();
    return RUV_OK;
}

void batch_destroy(BatchImpl* q) { delete q; }

int batch_solve_dev(BatchImpl* Q, const double* d_b, double* d_x, uint64_t max_iter, double tol,
                    ruv_solve_info* info, uint64_t* chart_iters, double* chart_resid) {
    HIPCHK(hipEventRecord(Q->ev0, Q->stream));
    int rc;
    const bool j = Q->precond == RUV_PC_JACOBI;
    if (Q->mode == RUV_MODE_MIXED) {
        if (Q->R == 2) rc = j ? batch_run<float, 2, 1>(*Q, d_b, d_x, max_iter, tol) : batch_run<float, 2, 0>(*Q, d_b, d_x, max_iter, tol);
        else rc = j ? batch_run<float, 1, 1>(*Q, d_b, d_x, max_iter, tol) : batch_run<float, 1, 0>(*Q, d_b, d_x, max_iter, tol);
    } else {
        if (Q->R == 2) rc = j ? batch_run<double, 2, 1>(*Q, d_b, d_x, max_iter, tol) : batch_run<double, 2, 0>(*Q, d_b, d_x, max_iter, tol);
        else rc = j ? batch_run<double, 1, 1>(*Q, d_b, d_x, max_iter, tol) : batch_run<double, 1, 0>(*Q, d_b, d_x, max_iter, tol);
    }
    if (rc) return rc;
    HIPCHK(hipEventRecord(Q->ev1, Q->stream));
    HIPCHK(hipEventSynchronize(Q->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, Q->ev0, Q->ev1));
    if (info) {
        uint64_t tot = 0, mx = 0;
        double worst = 0.0;
        for (int c = 0; c < Q->nchart; ++c) {
            tot += Q->iters_total[c];
            mx = std::max<uint64_t>(mx, Q->iters_total[c]);
            worst = std::max(worst, Q->h_true[c]);
        }
        info->iters = mx;
        info->refinements = 0;
        info->precond = Q->precond;
        info->true_rel_resid = worst;
        info->seconds = ms * 1e-3;
        (void)tot;
    }
    if (chart_iters)
        for (int c = 0; c < Q->nchart; ++c) chart_iters[c] = Q->iters_total[c];
    if (chart_resid)
        for (int c = 0; c < Q->nchart; ++c) chart_resid[c] = Q->h_true[c];
    return RUV_OK;
}

int batch_spmv_time(BatchImpl* Q, int reps, double* sec, double* bytes) {
    if (reps <= 0) return fail(RUV_ERR_INVALID, "reps must be positive");
    hipStream_t s = Q->stream;
    BDev B = Q->B;
    B.max_iter = 0xffffffffu;
    std::vector<int> zeros(Q->nchart, 0);
    HIPCHK(hipMemcpyAsync(B.c_done, zeros.data(), sizeof(int) * Q->nchart, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B.c_iters, zeros.data(), sizeof(int) * Q->nchart, hipMemcpyHostToDevice, s));
    int na = 0;
    int rc = set_active(*Q, [](int) { return true; }, &na);
    if (rc) return rc;
    const bool f32 = Q->mode == RUV_MODE_MIXED;
    const int64_t N = Q->nrows * Q->R;
    std::vector<double> ones_d;
    std::vector<float> ones_f;
    if (f32) {
        ones_f.assign(N, 1.0f);
        HIPCHK(hipMemcpyAsync(Q->pf, ones_f.data(), sizeof(float) * N, hipMemcpyHostToDevice, s));
    } else {
        ones_d.assign(N, 1.0);
        HIPCHK(hipMemcpyAsync(Q->pd, ones_d.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipMemsetAsync(B.part_c, 0, sizeof(double) * Q->ntile, s));
    HIPCHK(hipStreamSynchronize(s));
    auto launch = [&]() {
        if (f32) {
            if (Q->R == 2) hipLaunchKernelGGL((kb_spmv<float, 2>), dim3(na), dim3(BLK), 0, s, B, (const int*)Q->act, (const float*)Q->vals32, (const float*)Q->pf, Q->qf);
            else hipLaunchKernelGGL((kb_spmv<float, 1>), dim3(na), dim3(BLK), 0, s, B, (const int*)Q->act, (const float*)Q->vals32, (const float*)Q->pf, Q->qf);
        } else {
            if (Q->R == 2) hipLaunchKernelGGL((kb_spmv<double, 2>), dim3(na), dim3(BLK), 0, s, B, (const int*)Q->act, (const double*)Q->vals64, (const double*)Q->pd, Q->qd);
            else hipLaunchKernelGGL((kb_spmv<double, 1>), dim3(na), dim3(BLK), 0, s, B, (const int*)Q->act, (const double*)Q->vals64, (const double*)Q->pd, Q->qd);
        }
    };
    launch();
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(Q->ev0, s));
    for (int i = 0; i < reps; ++i) launch();
    HIPCHK(hipEventRecord(Q->ev1, s));
    HIPCHK(hipEventSynchronize(Q->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, Q->ev0, Q->ev1));
    *sec = ms * 1e-3 / reps;
    const double sv = f32 ? 4.0 : 8.0;
    *bytes = (double)Q->nnz * (sv + 4.0) + (double)(Q->nrows + 1) * 8.0 + 2.0 * (double)N * sv;
    return RUV_OK;
}

}  // namespace ruv
