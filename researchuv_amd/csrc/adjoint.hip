// adjoint.hip -- forward LSCM solve from positions and its vector-Jacobian product (the differentiable form).
//
// Forward: positions p (chart-local, nv x 3) -> cotan weights c_e(p) -> L(p) x = B(p) -> uv (free vertices
// from x, pins from their frozen targets), plus the solver's own quadratic energy
// E = sum_e c_e |uv_i - uv_j|^2 over the assembled edges.  Pins are treated as constants of p (the border
// pass, lscm.rs:280-413, moves them with arc lengths and areas through non-smooth branches).  The default
// path is assumed (angle_mix = 1, so B holds pin terms only, lscm.rs:596-597).
//
// Reverse: given an upstream gradient g on uv, L lam = g_free is one more PCG solve with the same SPD matrix
// (L^T = L, no transpose system is built); then dJ/dc_e = a_E |uv_i - uv_j|^2 - (lam_i - lam_j).(uv_i - uv_j)
// and, through c_e = w_e / |p_j - p_i|^2 and w_e = sum of clamped corner cotangents, dJ/dp is gathered per
// vertex on the device (assemble_kernels.hpp).  With a_E = 1 and g = dJ/duv of an extra term, this is the
// gradient of E + that term, since dE/duv vanishes at the solution (it is 2 (L x - B)).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#pragma clang fp contract(off)

#include "assemble_kernels.hpp"
#include "chart_topo.hpp"
#include "dev_assembly.hpp"
#include "ruv_internal.hpp"

namespace ruv {

#define HIPCHK(call)                                                                                         \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess) return fail(RUV_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

int prepare(const double* p3, int64_t nv, const uint32_t* tris, int64_t ntri, const int32_t* pin_ids,
            const double* pin_uv, int64_t npins, ChartTopo* T, DevAssembly* A, std::vector<double>* pin_at) {
    std::vector<char> is_pin(nv, 0);
    pin_at->assign(2 * nv, 0.0);
    for (int64_t q = 0; q < npins; ++q) {
        const int32_t v = pin_ids[q];
        if (v < 0 || v >= nv) return fail(RUV_ERR_BAD_DIMS, "pin id out of range");
        is_pin[v] = 1;
        (*pin_at)[2 * v] = pin_uv[2 * q];
        (*pin_at)[2 * v + 1] = pin_uv[2 * q + 1];
    }
    for (int64_t k = 0; k < 3 * ntri; ++k)
        if ((int64_t)tris[k] >= nv) return fail(RUV_ERR_BAD_DIMS, "triangle index out of range");
    build_topo(nv, tris, ntri, is_pin, T);
    std::vector<double> P(p3, p3 + 3 * nv), p2(2 * nv, 0.0);
    std::vector<uint32_t> tv(tris, tris + 3 * ntri);
    return dev_assemble(*T, P, p2, *pin_at, tv, 0.0, A);
}

}  // namespace

int lscm_forward(const double* p3, int64_t nv, const uint32_t* tris, int64_t ntri, const int32_t* pin_ids,
                 const double* pin_uv, int64_t npins, int precond, int mode, double tol, double* uv_out,
                 double* energy_out, ruv_solve_info* info) {
    ChartTopo T;
    DevAssembly A;
    std::vector<double> pin_at;
    int rc = prepare(p3, nv, tris, ntri, pin_ids, pin_uv, npins, &T, &A, &pin_at);
    if (rc) return rc;
    Csr L;
    std::vector<double> rhs;
    rc = dev_download(A, &L, &rhs, nullptr);
    if (rc) return rc;
    std::vector<double> x(2 * (size_t)A.nfree, 0.0);
    if (A.nfree > 0) {
        rc = device_solve(L, 2, rhs.data(), 16384, tol, precond, mode, x.data(), info);
        if (rc) return rc;
    }
    for (int64_t v = 0; v < nv; ++v) {
        const int32_t f = T.free_pos[v];
        uv_out[2 * v] = f >= 0 ? x[2 * (size_t)f] : pin_at[2 * v];
        uv_out[2 * v + 1] = f >= 0 ? x[2 * (size_t)f + 1] : pin_at[2 * v + 1];
    }
    if (energy_out) {
        std::vector<double> c(A.ne);
        if (A.ne) HIPCHK(hipMemcpyAsync(c.data(), A.c, sizeof(double) * A.ne, hipMemcpyDeviceToHost, A.stream));
        HIPCHK(hipStreamSynchronize(A.stream));
        double e = 0.0;
        for (int64_t k = 0; k < A.ne; ++k) {
            if (c[k] == 0.0) continue;
            const double du = uv_out[2 * (size_t)T.ei[k]] - uv_out[2 * (size_t)T.ej[k]];
            const double dv = uv_out[2 * (size_t)T.ei[k] + 1] - uv_out[2 * (size_t)T.ej[k] + 1];
            e += c[k] * (du * du + dv * dv);
        }
        *energy_out = e;
    }
    return RUV_OK;
}

int lscm_vjp(const double* p3, int64_t nv, const uint32_t* tris, int64_t ntri, const int32_t* pin_ids,
             const double* pin_uv, int64_t npins, const double* uv, const double* g_uv, double a_energy, int precond,
             int mode, double tol, double* grad_p3, ruv_solve_info* info) {
    ChartTopo T;
    DevAssembly A;
    std::vector<double> pin_at;
    int rc = prepare(p3, nv, tris, ntri, pin_ids, pin_uv, npins, &T, &A, &pin_at);
    if (rc) return rc;
    Csr L;
    std::vector<double> rhs;
    rc = dev_download(A, &L, &rhs, nullptr);
    if (rc) return rc;
    // adjoint solve with the same SPD L: L lam = g on the free vertices
    std::vector<double> gf(2 * (size_t)A.nfree), lf(2 * (size_t)A.nfree, 0.0);
    bool any = false;
    for (int32_t p = 0; p < A.nfree; ++p) {
        gf[2 * (size_t)p] = g_uv[2 * (size_t)T.free_ids[p]];
        gf[2 * (size_t)p + 1] = g_uv[2 * (size_t)T.free_ids[p] + 1];
        any = any || gf[2 * (size_t)p] != 0.0 || gf[2 * (size_t)p + 1] != 0.0;
    }
    if (A.nfree > 0 && any) {
        rc = device_solve(L, 2, gf.data(), 16384, tol, precond, mode, lf.data(), info);
        if (rc) return rc;
    }
    std::vector<double> lam(2 * nv, 0.0);
    for (int32_t p = 0; p < A.nfree; ++p) {
        lam[2 * (size_t)T.free_ids[p]] = lf[2 * (size_t)p];
        lam[2 * (size_t)T.free_ids[p] + 1] = lf[2 * (size_t)p + 1];
    }
    double *d_uv = nullptr, *d_lam = nullptr, *d_gw = nullptr, *d_gd = nullptr, *d_grad = nullptr;
    struct Bufs {
        std::vector<void*> m;
        ~Bufs() {
            for (void* p : m) (void)hipFree(p);
        }
    } bufs;
    auto alloc = [&](double** p, size_t n) -> hipError_t {
        hipError_t e = hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(double));
        if (e == hipSuccess) bufs.m.push_back(*p);
        return e;
    };
    HIPCHK(alloc(&d_uv, 2 * nv));
    HIPCHK(alloc(&d_lam, 2 * nv));
    HIPCHK(alloc(&d_gw, A.ne));
    HIPCHK(alloc(&d_gd, A.ne));
    HIPCHK(alloc(&d_grad, 3 * nv));
    HIPCHK(hipMemcpyAsync(d_uv, uv, sizeof(double) * 2 * nv, hipMemcpyHostToDevice, A.stream));
    HIPCHK(hipMemcpyAsync(d_lam, lam.data(), sizeof(double) * 2 * nv, hipMemcpyHostToDevice, A.stream));
    const int ge = (int)std::max<int64_t>(1, std::min<int64_t>((A.ne + BLK - 1) / BLK, 4096));
    const int gv = (int)std::max<int64_t>(1, std::min<int64_t>((nv + BLK - 1) / BLK, 4096));
    hipLaunchKernelGGL(k_edge_adj, dim3(ge), dim3(BLK), 0, A.stream, A.ne, (const int32_t*)A.ei, (const int32_t*)A.ej,
                       (const double*)A.c, (const double*)A.w, (const double*)A.dd, (const double*)d_uv,
                       (const double*)d_lam, a_energy, d_gw, d_gd);
    hipLaunchKernelGGL(k_vertex_grad, dim3(gv), dim3(BLK), 0, A.stream, nv, (const double*)A.p3,
                       (const uint32_t*)A.tris, (const int64_t*)A.tri_edge, (const int64_t*)A.aoff,
                       (const int32_t*)A.anb, (const int64_t*)A.aedge, (const int64_t*)A.toff,
                       (const int32_t*)A.tlist, (const double*)d_gw, (const double*)d_gd, d_grad);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(grad_p3, d_grad, sizeof(double) * 3 * nv, hipMemcpyDeviceToHost, A.stream));
    HIPCHK(hipStreamSynchronize(A.stream));
    return RUV_OK;
}

}  // namespace ruv
