// assemble_kernels.hpp -- device kernels of the LSCM assembly (cotan weights, free-block CSR, rhs) and of
// its adjoint (the derivative of every edge weight with respect to the vertex positions).
//
// Included by assemble.hip and adjoint.hip, each compiled with multiply-add contraction disabled so the
// forward weights repeat the reference's operation order and IEEE rounding (lscm.rs:252-274, 594-666).
#pragma once

#include <hip/hip_runtime.h>

#include "dev_common.hpp"

namespace ruv {
namespace {

constexpr double EPS_DEGENERATE = 1.1754943508222875e-38;  // lscm.rs:36

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 ld3(const double* p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ D3 sub3(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 add3(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 mul3(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross3(D3 a, D3 o) {
    return {a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x};
}

// Clamped cotangent of the angle at po between the edges to pi and pj, exactly as edge_weights computes it
// (a1 = p[e1] - p[o], a2 = p[e0] - p[o]; the value is symmetric in e0/e1).  valid = 0 when a guard skips it.
__device__ __forceinline__ double corner_cot(D3 pi, D3 pj, D3 po, int* valid) {
    const D3 a1 = sub3(pj, po), a2 = sub3(pi, po);
    const double la = sqrt(dot3(a1, a1)), lb = sqrt(dot3(a2, a2));
    *valid = 0;
    if (la <= EPS_DEGENERATE || lb <= EPS_DEGENERATE) return 0.0;
    const D3 cr = cross3(a1, a2);
    const double s = sqrt(dot3(cr, cr)) / (la * lb);
    if (s <= EPS_DEGENERATE) return 0.0;
    *valid = 1;
    const double cot = (dot3(a1, a2) / (la * lb)) / s;
    return cot > 0.0 ? cot : 0.0;
}

// per undirected edge: w (sum of clamped cotangents in triangle order), D = |p_j - p_i|^2 and
// c = w / max(D, EPS) when w > EPS, else 0 (not assembled)
__global__ __launch_bounds__(BLK) void k_edge_c(int64_t ne, const int32_t* __restrict__ ei,
                                                const int32_t* __restrict__ ej, const int64_t* __restrict__ ooff,
                                                const int32_t* __restrict__ opp, const double* __restrict__ p3,
                                                double* __restrict__ c_out, double* __restrict__ w_out,
                                                double* __restrict__ d_out) {
    for (int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x; e < ne; e += (int64_t)gridDim.x * BLK) {
        const int32_t i = ei[e], j = ej[e];
        const D3 pi = ld3(p3, i), pj = ld3(p3, j);
        double w = 0.0;
        for (int64_t k = ooff[e]; k < ooff[e + 1]; ++k) {
            int valid;
            const double cot = corner_cot(pi, pj, ld3(p3, opp[k]), &valid);
            if (valid) w += cot;
        }
        const D3 d = sub3(pj, pi);
        const double dd = dot3(d, d);
        double c = 0.0;
        if (w > EPS_DEGENERATE) c = w / (dd > EPS_DEGENERATE ? dd : EPS_DEGENERATE);
        c_out[e] = c;
        if (w_out) w_out[e] = w;
        if (d_out) d_out[e] = dd;
    }
}

// entries of free row p: weighted free neighbours plus the diagonal (when any weighted edge exists)
__global__ __launch_bounds__(BLK) void k_row_count(int32_t nfree, const int32_t* __restrict__ free_ids,
                                                   const int64_t* __restrict__ aoff, const int32_t* __restrict__ anb,
                                                   const int64_t* __restrict__ aedge, const double* __restrict__ c,
                                                   const int32_t* __restrict__ free_pos, int32_t* __restrict__ cnt) {
    for (int32_t p = blockIdx.x * BLK + threadIdx.x; p < nfree; p += gridDim.x * BLK) {
        const int32_t i = free_ids[p];
        int32_t n = 0, any = 0;
        for (int64_t k = aoff[i]; k < aoff[i + 1]; ++k) {
            if (c[aedge[k]] != 0.0) {
                any = 1;
                if (free_pos[anb[k]] >= 0) ++n;
            }
        }
        cnt[p] = n + any;
    }
}

// row p of L (ascending columns, diagonal at its sorted position) and the interleaved rhs: diagonal = sum of
// c over weighted edges in ascending neighbour order (the Sparse::add accumulation order), rhs += B c dv
// (+ c pin for a pinned neighbour), lscm.rs:605-640
__global__ __launch_bounds__(BLK) void k_row_fill(int32_t nfree, const int32_t* __restrict__ free_ids,
                                                  const int64_t* __restrict__ aoff, const int32_t* __restrict__ anb,
                                                  const int64_t* __restrict__ aedge, const double* __restrict__ c,
                                                  const int32_t* __restrict__ free_pos, const double* __restrict__ p2,
                                                  const double* __restrict__ pin_at, double mix_b,
                                                  const int32_t* __restrict__ rp, int32_t* __restrict__ cols,
                                                  double* __restrict__ vals, double* __restrict__ rhs) {
    for (int32_t p = blockIdx.x * BLK + threadIdx.x; p < nfree; p += gridDim.x * BLK) {
        const int32_t i = free_ids[p];
        int32_t w = rp[p], dslot = -1;
        double diag = 0.0, ru = 0.0, rv = 0.0;
        for (int64_t k = aoff[i]; k < aoff[i + 1]; ++k) {
            const double ce = c[aedge[k]];
            if (ce == 0.0) continue;
            const int32_t j = anb[k];
            const double du = p2[2 * (int64_t)j] - p2[2 * (int64_t)i];
            const double dv = p2[2 * (int64_t)j + 1] - p2[2 * (int64_t)i + 1];
            diag += ce;
            const int32_t fj = free_pos[j];
            if (fj >= 0) {
                if (dslot < 0 && j > i) dslot = w++;
                cols[w] = fj;
                vals[w] = -ce;
                ++w;
                ru += mix_b * ce * du;
                rv += mix_b * ce * dv;
            } else {
                ru += mix_b * ce * du + ce * pin_at[2 * (int64_t)j];
                rv += mix_b * ce * dv + ce * pin_at[2 * (int64_t)j + 1];
            }
        }
        if (diag != 0.0) {
            if (dslot < 0) dslot = w++;
            cols[dslot] = p;
            vals[dslot] = diag;
        }
        rhs[2 * (int64_t)p] = ru;
        rhs[2 * (int64_t)p + 1] = rv;
    }
}

// ------------------------------------------------------------------------------------------------ adjoint
// dJ/dc_e = a_E |uv_i - uv_j|^2 - (lam_i - lam_j).(uv_i - uv_j)  (lam = 0 on pins), then split through
// c = w / D into gw_e = dJ/dw_e and gd_e = dJ/dD_e (zero for an edge that is not assembled)
__global__ __launch_bounds__(BLK) void k_edge_adj(int64_t ne, const int32_t* __restrict__ ei,
                                                  const int32_t* __restrict__ ej, const double* __restrict__ c,
                                                  const double* __restrict__ w, const double* __restrict__ dd,
                                                  const double* __restrict__ uv, const double* __restrict__ lam,
                                                  double a_energy, double* __restrict__ gw, double* __restrict__ gd) {
    for (int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x; e < ne; e += (int64_t)gridDim.x * BLK) {
        const int32_t i = ei[e], j = ej[e];
        double g_w = 0.0, g_d = 0.0;
        if (c[e] != 0.0) {
            const double du = uv[2 * (int64_t)i] - uv[2 * (int64_t)j];
            const double dv = uv[2 * (int64_t)i + 1] - uv[2 * (int64_t)j + 1];
            const double lu = lam[2 * (int64_t)i] - lam[2 * (int64_t)j];
            const double lv = lam[2 * (int64_t)i + 1] - lam[2 * (int64_t)j + 1];
            const double gc = a_energy * (du * du + dv * dv) - (lu * du + lv * dv);
            const double D = dd[e] > EPS_DEGENERATE ? dd[e] : EPS_DEGENERATE;
            g_w = gc / D;
            if (dd[e] > EPS_DEGENERATE) g_d = -gc * w[e] / (D * D);
        }
        gw[e] = g_w;
        gd[e] = g_d;
    }
}

// d cot(angle at o between a = p_i - p_o and b = p_j - p_o) / d p_v for v in {i, j, o}:
// cot = (a.b)/|a x b|; d/da = b/s - (a.b)(b x n)/s^2, d/db = a/s - (a.b)(n x a)/s^2, d/dp_o = -(d/da + d/db),
// n the unit normal a x b / s.  Zero where the reference clamps (cot <= 0) or skips (degenerate corner).
__device__ __forceinline__ D3 dcot_dp(D3 pi, D3 pj, D3 po, int which) {
    int valid;
    const double cot = corner_cot(pi, pj, po, &valid);
    if (!valid || !(cot > 0.0)) return {0.0, 0.0, 0.0};
    const D3 a = sub3(pi, po), b = sub3(pj, po);
    const D3 cr = cross3(a, b);
    const double s = sqrt(dot3(cr, cr));
    const D3 n = mul3(cr, 1.0 / s);
    const double ab = dot3(a, b);
    const D3 da = sub3(mul3(b, 1.0 / s), mul3(cross3(b, n), ab / (s * s)));
    const D3 db = sub3(mul3(a, 1.0 / s), mul3(cross3(n, a), ab / (s * s)));
    if (which == 0) return da;
    if (which == 1) return db;
    return mul3(add3(da, db), -1.0);
}

// grad_v = sum_{edges e at v} gd_e dD_e/dp_v + sum_{triangles at v} sum_{corners} gw_{edge} dcot/dp_v,
// gathered per vertex in ascending edge / triangle order (deterministic, no atomics)
__global__ __launch_bounds__(BLK) void k_vertex_grad(int64_t nv, const double* __restrict__ p3,
                                                     const uint32_t* __restrict__ tris,
                                                     const int64_t* __restrict__ tri_edge,
                                                     const int64_t* __restrict__ aoff, const int32_t* __restrict__ anb,
                                                     const int64_t* __restrict__ aedge,
                                                     const int64_t* __restrict__ toff, const int32_t* __restrict__ tlist,
                                                     const double* __restrict__ gw, const double* __restrict__ gd,
                                                     double* __restrict__ grad) {
    for (int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x; v < nv; v += (int64_t)gridDim.x * BLK) {
        const D3 pv = ld3(p3, v);
        D3 g{0.0, 0.0, 0.0};
        for (int64_t k = aoff[v]; k < aoff[v + 1]; ++k) {
            const double s = gd[aedge[k]];
            if (s != 0.0) g = add3(g, mul3(sub3(pv, ld3(p3, anb[k])), 2.0 * s));  // D = |p_v - p_u|^2
        }
        int32_t last = -1;
        for (int64_t k = toff[v]; k < toff[v + 1]; ++k) {
            const int32_t f = tlist[k];
            if (f == last) continue;
            last = f;
            for (int q = 0; q < 3; ++q) {
                const double s = gw[tri_edge[3 * (int64_t)f + q]];
                if (s == 0.0) continue;
                const uint32_t e0 = tris[3 * (int64_t)f + q], e1 = tris[3 * (int64_t)f + (q + 1) % 3];
                const uint32_t o = tris[3 * (int64_t)f + (q + 2) % 3];
                const int which = ((int64_t)e0 == v) ? 0 : ((int64_t)e1 == v) ? 1 : ((int64_t)o == v) ? 2 : -1;
                if (which < 0) continue;
                g = add3(g, mul3(dcot_dp(ld3(p3, e0), ld3(p3, e1), ld3(p3, o), which), s));
            }
        }
        grad[3 * v] = g.x;
        grad[3 * v + 1] = g.y;
        grad[3 * v + 2] = g.z;
    }
}

}  // namespace
}  // namespace ruv
