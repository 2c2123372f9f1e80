// assemble.hip -- device assembly of the LSCM free block (cotan weights -> L in CSR + right-hand side).
//
// Restates edge_weights (crates/researchuv-unwrap/src/lscm.rs:252-274) and the assembly loop (:594-666) with
// Sparse::add / to_csr (sparse.rs:47-84) on the device.  The host supplies only integer topology (the
// undirected edges, each edge's opposite vertices in triangle order, the per-vertex neighbour lists in
// ascending order) and the pins; the floating-point work runs here, one thread per edge and one per free
// row.  Multiply-add contraction is disabled in this file, so each weight, diagonal and right-hand-side
// entry is computed with the reference's operation order and IEEE rounding (sqrt and division are correctly
// rounded in FP64), i.e. bitwise as the host assembly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "dev_common.hpp"
#include "host_lscm.hpp"
#include "ruv_internal.hpp"

#pragma clang fp contract(off)

namespace ruv {

#define HIPCHK(call)                                                                                         \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess) return fail(RUV_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

constexpr double EPS_DEGENERATE = 1.1754943508222875e-38;  // lscm.rs:36

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 ld3(const double* p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ D3 sub3(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross3(D3 a, D3 o) {
    return {a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x};
}

// per undirected edge (i < j): w = sum of the clamped opposite-angle cotangents in triangle order, and
// c = w / max(|p_j - p_i|^2, EPS) when w > EPS (else 0: the edge is not assembled)
__global__ __launch_bounds__(BLK) void k_edge_c(int64_t ne, const int32_t* __restrict__ ei,
                                                const int32_t* __restrict__ ej, const int64_t* __restrict__ ooff,
                                                const int32_t* __restrict__ opp, const double* __restrict__ p3,
                                                double* __restrict__ c_out) {
    for (int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x; e < ne; e += (int64_t)gridDim.x * BLK) {
        const int32_t i = ei[e], j = ej[e];
        const D3 pi = ld3(p3, i), pj = ld3(p3, j);
        double w = 0.0;
        for (int64_t k = ooff[e]; k < ooff[e + 1]; ++k) {
            const D3 po = ld3(p3, opp[k]);
            const D3 a1 = sub3(pj, po), a2 = sub3(pi, po);
            const double la = sqrt(dot3(a1, a1)), lb = sqrt(dot3(a2, a2));
            if (la <= EPS_DEGENERATE || lb <= EPS_DEGENERATE) continue;
            const D3 cr = cross3(a1, a2);
            const double s = sqrt(dot3(cr, cr)) / (la * lb);
            if (s <= EPS_DEGENERATE) continue;
            double cot = (dot3(a1, a2) / (la * lb)) / s;
            cot = cot > 0.0 ? cot : 0.0;
            w += cot;
        }
        double c = 0.0;
        if (w > EPS_DEGENERATE) {
            const D3 d = sub3(pj, pi);
            const double dd = dot3(d, d);
            c = w / (dd > EPS_DEGENERATE ? dd : EPS_DEGENERATE);
        }
        c_out[e] = c;
    }
}

// entries of free row p: its weighted free neighbours plus the diagonal (when any weighted edge exists)
__global__ __launch_bounds__(BLK) void k_row_count(int32_t nfree, const int32_t* __restrict__ free_ids,
                                                   const int64_t* __restrict__ aoff, const int32_t* __restrict__ anb,
                                                   const int64_t* __restrict__ aedge, const double* __restrict__ c,
                                                   const int32_t* __restrict__ free_pos, int32_t* __restrict__ cnt) {
    for (int32_t p = blockIdx.x * BLK + threadIdx.x; p < nfree; p += gridDim.x * BLK) {
        const int32_t i = free_ids[p];
        int32_t n = 0, any = 0;
        for (int64_t k = aoff[i]; k < aoff[i + 1]; ++k) {
            if (c[aedge[k]] > 0.0) {
                any = 1;
                if (free_pos[anb[k]] >= 0) ++n;
            }
        }
        cnt[p] = n + any;
    }
}

// row p of L (ascending columns, diagonal at its sorted position) and the interleaved right-hand side:
// diagonal = sum of c over weighted edges in ascending neighbour order (Sparse::add accumulation order),
// rhs += B c dv (+ c pin for a pinned neighbour), lscm.rs:605-640
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
            if (!(ce > 0.0)) continue;
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

}  // namespace

// Assemble chart `ch` of `pos` on the device.  Host: tangent frame, pins, free map and topology (integer
// work, build_system's first half); device: every floating-point entry of L and the rhs.  Outputs are
// copied back to host buffers (L and rhs of the same shapes as build_system's).
int device_assemble(const double* pos, const Chart& ch, double angle_mix, ChartSystem* S) {
    // pins, frame and free map from the host builder, without its weights/assembly pass
    build_system_pins(pos, ch, S);
    const int64_t nv = S->nv;
    const int32_t nfree = S->nfree;
    // topology: undirected edges ascending, opposite vertices in triangle order
    const size_t ntri = ch.tris.size() / 3;
    std::vector<std::pair<uint64_t, std::pair<uint32_t, int32_t>>> con;  // key, (order, opposite)
    con.reserve(3 * ntri);
    for (size_t f = 0; f < ntri; ++f)
        for (int q = 0; q < 3; ++q) {
            const uint32_t e0 = ch.tris[3 * f + q], e1 = ch.tris[3 * f + (q + 1) % 3], o = ch.tris[3 * f + (q + 2) % 3];
            const uint32_t i = std::min(e0, e1), j = std::max(e0, e1);
            con.push_back({((uint64_t)i << 32) | j, {(uint32_t)(3 * f + q), (int32_t)o}});
        }
    std::sort(con.begin(), con.end());
    std::vector<int32_t> ei, ej, opp;
    std::vector<int64_t> ooff{0};
    for (size_t q = 0; q < con.size(); ++q) {
        if (q == 0 || con[q].first != con[q - 1].first) {
            if (q) ooff.push_back((int64_t)opp.size());
            ei.push_back((int32_t)(con[q].first >> 32));
            ej.push_back((int32_t)(con[q].first & 0xffffffffu));
        }
        opp.push_back(con[q].second.second);
    }
    ooff.push_back((int64_t)opp.size());
    const int64_t ne = (int64_t)ei.size();
    // per-vertex ascending neighbour lists with edge ids
    std::vector<int64_t> aoff(nv + 1, 0);
    for (int64_t e = 0; e < ne; ++e) {
        aoff[ei[e] + 1]++;
        aoff[ej[e] + 1]++;
    }
    for (int64_t i = 0; i < nv; ++i) aoff[i + 1] += aoff[i];
    std::vector<int32_t> anb(aoff[nv]);
    std::vector<int64_t> aedge(aoff[nv]);
    {
        std::vector<int64_t> cur(aoff.begin(), aoff.end() - 1);
        for (int64_t e = 0; e < ne; ++e) {  // edges ascending -> lists ascending
            anb[cur[ei[e]]] = ej[e];
            aedge[cur[ei[e]]++] = e;
            anb[cur[ej[e]]] = ei[e];
            aedge[cur[ej[e]]++] = e;
        }
    }
    std::vector<double> pin_at(2 * nv, 0.0);
    for (size_t q = 0; q < S->pin_ids.size(); ++q) {
        pin_at[2 * S->pin_ids[q]] = S->pin_uv[2 * q];
        pin_at[2 * S->pin_ids[q] + 1] = S->pin_uv[2 * q + 1];
    }
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct Guard {
        hipStream_t s;
        std::vector<void*> m;
        ~Guard() {
            (void)hipStreamSynchronize(s);
            for (void* p : m) (void)hipFree(p);
            (void)hipStreamDestroy(s);
        }
    } g{s, {}};
    auto up = [&](auto** dst, const auto& v) -> hipError_t {
        using E = std::remove_reference_t<decltype(v[0])>;
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(v.size(), 1) * sizeof(E));
        if (e != hipSuccess) return e;
        g.m.push_back(q);
        *dst = (std::remove_pointer_t<std::remove_reference_t<decltype(dst)>>)q;
        return v.empty() ? hipSuccess : hipMemcpyAsync(q, v.data(), v.size() * sizeof(E), hipMemcpyHostToDevice, s);
    };
    int32_t *d_ei, *d_ej, *d_opp, *d_anb, *d_free_ids, *d_free_pos, *d_cnt, *d_rp, *d_cols;
    int64_t *d_ooff, *d_aoff, *d_aedge;
    double *d_p3, *d_p2, *d_pin, *d_c, *d_vals, *d_rhs;
    HIPCHK(up(&d_ei, ei));
    HIPCHK(up(&d_ej, ej));
    HIPCHK(up(&d_opp, opp));
    HIPCHK(up(&d_ooff, ooff));
    HIPCHK(up(&d_anb, anb));
    HIPCHK(up(&d_aoff, aoff));
    HIPCHK(up(&d_aedge, aedge));
    HIPCHK(up(&d_free_ids, S->free_ids));
    HIPCHK(up(&d_free_pos, S->free_pos));
    HIPCHK(up(&d_p3, S->p3));
    HIPCHK(up(&d_p2, S->p2));
    HIPCHK(up(&d_pin, pin_at));
    std::vector<double> zc(ne, 0.0);
    std::vector<int32_t> zcnt(nfree, 0);
    HIPCHK(up(&d_c, zc));
    HIPCHK(up(&d_cnt, zcnt));
    const int ge = (int)std::max<int64_t>(1, std::min<int64_t>((ne + BLK - 1) / BLK, 4096));
    const int gr = (int)std::max<int64_t>(1, std::min<int64_t>((nfree + BLK - 1) / BLK, 4096));
    hipLaunchKernelGGL(k_edge_c, dim3(ge), dim3(BLK), 0, s, ne, (const int32_t*)d_ei, (const int32_t*)d_ej,
                       (const int64_t*)d_ooff, (const int32_t*)d_opp, (const double*)d_p3, d_c);
    if (nfree > 0)
        hipLaunchKernelGGL(k_row_count, dim3(gr), dim3(BLK), 0, s, nfree, (const int32_t*)d_free_ids,
                           (const int64_t*)d_aoff, (const int32_t*)d_anb, (const int64_t*)d_aedge, (const double*)d_c,
                           (const int32_t*)d_free_pos, d_cnt);
    HIPCHK(hipGetLastError());
    // row pointers (host scan of the counts)
    std::vector<int32_t> cnt(nfree);
    if (nfree) HIPCHK(hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int32_t) * nfree, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<int32_t> rp(nfree + 1, 0);
    for (int32_t p = 0; p < nfree; ++p) rp[p + 1] = rp[p] + cnt[p];
    const int64_t nnz = rp[nfree];
    HIPCHK(up(&d_rp, rp));
    std::vector<int32_t> zcols(nnz, 0);
    std::vector<double> zvals(nnz, 0.0), zrhs(2 * (size_t)nfree, 0.0);
    HIPCHK(up(&d_cols, zcols));
    HIPCHK(up(&d_vals, zvals));
    HIPCHK(up(&d_rhs, zrhs));
    if (nfree > 0)
        hipLaunchKernelGGL(k_row_fill, dim3(gr), dim3(BLK), 0, s, nfree, (const int32_t*)d_free_ids,
                           (const int64_t*)d_aoff, (const int32_t*)d_anb, (const int64_t*)d_aedge, (const double*)d_c,
                           (const int32_t*)d_free_pos, (const double*)d_p2, (const double*)d_pin, 1.0 - angle_mix,
                           (const int32_t*)d_rp, d_cols, d_vals, d_rhs);
    HIPCHK(hipGetLastError());
    S->l.n = nfree;
    S->l.ptr = rp;
    S->l.cols.assign(nnz, 0);
    S->l.vals.assign(nnz, 0.0);
    S->rhs.assign(2 * (size_t)nfree, 0.0);
    if (nnz) {
        HIPCHK(hipMemcpyAsync(S->l.cols.data(), d_cols, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(S->l.vals.data(), d_vals, sizeof(double) * nnz, hipMemcpyDeviceToHost, s));
    }
    if (nfree) HIPCHK(hipMemcpyAsync(S->rhs.data(), d_rhs, sizeof(double) * 2 * nfree, hipMemcpyDeviceToHost, s));
    std::vector<double> cc(ne);
    if (ne) HIPCHK(hipMemcpyAsync(cc.data(), d_c, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    S->edges = (size_t)std::count_if(cc.begin(), cc.end(), [](double v) { return v > 0.0; });
    return RUV_OK;
}

}  // namespace ruv
