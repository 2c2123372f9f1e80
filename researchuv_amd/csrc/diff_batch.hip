// diff_batch.hip -- the differentiable batch kept on the device: forward LSCM solves of many charts from their
// positions and the vector-Jacobian product, with the topology built once.
//
// The charts form one disjoint union (chart c owns vertices [voff[c], voff[c+1])), so one assembly, one batched
// solve and one adjoint gather serve them all, and each chart still runs its own PCG in the batched engine.
// The operator's structure is fixed at create time: every edge with a free end is an entry, including edges
// whose clamped cotangent weight is zero (they hold an explicit 0), so moving the positions only changes
// values.  Per forward call: positions -> cotan weights -> L and the pin right-hand side (the device assembly of
// lscm.rs:594-666) -> the batch's value arrays (and its AMG hierarchy) -> batched solve -> uv and the
// solver's quadratic energy.  Per VJP call (on the state of the last forward): L lam = g on the free rows with
// the same SPD operator, then dJ/dp through the assembly (adjoint.hip's method, assemble_kernels.hpp).
// Pins are constants of the positions, as in adjoint.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#pragma clang fp contract(off)

#include "assemble_kernels.hpp"
#include "batch_impl.hpp"
#include "chart_topo.hpp"
#include "ruv_internal.hpp"

namespace ruv {

int batch_create(int64_t nrows, const int64_t* ptr, const int32_t* cols, const double* vals, int nchart,
                 const int64_t* crow, int R, int precond, int mode, BatchImpl** out);
void batch_destroy(BatchImpl* q);
int batch_solve_dev(BatchImpl* Q, const double* d_b, double* d_x, uint64_t max_iter, double tol,
                    ruv_solve_info* info, uint64_t* chart_iters, double* chart_resid);
int batch_value_map(const BatchImpl& Q, const int64_t* ptr, const int32_t* cols, std::vector<int64_t>* map);
int batch_update_values(BatchImpl* Q, const double* d_vals, const int64_t* d_map, const int64_t* d_diagpos);

namespace {

// entries of free row p: every free neighbour plus the diagonal
__global__ __launch_bounds__(BLK) void kd_row_count(int32_t nfree, const int32_t* __restrict__ free_ids,
                                                    const int64_t* __restrict__ aoff, const int32_t* __restrict__ anb,
                                                    const int32_t* __restrict__ free_pos, int64_t* __restrict__ cnt) {
    const int32_t p = blockIdx.x * BLK + threadIdx.x;
    if (p >= nfree) return;
    const int32_t i = free_ids[p];
    int64_t n = 1;
    for (int64_t k = aoff[i]; k < aoff[i + 1]; ++k)
        if (free_pos[anb[k]] >= 0) ++n;
    cnt[p] = n;
}

// row p (ascending columns, the diagonal at its sorted position): -c per free neighbour, the diagonal = sum of
// c over all edges of the row in ascending neighbour order, the pin right-hand side (lscm.rs:605-640 with
// B = 0); structure written only when cols is given
__global__ __launch_bounds__(BLK) void kd_row_fill(int32_t nfree, const int32_t* __restrict__ free_ids,
                                                   const int64_t* __restrict__ aoff, const int32_t* __restrict__ anb,
                                                   const int64_t* __restrict__ aedge, const double* __restrict__ c,
                                                   const int32_t* __restrict__ free_pos,
                                                   const double* __restrict__ pin_at, const int64_t* __restrict__ rp,
                                                   int32_t* __restrict__ cols, double* __restrict__ vals,
                                                   double* __restrict__ rhs, int64_t* __restrict__ diagpos) {
    const int32_t p = blockIdx.x * BLK + threadIdx.x;
    if (p >= nfree) return;
    const int32_t i = free_ids[p];
    int64_t w = rp[p], dslot = -1;
    double diag = 0.0, ru = 0.0, rv = 0.0;
    for (int64_t k = aoff[i]; k < aoff[i + 1]; ++k) {
        const double ce = c[aedge[k]];
        const int32_t j = anb[k];
        diag += ce;
        const int32_t fj = free_pos[j];
        if (fj >= 0) {
            if (dslot < 0 && j > i) dslot = w++;
            if (cols) cols[w] = fj;
            vals[w] = -ce;
            ++w;
        } else {
            ru += ce * pin_at[2 * (int64_t)j];
            rv += ce * pin_at[2 * (int64_t)j + 1];
        }
    }
    if (dslot < 0) dslot = w++;
    if (cols) cols[dslot] = p;
    vals[dslot] = diag;
    if (diagpos) diagpos[p] = dslot;
    rhs[2 * (int64_t)p] = ru;
    rhs[2 * (int64_t)p + 1] = rv;
}

// uv of every vertex: the solution on free vertices, the frozen targets on pins
__global__ __launch_bounds__(BLK) void kd_uv(int64_t nv, const int32_t* __restrict__ free_pos,
                                             const double* __restrict__ x, const double* __restrict__ pin_at,
                                             double* __restrict__ uv) {
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= nv) return;
    const int32_t f = free_pos[v];
    uv[2 * v] = f >= 0 ? x[2 * (int64_t)f] : pin_at[2 * v];
    uv[2 * v + 1] = f >= 0 ? x[2 * (int64_t)f + 1] : pin_at[2 * v + 1];
}

// per-block partial sums of E = sum_e c_e |uv_i - uv_j|^2
__global__ __launch_bounds__(BLK) void kd_energy(int64_t ne, const int32_t* __restrict__ ei,
                                                 const int32_t* __restrict__ ej, const double* __restrict__ c,
                                                 const double* __restrict__ uv, double* __restrict__ part) {
    __shared__ double sh[WAVES];
    double acc = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x; e < ne; e += (int64_t)gridDim.x * BLK) {
        const double ce = c[e];
        if (ce == 0.0) continue;
        const double du = uv[2 * (int64_t)ei[e]] - uv[2 * (int64_t)ej[e]];
        const double dv = uv[2 * (int64_t)ei[e] + 1] - uv[2 * (int64_t)ej[e] + 1];
        acc += ce * (du * du + dv * dv);
    }
    const double s = block_sum_all(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// g on the free rows (interleaved) from g on every vertex; lam on every vertex from the free-row solution
__global__ __launch_bounds__(BLK) void kd_gather_free(int32_t nfree, const int32_t* __restrict__ free_ids,
                                                      const double* __restrict__ g, double* __restrict__ gf) {
    const int32_t p = blockIdx.x * BLK + threadIdx.x;
    if (p >= nfree) return;
    gf[2 * (int64_t)p] = g[2 * (int64_t)free_ids[p]];
    gf[2 * (int64_t)p + 1] = g[2 * (int64_t)free_ids[p] + 1];
}
__global__ __launch_bounds__(BLK) void kd_scatter_free(int64_t nv, const int32_t* __restrict__ free_pos,
                                                       const double* __restrict__ lf, double* __restrict__ lam) {
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= nv) return;
    const int32_t f = free_pos[v];
    lam[2 * v] = f >= 0 ? lf[2 * (int64_t)f] : 0.0;
    lam[2 * v + 1] = f >= 0 ? lf[2 * (int64_t)f + 1] : 0.0;
}

inline dim3 grid1(int64_t n) { return dim3((unsigned)std::max<int64_t>(1, (n + BLK - 1) / BLK)); }

}  // namespace

struct DiffBatchImpl {
    hipStream_t stream = nullptr;  // the batch's stream
    std::vector<void*> mem;
    ChartTopo T;
    int64_t nv = 0, ne = 0, nnz = 0;
    int32_t nfree = 0;
    int mode = 0;
    double tol = 1e-12;
    BatchImpl* batch = nullptr;
    int32_t *ei = nullptr, *ej = nullptr, *opp = nullptr, *anb = nullptr, *tlist = nullptr, *free_ids = nullptr,
            *free_pos = nullptr, *cols = nullptr;
    int64_t *ooff = nullptr, *aoff = nullptr, *aedge = nullptr, *toff = nullptr, *tri_edge = nullptr, *rp = nullptr,
            *map = nullptr, *diagpos = nullptr;
    uint32_t* tris = nullptr;
    double *p3 = nullptr, *pin_at = nullptr, *c = nullptr, *w = nullptr, *dd = nullptr, *vals = nullptr,
           *rhs = nullptr, *x = nullptr, *uv = nullptr, *gf = nullptr, *lf = nullptr, *lam = nullptr, *g = nullptr,
           *gw = nullptr, *gd = nullptr, *grad = nullptr, *part = nullptr;
    bool have_forward = false;
    ~DiffBatchImpl() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : mem) (void)hipFree(p);
        if (batch) batch_destroy(batch);
    }
    template <typename E>
    hipError_t alloc(E** p, size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(E));
        if (e == hipSuccess) mem.push_back(q);
        *p = (E*)q;
        return e;
    }
    template <typename E>
    hipError_t upload(E** p, const E* h, size_t n, hipStream_t s) {
        hipError_t e = alloc(p, n);
        if (e == hipSuccess && n) e = hipMemcpyAsync(*p, h, n * sizeof(E), hipMemcpyHostToDevice, s);
        return e;
    }
};

namespace {

// positions -> weights -> values and rhs (structure fixed)
int assemble(DiffBatchImpl& D, const double* p3, hipStream_t s) {
    HIPCHK(hipMemcpyAsync(D.p3, p3, sizeof(double) * 3 * D.nv, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_edge_c, grid1(D.ne), dim3(BLK), 0, s, D.ne, (const int32_t*)D.ei, (const int32_t*)D.ej,
                       (const int64_t*)D.ooff, (const int32_t*)D.opp, (const double*)D.p3, D.c, D.w, D.dd);
    hipLaunchKernelGGL(kd_row_fill, grid1(D.nfree), dim3(BLK), 0, s, D.nfree, (const int32_t*)D.free_ids,
                       (const int64_t*)D.aoff, (const int32_t*)D.anb, (const int64_t*)D.aedge, (const double*)D.c,
                       (const int32_t*)D.free_pos, (const double*)D.pin_at, (const int64_t*)D.rp, (int32_t*)nullptr,
                       D.vals, D.rhs, (int64_t*)nullptr);
    HIPCHK(hipGetLastError());
    return RUV_OK;
}

}  // namespace

int diff_create(const double* p3, int64_t nv, const uint32_t* tris, int64_t ntri, const int32_t* pin_ids,
                const double* pin_uv, int64_t npins, const int64_t* voff, int32_t nchart, int precond, int mode,
                double tol, DiffBatchImpl** out) {
    if (nchart <= 0 || voff[0] != 0 || voff[nchart] != nv) return fail(RUV_ERR_BAD_DIMS, "chart vertex offsets");
    for (int32_t k = 0; k < nchart; ++k)
        if (voff[k + 1] < voff[k]) return fail(RUV_ERR_BAD_DIMS, "chart vertex offsets not monotone");
    std::unique_ptr<DiffBatchImpl> D(new DiffBatchImpl());
    D->nv = nv;
    D->mode = mode;
    D->tol = tol;
    std::vector<char> is_pin(nv, 0);
    std::vector<double> pin_at(2 * nv, 0.0);
    for (int64_t q = 0; q < npins; ++q) {
        const int32_t v = pin_ids[q];
        if (v < 0 || v >= nv) return fail(RUV_ERR_BAD_DIMS, "pin id out of range");
        is_pin[v] = 1;
        pin_at[2 * v] = pin_uv[2 * q];
        pin_at[2 * v + 1] = pin_uv[2 * q + 1];
    }
    for (int64_t k = 0; k < 3 * ntri; ++k)
        if ((int64_t)tris[k] >= nv) return fail(RUV_ERR_BAD_DIMS, "triangle index out of range");
    // every triangle inside one chart, so the union's operator is block-diagonal by chart
    for (int64_t f = 0; f < ntri; ++f) {
        const int64_t* hi = std::upper_bound(voff, voff + nchart + 1, (int64_t)tris[3 * f]);
        const int64_t c = (hi - voff) - 1;
        for (int q = 1; q < 3; ++q)
            if ((int64_t)tris[3 * f + q] < voff[c] || (int64_t)tris[3 * f + q] >= voff[c + 1])
                return fail(RUV_ERR_BAD_DIMS, "a triangle spans two charts");
    }
    ChartTopo& T = D->T;
    build_topo(nv, tris, ntri, is_pin, &T);
    D->ne = T.ne;
    D->nfree = (int32_t)T.free_ids.size();
    if (D->nfree == 0) return fail(RUV_ERR_BAD_DIMS, "no free vertex");
    // a temporary stream for the structure; the batch's own stream afterwards
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() {
            if (s) {
                (void)hipStreamSynchronize(s);
                (void)hipStreamDestroy(s);
            }
        }
    } guard{s};
    std::vector<uint32_t> tv(tris, tris + 3 * ntri);
    HIPCHK(D->upload(&D->ei, T.ei.data(), T.ei.size(), s));
    HIPCHK(D->upload(&D->ej, T.ej.data(), T.ej.size(), s));
    HIPCHK(D->upload(&D->opp, T.opp.data(), T.opp.size(), s));
    HIPCHK(D->upload(&D->ooff, T.ooff.data(), T.ooff.size(), s));
    HIPCHK(D->upload(&D->aoff, T.aoff.data(), T.aoff.size(), s));
    HIPCHK(D->upload(&D->anb, T.anb.data(), T.anb.size(), s));
    HIPCHK(D->upload(&D->aedge, T.aedge.data(), T.aedge.size(), s));
    HIPCHK(D->upload(&D->toff, T.toff.data(), T.toff.size(), s));
    HIPCHK(D->upload(&D->tlist, T.tlist.data(), T.tlist.size(), s));
    HIPCHK(D->upload(&D->tri_edge, T.tri_edge.data(), T.tri_edge.size(), s));
    HIPCHK(D->upload(&D->tris, tv.data(), tv.size(), s));
    HIPCHK(D->upload(&D->free_ids, T.free_ids.data(), T.free_ids.size(), s));
    HIPCHK(D->upload(&D->free_pos, T.free_pos.data(), T.free_pos.size(), s));
    HIPCHK(D->upload(&D->pin_at, pin_at.data(), pin_at.size(), s));
    HIPCHK(D->alloc(&D->p3, 3 * (size_t)nv));
    HIPCHK(D->alloc(&D->c, (size_t)D->ne));
    HIPCHK(D->alloc(&D->w, (size_t)D->ne));
    HIPCHK(D->alloc(&D->dd, (size_t)D->ne));
    // structure: row counts, pointers, columns, diagonal positions
    int64_t* cnt = nullptr;
    HIPCHK(D->alloc(&cnt, (size_t)D->nfree));
    hipLaunchKernelGGL(kd_row_count, grid1(D->nfree), dim3(BLK), 0, s, D->nfree, (const int32_t*)D->free_ids,
                       (const int64_t*)D->aoff, (const int32_t*)D->anb, (const int32_t*)D->free_pos, cnt);
    HIPCHK(hipGetLastError());
    std::vector<int64_t> hc(D->nfree), rp(D->nfree + 1, 0);
    HIPCHK(hipMemcpyAsync(hc.data(), cnt, sizeof(int64_t) * D->nfree, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int32_t p = 0; p < D->nfree; ++p) rp[p + 1] = rp[p] + hc[p];
    D->nnz = rp[D->nfree];
    HIPCHK(D->upload(&D->rp, rp.data(), rp.size(), s));
    HIPCHK(D->alloc(&D->cols, (size_t)D->nnz));
    HIPCHK(D->alloc(&D->vals, (size_t)D->nnz));
    HIPCHK(D->alloc(&D->rhs, 2 * (size_t)D->nfree));
    HIPCHK(D->alloc(&D->diagpos, (size_t)D->nfree));
    HIPCHK(hipMemcpyAsync(D->p3, p3, sizeof(double) * 3 * nv, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_edge_c, grid1(D->ne), dim3(BLK), 0, s, D->ne, (const int32_t*)D->ei, (const int32_t*)D->ej,
                       (const int64_t*)D->ooff, (const int32_t*)D->opp, (const double*)D->p3, D->c, D->w, D->dd);
    hipLaunchKernelGGL(kd_row_fill, grid1(D->nfree), dim3(BLK), 0, s, D->nfree, (const int32_t*)D->free_ids,
                       (const int64_t*)D->aoff, (const int32_t*)D->anb, (const int64_t*)D->aedge,
                       (const double*)D->c, (const int32_t*)D->free_pos, (const double*)D->pin_at,
                       (const int64_t*)D->rp, D->cols, D->vals, D->rhs, D->diagpos);
    HIPCHK(hipGetLastError());
    std::vector<int32_t> hcols(D->nnz);
    std::vector<double> hvals(D->nnz);
    HIPCHK(hipMemcpyAsync(hcols.data(), D->cols, sizeof(int32_t) * D->nnz, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hvals.data(), D->vals, sizeof(double) * D->nnz, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // chart row offsets: free rows follow vertex order, so chart c owns the free rows of its vertex range
    std::vector<int64_t> crow(nchart + 1, 0);
    {
        int64_t f = 0;
        int32_t c = 0;
        for (int64_t v = 0; v <= nv; ++v) {
            while (c <= nchart && voff[c] == v) crow[c++] = f;
            if (v < nv && T.free_pos[v] >= 0) ++f;
        }
    }
    BatchImpl* B = nullptr;
    int rc = batch_create(D->nfree, rp.data(), hcols.data(), hvals.data(), nchart, crow.data(), 2, precond, mode, &B);
    if (rc) return rc;
    D->batch = B;
    D->stream = B->stream;
    std::vector<int64_t> map;
    rc = batch_value_map(*B, rp.data(), hcols.data(), &map);
    if (rc) return rc;
    HIPCHK(D->upload(&D->map, map.data(), map.size(), s));
    HIPCHK(D->alloc(&D->x, 2 * (size_t)D->nfree));
    HIPCHK(D->alloc(&D->uv, 2 * (size_t)nv));
    HIPCHK(D->alloc(&D->g, 2 * (size_t)nv));
    HIPCHK(D->alloc(&D->gf, 2 * (size_t)D->nfree));
    HIPCHK(D->alloc(&D->lf, 2 * (size_t)D->nfree));
    HIPCHK(D->alloc(&D->lam, 2 * (size_t)nv));
    HIPCHK(D->alloc(&D->gw, (size_t)D->ne));
    HIPCHK(D->alloc(&D->gd, (size_t)D->ne));
    HIPCHK(D->alloc(&D->grad, 3 * (size_t)nv));
    HIPCHK(D->alloc(&D->part, 1024));
    HIPCHK(hipStreamSynchronize(s));
    *out = D.release();
    return RUV_OK;
}

void diff_destroy(DiffBatchImpl* d) { delete d; }

int diff_forward(DiffBatchImpl* D, const double* p3, double* uv_out, double* energy_out, ruv_solve_info* info) {
    hipStream_t s = D->stream;
    int rc = assemble(*D, p3, s);
    if (rc) return rc;
    rc = batch_update_values(D->batch, D->vals, D->map, D->diagpos);
    if (rc) return rc;
    rc = batch_solve_dev(D->batch, D->rhs, D->x, 1000000, D->tol, info, nullptr, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(kd_uv, grid1(D->nv), dim3(BLK), 0, s, D->nv, (const int32_t*)D->free_pos, (const double*)D->x,
                       (const double*)D->pin_at, D->uv);
    const int ge = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (D->ne + BLK - 1) / BLK));
    hipLaunchKernelGGL(kd_energy, dim3(ge), dim3(BLK), 0, s, D->ne, (const int32_t*)D->ei, (const int32_t*)D->ej,
                       (const double*)D->c, (const double*)D->uv, D->part);
    HIPCHK(hipGetLastError());
    std::vector<double> part(ge);
    HIPCHK(hipMemcpyAsync(uv_out, D->uv, sizeof(double) * 2 * D->nv, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(part.data(), D->part, sizeof(double) * ge, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (energy_out) {
        double e = 0.0;
        for (double v : part) e += v;
        *energy_out = e;
    }
    D->have_forward = true;
    return RUV_OK;
}

int diff_vjp(DiffBatchImpl* D, const double* g_uv, double a_energy, double* grad_out, ruv_solve_info* info) {
    if (!D->have_forward) return fail(RUV_ERR_INVALID, "vjp before any forward");
    hipStream_t s = D->stream;
    HIPCHK(hipMemcpyAsync(D->g, g_uv, sizeof(double) * 2 * D->nv, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(kd_gather_free, grid1(D->nfree), dim3(BLK), 0, s, D->nfree, (const int32_t*)D->free_ids,
                       (const double*)D->g, D->gf);
    HIPCHK(hipGetLastError());
    int rc = batch_solve_dev(D->batch, D->gf, D->lf, 1000000, D->tol, info, nullptr, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(kd_scatter_free, grid1(D->nv), dim3(BLK), 0, s, D->nv, (const int32_t*)D->free_pos,
                       (const double*)D->lf, D->lam);
    const int ge = (int)std::max<int64_t>(1, std::min<int64_t>((D->ne + BLK - 1) / BLK, 4096));
    const int gv = (int)std::max<int64_t>(1, std::min<int64_t>((D->nv + BLK - 1) / BLK, 4096));
    hipLaunchKernelGGL(k_edge_adj, dim3(ge), dim3(BLK), 0, s, D->ne, (const int32_t*)D->ei, (const int32_t*)D->ej,
                       (const double*)D->c, (const double*)D->w, (const double*)D->dd, (const double*)D->uv,
                       (const double*)D->lam, a_energy, D->gw, D->gd);
    hipLaunchKernelGGL(k_vertex_grad, dim3(gv), dim3(BLK), 0, s, D->nv, (const double*)D->p3,
                       (const uint32_t*)D->tris, (const int64_t*)D->tri_edge, (const int64_t*)D->aoff,
                       (const int32_t*)D->anb, (const int64_t*)D->aedge, (const int64_t*)D->toff,
                       (const int32_t*)D->tlist, (const double*)D->gw, (const double*)D->gd, D->grad);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(grad_out, D->grad, sizeof(double) * 3 * D->nv, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return RUV_OK;
}

}  // namespace ruv
