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

#pragma clang fp contract(off)

#include "assemble_kernels.hpp"
#include "chart_topo.hpp"
#include "dev_assembly.hpp"
#include "host_lscm.hpp"
#include "ruv_internal.hpp"

namespace ruv {

#define HIPCHK(call)                                                                                         \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess) return fail(RUV_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

DevAssembly::~DevAssembly() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : mem) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
}

template <typename V>
static hipError_t dev_upload(DevAssembly& A, typename V::value_type** dst, const V& v) {
    using E = typename V::value_type;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(v.size(), 1) * sizeof(E));
    if (e != hipSuccess) return e;
    A.mem.push_back(q);
    *dst = (E*)q;
    return v.empty() ? hipSuccess : hipMemcpyAsync(q, v.data(), v.size() * sizeof(E), hipMemcpyHostToDevice, A.stream);
}

template <typename E>
static hipError_t dev_alloc(DevAssembly& A, E** dst, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(E));
    if (e == hipSuccess) {
        A.mem.push_back(q);
        *dst = (E*)q;
    }
    return e;
}

// Upload topology + geometry, compute the per-edge weights and assemble L / rhs on the device.
int dev_assemble(const ChartTopo& T, const std::vector<double>& p3, const std::vector<double>& p2,
                 const std::vector<double>& pin_at, const std::vector<uint32_t>& tris, double mix_b, DevAssembly* A) {
    HIPCHK(hipStreamCreateWithFlags(&A->stream, hipStreamNonBlocking));
    A->nv = T.nv;
    A->ne = T.ne;
    A->ntri = T.ntri;
    A->nfree = (int32_t)T.free_ids.size();
    HIPCHK(dev_upload(*A, &A->ei, T.ei));
    HIPCHK(dev_upload(*A, &A->ej, T.ej));
    HIPCHK(dev_upload(*A, &A->opp, T.opp));
    HIPCHK(dev_upload(*A, &A->ooff, T.ooff));
    HIPCHK(dev_upload(*A, &A->aoff, T.aoff));
    HIPCHK(dev_upload(*A, &A->anb, T.anb));
    HIPCHK(dev_upload(*A, &A->aedge, T.aedge));
    HIPCHK(dev_upload(*A, &A->toff, T.toff));
    HIPCHK(dev_upload(*A, &A->tlist, T.tlist));
    HIPCHK(dev_upload(*A, &A->tri_edge, T.tri_edge));
    HIPCHK(dev_upload(*A, &A->tris, tris));
    HIPCHK(dev_upload(*A, &A->free_ids, T.free_ids));
    HIPCHK(dev_upload(*A, &A->free_pos, T.free_pos));
    HIPCHK(dev_upload(*A, &A->p3, p3));
    HIPCHK(dev_upload(*A, &A->p2, p2));
    HIPCHK(dev_upload(*A, &A->pin_at, pin_at));
    HIPCHK(dev_alloc(*A, &A->c, (size_t)T.ne));
    HIPCHK(dev_alloc(*A, &A->w, (size_t)T.ne));
    HIPCHK(dev_alloc(*A, &A->dd, (size_t)T.ne));
    HIPCHK(dev_alloc(*A, &A->cnt, (size_t)A->nfree));
    const int ge = (int)std::max<int64_t>(1, std::min<int64_t>((T.ne + BLK - 1) / BLK, 4096));
    const int gr = (int)std::max<int64_t>(1, std::min<int64_t>((A->nfree + BLK - 1) / BLK, 4096));
    hipLaunchKernelGGL(k_edge_c, dim3(ge), dim3(BLK), 0, A->stream, T.ne, (const int32_t*)A->ei, (const int32_t*)A->ej,
                       (const int64_t*)A->ooff, (const int32_t*)A->opp, (const double*)A->p3, A->c, A->w, A->dd);
    if (A->nfree > 0)
        hipLaunchKernelGGL(k_row_count, dim3(gr), dim3(BLK), 0, A->stream, A->nfree, (const int32_t*)A->free_ids,
                           (const int64_t*)A->aoff, (const int32_t*)A->anb, (const int64_t*)A->aedge,
                           (const double*)A->c, (const int32_t*)A->free_pos, A->cnt);
    HIPCHK(hipGetLastError());
    std::vector<int32_t> cnt(A->nfree);
    if (A->nfree) HIPCHK(hipMemcpyAsync(cnt.data(), A->cnt, sizeof(int32_t) * A->nfree, hipMemcpyDeviceToHost, A->stream));
    HIPCHK(hipStreamSynchronize(A->stream));
    A->rp.assign(A->nfree + 1, 0);
    for (int32_t p = 0; p < A->nfree; ++p) A->rp[p + 1] = A->rp[p] + cnt[p];
    A->nnz = A->rp[A->nfree];
    HIPCHK(dev_upload(*A, &A->d_rp, A->rp));
    HIPCHK(dev_alloc(*A, &A->cols, (size_t)A->nnz));
    HIPCHK(dev_alloc(*A, &A->vals, (size_t)A->nnz));
    HIPCHK(dev_alloc(*A, &A->rhs, 2 * (size_t)A->nfree));
    if (A->nfree > 0)
        hipLaunchKernelGGL(k_row_fill, dim3(gr), dim3(BLK), 0, A->stream, A->nfree, (const int32_t*)A->free_ids,
                           (const int64_t*)A->aoff, (const int32_t*)A->anb, (const int64_t*)A->aedge,
                           (const double*)A->c, (const int32_t*)A->free_pos, (const double*)A->p2,
                           (const double*)A->pin_at, mix_b, (const int32_t*)A->d_rp, A->cols, A->vals, A->rhs);
    HIPCHK(hipGetLastError());
    return RUV_OK;
}

// L (host CSR) and rhs back from a device assembly
int dev_download(DevAssembly& A, Csr* L, std::vector<double>* rhs, size_t* edges) {
    L->n = A.nfree;
    L->ptr = A.rp;
    L->cols.assign(A.nnz, 0);
    L->vals.assign(A.nnz, 0.0);
    rhs->assign(2 * (size_t)A.nfree, 0.0);
    if (A.nnz) {
        HIPCHK(hipMemcpyAsync(L->cols.data(), A.cols, sizeof(int32_t) * A.nnz, hipMemcpyDeviceToHost, A.stream));
        HIPCHK(hipMemcpyAsync(L->vals.data(), A.vals, sizeof(double) * A.nnz, hipMemcpyDeviceToHost, A.stream));
    }
    if (A.nfree)
        HIPCHK(hipMemcpyAsync(rhs->data(), A.rhs, sizeof(double) * 2 * A.nfree, hipMemcpyDeviceToHost, A.stream));
    std::vector<double> cc(A.ne);
    if (A.ne) HIPCHK(hipMemcpyAsync(cc.data(), A.c, sizeof(double) * A.ne, hipMemcpyDeviceToHost, A.stream));
    HIPCHK(hipStreamSynchronize(A.stream));
    if (edges) *edges = (size_t)std::count_if(cc.begin(), cc.end(), [](double v) { return v != 0.0; });
    return RUV_OK;
}

// Chart `ch` of `pos`: pins and frame on the host, every floating-point entry of L and the rhs on the device.
int device_assemble(const double* pos, const Chart& ch, double angle_mix, ChartSystem* S) {
    build_system_pins(pos, ch, S);
    std::vector<char> is_pin(S->nv, 0);
    std::vector<double> pin_at(2 * S->nv, 0.0);
    for (size_t q = 0; q < S->pin_ids.size(); ++q) {
        is_pin[S->pin_ids[q]] = 1;
        pin_at[2 * S->pin_ids[q]] = S->pin_uv[2 * q];
        pin_at[2 * S->pin_ids[q] + 1] = S->pin_uv[2 * q + 1];
    }
    ChartTopo T;
    build_topo(S->nv, ch.tris.data(), (int64_t)ch.tris.size() / 3, is_pin, &T);
    DevAssembly A;
    int rc = dev_assemble(T, S->p3, S->p2, pin_at, ch.tris, 1.0 - angle_mix, &A);
    if (rc) return rc;
    return dev_download(A, &S->l, &S->rhs, &S->edges);
}

}  // namespace ruv
