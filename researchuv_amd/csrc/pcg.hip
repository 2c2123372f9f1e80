// pcg.hip -- device PCG for the LSCM free block on MI355X (gfx950, wave64).
//
// Replaces GpuSolver::cg_solve_precond (crates/researchuv-gpu/src/solver.rs:284-565) and its kernels
// (kernels.cu:7-102).  Design differences from the reference, all MI355X-first:
//  * every scalar of the recurrence (alpha, beta, r.r, r.z, p.Ap) lives on the device: a kernel that
//    produces a dot product writes one partial per workgroup, and every workgroup of the kernel that
//    consumes it sums those partials itself in a fixed order (deterministic, no host round trip and no
//    extra finalize launch); the reference synchronised and copied partials to the host on every dot
//    (solver.rs:569-615);
//  * the operator is streamed as L (x) I_R: for the LSCM system (A = L (x) I2, lscm.rs:605-640) each row of
//    L is read once and applied to the interleaved (u, v) pair, halving matrix traffic;
//  * the x/r/z updates of an iteration are one fused kernel (the reference launched axpy, axpy, mul and
//    copy separately);
//  * the host only polls a small status block every few iterations (chunks of iterations in flight).
// The plain form keeps the reference's recurrence and stopping rule exactly: x0 = 0,
// ||r_k|| <= tol*||b|| on the recursive residual, breakdown when |p.Ap| <= 1e-300 or non-finite.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <tuple>
#include <utility>
#include <vector>

#include "amg_kernels.hpp"
#include "ruv_internal.hpp"
#include "pcg_kernels.hpp"

namespace ruv {

#define HIPCHK(call)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return fail(RUV_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

// Device buffer owner (freed on scope exit; the reference's DeviceMem, solver.rs:671-729).
struct DevArena {
    std::vector<void*> ptrs;
    ~DevArena() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    hipError_t alloc(T** out, size_t count) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(p);
        *out = (T*)p;
        return e;
    }
    template <typename T>
    hipError_t upload(T** out, const T* host, size_t count, hipStream_t s) {
        hipError_t e = alloc(out, count);
        if (e != hipSuccess) return e;
        if (count == 0) return hipSuccess;
        return hipMemcpyAsync(*out, host, count * sizeof(T), hipMemcpyHostToDevice, s);
    }
};

int grid_for(int64_t n) {
    int64_t g = (n + BLK - 1) / BLK;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, MAXG));
}

// A triangular factor on the device (DeviceIc0, precond.rs:502-518).
template <typename T>
struct DevTri {
    int n = 0;
    int *l_ptr = nullptr, *l_cols = nullptr, *lt_ptr = nullptr, *lt_cols = nullptr;
    T *l_vals = nullptr, *lt_vals = nullptr;
    int *fwd_rows = nullptr, *fwd_off = nullptr, *bwd_rows = nullptr, *bwd_off = nullptr;
    int n_fwd = 0, n_bwd = 0;
    // Sweep schedule: a level of at least `wide` rows runs as its own many-workgroup launch (k_tri_level), the
    // runs of narrower levels in single-workgroup launches (k_tri_sweep).  Each segment: first level, end level,
    // and for a wide level its row range.
    struct Seg {
        int l0, l1, a, b;
        bool wide;
    };
    std::vector<Seg> fwd_seg, bwd_seg;
};

// cut the levels (offsets off, nlev + 1 entries) into wide single levels and runs of narrow ones
inline std::vector<typename DevTri<float>::Seg> tri_segments(const std::vector<int32_t>& off, int wide) {
    std::vector<typename DevTri<float>::Seg> out;
    const int nlev = (int)off.size() - 1;
    int l = 0;
    while (l < nlev) {
        if (off[l + 1] - off[l] >= wide) {
            out.push_back({l, l + 1, off[l], off[l + 1], true});
            ++l;
            continue;
        }
        int e = l;
        while (e < nlev && off[e + 1] - off[e] < wide) ++e;
        out.push_back({l, e, off[l], off[e], false});
        l = e;
    }
    return out;
}

// the factor's structure and levels; its values are computed on the device (ic0_numeric)
template <typename T>
int upload_tri(DevArena& m, const TriFactor& f, hipStream_t s, DevTri<T>* d) {
    d->n = f.l.n;
    HIPCHK(m.upload(&d->l_ptr, f.l.ptr.data(), f.l.ptr.size(), s));
    HIPCHK(m.upload(&d->l_cols, f.l.cols.data(), f.l.cols.size(), s));
    HIPCHK(m.alloc(&d->l_vals, f.l.cols.size()));
    HIPCHK(m.upload(&d->lt_ptr, f.lt.ptr.data(), f.lt.ptr.size(), s));
    HIPCHK(m.upload(&d->lt_cols, f.lt.cols.data(), f.lt.cols.size(), s));
    HIPCHK(m.alloc(&d->lt_vals, f.lt.cols.size()));
    HIPCHK(m.upload(&d->fwd_rows, f.fwd_rows.data(), f.fwd_rows.size(), s));
    HIPCHK(m.upload(&d->fwd_off, f.fwd_off.data(), f.fwd_off.size(), s));
    HIPCHK(m.upload(&d->bwd_rows, f.bwd_rows.data(), f.bwd_rows.size(), s));
    HIPCHK(m.upload(&d->bwd_off, f.bwd_off.data(), f.bwd_off.size(), s));
    d->n_fwd = (int)f.fwd_off.size() - 1;
    d->n_bwd = (int)f.bwd_off.size() - 1;
    int wide = 16384;  // rows from which one level fills enough workgroups to beat the single-workgroup sweep
    if (const char* e = std::getenv("RUV_TRI_WIDE")) wide = std::max(1, std::atoi(e));
    d->fwd_seg.clear();
    d->bwd_seg.clear();
    for (const auto& g : tri_segments(f.fwd_off, wide)) d->fwd_seg.push_back({g.l0, g.l1, g.a, g.b, g.wide});
    for (const auto& g : tri_segments(f.bwd_off, wide)) d->bwd_seg.push_back({g.l0, g.l1, g.a, g.b, g.wide});
    // the host vectors must outlive the async copies
    HIPCHK(hipStreamSynchronize(s));
    return RUV_OK;
}

// One AMG level on the device (amg.rs:74-102, V-cycle buffers).
template <typename T>
struct DevAmgLevel {
    int n = 0, nc = 0;
    int *a_ptr = nullptr, *a_cols = nullptr, *p_ptr = nullptr, *p_cols = nullptr, *r_ptr = nullptr, *r_cols = nullptr;
    T *a_vals = nullptr, *dinv = nullptr, *p_vals = nullptr, *r_vals = nullptr;
    T *rin = nullptr, *t = nullptr, *e1 = nullptr, *e = nullptr;  // level input residual, temps, output
    double* dense_inv = nullptr;                                  // coarsest: FP64 inverse
    bool coarsest = false, diag_only = false;
};

template <typename T>
struct DevAmg {
    std::vector<DevAmgLevel<T>> lv;
    T omega = T(2.0 / 3.0);
};

template <typename T>
int upload_amg(DevArena& m, const std::vector<AmgHostLevel>& H, int R, hipStream_t s, DevAmg<T>* d) {
    std::vector<std::vector<T>> keep;  // staging that must outlive the async copies
    auto up = [&](T** dst, const std::vector<double>& src) {
        keep.emplace_back(src.begin(), src.end());
        return m.upload(dst, keep.back().data(), keep.back().size(), s);
    };
    d->lv.resize(H.size());
    for (size_t l = 0; l < H.size(); ++l) {
        const AmgHostLevel& h = H[l];
        DevAmgLevel<T>& L = d->lv[l];
        L.n = h.op.n;
        L.nc = h.n_coarse;
        L.coarsest = (l + 1 == H.size());
        L.diag_only = h.diag_only;
        std::vector<double> dinv(h.op.n);
        for (int32_t i = 0; i < h.op.n; ++i) dinv[i] = 1.0 / h.op.get(i, i);
        HIPCHK(m.upload(&L.a_ptr, h.op.ptr.data(), h.op.ptr.size(), s));
        HIPCHK(m.upload(&L.a_cols, h.op.cols.data(), h.op.cols.size(), s));
        HIPCHK(up(&L.a_vals, h.op.vals));
        HIPCHK(up(&L.dinv, dinv));
        const size_t N = (size_t)h.op.n * R;
        if (l > 0) HIPCHK(m.alloc(&L.rin, N));
        HIPCHK(m.alloc(&L.t, N));
        HIPCHK(m.alloc(&L.e1, N));
        HIPCHK(m.alloc(&L.e, N));
        if (!L.coarsest) {
            HIPCHK(m.upload(&L.p_ptr, h.p.ptr.data(), h.p.ptr.size(), s));
            HIPCHK(m.upload(&L.p_cols, h.p.cols.data(), h.p.cols.size(), s));
            HIPCHK(up(&L.p_vals, h.p.vals));
            HIPCHK(m.upload(&L.r_ptr, h.r.ptr.data(), h.r.ptr.size(), s));
            HIPCHK(m.upload(&L.r_cols, h.r.cols.data(), h.r.cols.size(), s));
            HIPCHK(up(&L.r_vals, h.r.vals));
        } else if (h.has_factor) {
            HIPCHK(m.upload(&L.dense_inv, h.dense_inv.data(), h.dense_inv.size(), s));
        }
    }
    HIPCHK(hipStreamSynchronize(s));
    return RUV_OK;
}

}  // namespace

// ------------------------------------------------------------------------------------------------------
// Plan: a device-resident system L (x) I_R with its preconditioner and work vectors.
// ------------------------------------------------------------------------------------------------------
struct PlanImpl {
    int R = 1, n = 0, mode = 0, precond = 0, G = 1;
    int64_t nnz = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevArena mem;
    int *ptr = nullptr, *cols = nullptr;
    double* vals64 = nullptr;
    float* vals32 = nullptr;
    double* dinv64 = nullptr;
    float* dinv32 = nullptr;
    DevTri<double> tri64;
    DevTri<float> tri32;
    bool has_tri = false;
    DevAmg<double> amg64;
    DevAmg<float> amg32;
    int amg_levels = 0;
    // permutation round trip for the colour-ordered IC(0) (precond.rs:613-630)
    int *to_orig = nullptr, *to_perm = nullptr;
    bool permuted = false;
    // IC(0) computed on the device from vals64: each strict-lower L entry's and each diagonal's position in the
    // values, each L^T entry's L entry, the FP64 workspace (mixed form), a failure flag
    int64_t *ic0_src = nullptr, *ic0_dsrc = nullptr;
    int32_t* ic0_ltsrc = nullptr;
    int64_t ic0_nl = 0;
    double* ic0_work = nullptr;
    int* ic0_fail = nullptr;
    std::vector<int64_t> jacobi_diag;  // Jacobi: each row's diagonal entry (-1: none)
    // work vectors
    double *x64 = nullptr, *r64 = nullptr, *z64 = nullptr, *p64 = nullptr, *q64 = nullptr, *y64 = nullptr,
           *t64 = nullptr;
    float *x32 = nullptr, *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *q32 = nullptr, *y32 = nullptr,
          *t32 = nullptr;
    double *part_a = nullptr, *part_b = nullptr, *part_c = nullptr;
    SolveState* st = nullptr;
    SolveState* h_st = nullptr;
    double* h_part = nullptr;
    double inner_tol = 1e-4;
    int max_refine = 40;
    // mixed refinement: also stop only once the estimated forward error is at most uv_tol max|x| (a tenth of the
    // 1e-6 UV bound, as the batched solver; 0 turns the test off)
    double uv_tol = 1e-7;
    bool use_graph = true;
    hipGraphExec_t gexec = nullptr;
    // While a graph of iterations is built: the graph, its last node and the first error (plaunch adds kernel
    // nodes instead of launching)
    struct GraphRec {
        hipGraph_t g = nullptr;
        hipGraphNode_t last = nullptr;
        hipError_t err = hipSuccess;
    };
    GraphRec* rec = nullptr;

    ~PlanImpl() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (h_st) (void)hipHostFree(h_st);
        if (h_part) (void)hipHostFree(h_part);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Launch kernel k on the plan's stream, or, while the plan records a graph, append it as a kernel node after the
// last one (the node keeps a copy of the arguments).  Building the graph node by node instead of capturing the
// stream keeps the plan clear of stream-capture rules, which other threads' work on the legacy stream would
// otherwise break (a capture in progress makes such work fail, and fails itself).
template <typename... KA, typename... A, size_t... I>
void plaunch_rec(PlanImpl& P, void (*k)(KA...), dim3 grid, dim3 block, std::index_sequence<I...>, A&&... args) {
    std::tuple<std::decay_t<KA>...> t(std::forward<A>(args)...);
    void* params[] = {(void*)&std::get<I>(t)...};
    hipKernelNodeParams kp{};
    kp.func = reinterpret_cast<void*>(k);
    kp.gridDim = grid;
    kp.blockDim = block;
    kp.sharedMemBytes = 0;
    kp.kernelParams = params;
    kp.extra = nullptr;
    PlanImpl::GraphRec& G = *P.rec;
    hipGraphNode_t node = nullptr;
    const hipError_t e = hipGraphAddKernelNode(&node, G.g, G.last ? &G.last : nullptr, G.last ? 1 : 0, &kp);
    if (e != hipSuccess) {
        if (G.err == hipSuccess) G.err = e;
        return;
    }
    G.last = node;
}
template <typename... KA, typename... A>
void plaunch(PlanImpl& P, void (*k)(KA...), dim3 grid, dim3 block, A&&... args) {
    static_assert(sizeof...(KA) == sizeof...(A), "kernel argument count");
    if (P.rec) plaunch_rec(P, k, grid, block, std::index_sequence_for<KA...>{}, std::forward<A>(args)...);
    else hipLaunchKernelGGL(k, grid, block, 0, P.stream, std::forward<A>(args)...);
}

// The IC(0) factor's values on the device from the plan's FP64 values (k_ic0_sweep / k_ic0_level over the forward
// levels, then the sweeps' copies); RUV_ERR_PRECOND on a vanishing pivot or diagonal (factor_ic0 returning None,
// precond.rs:177-187).
template <typename T>
int ic0_numeric(PlanImpl& P, DevTri<T>& d) {
    hipStream_t s = P.stream;
    double* lv = std::is_same<T, double>::value ? (double*)(void*)d.l_vals : P.ic0_work;
    HIPCHK(hipMemsetAsync(P.ic0_fail, 0, sizeof(int), s));
    for (const auto& g : d.fwd_seg) {
        if (g.wide)
            hipLaunchKernelGGL(k_ic0_level, dim3((unsigned)((g.b - g.a + BLK - 1) / BLK)), dim3(BLK), 0, s, g.a, g.b,
                               (const int*)d.fwd_rows, (const double*)P.vals64, (const int64_t*)P.ic0_src,
                               (const int64_t*)P.ic0_dsrc, (const int*)d.l_ptr, (const int*)d.l_cols, lv, P.ic0_fail);
        else
            hipLaunchKernelGGL(k_ic0_sweep, dim3(1), dim3(TRI_BLK), 0, s, g.l1 - g.l0, (const int*)d.fwd_off + g.l0,
                               (const int*)d.fwd_rows, (const double*)P.vals64, (const int64_t*)P.ic0_src,
                               (const int64_t*)P.ic0_dsrc, (const int*)d.l_ptr, (const int*)d.l_cols, lv, P.ic0_fail);
    }
    if (P.ic0_nl > 0)
        hipLaunchKernelGGL((k_ic0_store<T>), dim3((unsigned)((P.ic0_nl + BLK - 1) / BLK)), dim3(BLK), 0, s, P.ic0_nl,
                           (const double*)lv, (const int32_t*)P.ic0_ltsrc, d.l_vals, d.lt_vals);
    HIPCHK(hipGetLastError());
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, P.ic0_fail, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (bad) return fail(RUV_ERR_PRECOND, "ic0: vanishing pivot or diagonal");
    return RUV_OK;
}

int ic0_refactor(PlanImpl& P) {
    return P.mode == RUV_MODE_MIXED ? ic0_numeric<float>(P, P.tri32) : ic0_numeric<double>(P, P.tri64);
}

// Jacobi inverse diagonal (precond.rs:595-611) from the host values: a vanishing diagonal is a setup failure.
int jacobi_upload(PlanImpl& P, const double* vals) {
    hipStream_t s = P.stream;
    std::vector<double> dinv((size_t)P.n, 1.0);
    for (int32_t i = 0; i < P.n; ++i) {
        const double d = P.jacobi_diag[(size_t)i] >= 0 ? vals[P.jacobi_diag[(size_t)i]] : 0.0;
        if (std::fabs(d) < 1e-300) return fail(RUV_ERR_PRECOND, "jacobi: zero diagonal at row " + std::to_string(i));
        dinv[(size_t)i] = 1.0 / d;
    }
    std::vector<float> dinv32;
    if (!P.dinv64) HIPCHK(P.mem.alloc(&P.dinv64, dinv.size()));
    HIPCHK(hipMemcpyAsync(P.dinv64, dinv.data(), sizeof(double) * dinv.size(), hipMemcpyHostToDevice, s));
    if (P.mode == RUV_MODE_MIXED) {
        dinv32.assign(dinv.begin(), dinv.end());
        if (!P.dinv32) HIPCHK(P.mem.alloc(&P.dinv32, dinv32.size()));
        HIPCHK(hipMemcpyAsync(P.dinv32, dinv32.data(), sizeof(float) * dinv32.size(), hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    return RUV_OK;
}

int plan_setup(PlanImpl& P, const Csr& a, int R, int precond, int mode) {
    const char* ng = std::getenv("RUV_NO_GRAPH");  // direct launches instead of graph replay (A/B checks)
    P.use_graph = !(ng && ng[0] == '1');
    if (const char* e = std::getenv("RUV_UV_TOL")) P.uv_tol = std::atof(e);
    P.R = R;
    P.n = a.n;
    P.nnz = a.nnz();
    P.mode = mode;
    P.precond = precond;
    P.G = grid_for(a.n);
    HIPCHK(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&P.ev0));
    HIPCHK(hipEventCreate(&P.ev1));
    hipStream_t s = P.stream;
    HIPCHK(P.mem.upload(&P.ptr, a.ptr.data(), a.ptr.size(), s));
    HIPCHK(P.mem.upload(&P.cols, a.cols.data(), a.cols.size(), s));
    HIPCHK(P.mem.upload(&P.vals64, a.vals.data(), a.vals.size(), s));
    std::vector<float> v32;
    if (mode == RUV_MODE_MIXED) {
        v32.assign(a.vals.begin(), a.vals.end());
        HIPCHK(P.mem.upload(&P.vals32, v32.data(), v32.size(), s));
    }
    const size_t N = (size_t)a.n * R;
    if (precond == RUV_PC_JACOBI) {
        P.jacobi_diag.assign((size_t)a.n, -1);
        for (int32_t i = 0; i < a.n; ++i)
            for (int32_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k)
                if (a.cols[k] == i) {  // the first entry, as Csr::get
                    P.jacobi_diag[(size_t)i] = k;
                    break;
                }
        int rc = jacobi_upload(P, a.vals.data());
        if (rc) return rc;
    } else if (precond == RUV_PC_IC0 || precond == RUV_PC_IC0COLOR) {
        // the structure on the host, the values on the device (a values update refactors there)
        TriFactor f;
        std::vector<int64_t> src, dsrc;
        std::vector<int32_t> lt_src;
        if (precond == RUV_PC_IC0) {
            ic0_pattern(a, nullptr, &f, &src, &dsrc, &lt_src);
        } else {
            std::vector<uint32_t> color = graph_coloring(a);
            std::vector<int32_t> to_perm, to_orig;
            color_permutation(color, &to_perm, &to_orig);
            const Csr pa = permute_csr(a, to_perm, to_orig);
            const std::vector<int64_t> val_of = permuted_positions(a, to_perm, to_orig);
            ic0_pattern(pa, &val_of, &f, &src, &dsrc, &lt_src);
            HIPCHK(P.mem.upload(&P.to_orig, to_orig.data(), to_orig.size(), s));
            HIPCHK(P.mem.upload(&P.to_perm, to_perm.data(), to_perm.size(), s));
            P.permuted = true;
        }
        P.ic0_nl = (int64_t)lt_src.size();
        HIPCHK(P.mem.upload(&P.ic0_src, src.data(), src.size(), s));
        HIPCHK(P.mem.upload(&P.ic0_dsrc, dsrc.data(), dsrc.size(), s));
        HIPCHK(P.mem.upload(&P.ic0_ltsrc, lt_src.data(), lt_src.size(), s));
        HIPCHK(P.mem.alloc(&P.ic0_fail, 1));
        if (mode == RUV_MODE_MIXED) HIPCHK(P.mem.alloc(&P.ic0_work, (size_t)P.ic0_nl));
        int rc = (mode == RUV_MODE_MIXED) ? upload_tri(P.mem, f, s, &P.tri32) : upload_tri(P.mem, f, s, &P.tri64);
        if (!rc) rc = ic0_refactor(P);
        if (rc) return rc;
        P.has_tri = true;
    } else if (precond == RUV_PC_AMG || precond == RUV_PC_AMGSA) {
        std::vector<AmgHostLevel> H;
        std::string why;
        if (!amg_hierarchy(a, precond == RUV_PC_AMGSA, &H, &why)) return fail(RUV_ERR_PRECOND, "amg: " + why);
        int rc = (mode == RUV_MODE_MIXED) ? upload_amg(P.mem, H, R, s, &P.amg32) : upload_amg(P.mem, H, R, s, &P.amg64);
        if (rc) return rc;
        P.amg_levels = (int)H.size();
    } else if (precond != RUV_PC_NONE) {
        return fail(RUV_ERR_INVALID, "precond " + std::to_string(precond) + " is not available in this build");
    }
    HIPCHK(P.mem.alloc(&P.x64, N));
    HIPCHK(P.mem.alloc(&P.r64, N));
    if (mode == RUV_MODE_FP64) {
        HIPCHK(P.mem.alloc(&P.z64, N));
        HIPCHK(P.mem.alloc(&P.p64, N));
        HIPCHK(P.mem.alloc(&P.q64, N));
        HIPCHK(P.mem.alloc(&P.y64, N));
        HIPCHK(P.mem.alloc(&P.t64, N));
    } else {
        HIPCHK(P.mem.alloc(&P.x32, N));
        HIPCHK(P.mem.alloc(&P.r32, N));
        HIPCHK(P.mem.alloc(&P.z32, N));
        HIPCHK(P.mem.alloc(&P.p32, N));
        HIPCHK(P.mem.alloc(&P.q32, N));
        HIPCHK(P.mem.alloc(&P.y32, N));
        HIPCHK(P.mem.alloc(&P.t32, N));
    }
    HIPCHK(P.mem.alloc(&P.part_a, MAXG));
    HIPCHK(P.mem.alloc(&P.part_b, MAXG));
    HIPCHK(P.mem.alloc(&P.part_c, MAXG));
    HIPCHK(P.mem.alloc(&P.st, 1));
    HIPCHK(hipHostMalloc((void**)&P.h_st, sizeof(SolveState), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&P.h_part, sizeof(double) * MAXG, hipHostMallocDefault));
    HIPCHK(hipStreamSynchronize(s));  // the staging vectors above must outlive their copies
    return RUV_OK;
}

// Sum of the G partials of part (host side, block order: deterministic).
int host_sum(PlanImpl& P, const double* part, double* out) {
    HIPCHK(hipMemcpyAsync(P.h_part, part, sizeof(double) * P.G, hipMemcpyDeviceToHost, P.stream));
    HIPCHK(hipStreamSynchronize(P.stream));
    double s = 0.0;
    for (int i = 0; i < P.G; ++i) s += P.h_part[i];
    *out = s;
    return RUV_OK;
}

int host_max(PlanImpl& P, const double* part, double* out) {
    HIPCHK(hipMemcpyAsync(P.h_part, part, sizeof(double) * P.G, hipMemcpyDeviceToHost, P.stream));
    HIPCHK(hipStreamSynchronize(P.stream));
    double m = 0.0;
    for (int i = 0; i < P.G; ++i) m = std::max(m, P.h_part[i]);
    *out = m;
    return RUV_OK;
}

// z = M^-1 r for the triangular preconditioners (forward over L, backward over L^T).
template <typename T, int R>
void apply_tri(PlanImpl& P, const DevTri<T>& f, const T* r, T* z, T* y, T* t, const SolveState* st) {
    const T* in = r;
    T* out = z;
    if (P.permuted) {
        plaunch(P, k_gather<T, R>, dim3(P.G), dim3(BLK), P.n, P.to_orig, r, t, st);
        in = t;
    }
    auto sweep = [&](const std::vector<typename DevTri<T>::Seg>& segs, const int* off, const int* rows,
                     const int* ptr, const int* cols, const T* vals, const T* x, T* o) {
        for (const auto& g : segs) {
            if (g.wide)
                plaunch(P, k_tri_level<T, R>, dim3((unsigned)((g.b - g.a + BLK - 1) / BLK)), dim3(BLK), g.a, g.b, rows,
                        ptr, cols, vals, x, o, st);
            else
                plaunch(P, k_tri_sweep<T, R>, dim3(1), dim3(TRI_BLK), g.l1 - g.l0, off + g.l0, rows, ptr, cols, vals, x,
                        o, st);
        }
    };
    sweep(f.fwd_seg, f.fwd_off, f.fwd_rows, f.l_ptr, f.l_cols, f.l_vals, in, y);
    sweep(f.bwd_seg, f.bwd_off, f.bwd_rows, f.lt_ptr, f.lt_cols, f.lt_vals, (const T*)y, P.permuted ? t : out);
    if (P.permuted) plaunch(P, k_gather<T, R>, dim3(P.G), dim3(BLK), P.n, P.to_perm, (const T*)t, out, st);
}

template <typename T> DevTri<T>& tri_of(PlanImpl& P);
template <> DevTri<double>& tri_of<double>(PlanImpl& P) { return P.tri64; }
template <> DevTri<float>& tri_of<float>(PlanImpl& P) { return P.tri32; }
template <typename T> DevAmg<T>& amg_of(PlanImpl& P);
template <> DevAmg<double>& amg_of<double>(PlanImpl& P) { return P.amg64; }
template <> DevAmg<float>& amg_of<float>(PlanImpl& P) { return P.amg32; }

int grid_rows(int n) { return grid_for(n); }

// One V(1,1) cycle from level l: e = M_l^-1 r (see amg_kernels.hpp).
template <typename T, int R>
void amg_cycle(PlanImpl& P, DevAmg<T>& A, size_t l, const T* r, T* e, const int32_t* done) {
    DevAmgLevel<T>& L = A.lv[l];
    const int g = grid_rows(L.n);
    if (L.coarsest) {
        if (L.dense_inv) {
            plaunch(P, k_amg_dense<T, R>, dim3((L.n + WAVES - 1) / WAVES), dim3(BLK), L.n, (const double*)L.dense_inv, r,
                    e, done);
        } else {
            plaunch(P, k_amg_diag<T, R>, dim3(g), dim3(BLK), L.n, (const T*)L.dinv, r, e, done);
        }
        return;
    }
    DevAmgLevel<T>& C = A.lv[l + 1];
    plaunch(P, k_amg_pre<T, R>, dim3(g), dim3(BLK), L.n, L.a_ptr, L.a_cols, (const T*)L.a_vals, (const T*)L.dinv,
            A.omega, r, L.t, done);
    plaunch(P, k_amg_spmv<T, R>, dim3(grid_rows(L.nc)), dim3(BLK), L.nc, L.r_ptr, L.r_cols, (const T*)L.r_vals,
            (const T*)L.t, C.rin, done);
    amg_cycle<T, R>(P, A, l + 1, C.rin, C.e, done);
    plaunch(P, k_amg_prolong<T, R>, dim3(g), dim3(BLK), L.n, L.p_ptr, L.p_cols, (const T*)L.p_vals, (const T*)C.e,
            (const T*)L.dinv, A.omega, r, L.e1, done);
    plaunch(P, k_amg_post<T, R>, dim3(g), dim3(BLK), L.n, L.a_ptr, L.a_cols, (const T*)L.a_vals, (const T*)L.dinv,
            A.omega, r, (const T*)L.e1, e, done);
}

// z = M^-1 r for the preconditioners that are not element-wise (IC(0) variants, AMG)
template <typename T, int R>
void apply_external(PlanImpl& P, const T* r, T* z, T* y, T* t, const SolveState* st) {
    if (P.precond == RUV_PC_AMG || P.precond == RUV_PC_AMGSA)
        amg_cycle<T, R>(P, amg_of<T>(P), 0, r, z, st ? &st->done : nullptr);
    else
        apply_tri<T, R>(P, tri_of<T>(P), r, z, y, t, st);
}

// PCG on the device.  On entry r holds the right-hand side (x0 = 0), x is overwritten.
// tol_abs_from: tol relative to ||r0|| (tol absolute when ||r0|| <= 1e-300), as solver.rs:465-468.
template <typename T, int R>
int run_pcg(PlanImpl& P, const T* vals, const T* dinv, T* x, T* r, T* z, T* p, T* q,
            T* y, T* t, uint64_t max_iter, double tol, uint64_t* iters_out, int* done_out) {
    const int G = P.G, n = P.n;
    hipStream_t s = P.stream;
    const int pc = P.precond;
    // x = 0; bb = r.r
    hipLaunchKernelGGL((k_fill0<T>), dim3(G), dim3(BLK), 0, s, (int64_t)n * R, x);
    hipLaunchKernelGGL((k_dot<T, R>), dim3(G), dim3(BLK), 0, s, n, (const T*)r, (const T*)r, P.part_a, (const SolveState*)nullptr);
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(BLK), 0, s, P.part_a, G, tol, max_iter, P.st);
    // z0 = M^-1 r0; p = z0; part_c = r.z
    if (pc == RUV_PC_JACOBI) {
        hipLaunchKernelGGL((k_jacobi<T, R>), dim3(G), dim3(BLK), 0, s, n, dinv, (const T*)r, z);
    } else if (pc != RUV_PC_NONE) {
        apply_external<T, R>(P, r, z, y, t, nullptr);
    }
    const T* zz = (pc == RUV_PC_NONE) ? (const T*)r : (const T*)z;
    hipLaunchKernelGGL((k_copy<T>), dim3(G), dim3(BLK), 0, s, (int64_t)n * R, zz, p);
    hipLaunchKernelGGL((k_dot<T, R>), dim3(G), dim3(BLK), 0, s, n, (const T*)r, zz, P.part_c, (const SolveState*)nullptr);
    HIPCHK(hipGetLastError());

    auto body = [&](int iters) {
        for (int it = 0; it < iters; ++it) {
            plaunch(P, k_spmv_pq<T, R>, dim3(G), dim3(BLK), n, P.ptr, P.cols, vals, (const T*)p, q, P.part_a,
                    (const double*)P.part_c, G, P.st);
            if (pc == RUV_PC_JACOBI) {
                plaunch(P, k_update<T, R, 1>, dim3(G), dim3(BLK), n, (const T*)p, (const T*)q, x, r, z, dinv,
                        (const double*)P.part_a, P.part_b, P.part_c, G, P.st);
                plaunch(P, k_pupdate<T, R, 1>, dim3(G), dim3(BLK), n, (const T*)z, p, (const double*)P.part_b,
                        (const double*)P.part_c, G, P.st);
            } else if (pc == RUV_PC_NONE) {
                plaunch(P, k_update<T, R, 0>, dim3(G), dim3(BLK), n, (const T*)p, (const T*)q, x, r, z, dinv,
                        (const double*)P.part_a, P.part_b, P.part_c, G, P.st);
                plaunch(P, k_pupdate<T, R, 0>, dim3(G), dim3(BLK), n, (const T*)r, p, (const double*)P.part_b,
                        (const double*)P.part_c, G, P.st);
            } else {
                plaunch(P, k_update<T, R, 2>, dim3(G), dim3(BLK), n, (const T*)p, (const T*)q, x, r, z, dinv,
                        (const double*)P.part_a, P.part_b, P.part_c, G, P.st);
                apply_external<T, R>(P, r, z, y, t, P.st);
                plaunch(P, k_dot<T, R>, dim3(G), dim3(BLK), n, (const T*)r, (const T*)z, P.part_c,
                        (const SolveState*)P.st);
                plaunch(P, k_pupdate<T, R, 2>, dim3(G), dim3(BLK), n, (const T*)z, p, (const double*)P.part_b,
                        (const double*)P.part_c, G, P.st);
            }
        }
    };
    // A block of GRAPH_ITERS iterations is built once per plan into a HIP graph (one chain of kernel nodes, no
    // memset / memcpy nodes) and replayed: the kernel arguments (device pointers, the state block) never change
    // between solves of a plan.
    hipGraphExec_t gexec = nullptr;
    if (P.use_graph) {
        if (!P.gexec) {
            PlanImpl::GraphRec rec;
            HIPCHK(hipGraphCreate(&rec.g, 0));
            P.rec = &rec;
            body(GRAPH_ITERS);
            P.rec = nullptr;
            hipError_t e = rec.err;
            if (e == hipSuccess) e = hipGraphInstantiate(&P.gexec, rec.g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(rec.g);
            if (e != hipSuccess) return fail(RUV_ERR_DEVICE, std::string("graph build: ") + hipGetErrorString(e));
        }
        gexec = P.gexec;
    }
    // Iterations in flight between polls: start small, then aim at the predicted remainder from the decay of
    // the recursive residual so that the post-convergence overshoot (early-exiting kernels) stays small.
    uint64_t launched = 0;
    uint64_t next = gexec ? GRAPH_ITERS : 4;
    double prev_rr = -1.0;
    uint64_t prev_it = 0;
    for (;;) {
        if (gexec) {
            const uint64_t reps = (next + GRAPH_ITERS - 1) / GRAPH_ITERS;
            for (uint64_t k = 0; k < reps; ++k) HIPCHK(hipGraphLaunch(gexec, s));
            launched += reps * GRAPH_ITERS;
        } else {
            body((int)next);
            launched += next;
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(P.h_st, P.st, sizeof(SolveState), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (P.h_st->done != RUNNING) break;
        if (launched > max_iter + 1) {  // cannot happen: the device stops at max_iter
            return fail(RUV_ERR_NOT_CONVERGED, "pcg: iteration accounting overrun");
        }
        const double rr = P.h_st->rr, tol2 = P.h_st->tol_abs * P.h_st->tol_abs;
        const uint64_t it = P.h_st->iters;
        uint64_t want = std::min<uint64_t>(2 * next, 256);
        if (prev_rr > 0.0 && rr > 0.0 && rr < prev_rr && it > prev_it && tol2 > 0.0) {
            const double rate = std::log(rr / prev_rr) / (double)(it - prev_it);  // < 0
            const double need = std::log(tol2 / rr) / rate;
            if (std::isfinite(need) && need > 0.0) want = (uint64_t)(need * 1.05) + 2;
        }
        prev_rr = rr;
        prev_it = it;
        next = std::max<uint64_t>(4, std::min<uint64_t>(want, std::max<uint64_t>(8, 2 * it)));
    }
    *iters_out = P.h_st->iters;
    *done_out = P.h_st->done;
    return RUV_OK;
}

int status_of(int done, uint64_t iters) {
    if (done == CONVERGED) return RUV_OK;
    if (done == BREAKDOWN) return fail(RUV_ERR_BREAKDOWN, "pcg: breakdown (|p.Ap| <= 1e-300 or non-finite) at iteration " + std::to_string(iters));
    return fail(RUV_ERR_NOT_CONVERGED, "pcg: not converged after " + std::to_string(iters) + " iterations");
}

template <int R>
int true_residual(PlanImpl& P, const double* b, const double* x, double* rel) {
    hipLaunchKernelGGL((k_residual<R>), dim3(P.G), dim3(BLK), 0, P.stream, P.n, P.ptr, P.cols, P.vals64, x, b, P.r64,
                       P.part_a);
    hipLaunchKernelGGL((k_dot<double, R>), dim3(P.G), dim3(BLK), 0, P.stream, P.n, b, b, P.part_b, (const SolveState*)nullptr);
    HIPCHK(hipGetLastError());
    double rr = 0, bb = 0;
    int rc = host_sum(P, P.part_a, &rr);
    if (rc) return rc;
    rc = host_sum(P, P.part_b, &bb);
    if (rc) return rc;
    *rel = bb > 0 ? std::sqrt(rr) / std::sqrt(bb) : std::sqrt(rr);
    return RUV_OK;
}

template <int R>
int solve_fp64(PlanImpl& P, const double* d_b, double* d_x, uint64_t max_iter, double tol, ruv_solve_info* info) {
    hipStream_t s = P.stream;
    HIPCHK(hipMemcpyAsync(P.r64, d_b, sizeof(double) * (size_t)P.n * R, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipEventRecord(P.ev0, s));
    uint64_t iters = 0;
    int done = 0;
    // the iterate lives in the plan's own x64: the replayed graph's kernels keep the pointers of its capture, and
    // the caller's d_x may differ from one solve to the next
    int rc = run_pcg<double, R>(P, P.vals64, P.dinv64, P.x64, P.r64, P.z64, P.p64, P.q64, P.y64, P.t64,
                                max_iter, tol, &iters, &done);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(d_x, P.x64, sizeof(double) * (size_t)P.n * R, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipEventRecord(P.ev1, s));
    HIPCHK(hipEventSynchronize(P.ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, P.ev0, P.ev1));
    info->iters = iters;
    info->refinements = 0;
    info->precond = P.precond;
    info->seconds = ms * 1e-3;
    rc = status_of(done, iters);
    if (rc) return rc;
    return true_residual<R>(P, d_b, d_x, &info->true_rel_resid);
}

template <int R>
int solve_mixed(PlanImpl& P, const double* d_b, double* d_x, uint64_t max_iter, double tol, ruv_solve_info* info) {
    hipStream_t s = P.stream;
    const int G = P.G, n = P.n;
    const int64_t N = (int64_t)n * R;
    HIPCHK(hipEventRecord(P.ev0, s));
    hipLaunchKernelGGL((k_fill0<double>), dim3(G), dim3(BLK), 0, s, N, d_x);
    HIPCHK(hipMemcpyAsync(P.r64, d_b, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL((k_dot<double, R>), dim3(G), dim3(BLK), 0, s, n, d_b, d_b, P.part_a, (const SolveState*)nullptr);
    double bb = 0;
    int rc = host_sum(P, P.part_a, &bb);
    if (rc) return rc;
    const double bnorm = std::sqrt(bb);
    const double target = (bnorm <= 1e-300) ? tol : tol * bnorm;
    double rr = bb;
    uint64_t total = 0;
    uint32_t refs = 0;
    // Outer test: ||r|| <= tol ||b|| and, when uv_tol > 0, the forward-error estimate gamma ||r|| <= uv_tol max|x|,
    // gamma the largest ratio max|correction| / ||r_start|| of a round so far (the correction of a round solved
    // A d = r_start); i.e. ||r|| <= goal = min(target, uv_tol max|x| / gamma), never below 1e-4 target.  A round
    // that fails to halve a residual already at the target ends the loop (the FP64 residual has stagnated).
    double goal = target, gamma = 0.0, rstart = std::sqrt(bb);
    for (;;) {
        const double rn = std::sqrt(rr);
        if (rn <= goal) break;
        if (refs > 0 && rn <= target && rn > 0.5 * rstart) break;
        if (refs >= (uint32_t)P.max_refine) {
            return fail(RUV_ERR_NOT_CONVERGED, "mixed: refinement stalled at rel. residual " +
                                                   std::to_string(std::sqrt(rr) / std::max(bnorm, 1e-300)));
        }
        if (total >= max_iter) return fail(RUV_ERR_NOT_CONVERGED, "mixed: inner iteration budget exhausted");
        rstart = rn;
        // r32 = fp32(r64); inner FP32 PCG to inner_tol relative to ||r32||
        hipLaunchKernelGGL((k_to_f32), dim3(G), dim3(BLK), 0, s, N, (const double*)P.r64, P.r32);
        uint64_t it = 0;
        int done = 0;
        double itol = std::max(P.inner_tol, 0.25 * goal / std::max(rn, 1e-300));
        rc = run_pcg<float, R>(P, P.vals32, P.dinv32, P.x32, P.r32, P.z32, P.p32, P.q32, P.y32, P.t32,
                               max_iter - total, itol, &it, &done);
        if (rc) return rc;
        total += it;
        refs++;
        if (done == BREAKDOWN && it <= 1) return status_of(done, it);
        // x += d; r = b - A x (FP64) ; rr
        hipLaunchKernelGGL((k_axpy_f32), dim3(G), dim3(BLK), 0, s, N, (const float*)P.x32, d_x, P.part_b, P.part_c);
        hipLaunchKernelGGL((k_residual<R>), dim3(G), dim3(BLK), 0, s, n, P.ptr, P.cols, P.vals64, (const double*)d_x,
                           d_b, P.r64, P.part_a);
        HIPCHK(hipGetLastError());
        rc = host_sum(P, P.part_a, &rr);
        if (rc) return rc;
        if (!std::isfinite(rr)) return fail(RUV_ERR_NONFINITE, "mixed: non-finite residual");
        if (P.uv_tol > 0.0) {
            double dmax = 0.0, xmax = 0.0;
            rc = host_max(P, P.part_b, &dmax);
            if (!rc) rc = host_max(P, P.part_c, &xmax);
            if (rc) return rc;
            if (rstart > 0.0) gamma = std::max(gamma, dmax / rstart);
            if (gamma > 0.0) goal = std::max(std::min(target, P.uv_tol * xmax / gamma), 1e-4 * target);
        }
    }
    HIPCHK(hipEventRecord(P.ev1, s));
    HIPCHK(hipEventSynchronize(P.ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, P.ev0, P.ev1));
    info->iters = total;
    info->refinements = refs;
    info->precond = P.precond;
    info->seconds = ms * 1e-3;
    info->true_rel_resid = bnorm > 0 ? std::sqrt(rr) / bnorm : std::sqrt(rr);
    return RUV_OK;
}

}  // namespace

int plan_create(const Csr& a, int R, int precond, int mode, PlanImpl** out) {
    if (R != 1 && R != 2) return fail(RUV_ERR_INVALID, "rhs_per_row must be 1 or 2");
    if (mode != RUV_MODE_FP64 && mode != RUV_MODE_MIXED) return fail(RUV_ERR_INVALID, "unknown mode");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(RUV_ERR_UNAVAILABLE, "no HIP device");
    std::unique_ptr<PlanImpl> P(new PlanImpl());
    int rc = plan_setup(*P, a, R, precond, mode);
    if (rc) return rc;
    *out = P.release();
    return RUV_OK;
}

void plan_destroy(PlanImpl* p) { delete p; }

// The IC(0) factor of a (natural order) as the device computes it: L's rows (strict-lower entries ascending, the
// diagonal last).
int ic0_factor_device(const Csr& a, std::vector<int32_t>* ptr, std::vector<int32_t>* cols, std::vector<double>* vals) {
    PlanImpl* P = nullptr;
    int rc = plan_create(a, 1, RUV_PC_IC0, RUV_MODE_FP64, &P);
    if (rc) return rc;
    std::unique_ptr<PlanImpl> hold(P);
    const DevTri<double>& d = P->tri64;
    const size_t nl = (size_t)P->ic0_nl;
    ptr->resize((size_t)d.n + 1);
    cols->resize(nl);
    vals->resize(nl);
    HIPCHK(hipMemcpyAsync(ptr->data(), d.l_ptr, sizeof(int32_t) * ptr->size(), hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipMemcpyAsync(cols->data(), d.l_cols, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipMemcpyAsync(vals->data(), d.l_vals, sizeof(double) * nl, hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    return RUV_OK;
}

int plan_update_values(PlanImpl* P, const double* vals) {
    if (P->precond == RUV_PC_AMG || P->precond == RUV_PC_AMGSA)
        return fail(RUV_ERR_INVALID, "the amg preconditioners are built on the host: create a new plan");
    hipStream_t s = P->stream;
    const size_t nz = (size_t)P->nnz;
    HIPCHK(hipStreamSynchronize(s));  // no solve of this plan is in flight
    HIPCHK(hipMemcpyAsync(P->vals64, vals, sizeof(double) * nz, hipMemcpyHostToDevice, s));
    std::vector<float> v32;
    if (P->mode == RUV_MODE_MIXED) {
        v32.assign(vals, vals + nz);
        HIPCHK(hipMemcpyAsync(P->vals32, v32.data(), sizeof(float) * nz, hipMemcpyHostToDevice, s));
    }
    int rc = RUV_OK;
    if (P->precond == RUV_PC_JACOBI) rc = jacobi_upload(*P, vals);
    else if (P->has_tri) rc = ic0_refactor(*P);
    HIPCHK(hipStreamSynchronize(s));
    return rc;
}

int plan_solve_dev(PlanImpl* P, const double* d_b, double* d_x, uint64_t max_iter, double tol, ruv_solve_info* info) {
    ruv_solve_info tmp{};
    if (!info) info = &tmp;
    if (P->mode == RUV_MODE_FP64) return P->R == 2 ? solve_fp64<2>(*P, d_b, d_x, max_iter, tol, info)
                                                   : solve_fp64<1>(*P, d_b, d_x, max_iter, tol, info);
    return P->R == 2 ? solve_mixed<2>(*P, d_b, d_x, max_iter, tol, info) : solve_mixed<1>(*P, d_b, d_x, max_iter, tol, info);
}

int plan_solve_host(PlanImpl* P, const double* b, double* x, uint64_t max_iter, double tol, ruv_solve_info* info) {
    const size_t N = (size_t)P->n * P->R;
    double *db = nullptr, *dx = nullptr;
    HIPCHK(hipMallocAsync((void**)&db, N * sizeof(double), P->stream));
    HIPCHK(hipMallocAsync((void**)&dx, N * sizeof(double), P->stream));
    HIPCHK(hipMemcpyAsync(db, b, N * sizeof(double), hipMemcpyHostToDevice, P->stream));
    int rc = plan_solve_dev(P, db, dx, max_iter, tol, info);
    if (rc == RUV_OK) {
        hipError_t e = hipMemcpyAsync(x, dx, N * sizeof(double), hipMemcpyDeviceToHost, P->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(P->stream);
        if (e != hipSuccess) rc = fail(RUV_ERR_DEVICE, std::string("copy x: ") + hipGetErrorString(e));
    }
    (void)hipFreeAsync(db, P->stream);
    (void)hipFreeAsync(dx, P->stream);
    (void)hipStreamSynchronize(P->stream);
    if (rc == RUV_OK) {
        for (size_t i = 0; i < N; ++i)
            if (!std::isfinite(x[i])) return fail(RUV_ERR_NONFINITE, "non-finite solution");
    }
    return rc;
}

int plan_spmv_time(PlanImpl* P, int reps, double* sec, double* bytes) {
    if (reps <= 0) return fail(RUV_ERR_INVALID, "reps must be positive");
    hipStream_t s = P->stream;
    const int64_t N = (int64_t)P->n * P->R;
    const bool f32 = P->mode == RUV_MODE_MIXED;
    // a converged-free state so the kernel does its full work; p = 1
    hipLaunchKernelGGL((k_fill0<double>), dim3(P->G), dim3(BLK), 0, s, (int64_t)MAXG, P->part_c);
    SolveState init{};
    init.max_iter = ~0ull;
    HIPCHK(hipMemcpyAsync(P->st, &init, sizeof(init), hipMemcpyHostToDevice, s));
    if (f32) hipLaunchKernelGGL((k_fill<float>), dim3(P->G), dim3(BLK), 0, s, N, 1.0f, P->p32);
    else hipLaunchKernelGGL((k_fill<double>), dim3(P->G), dim3(BLK), 0, s, N, 1.0, P->p64);
    HIPCHK(hipStreamSynchronize(s));
    auto launch = [&]() {
        if (f32) {
            if (P->R == 2) hipLaunchKernelGGL((k_spmv_pq<float, 2>), dim3(P->G), dim3(BLK), 0, s, P->n, P->ptr, P->cols, (const float*)P->vals32, (const float*)P->p32, P->q32, P->part_a, (const double*)P->part_c, P->G, P->st);
            else hipLaunchKernelGGL((k_spmv_pq<float, 1>), dim3(P->G), dim3(BLK), 0, s, P->n, P->ptr, P->cols, (const float*)P->vals32, (const float*)P->p32, P->q32, P->part_a, (const double*)P->part_c, P->G, P->st);
        } else {
            if (P->R == 2) hipLaunchKernelGGL((k_spmv_pq<double, 2>), dim3(P->G), dim3(BLK), 0, s, P->n, P->ptr, P->cols, (const double*)P->vals64, (const double*)P->p64, P->q64, P->part_a, (const double*)P->part_c, P->G, P->st);
            else hipLaunchKernelGGL((k_spmv_pq<double, 1>), dim3(P->G), dim3(BLK), 0, s, P->n, P->ptr, P->cols, (const double*)P->vals64, (const double*)P->p64, P->q64, P->part_a, (const double*)P->part_c, P->G, P->st);
        }
    };
    launch();
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(P->ev0, s));
    for (int i = 0; i < reps; ++i) launch();
    HIPCHK(hipEventRecord(P->ev1, s));
    HIPCHK(hipEventSynchronize(P->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, P->ev0, P->ev1));
    *sec = ms * 1e-3 / reps;
    const double sv = f32 ? 4.0 : 8.0;
    *bytes = (double)P->nnz * (sv + 4.0) + (double)(P->n + 1) * 4.0 + 2.0 * (double)N * sv;
    return RUV_OK;
}

int device_solve(const Csr& a, int R, const double* b, uint64_t max_iter, double tol, int precond, int mode,
                 double* x, ruv_solve_info* info) {
    PlanImpl* P = nullptr;
    int rc = plan_create(a, R, precond, mode, &P);
    if (rc) return rc;
    rc = plan_solve_host(P, b, x, max_iter, tol, info);
    plan_destroy(P);
    return rc;
}

int device_apply_ic0(const Csr& a, int R, const double* r, double* z) {
    PlanImpl* P = nullptr;
    int rc = plan_create(a, R, RUV_PC_IC0, RUV_MODE_FP64, &P);
    if (rc) return rc;
    std::unique_ptr<PlanImpl> hold(P);
    const size_t N = (size_t)P->n * R;
    HIPCHK(hipMemcpyAsync(P->r64, r, N * sizeof(double), hipMemcpyHostToDevice, P->stream));
    if (R == 2) apply_tri<double, 2>(*P, P->tri64, P->r64, P->z64, P->y64, P->t64, nullptr);
    else apply_tri<double, 1>(*P, P->tri64, P->r64, P->z64, P->y64, P->t64, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(z, P->z64, N * sizeof(double), hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    return RUV_OK;
}

}  // namespace ruv
