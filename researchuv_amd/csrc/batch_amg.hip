// batch_amg.hip -- builds the batched aggregation AMG on the device and applies its V(1,1) cycle
// (kernels and the method in bamg_kernels.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "bamg_kernels.hpp"
#include "batch_impl.hpp"

namespace ruv {

struct BamgLevel {
    int64_t n = 0, nell = 0, nsl = 0;
    int64_t nnz = 0;          // the operator's nonzeros (levels between the finest and the coarsest)
    int64_t* eoff = nullptr;  // sliced ELL of the operator (levels between the finest and the coarsest)
    int32_t* ew = nullptr;
    int32_t* srow = nullptr;  // slice row starts, at aggregate starts
    int32_t* sagg = nullptr;  // first next-level aggregate of every slice
    int2* sblk = nullptr;     // per slice: first and last chart
    void* ecols = nullptr;
    int ci = 1;               // 0: 16-bit column offsets
    void* evals = nullptr;    // T
    void* dinv = nullptr;     // T (level 0: the batch's Jacobi diagonal; the coarsest level: its inverse diagonal)
    void* q = nullptr;        // T, scaled levels (between the finest and the coarsest): sqrt(d_i), 1 / sqrt(d_i)
    void* rs = nullptr;
    int32_t* agg = nullptr;   // row -> next-level aggregate
    int32_t* mptr = nullptr;  // next level's aggregates -> member rows here (levels >= 1: first stored row)
    int32_t* mrows = nullptr; // (the finest level: the member rows)
    void* b = nullptr;
    void* z = nullptr;
    void* t = nullptr;
    const int32_t* chart = nullptr;  // row -> chart
    int2* blk = nullptr;             // per BLK rows: first and last chart
    double omega = 2.0 / 3.0;        // Jacobi weight of this level's smoothing steps
};

struct Bamg {
    std::vector<BamgLevel> lv;  // lv.back() is the coarsest
    void* ev0 = nullptr;        // the finest operator's values for the smoothing legs (CoarseVal<T>::type)
    void* bv0 = nullptr;        // ... and their band form (band charts)
    double* cinv = nullptr;     // each chart's inverted block of the coarsest operator, FP64 (null: Jacobi there)
    int64_t* cinv_off = nullptr;  // per chart: its block's first entry, -1 for a chart kept on its diagonal
    int64_t* crow_c = nullptr;  // chart row offsets on the coarsest level
    std::vector<int64_t> h_cinv_off, h_crow_c;  // their host sources (alive until the build's last wait)
    unsigned long long* chol_fallback = nullptr;  // [2] charts left on their diagonal: not SPD, coupled
    SetupCsr coarse{};  // the coarsest operator (its memory owned here, for the inspection hook)
    int64_t chol_too_large = 0;  // charts whose coarsest block exceeds the device factorisation
    // Jacobi weight of the smoothing steps: omega_max, or less where a level's Gershgorin bound g on D^-1 A
    // asks for it (omega <= 1.9 / g keeps each smoother an A-norm contraction, so the cycle stays SPD)
    double omega_max = 0.9;
    // coarse-grid corrections enter scaled by alpha (over-correction; the cycle stays SPD for any alpha > 0, since
    // its error operator S (I - alpha P Bc P^T A) S has every eigenvalue below 1 when S is a strict A-contraction):
    // with aggregates of about two rows the plain correction is too short, and alpha = 1.2 cut the PCG
    // iterations by about a third on height fields (49 -> 34 on a 256^2 patch in a CPU model of the cycle)
    double alpha = 1.2;
    // levels of at most this many rows (live rows, late in a solve) fold the prolongation into the post-smoothing
    // pass (kc_post<FUSED>)
    int64_t fuse_rows = 65536;
    double setup_seconds = 0.0;
    // Lists of the slices and BLK-row blocks of the running charts per level, for launches late in a solve
    // (empty vectors / null pointers: launches cover the whole level).  cs0/cs1 (cb0/cb1): per level, each chart's
    // slices (blocks) [first, last + 1); the lists live in one device buffer filled from pinned host memory.
    std::vector<std::vector<int32_t>> cs0, cs1, cb0, cb1;
    std::vector<int32_t*> slist, blist;
    std::vector<int64_t> nslist, nblist;
    int32_t* d_lists = nullptr;
    int32_t* h_lists = nullptr;  // the batch's pinned staging buffer (not owned): two halves of lists_cap
    size_t lists_cap = 0;
    int lists_half = 0;
    DevPool* pool = nullptr;
    std::vector<void*> mem;
    explicit Bamg(DevPool* p) : pool(p) {}
    ~Bamg() {
        for (void* p : mem) pool->put(p);
    }
    template <typename T>
    hipError_t alloc(T** p, size_t n) {
        void* q = nullptr;
        hipError_t e = pool->get(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) mem.push_back(q);
        *p = (T*)q;
        return e;
    }
};

void bamg_destroy(Bamg* a) { delete a; }

namespace {

inline dim3 rows_grid(int64_t n) { return dim3((unsigned)std::max<int64_t>(1, (n + BLK - 1) / BLK)); }

// temporary device buffers of one setup step, returned to the pool together
struct Scratch {
    DevPool* pool;
    std::vector<void*> mem;
    explicit Scratch(DevPool* p) : pool(p) {}
    ~Scratch() {
        for (void* p : mem) pool->put(p);
    }
    template <typename T>
    hipError_t alloc(T** p, size_t n) {
        void* q = nullptr;
        hipError_t e = pool->get(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) mem.push_back(q);
        *p = (T*)q;
        return e;
    }
};

// move the buffer p from one owner's list to another's
void hand_over(Scratch& from, void* p, std::vector<void*>& to) {
    from.mem.erase(std::remove(from.mem.begin(), from.mem.end(), p), from.mem.end());
    to.push_back(p);
}

constexpr int64_t SCAN_FEW = 4096;
// out[0..n) = exclusive scan of in, out[n] = total
int scan(hipStream_t s, int64_t n, const int64_t* in, int64_t* out, Scratch& S) {
    const int64_t nb = std::max<int64_t>(1, (n + SCAN_BLK - 1) / SCAN_BLK);
    int64_t* tot = nullptr;
    HIPCHK(S.alloc(&tot, (size_t)nb + 1));
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(BLK), 0, s, n, in, tot);
    // up to SCAN_FEW blocks each add up the block sums before their own; more get them scanned in a launch
    const int scanned = nb > SCAN_FEW ? 1 : 0;
    if (scanned) hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(TOT_BLK), 0, s, nb, tot);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb), dim3(BLK), 0, s, n, nb, in, (const int64_t*)tot, out,
                       scanned);
    HIPCHK(hipGetLastError());
    return RUV_OK;
}

// One matching of A within groups (pairs, then lone rows joining pairs) and its Galerkin operator.
struct Matched {
    int64_t nc = 0;
    int64_t* cid = nullptr;  // exclusive scan of root flags, n + 1
    int32_t* agg = nullptr;
    int64_t* mptr = nullptr;
    int64_t* mrows = nullptr;
    int32_t* group = nullptr;  // of the aggregates
    SetupCsr C{};
    double* cdiag = nullptr;
    double gersh = 0.0;  // Gershgorin bound on D^-1 A of the matched operator
};

int match_and_coarsen(hipStream_t s, const SetupCsr& A, const BandSrc& band, const int32_t* group, double theta,
                      int rounds, Scratch& S, Matched* M) {
    const int64_t n = A.n;
    int32_t *cand = nullptr, *ncand = nullptr;
    int64_t *prop = nullptr, *flag = nullptr, *mate = nullptr, *join = nullptr;
    unsigned long long* gmax = nullptr;
    HIPCHK(S.alloc(&gmax, 1));
    HIPCHK(hipMemsetAsync(gmax, 0, sizeof(unsigned long long), s));
    HIPCHK(S.alloc(&cand, (size_t)n * MCAND));
    HIPCHK(S.alloc(&ncand, n));
    HIPCHK(S.alloc(&prop, n));
    HIPCHK(S.alloc(&flag, n));
    HIPCHK(S.alloc(&mate, n));
    HIPCHK(S.alloc(&join, n));
    HIPCHK(S.alloc(&M->cid, n + 1));
    HIPCHK(S.alloc(&M->agg, n));
    HIPCHK(hipMemsetAsync(mate, 0xff, sizeof(int64_t) * n, s));  // -1: unmatched
    hipLaunchKernelGGL(k_candidates, rows_grid(n), dim3(BLK), 0, s, A, band, group, theta, cand, ncand, prop, gmax);
    for (int r = 0; r < rounds; ++r) {
        if (r > 0)
            hipLaunchKernelGGL(k_propose, rows_grid(n), dim3(BLK), 0, s, n, (const int32_t*)cand,
                               (const int32_t*)ncand, (const int64_t*)mate, prop);
        hipLaunchKernelGGL(k_accept, rows_grid(n), dim3(BLK), 0, s, n, (const int64_t*)prop, mate);
    }
    hipLaunchKernelGGL(k_join, rows_grid(n), dim3(BLK), 0, s, n, (const int32_t*)cand, (const int32_t*)ncand,
                       (const int64_t*)mate, join);
    hipLaunchKernelGGL(k_roots, rows_grid(n), dim3(BLK), 0, s, n, (const int64_t*)mate, (const int64_t*)join, flag);
    HIPCHK(hipGetLastError());
    int rc = scan(s, n, flag, M->cid, S);
    if (rc) return rc;
    unsigned long long hg = 0;
    HIPCHK(hipMemcpyAsync(&M->nc, M->cid + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&hg, gmax, sizeof(hg), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::memcpy(&M->gersh, &hg, sizeof(double));
    const int64_t nc = M->nc;
    {  // aggregate map, sizes, member lists (ascending) and each aggregate's group
        unsigned long long *cnt = nullptr, *cursor = nullptr;
        Scratch T(S.pool);
        HIPCHK(T.alloc(&cnt, nc));
        HIPCHK(T.alloc(&cursor, nc));
        HIPCHK(S.alloc(&M->group, nc));
        HIPCHK(S.alloc(&M->mptr, nc + 1));
        HIPCHK(S.alloc(&M->mrows, n));
        HIPCHK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * nc, s));
        HIPCHK(hipMemsetAsync(cursor, 0, sizeof(unsigned long long) * nc, s));
        hipLaunchKernelGGL(k_root_agg, rows_grid(n), dim3(BLK), 0, s, n, (const int64_t*)mate, (const int64_t*)join,
                           (const int64_t*)M->cid, M->agg, cnt);
        HIPCHK(hipGetLastError());
        rc = scan(s, nc, (const int64_t*)cnt, M->mptr, T);
        if (rc) return rc;
        hipLaunchKernelGGL(k_member_fill, rows_grid(n), dim3(BLK), 0, s, n, (const int64_t*)mate,
                           (const int64_t*)join, (const int32_t*)M->agg, (const int64_t*)M->mptr, cursor, M->mrows);
        hipLaunchKernelGGL(k_member_sort, rows_grid(nc), dim3(BLK), 0, s, nc, (const int64_t*)M->mptr, M->mrows,
                           group, M->group);
        HIPCHK(hipGetLastError());
    }  // the counters go back to the pool; their next user is ordered after these kernels
    // Galerkin operator: rows of few distinct columns in registers (16 slots, then 32 for the rows listed by the
    // first pass), the others (listed again) through LDS; the register passes stage their rows while counting
    // them (columns in first-appearance order), and the staged rows are written out once the offsets are known
    int64_t *cnt = nullptr, *cptr = nullptr, *over1 = nullptr, *over2 = nullptr;
    unsigned long long* nover = nullptr;
    int32_t *t1c = nullptr, *t2c = nullptr;
    float *t1v = nullptr, *t2v = nullptr;
    HIPCHK(S.alloc(&cnt, nc));
    HIPCHK(S.alloc(&cptr, nc + 1));
    HIPCHK(S.alloc(&over1, nc));
    HIPCHK(S.alloc(&over2, nc));
    HIPCHK(S.alloc(&nover, 2));
    HIPCHK(S.alloc(&M->cdiag, nc));
    Scratch TS(S.pool);  // staging, back to the pool when this level is built
    HIPCHK(TS.alloc(&t1c, (size_t)GR_SLOTS_FIRST * nc));
    HIPCHK(TS.alloc(&t1v, (size_t)GR_SLOTS_FIRST * nc));
    HIPCHK(hipMemsetAsync(nover, 0, 2 * sizeof(unsigned long long), s));
    unsigned long long hover[2] = {0, 0};
    hipLaunchKernelGGL((k_galerkin_stage<GR_SLOTS_FIRST>), rows_grid(nc), dim3(BLK), 0, s, nc,
                       (const int64_t*)nullptr, A, (const int64_t*)M->mptr, (const int64_t*)M->mrows,
                       (const int32_t*)M->agg, cnt, M->cdiag, nover, over1, t1c, t1v);
    HIPCHK(hipMemcpyAsync(&hover[0], nover, sizeof(hover[0]), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const int64_t n1 = (int64_t)hover[0];
    if (n1 > 0) {
        HIPCHK(TS.alloc(&t2c, (size_t)GR_SLOTS * n1));
        HIPCHK(TS.alloc(&t2v, (size_t)GR_SLOTS * n1));
        hipLaunchKernelGGL((k_galerkin_stage<GR_SLOTS>), rows_grid(n1), dim3(BLK), 0, s, n1,
                           (const int64_t*)over1, A, (const int64_t*)M->mptr, (const int64_t*)M->mrows,
                           (const int32_t*)M->agg, cnt, M->cdiag, nover + 1, over2, t2c, t2v);
        HIPCHK(hipMemcpyAsync(&hover[1], nover + 1, sizeof(hover[1]), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    const int64_t no = (int64_t)hover[1];
    const dim3 go((unsigned)std::max<int64_t>(1, (no + GAL_BLK - 1) / GAL_BLK));
    if (no > 0)
        hipLaunchKernelGGL((k_galerkin<false>), go, dim3(GAL_BLK), 0, s, no, A, (const int64_t*)M->mptr,
                           (const int64_t*)M->mrows, (const int32_t*)M->agg, cnt, (const int64_t*)nullptr,
                           (int32_t*)nullptr, (float*)nullptr, (double*)nullptr, (const int64_t*)over2);
    HIPCHK(hipGetLastError());
    rc = scan(s, nc, cnt, cptr, S);
    if (rc) return rc;
    int64_t cnnz = 0;
    HIPCHK(hipMemcpyAsync(&cnnz, cptr + nc, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (cnnz >= INT32_MAX) return fail(RUV_ERR_PRECOND, "batch amg: coarse operator too large");
    int32_t* ccols = nullptr;
    float* cvals = nullptr;
    HIPCHK(S.alloc(&ccols, cnnz));
    HIPCHK(S.alloc(&cvals, cnnz));
    hipLaunchKernelGGL(k_galerkin_unstage, rows_grid(nc), dim3(BLK), 0, s, nc, (const int64_t*)nullptr,
                       (const int32_t*)t1c, (const float*)t1v, (const int64_t*)cptr, ccols, cvals);
    if (n1 > 0)
        hipLaunchKernelGGL(k_galerkin_unstage, rows_grid(n1), dim3(BLK), 0, s, n1, (const int64_t*)over1,
                           (const int32_t*)t2c, (const float*)t2v, (const int64_t*)cptr, ccols, cvals);
    if (no > 0)
        hipLaunchKernelGGL((k_galerkin<true>), go, dim3(GAL_BLK), 0, s, no, A, (const int64_t*)M->mptr,
                           (const int64_t*)M->mrows, (const int32_t*)M->agg, (int64_t*)nullptr,
                           (const int64_t*)cptr, ccols, cvals, M->cdiag, (const int64_t*)over2);
    HIPCHK(hipGetLastError());
    M->C = SetupCsr{nc, cptr, ccols, cvals};
    return RUV_OK;
}

// chart row offsets on the next level
__global__ void k_gather_offsets(int64_t m, const int64_t* __restrict__ at, const int64_t* __restrict__ cid,
                                 int64_t* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (k < m) out[k] = cid[at[k]];
}

// Factor and invert each chart's block of the coarsest operator on the device (k_coarsest_chol, FP64): offsets of
// the blocks and of the chart rows from the host's copy of the level's chart offsets, nothing read back.
int coarsest_device(BatchImpl& Q, Bamg& A, const SetupCsr& cur, const std::vector<int64_t>& crow) {
    hipStream_t s = Q.stream;
    const int C = Q.nchart;
    A.h_cinv_off.assign((size_t)C + 1, 0);
    A.h_crow_c.assign(crow.begin(), crow.end());
    int64_t total = 0;
    int maxm = 0;
    A.chol_too_large = 0;
    for (int c = 0; c < C; ++c) {
        const int64_t m = crow[c + 1] - crow[c];
        if (m > CHOL_MAX) {
            A.h_cinv_off[c] = -1;
            ++A.chol_too_large;
            continue;
        }
        A.h_cinv_off[c] = total;
        total += m * m;
        maxm = std::max(maxm, (int)m);
    }
    A.h_cinv_off[C] = total;
    HIPCHK(A.alloc(&A.cinv, (size_t)std::max<int64_t>(total, 1)));
    HIPCHK(A.alloc(&A.cinv_off, (size_t)C + 1));
    HIPCHK(A.alloc(&A.crow_c, (size_t)C + 1));
    HIPCHK(A.alloc(&A.chol_fallback, 2));
    HIPCHK(hipMemcpyAsync(A.cinv_off, A.h_cinv_off.data(), sizeof(int64_t) * (C + 1), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(A.crow_c, A.h_crow_c.data(), sizeof(int64_t) * (C + 1), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(A.chol_fallback, 0, 2 * sizeof(unsigned long long), s));
    if (C > 0 && maxm > 0) {
        const size_t lds = sizeof(double) * ((size_t)maxm * maxm + 2 * (size_t)maxm);
        hipLaunchKernelGGL(k_coarsest_chol, dim3((unsigned)C), dim3(BLK), lds, s, cur, (const int64_t*)A.crow_c,
                           (const int64_t*)A.cinv_off, A.cinv, A.chol_fallback);
    }
    HIPCHK(hipGetLastError());
    return RUV_OK;
}

}  // namespace

// Keep the finest operator (FP64 CSR) and each row's chart on the device for (re)builds of the hierarchy.
int bamg_upload_fine(BatchImpl& Q, const int64_t* ptr, const int32_t* cols, const double* vals) {
    hipStream_t s = Q.stream;
    const int64_t n0 = Q.nrows;
    HIPCHK(Q.upload(&Q.amg_ptr, ptr, (size_t)n0 + 1));
    HIPCHK(Q.upload(&Q.amg_cols, cols, (size_t)Q.nnz));
    std::vector<float> v32(vals, vals + Q.nnz);  // the setup reads FP32 values (SetupCsr)
    HIPCHK(Q.upload(&Q.amg_vals, v32.data(), v32.size()));
    std::vector<int32_t> g(n0);
    for (int c = 0; c < Q.nchart; ++c)
        for (int64_t i = Q.crow[c]; i < Q.crow[c + 1]; ++i) g[i] = c;
    HIPCHK(Q.upload(&Q.amg_group, g.data(), g.size()));
    HIPCHK(hipStreamSynchronize(s));
    return RUV_OK;
}

// New finest values (FP64, device, CSR order of bamg_upload_fine) for the next build.
int bamg_set_fine_values(BatchImpl& Q, const double* d_vals) {
    hipLaunchKernelGGL((k_cast<float>), rows_grid(Q.nnz), dim3(BLK), 0, Q.stream, Q.nnz, d_vals, Q.amg_vals);
    HIPCHK(hipGetLastError());
    return RUV_OK;
}

// Build the hierarchy on the device from the finest operator kept by bamg_upload_fine (tiles fixed).
int bamg_build(BatchImpl& Q) {
    hipStream_t s = Q.stream;
    if (!Q.amg_ptr) return fail(RUV_ERR_INVALID, "batch amg: no finest operator on the device");
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s));
    std::unique_ptr<Bamg> A(new Bamg(&Q.pool));
    if (const char* e = std::getenv("RUV_BATCH_AMG_OMEGA")) {
        const double v = std::atof(e);
        if (v > 0.0 && v < 1.0) A->omega_max = v;
    }
    if (const char* e = std::getenv("RUV_BATCH_AMG_FUSE_ROWS")) A->fuse_rows = std::atoll(e);
    if (const char* e = std::getenv("RUV_BATCH_AMG_ALPHA")) {
        const double v = std::atof(e);
        if (v > 0.0 && v < 2.0) A->alpha = v;
    }
    const bool f32 = Q.mode == RUV_MODE_MIXED;
    const size_t sv = f32 ? sizeof(float) : sizeof(double);
    const int R = Q.R;
    // matching: coupling threshold relative to a row's strongest, handshake rounds (overridable for tuning)
    // and matchings per level: one (aggregates of about two rows) converges in about half the iterations of two
    // (aggregates of about four) on the heterogeneous weights of jittered meshes, for about 1.5x the coarse work
    double theta = 0.25;
    int rounds = 3;
    if (const char* e = std::getenv("RUV_BATCH_AMG_THETA")) theta = std::atof(e);
    if (const char* e = std::getenv("RUV_BATCH_AMG_ROUNDS")) rounds = std::max(1, std::atoi(e));
    // coarsening stops once every chart has at most dense_rows rows; the coarsest level is then solved exactly
    // with each chart's inverted block (up to CHOL_MAX rows a chart, else its diagonal)
    int64_t dense_rows = 64;
    if (const char* e = std::getenv("RUV_BATCH_AMG_DENSE")) dense_rows = std::max(0, std::atoi(e));
    auto max_chart_rows = [&](const std::vector<int64_t>& cr) {
        int64_t m = 0;
        for (int c = 0; c < Q.nchart; ++c) m = std::max(m, cr[c + 1] - cr[c]);
        return m;
    };
    const int64_t n0 = Q.nrows;
    // Bottom-up: the matchings and Galerkin operators, every level in the order of its construction.
    struct NatLevel {
        int64_t n = 0;
        SetupCsr A{};
        const double* diag = nullptr;     // levels >= 1 (level 0: the batch's Jacobi diagonal)
        const int32_t* group = nullptr;   // row -> chart
        int32_t* agg = nullptr;           // row -> next-level aggregate (null on the coarsest level)
        int64_t* mptr = nullptr;          // next level's aggregates -> member rows (ascending)
        int64_t* mrows = nullptr;
        double omega = 2.0 / 3.0;
        std::vector<int64_t> crow;        // chart row ranges
        std::unique_ptr<Scratch> own;     // the level's operator (levels >= 1) and matching
    };
    std::vector<NatLevel> nat;
    const int32_t* group = Q.amg_group;
    SetupCsr cur{n0, Q.amg_ptr, Q.amg_cols, Q.amg_vals};
    const double* cur_diag = nullptr;
    std::vector<int64_t> crow = Q.crow;
    std::unique_ptr<Scratch> keep(new Scratch(&Q.pool));  // owns the current level's setup CSR (levels >= 1)
    const int max_levels = 40;
    for (int l = 0;; ++l) {
        NatLevel N;
        N.n = cur.n;
        N.A = cur;
        N.diag = cur_diag;
        N.group = group;
        N.crow = crow;
        const bool last = (l + 1 >= max_levels) ||
                          (l > 0 && (cur.n <= (int64_t)Q.nchart || max_chart_rows(crow) <= dense_rows));
        if (last) {
            N.own = std::move(keep);
            nat.push_back(std::move(N));
            break;
        }
        // one matching makes the next level (aggregates of about two rows: a second matching of the Galerkin
        // operator, aggregates of about four, doubled the iterations on jittered meshes)
        std::unique_ptr<Scratch> S1(new Scratch(&Q.pool));
        Matched M1;
        // level 0: the band charts' rows ranked from their band form (the same entries, read coalesced)
        BandSrc band{};
        if (l == 0 && Q.nband > 0)
            band = BandSrc{Q.B.c_bw, Q.B.c_bmask, Q.B.c_bbase, Q.B.crow, Q.amg_group, Q.bv32, Q.bv64};
        int rc = match_and_coarsen(s, cur, band, group, theta, rounds, *S1, &M1);
        if (rc) return rc;
        const Matched& MF = M1;
        N.omega = (M1.gersh > 0.0) ? std::min(A->omega_max, 1.9 / M1.gersh) : A->omega_max;
        if (l > 0 && MF.nc > (int64_t)(0.9 * (double)cur.n)) {  // coarsening stalled: this level is the coarsest
            N.own = std::move(keep);
            nat.push_back(std::move(N));
            break;
        }
        N.agg = M1.agg;
        N.mptr = M1.mptr;
        N.mrows = M1.mrows;
        N.own = std::move(keep);
        for (const void* p : {(const void*)M1.agg, (const void*)M1.mptr, (const void*)M1.mrows})
            hand_over(*S1, (void*)p, N.own->mem);
        // chart offsets on the next level
        {
            std::vector<int64_t> at(crow.begin(), crow.end());
            int64_t *d_at = nullptr, *d_out = nullptr;
            HIPCHK(S1->alloc(&d_at, at.size()));
            HIPCHK(S1->alloc(&d_out, at.size()));
            HIPCHK(hipMemcpyAsync(d_at, at.data(), sizeof(int64_t) * at.size(), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_gather_offsets, rows_grid((int64_t)at.size()), dim3(BLK), 0, s, (int64_t)at.size(),
                               (const int64_t*)d_at, (const int64_t*)M1.cid, d_out);
            std::vector<int64_t> out(at.size());
            HIPCHK(hipMemcpyAsync(out.data(), d_out, sizeof(int64_t) * at.size(), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            crow.swap(out);
        }
        nat.push_back(std::move(N));
        // the next level's operator (kept), its diagonal, and chart groups
        std::unique_ptr<Scratch> nk(new Scratch(&Q.pool));
        for (const void* p : {(const void*)MF.C.ptr, (const void*)MF.C.cols, (const void*)MF.C.vals,
                              (const void*)MF.cdiag})
            hand_over(*S1, (void*)p, nk->mem);
        int32_t* ngroup = nullptr;  // kept: the cycle skips the rows of stopped charts by it
        HIPCHK(A->alloc(&ngroup, MF.nc));
        int64_t* d_crow = nullptr;
        HIPCHK(nk->alloc(&d_crow, crow.size()));
        HIPCHK(hipMemcpyAsync(d_crow, crow.data(), sizeof(int64_t) * crow.size(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_chart_groups, rows_grid(MF.nc), dim3(BLK), 0, s, MF.nc, (const int64_t*)d_crow,
                           Q.nchart, ngroup);
        HIPCHK(hipGetLastError());
        // (crow, the copy's source, changes next after the next level's matching, which waits on the stream)
        cur = MF.C;
        cur_diag = MF.cdiag;
        group = ngroup;
        keep = std::move(nk);
        S1.reset();
    }
    const size_t nl = nat.size();
    if (nl < 2) return fail(RUV_ERR_PRECOND, "batch amg: no coarse level");
    A->lv.resize(nl);
    for (auto* v : {&A->cs0, &A->cs1, &A->cb0, &A->cb1}) v->assign(nl, {});
    A->slist.assign(nl, nullptr);
    A->blist.assign(nl, nullptr);
    A->nslist.assign(nl, 0);
    A->nblist.assign(nl, 0);
    for (size_t l = 0; l < nl; ++l) {
        BamgLevel& L = A->lv[l];
        L.n = nat[l].n;
        L.chart = nat[l].group;
        L.omega = nat[l].omega;
        const int64_t nb = (L.n + BLK - 1) / BLK;
        HIPCHK(A->alloc(&L.blk, nb));
        if (nb > 0)
            hipLaunchKernelGGL(k_block_charts, rows_grid(nb), dim3(BLK), 0, s, L.n, L.chart, L.blk);
        HIPCHK(hipGetLastError());
    }
    // Top-down: the stored order of every level between the finest and the coarsest, and the cycle's data.
    Scratch P(&Q.pool);
    std::vector<int32_t*> pos(nl, nullptr);  // constructed row -> stored row (null: the identity)
    auto alloc_vectors = [&](BamgLevel& L) -> int {
        char *b = nullptr, *z = nullptr, *t = nullptr;
        const size_t bytes = (size_t)L.n * R * sv;
        HIPCHK(A->alloc(&b, bytes));
        HIPCHK(A->alloc(&z, bytes));
        HIPCHK(A->alloc(&t, bytes));
        HIPCHK(hipMemsetAsync(b, 0, bytes, s));
        HIPCHK(hipMemsetAsync(z, 0, bytes, s));
        HIPCHK(hipMemsetAsync(t, 0, bytes, s));
        L.b = b;
        L.z = z;
        L.t = t;
        return RUV_OK;
    };
    // member offsets, in the next level's stored order, of the aggregates of constructed level l
    auto stored_starts = [&](size_t l, int64_t** start) -> int {
        const int64_t nI = nat[l + 1].n;
        int64_t* sz = nullptr;
        HIPCHK(P.alloc(&sz, nI));
        HIPCHK(P.alloc(start, nI + 1));
        hipLaunchKernelGGL(k_perm_sizes, rows_grid(nI), dim3(BLK), 0, s, nI, (const int32_t*)pos[l + 1],
                           (const int64_t*)nat[l].mptr, sz);
        HIPCHK(hipGetLastError());
        return scan(s, nI, sz, *start, P);
    };
    {  // coarsest: constructed order
        BamgLevel& L = A->lv[nl - 1];
        const NatLevel& N = nat[nl - 1];
        char* d = nullptr;
        HIPCHK(A->alloc(&d, (size_t)L.n * sv));
        if (f32) hipLaunchKernelGGL((k_cast_inv<float>), rows_grid(L.n), dim3(BLK), 0, s, L.n, N.diag, (float*)d);
        else hipLaunchKernelGGL((k_cast_inv<double>), rows_grid(L.n), dim3(BLK), 0, s, L.n, N.diag, (double*)d);
        HIPCHK(hipGetLastError());
        L.dinv = d;
        int rc = alloc_vectors(L);
        if (rc) return rc;
    }
    for (size_t l = nl - 1; l-- > 1;) {
        BamgLevel& L = A->lv[l];
        const NatLevel& N = nat[l];
        const int64_t n = L.n, nI = nat[l + 1].n;
        int64_t* start = nullptr;
        int rc = stored_starts(l, &start);
        if (rc) return rc;
        int32_t *rpos = nullptr, *ipos = nullptr, *flags = nullptr;
        HIPCHK(P.alloc(&rpos, n));
        HIPCHK(P.alloc(&ipos, n));
        HIPCHK(P.alloc(&flags, 2));
        HIPCHK(hipMemsetAsync(flags, 0, 2 * sizeof(int32_t), s));
        hipLaunchKernelGGL(k_perm_expand, rows_grid(nI), dim3(BLK), 0, s, nI, (const int32_t*)pos[l + 1],
                           (const int64_t*)N.mptr, (const int64_t*)N.mrows, (const int64_t*)start, rpos,
                           (int32_t*)nullptr);
        hipLaunchKernelGGL(k_perm_inverse, rows_grid(n), dim3(BLK), 0, s, n, (const int32_t*)rpos, ipos);
        pos[l] = rpos;
        HIPCHK(A->alloc(&L.agg, n));
        hipLaunchKernelGGL(k_perm_agg, rows_grid(n), dim3(BLK), 0, s, n, (const int32_t*)N.agg, (const int32_t*)rpos,
                           (const int32_t*)pos[l + 1], L.agg);
        HIPCHK(A->alloc(&L.mptr, nI + 1));
        hipLaunchKernelGGL(k_narrow, rows_grid(nI + 1), dim3(BLK), 0, s, nI + 1, (const int64_t*)start, L.mptr);
        // slices aligned to aggregate starts, their ELL blocks
        L.nsl = (n + CSLICE_STEP - 1) / CSLICE_STEP;
        const int64_t nsl = L.nsl;
        int64_t* sz = nullptr;
        HIPCHK(A->alloc(&L.srow, nsl + 1));
        HIPCHK(A->alloc(&L.sagg, nsl + 1));
        HIPCHK(A->alloc(&L.sblk, nsl));
        HIPCHK(A->alloc(&L.ew, nsl));
        HIPCHK(A->alloc(&L.eoff, nsl + 1));
        HIPCHK(P.alloc(&sz, nsl));
        hipLaunchKernelGGL(k_slice_rows, rows_grid(nsl + 1), dim3(BLK), 0, s, nsl, n, nI, (const int64_t*)start,
                           L.srow, L.sagg, flags);
        if (nsl > 0)
            hipLaunchKernelGGL(k_pslice_sizes, dim3((unsigned)nsl), dim3(BLK), 0, s, (const int32_t*)L.srow,
                               (const int32_t*)ipos, N.A.ptr, N.group, L.ew, sz, L.sblk);
        hipLaunchKernelGGL(k_poffsets_wide, rows_grid(n), dim3(BLK), 0, s, n, N.A.ptr, N.A.cols, (const int32_t*)rpos,
                           flags + 1);
        HIPCHK(hipGetLastError());
        rc = scan(s, nsl, sz, L.eoff, P);
        if (rc) return rc;
        int32_t hf[2] = {0, 0};
        std::vector<int32_t> hsrow((size_t)nsl + 1);
        HIPCHK(hipMemcpyAsync(&L.nell, L.eoff + nsl, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(&L.nnz, N.A.ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(hsrow.data(), L.srow, sizeof(int32_t) * (nsl + 1), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        {  // each chart's slices: those holding its first and last rows, and all between
            const std::vector<int64_t>& cr = nat[l].crow;
            auto slice_of = [&](int64_t p) {
                int64_t b = std::min<int64_t>(p / CSLICE_STEP, nsl - 1);
                while (hsrow[b] > p) --b;
                return (int32_t)b;
            };
            A->cs0[l].assign(Q.nchart, 0);
            A->cs1[l].assign(Q.nchart, 0);
            for (int c = 0; c < Q.nchart; ++c)
                if (cr[c + 1] > cr[c]) {
                    A->cs0[l][c] = slice_of(cr[c]);
                    A->cs1[l][c] = slice_of(cr[c + 1] - 1) + 1;
                }
        }
        if (hf[0]) return fail(RUV_ERR_PRECOND, "batch amg: an aggregate holds more rows than a slice allows");
        // columns: 16-bit offsets from the row (packed with the bfloat16 value under the FP32 cycle: one load a
        // slot), or 32-bit stored rows when some offset does not fit
        L.ci = hf[1] ? 1 : (f32 ? 2 : 0);
        char *v = nullptr, *q = nullptr, *rs = nullptr, *ec = nullptr;
        HIPCHK(A->alloc(&ec, L.nell * (L.ci == 0 ? sizeof(int16_t) : sizeof(int32_t))));
        L.ecols = ec;
        if (L.ci != 2) HIPCHK(A->alloc(&v, L.nell * (f32 ? sizeof(CoarseVal<float>::type) : sizeof(double))));
        HIPCHK(A->alloc(&q, n * sv));
        HIPCHK(A->alloc(&rs, n * sv));
        auto fill = [&](auto tv) {
            using TT = decltype(tv);
            using VT = typename CoarseVal<TT>::type;
            if (L.ci == 1)
                hipLaunchKernelGGL((k_pell_fill<TT, 1>), rows_grid(n), dim3(BLK), 0, s, n, N.A, N.diag,
                                   (const int32_t*)rpos, (const int32_t*)ipos, nsl, (const int32_t*)L.srow,
                                   (const int64_t*)L.eoff, (const int32_t*)L.ew, (VT*)v, L.ecols);
            else if (L.ci == 2)
                hipLaunchKernelGGL((k_pell_fill<TT, 2>), rows_grid(n), dim3(BLK), 0, s, n, N.A, N.diag,
                                   (const int32_t*)rpos, (const int32_t*)ipos, nsl, (const int32_t*)L.srow,
                                   (const int64_t*)L.eoff, (const int32_t*)L.ew, (VT*)v, L.ecols);
            else
                hipLaunchKernelGGL((k_pell_fill<TT, 0>), rows_grid(n), dim3(BLK), 0, s, n, N.A, N.diag,
                                   (const int32_t*)rpos, (const int32_t*)ipos, nsl, (const int32_t*)L.srow,
                                   (const int64_t*)L.eoff, (const int32_t*)L.ew, (VT*)v, L.ecols);
            hipLaunchKernelGGL((k_pscale<TT>), rows_grid(n), dim3(BLK), 0, s, n, N.diag, (const int32_t*)rpos,
                               (TT*)q, (TT*)rs);
        };
        if (f32) fill(0.0f);
        else fill(0.0);
        HIPCHK(hipGetLastError());
        L.evals = v;
        L.q = q;
        L.rs = rs;
        rc = alloc_vectors(L);
        if (rc) return rc;
    }
    {  // finest level: member lists and aggregate map into level 1's stored order
        BamgLevel& L = A->lv[0];
        const NatLevel& N = nat[0];
        const int64_t n1 = nat[1].n;
        L.dinv = f32 ? (void*)Q.dinv32 : (void*)Q.dinv64;
        if (f32) {  // bfloat16 copy of the batch's FP32 values (they may have changed since the last build): the
                    // sliced ELL unless every chart is in band form, and the band form
            A->ev0 = nullptr;
            if (Q.n_ell_ranges > 0) {  // the entries of the charts not in band form (the others stay unwritten)
                Bf16* e16 = nullptr;
                HIPCHK(A->alloc(&e16, (size_t)Q.nell));
                hipLaunchKernelGGL(k_to_bf16_ranges, dim3((unsigned)Q.n_ell_ranges), dim3(BLK), 0, s,
                                   (const int64_t*)Q.ell_ranges, (const float*)Q.ev32, e16);
                A->ev0 = e16;
            }
            A->bv0 = nullptr;
            if (Q.nbv) {
                Bf16* b16 = nullptr;
                HIPCHK(A->alloc(&b16, (size_t)Q.nbv));
                hipLaunchKernelGGL(k_to_bf16, rows_grid((Q.nbv + BF16_VEC - 1) / BF16_VEC), dim3(BLK), 0, s, Q.nbv,
                                   (const float*)Q.bv32, b16);
                A->bv0 = b16;
            }
            HIPCHK(hipGetLastError());
        } else {
            A->ev0 = Q.ev64;
            A->bv0 = Q.bv64;
        }
        char* t = nullptr;
        HIPCHK(A->alloc(&t, L.n * R * sv));
        HIPCHK(hipMemsetAsync(t, 0, L.n * R * sv, s));  // rows of converged charts keep finite values
        L.t = t;
        int64_t* start = nullptr;
        int rc = stored_starts(0, &start);
        if (rc) return rc;
        HIPCHK(A->alloc(&L.mrows, L.n));
        hipLaunchKernelGGL(k_perm_expand, rows_grid(n1), dim3(BLK), 0, s, n1, (const int32_t*)pos[1],
                           (const int64_t*)N.mptr, (const int64_t*)N.mrows, (const int64_t*)start, (int32_t*)nullptr,
                           L.mrows);
        HIPCHK(A->alloc(&L.mptr, n1 + 1));
        hipLaunchKernelGGL(k_narrow, rows_grid(n1 + 1), dim3(BLK), 0, s, n1 + 1, (const int64_t*)start, L.mptr);
        HIPCHK(A->alloc(&L.agg, L.n));
        hipLaunchKernelGGL(k_perm_agg, rows_grid(L.n), dim3(BLK), 0, s, L.n, (const int32_t*)N.agg,
                           (const int32_t*)nullptr, (const int32_t*)pos[1], L.agg);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    }
    {
        int rc = coarsest_device(Q, *A, nat[nl - 1].A, crow);
        if (rc) return rc;
        A->coarse = nat[nl - 1].A;  // kept with the hierarchy
        if (nat[nl - 1].own) {
            for (void* p : nat[nl - 1].own->mem) A->mem.push_back(p);
            nat[nl - 1].own->mem.clear();
        }
    }
    {  // each chart's BLK-row blocks on levels >= 1, and room for the lists of running slices and blocks
        size_t cap = 0;
        for (size_t l = 1; l < nl; ++l) {
            const std::vector<int64_t>& cr = nat[l].crow;
            A->cb0[l].assign(Q.nchart, 0);
            A->cb1[l].assign(Q.nchart, 0);
            for (int c = 0; c < Q.nchart; ++c)
                if (cr[c + 1] > cr[c]) {
                    A->cb0[l][c] = (int32_t)(cr[c] / BLK);
                    A->cb1[l][c] = (int32_t)((cr[c + 1] - 1) / BLK + 1);
                }
            cap += (size_t)A->lv[l].nsl + (size_t)((A->lv[l].n + BLK - 1) / BLK);
        }
        HIPCHK(A->alloc(&A->d_lists, cap));
        if (2 * cap > Q.h_lists_cap) {  // no copy from it is in flight: every solve ends with a wait on the stream
            if (Q.h_lists) (void)hipHostFree(Q.h_lists);
            Q.h_lists = nullptr;
            Q.h_lists_cap = 0;
            HIPCHK(hipHostMalloc((void**)&Q.h_lists, sizeof(int32_t) * std::max<size_t>(2 * cap, 1),
                                 hipHostMallocDefault));
            Q.h_lists_cap = 2 * cap;
        }
        A->h_lists = Q.h_lists;
        A->lists_cap = cap;
    }
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    A->setup_seconds = ms * 1e-3;
    if (Q.amg) bamg_destroy(Q.amg);
    Q.amg = A.release();
    return RUV_OK;
}

// Launches of the coarse levels cover every slice and block (done == nullptr) or only those of the charts with
// done[c] == B_RUNNING (the lists are uploaded on the solver's stream).
int bamg_set_live(BatchImpl& Q, const int32_t* done) {
    Bamg& A = *Q.amg;
    const size_t nl = A.lv.size();
    std::fill(A.slist.begin(), A.slist.end(), nullptr);
    std::fill(A.blist.begin(), A.blist.end(), nullptr);
    if (!done) return RUV_OK;
    // staged in the half not used by the previous lists, whose copy may still be queued (the copy before that has
    // completed: the solver waits for a look recorded after it before it builds the next lists)
    A.lists_half ^= 1;
    int32_t* h = A.h_lists + (size_t)A.lists_half * A.lists_cap;
    size_t pos = 0;
    auto add = [&](const std::vector<int32_t>& f, const std::vector<int32_t>& e, int32_t** list, int64_t* count) {
        const size_t start = pos;
        int32_t last = -1;
        for (int c = 0; c < Q.nchart; ++c)
            if (done[c] == B_RUNNING)
                for (int32_t b = std::max(f[c], last + 1); b < e[c]; ++b) h[pos++] = last = b;
        *list = A.d_lists + start;
        *count = (int64_t)(pos - start);
    };
    for (size_t l = 1; l < nl; ++l) {
        if (l + 1 < nl) add(A.cs0[l], A.cs1[l], &A.slist[l], &A.nslist[l]);
        add(A.cb0[l], A.cb1[l], &A.blist[l], &A.nblist[l]);
    }
    if (pos) HIPCHK(hipMemcpyAsync(A.d_lists, h, sizeof(int32_t) * pos, hipMemcpyHostToDevice, Q.stream));
    return RUV_OK;
}

// z = M^-1 r (V(1,1) cycle) over the live tiles; partial r.z of every live tile into part_c[par].
// With Q.marks set (profiling, every chart running, whole levels) each launch is marked with its kind and bytes:
// algorithmic (every nonzero's value and stored column, every vector element read or written once -- neighbour
// gathers are served by the caches -- and the level's index arrays) and as stored (padding slots included).
template <typename T, int R, int CI>
int bamg_apply(BatchImpl& Q, int na, int par, const T* r, T* z) {
    Bamg& A = *Q.amg;
    hipStream_t s = Q.stream;
    const size_t nl = A.lv.size();
    BamgLevel& L0 = A.lv[0];
    BamgLevel& L1 = A.lv[1];
    using VT = typename CoarseVal<T>::type;
    constexpr int PACKED = sizeof(T) == 4 ? 2 : 0;  // the packed column mode exists under the FP32 cycle only
    const VT* vals = (const VT*)A.ev0;
    const VT* bvals = (const VT*)A.bv0;
    const dim3 gt((unsigned)(((na + XCDS - 1) / XCDS) * XCDS));
    const T w0 = (T)L0.omega;
    const bool prof = Q.marks != nullptr;
    const double vb = (double)R * sizeof(T), db = (double)sizeof(T), cv = (double)sizeof(VT);
    const double n0 = (double)L0.n, n1 = (double)L1.n;
    const double fin = (double)Q.nnz * cv + Q.fin_colb + Q.fin_meta;
    const double fin_st = Q.fin_slots * cv + Q.fin_colb_st + Q.fin_meta;
    auto mark = [&](int kind, int level, double bytes, double stored) -> int {
        HIPCHK(Q.mark(kind, level, bytes, stored));
        return RUV_OK;
    };
    int rc = RUV_OK;
    {  // the tiles of band charts (listed first) and the others, one launch each
        const int nb = std::min(std::max(Q.na_band, 0), na);
        const auto g = [](int n) { return dim3((unsigned)(((n + XCDS - 1) / XCDS) * XCDS)); };
        if (nb > 0)
            hipLaunchKernelGGL((kc_resid0<T, R, CI, 1>), g(nb), dim3(BLK), 0, s, Q.B, (const int*)Q.act, nb, vals,
                               bvals, (const T*)L0.dinv, w0, r, (T*)L0.t);
        if (na > nb)
            hipLaunchKernelGGL((kc_resid0<T, R, CI, 2>), g(na - nb), dim3(BLK), 0, s, Q.B, (const int*)Q.act + nb,
                               na - nb, vals, bvals, (const T*)L0.dinv, w0, r, (T*)L0.t);
    }
    if (prof) {  // r, dinv read, t written
        const double v = n0 * (2.0 * vb + db);
        if ((rc = mark(BK_RESID0, 0, fin + v, fin_st + v))) return rc;
    }
    auto cyc = [&](size_t l) {
        BamgLevel& L = A.lv[l];
        return CycLevel<T>{L.n, L.eoff, L.ew, L.ecols, (const typename CoarseVal<T>::type*)L.evals,
                           L.srow, L.sagg, L.sblk, A.slist[l], A.blist[l], (const T*)L.dinv, L.agg, L.mptr, L.mrows,
                           (T*)L.b, (T*)L.z, (T*)L.t, L.chart, L.blk, Q.B.c_done, (const T*)L.q, (const T*)L.rs,
                           (l + 2 < nl) ? (const T*)A.lv[l + 1].rs : nullptr};
    };
    // grids: every slice / block of a level, or the listed ones of the running charts
    auto sgrid = [&](size_t l) -> int64_t { return A.slist[l] ? A.nslist[l] : A.lv[l].nsl; };
    auto bgrid = [&](size_t l) -> int64_t { return A.blist[l] ? A.nblist[l] : (A.lv[l].n + BLK - 1) / BLK; };
    auto g1 = [](int64_t n) { return dim3((unsigned)n); };
    // a coarse level's operator: values and 16- or 32-bit columns, per nonzero or per stored slot, and the slice
    // tables (row start, aggregate start, chart range, width, offset)
    auto lmat = [&](size_t l, bool stored) {
        const BamgLevel& L = A.lv[l];
        const double e = L.ci == 2 ? 4.0 : cv + (L.ci == 1 ? 4.0 : 2.0);  // value and column bytes a slot
        return (stored ? (double)L.nell : (double)L.nnz) * e + (double)L.nsl * 28.0;
    };
    if (bgrid(1) > 0) {
        hipLaunchKernelGGL((kc_restrict<T, R>), g1(bgrid(1)), dim3(BLK), 0, s, cyc(0), L1.n, L1.chart,
                           (const int2*)L1.blk, (const int32_t*)A.blist[1], (T*)L1.b);
        if (prof) {  // t and the member rows read, member offsets read, b_1 written (and its scale read)
            const double v = n0 * (vb + 4.0) + n1 * (4.0 + vb + (nl > 2 ? db : 0.0));
            if ((rc = mark(BK_RESTRICT, 0, v, v))) return rc;
        }
    }
    for (size_t l = 1; l + 1 < nl; ++l) {
        const T w = (T)A.lv[l].omega;
        const int64_t ns = sgrid(l);
        if (ns == 0) continue;
        if (A.lv[l].ci == 1) hipLaunchKernelGGL((kc_down<T, R, 1>), g1(ns), dim3(BLK), 0, s, cyc(l), w, (T*)A.lv[l + 1].b);
        else if (A.lv[l].ci == 2) hipLaunchKernelGGL((kc_down<T, R, PACKED>), g1(ns), dim3(BLK), 0, s, cyc(l), w, (T*)A.lv[l + 1].b);
        else hipLaunchKernelGGL((kc_down<T, R, 0>), g1(ns), dim3(BLK), 0, s, cyc(l), w, (T*)A.lv[l + 1].b);
        if (prof) {  // b, dinv read; next level's member offsets read, its b written
            const double v = (double)A.lv[l].n * (vb + db) + (double)A.lv[l + 1].n * (4.0 + vb);
            if ((rc = mark(BK_DOWN, (int)l, lmat(l, false) + v, lmat(l, true) + v))) return rc;
        }
    }
    const int64_t nbc = bgrid(nl - 1);
    if (nbc > 0) {
        if (A.cinv)
            hipLaunchKernelGGL((kc_coarsest_dense<T, R>), g1(nbc), dim3(BLK), 0, s, cyc(nl - 1),
                               (const int64_t*)A.crow_c, (const int64_t*)A.cinv_off, (const double*)A.cinv);
        else
            hipLaunchKernelGGL((kc_coarsest<T, R>), g1(nbc), dim3(BLK), 0, s, cyc(nl - 1));
        if (prof) {  // b read, z written, each chart's FP64 inverse (or the diagonal) read
            const double nc = (double)A.lv[nl - 1].n;
            const double inv = A.cinv ? 8.0 * (double)A.h_cinv_off.back() : nc * db;
            const double v = nc * (2.0 * vb + 4.0) + inv;
            if ((rc = mark(BK_COARSEST, (int)nl - 1, v, v))) return rc;
        }
    }
    for (size_t l = nl - 1; l-- > 1;) {
        const T w = (T)A.lv[l].omega;
        const int64_t ns = sgrid(l), nb = bgrid(l);
        if (ns == 0 || nb == 0) continue;
        const T* zn = (const T*)A.lv[l + 1].z;
        const double nL = (double)A.lv[l].n, nN = (double)A.lv[l + 1].n;
        // small level, or few rows of it live late in a solve: the prolongation inside the post-smoothing pass
        const int64_t live = A.slist[l] ? A.nslist[l] * (int64_t)SLICE : A.lv[l].n;
        if (live <= A.fuse_rows) {
            if (A.lv[l].ci == 1)
                hipLaunchKernelGGL((kc_post<T, R, 1, 1>), g1(ns), dim3(BLK), 0, s, cyc(l), w, (T)A.alpha, zn);
            else if (A.lv[l].ci == 2)
                hipLaunchKernelGGL((kc_post<T, R, PACKED, 1>), g1(ns), dim3(BLK), 0, s, cyc(l), w, (T)A.alpha, zn);
            else
                hipLaunchKernelGGL((kc_post<T, R, 0, 1>), g1(ns), dim3(BLK), 0, s, cyc(l), w, (T)A.alpha, zn);
            if (prof) {  // b, dinv, agg read, z written; the next level's z read
                const double v = nL * (2.0 * vb + db + 4.0) + nN * vb;
                if ((rc = mark(BK_POST, (int)l, lmat(l, false) + v, lmat(l, true) + v))) return rc;
            }
            continue;
        }
        hipLaunchKernelGGL((kc_prolong<T, R>), g1(nb), dim3(BLK), 0, s, cyc(l), w, (T)A.alpha, zn);
        if (prof) {  // b, dinv, agg read, t written; the next level's z read
            const double v = nL * (2.0 * vb + db + 4.0) + nN * vb;
            if ((rc = mark(BK_PROLONG, (int)l, v, v))) return rc;
        }
        if (A.lv[l].ci == 1) hipLaunchKernelGGL((kc_post<T, R, 1>), g1(ns), dim3(BLK), 0, s, cyc(l), w, T(0), zn);
        else if (A.lv[l].ci == 2) hipLaunchKernelGGL((kc_post<T, R, PACKED>), g1(ns), dim3(BLK), 0, s, cyc(l), w, T(0), zn);
        else hipLaunchKernelGGL((kc_post<T, R, 0>), g1(ns), dim3(BLK), 0, s, cyc(l), w, T(0), zn);
        if (prof) {  // b, t, dinv read, z written
            const double v = nL * (3.0 * vb + db);
            if ((rc = mark(BK_POST, (int)l, lmat(l, false) + v, lmat(l, true) + v))) return rc;
        }
    }
    // up: e1 in level 0's work vector (free once restricted), then the post-smoothing over it (two passes:
    // recomputing e1 for every gathered neighbour inside the post-smoothing measured slower)
    hipLaunchKernelGGL((kc_prolong0<T, R>), gt, dim3(BLK), 0, s, Q.B, (const int*)Q.act, na, (const T*)L0.dinv, w0,
                       (T)A.alpha, r, (const int32_t*)L0.agg, (const T*)L1.z, (T*)L0.t);
    if (prof) {  // r, dinv, agg read, e written; level 1's z read
        const double v = n0 * (2.0 * vb + db + 4.0) + n1 * vb;
        if ((rc = mark(BK_PROLONG0, 0, v, v))) return rc;
    }
    hipLaunchKernelGGL((kc_post0<T, R, CI>), gt, dim3(BLK), 0, s, Q.B, (const int*)Q.act, na, par, vals, bvals,
                       (const T*)L0.dinv, w0, r, (const T*)L0.t, z);
    if (prof) {  // r, e, dinv read, z written, a partial per tile
        const double v = n0 * (3.0 * vb + db) + 8.0 * Q.ntile;
        if ((rc = mark(BK_POST0, 0, fin + v, fin_st + v))) return rc;
    }
    HIPCHK(hipGetLastError());
    return RUV_OK;
}

template int bamg_apply<float, 1, 0>(BatchImpl&, int, int, const float*, float*);
template int bamg_apply<float, 1, 1>(BatchImpl&, int, int, const float*, float*);
template int bamg_apply<float, 2, 0>(BatchImpl&, int, int, const float*, float*);
template int bamg_apply<float, 2, 1>(BatchImpl&, int, int, const float*, float*);
template int bamg_apply<double, 1, 0>(BatchImpl&, int, int, const double*, double*);
template int bamg_apply<double, 1, 1>(BatchImpl&, int, int, const double*, double*);
template int bamg_apply<double, 2, 0>(BatchImpl&, int, int, const double*, double*);
template int bamg_apply<double, 2, 1>(BatchImpl&, int, int, const double*, double*);

// Charts whose coarsest block kept its diagonal instead of the exact inverse: not numerically positive definite,
// coupled to another chart, or larger than the device factorisation takes (reads two counters back).
int bamg_fallbacks(BatchImpl& Q, int64_t* not_spd, int64_t* coupled, int64_t* too_large) {
    unsigned long long h[2] = {0, 0};
    if (Q.amg && Q.amg->chol_fallback) {
        HIPCHK(hipMemcpyAsync(h, Q.amg->chol_fallback, sizeof(h), hipMemcpyDeviceToHost, Q.stream));
        HIPCHK(hipStreamSynchronize(Q.stream));
    }
    if (not_spd) *not_spd = (int64_t)h[0];
    if (coupled) *coupled = (int64_t)h[1];
    if (too_large) *too_large = Q.amg ? Q.amg->chol_too_large : 0;
    return RUV_OK;
}

// Chart c's block of the coarsest operator (dense, FP64 from the stored values) and the inverse the device
// factorisation produced for it (row-major, m x m each; m alone when both pointers are null).  An inspection hook.
int bamg_coarsest_block(BatchImpl& Q, int c, int64_t* m_out, double* block, double* inverse) {
    if (!Q.amg || !Q.amg->cinv) return fail(RUV_ERR_INVALID, "batch amg: no coarsest inverse");
    Bamg& A = *Q.amg;
    if (c < 0 || c >= Q.nchart) return fail(RUV_ERR_INVALID, "batch amg: chart out of range");
    const int64_t r0 = A.h_crow_c[c], m = A.h_crow_c[c + 1] - r0, mo = A.h_cinv_off[c];
    if (m_out) *m_out = m;
    if ((!block && !inverse) || m == 0) return RUV_OK;
    if (mo < 0) return fail(RUV_ERR_PRECOND, "batch amg: this chart's coarsest block keeps its diagonal");
    hipStream_t s = Q.stream;
    if (inverse) HIPCHK(hipMemcpyAsync(inverse, A.cinv + mo, sizeof(double) * m * m, hipMemcpyDeviceToHost, s));
    if (block) {
        std::vector<int64_t> ptr((size_t)m + 1);
        HIPCHK(hipMemcpyAsync(ptr.data(), A.coarse.ptr + r0, sizeof(int64_t) * (m + 1), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const int64_t nz = ptr[m] - ptr[0];
        std::vector<int32_t> cols((size_t)std::max<int64_t>(nz, 1));
        std::vector<float> vals((size_t)std::max<int64_t>(nz, 1));
        HIPCHK(hipMemcpyAsync(cols.data(), A.coarse.cols + ptr[0], sizeof(int32_t) * nz, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(vals.data(), A.coarse.vals + ptr[0], sizeof(float) * nz, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        std::fill(block, block + m * m, 0.0);
        for (int64_t i = 0; i < m; ++i)
            for (int64_t k = ptr[i] - ptr[0]; k < ptr[i + 1] - ptr[0]; ++k) {
                const int64_t j = cols[k] - r0;
                if (j >= 0 && j < m) block[i * m + j] += (double)vals[k];
            }
    }
    HIPCHK(hipStreamSynchronize(s));
    return RUV_OK;
}

// Level l of the hierarchy in its stored order (the inspection hook): sizes, the operator as CSR from the sliced
// ELL (levels between the finest and the coarsest) or the kept coarsest CSR, the aggregate map and row charts.
int bamg_level(BatchImpl& Q, int l, int64_t* sizes, int64_t* row_ptr, int32_t* cols, double* vals, int32_t* agg,
               int32_t* chart) {
    if (!Q.amg) return fail(RUV_ERR_INVALID, "batch amg: no hierarchy");
    Bamg& A = *Q.amg;
    const int nl = (int)A.lv.size();
    if (l < 0 || l >= nl) return fail(RUV_ERR_INVALID, "batch amg: level out of range");
    const BamgLevel& L = A.lv[(size_t)l];
    hipStream_t s = Q.stream;
    const int64_t n = L.n;
    std::vector<int64_t> ptr(1, 0);
    std::vector<int32_t> cc;
    std::vector<double> vv;
    const bool f32 = Q.mode == RUV_MODE_MIXED;
    if (l == nl - 1) {  // the coarsest: the setup CSR (FP32 values)
        std::vector<int64_t> p((size_t)n + 1);
        HIPCHK(hipMemcpyAsync(p.data(), A.coarse.ptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const int64_t nz = p[n];
        cc.resize((size_t)std::max<int64_t>(nz, 1));
        std::vector<float> fv((size_t)std::max<int64_t>(nz, 1));
        HIPCHK(hipMemcpyAsync(cc.data(), A.coarse.cols, sizeof(int32_t) * nz, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(fv.data(), A.coarse.vals, sizeof(float) * nz, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        cc.resize((size_t)nz);
        vv.assign(fv.begin(), fv.begin() + nz);
        ptr = p;
    } else if (l > 0) {  // sliced ELL, padding (0, the row itself) dropped
        const int64_t nsl = L.nsl, ne = L.nell;
        std::vector<int64_t> eoff((size_t)nsl + 1);
        std::vector<int32_t> ew((size_t)std::max<int64_t>(nsl, 1)), srow((size_t)nsl + 1);
        HIPCHK(hipMemcpyAsync(eoff.data(), L.eoff, sizeof(int64_t) * (nsl + 1), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(ew.data(), L.ew, sizeof(int32_t) * nsl, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(srow.data(), L.srow, sizeof(int32_t) * (nsl + 1), hipMemcpyDeviceToHost, s));
        // columns: 32-bit rows (ci 1) or 32-bit packed words (ci 2: bfloat16 value high, offset low) in c32,
        // 16-bit offsets (ci 0) in c16; values of ci 0 and 1 in b16 (FP32 cycle) or d64
        std::vector<int32_t> c32(L.ci ? (size_t)std::max<int64_t>(ne, 1) : 0);
        std::vector<int16_t> c16(L.ci ? 0 : (size_t)std::max<int64_t>(ne, 1));
        if (L.ci) HIPCHK(hipMemcpyAsync(c32.data(), L.ecols, sizeof(int32_t) * ne, hipMemcpyDeviceToHost, s));
        else HIPCHK(hipMemcpyAsync(c16.data(), L.ecols, sizeof(int16_t) * ne, hipMemcpyDeviceToHost, s));
        std::vector<uint16_t> b16(f32 && L.ci != 2 ? (size_t)std::max<int64_t>(ne, 1) : 0);
        std::vector<double> d64(f32 ? 0 : (size_t)std::max<int64_t>(ne, 1));
        if (f32 && L.ci != 2)
            HIPCHK(hipMemcpyAsync(b16.data(), L.evals, sizeof(uint16_t) * ne, hipMemcpyDeviceToHost, s));
        else if (!f32)
            HIPCHK(hipMemcpyAsync(d64.data(), L.evals, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
        // the stored values are scaled by 1 / sqrt(d_i d_j): undone with the stored sqrt(d)
        std::vector<float> q32(f32 ? (size_t)std::max<int64_t>(n, 1) : 0);
        std::vector<double> q64(f32 ? 0 : (size_t)std::max<int64_t>(n, 1));
        if (f32) HIPCHK(hipMemcpyAsync(q32.data(), L.q, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        else HIPCHK(hipMemcpyAsync(q64.data(), L.q, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        auto qv = [&](int64_t r) { return f32 ? (double)q32[(size_t)r] : q64[(size_t)r]; };
        for (int64_t b = 0; b < nsl; ++b) {
            const int64_t r0 = srow[b], ns = srow[b + 1] - r0;
            for (int64_t p = r0; p < r0 + ns; ++p) {
                for (int k = 0; k < ew[b]; ++k) {
                    const int64_t e = eoff[b] + (p - r0) + (int64_t)k * ns;
                    const uint32_t w = (uint32_t)(L.ci ? c32[e] : 0);
                    const int64_t c = L.ci == 1 ? (int64_t)c32[e]
                                                : p + (L.ci == 2 ? (int64_t)(int16_t)(w & 0xFFFFu) : (int64_t)c16[e]);
                    double v;
                    if (f32) {
                        const uint32_t u = L.ci == 2 ? (w & 0xFFFF0000u) : (uint32_t)b16[e] << 16;
                        float f;
                        std::memcpy(&f, &u, sizeof(f));
                        v = f;
                    } else {
                        v = d64[e];
                    }
                    if (c == p && v == 0.0) continue;  // padding
                    cc.push_back((int32_t)c);
                    vv.push_back(v * qv(p) * qv(c));
                }
                ptr.push_back((int64_t)cc.size());
            }
        }
    } else {
        ptr.assign((size_t)n + 1, 0);
    }
    if (sizes) {
        sizes[0] = n;
        sizes[1] = (int64_t)cc.size();
        sizes[2] = l + 1 < nl ? A.lv[(size_t)l + 1].n : 0;
    }
    if (row_ptr && l > 0) std::copy(ptr.begin(), ptr.end(), row_ptr);
    if (cols) std::copy(cc.begin(), cc.end(), cols);
    if (vals) std::copy(vv.begin(), vv.end(), vals);
    if (agg && l + 1 < nl) HIPCHK(hipMemcpyAsync(agg, L.agg, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    if (chart) HIPCHK(hipMemcpyAsync(chart, L.chart, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return RUV_OK;
}

int bamg_level_rows(const BatchImpl& Q, int64_t* rows, int cap) {
    if (!Q.amg) return 0;
    const int nl = (int)Q.amg->lv.size();
    for (int l = 0; l < nl && l < cap; ++l) rows[l] = Q.amg->lv[l].n;
    return nl;
}

int bamg_info(const BatchImpl& Q, int* levels, double* setup_seconds, int64_t* coarse_rows) {
    if (!Q.amg) {
        if (levels) *levels = 0;
        if (setup_seconds) *setup_seconds = 0.0;
        if (coarse_rows) *coarse_rows = 0;
        return RUV_OK;
    }
    if (levels) *levels = (int)Q.amg->lv.size();
    if (setup_seconds) *setup_seconds = Q.amg->setup_seconds;
    if (coarse_rows) {
        int64_t t = 0;
        for (size_t l = 1; l < Q.amg->lv.size(); ++l) t += Q.amg->lv[l].n;
        *coarse_rows = t;
    }
    return RUV_OK;
}

}  // namespace ruv
