// api.cpp -- the extern "C" entry points declared in include/ruv_lscm.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "ruv_internal.hpp"

namespace ruv {
struct PlanImpl;
int plan_create(const Csr& a, int R, int precond, int mode, PlanImpl** out);
void plan_destroy(PlanImpl* p);
int plan_solve_dev(PlanImpl* P, const double* d_b, double* d_x, uint64_t max_iter, double tol, ruv_solve_info* info);
int plan_solve_host(PlanImpl* P, const double* b, double* x, uint64_t max_iter, double tol, ruv_solve_info* info);
int plan_spmv_time(PlanImpl* P, int reps, double* sec, double* bytes);
}  // namespace ruv

struct ruv_plan {
    ruv::PlanImpl* impl;
};

namespace {

// Validate and copy a caller CSR (solver.rs:296-299: n == 0, b.len() != n, cols/vals length and row_ptr
// length mismatches are rejected).
int take_csr(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz, ruv::Csr* a) {
    if (n <= 0) return ruv::fail(RUV_ERR_BAD_DIMS, "n must be positive");
    if (!vals || !cols || !row_ptr) return ruv::fail(RUV_ERR_INVALID, "null CSR pointer");
    if (nnz < 0 || row_ptr[0] != 0 || (int64_t)row_ptr[n] != nnz)
        return ruv::fail(RUV_ERR_BAD_DIMS, "row_ptr does not describe nnz entries");
    for (int32_t i = 0; i < n; ++i)
        if (row_ptr[i + 1] < row_ptr[i]) return ruv::fail(RUV_ERR_BAD_DIMS, "row_ptr not monotone");
    for (int64_t k = 0; k < nnz; ++k)
        if (cols[k] < 0 || cols[k] >= n) return ruv::fail(RUV_ERR_BAD_DIMS, "column index out of range");
    a->n = n;
    a->vals.assign(vals, vals + nnz);
    a->cols.assign(cols, cols + nnz);
    a->ptr.assign(row_ptr, row_ptr + n + 1);
    return RUV_OK;
}

bool valid_precond(int32_t pc) { return pc >= RUV_PC_NONE && pc <= RUV_PC_AMGSA; }

int solve_common(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                 const double* b, int64_t b_len, uint64_t max_iter, double tol, int32_t precond, int mode,
                 double* x_out, ruv_solve_info* info) {
    if (!valid_precond(precond)) return ruv::fail(RUV_ERR_INVALID, "unknown precond");
    if (!b || !x_out) return ruv::fail(RUV_ERR_INVALID, "null vector pointer");
    if (b_len != n) return ruv::fail(RUV_ERR_BAD_DIMS, "b length != n");
    ruv::Csr a;
    int rc = take_csr(vals, cols, row_ptr, n, nnz, &a);
    if (rc) return rc;
    ruv::Csr l;
    if (ruv::detect_kron2(a, &l)) return ruv::device_solve(l, 2, b, max_iter, tol, precond, mode, x_out, info);
    return ruv::device_solve(a, 1, b, max_iter, tol, precond, mode, x_out, info);
}

}  // namespace

extern "C" {

int ruv_available(void) {
    int n = 0;
    return (hipGetDeviceCount(&n) == hipSuccess && n > 0) ? 1 : 0;
}

const char* ruv_last_error(void) { return ruv::last_error(); }

int ruv_cg_solve_precond(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                         const double* b, int64_t b_len, uint64_t max_iter, double tol, int32_t precond,
                         double* x_out, uint64_t* iters_out) {
    ruv_solve_info info{};
    int rc = solve_common(vals, cols, row_ptr, n, nnz, b, b_len, max_iter, tol, precond, RUV_MODE_FP64, x_out, &info);
    if (rc == RUV_OK && iters_out) *iters_out = info.iters;
    return rc;
}

int ruv_cg_solve(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                 const double* b, int64_t b_len, uint64_t max_iter, double tol, double* x_out) {
    return ruv_cg_solve_precond(vals, cols, row_ptr, n, nnz, b, b_len, max_iter, tol, RUV_PC_NONE, x_out, nullptr);
}

int ruv_cg_solve_mixed(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                       const double* b, int64_t b_len, uint64_t max_iter, double tol, int32_t precond,
                       double* x_out, ruv_solve_info* info) {
    ruv_solve_info tmp{};
    return solve_common(vals, cols, row_ptr, n, nnz, b, b_len, max_iter, tol, precond, RUV_MODE_MIXED, x_out,
                        info ? info : &tmp);
}

int ruv_apply_ic0(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                  const double* r, double* z_out) {
    if (!r || !z_out) return ruv::fail(RUV_ERR_INVALID, "null vector pointer");
    ruv::Csr a;
    int rc = take_csr(vals, cols, row_ptr, n, nnz, &a);
    if (rc) return rc;
    ruv::Csr l;
    if (ruv::detect_kron2(a, &l)) return ruv::device_apply_ic0(l, 2, r, z_out);
    return ruv::device_apply_ic0(a, 1, r, z_out);
}

int ruv_plan_create(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t nrows, int64_t nnz,
                    int32_t rhs_per_row, int32_t precond, int32_t mode, ruv_plan** out) {
    if (!out) return ruv::fail(RUV_ERR_INVALID, "null plan pointer");
    *out = nullptr;
    if (!valid_precond(precond)) return ruv::fail(RUV_ERR_INVALID, "unknown precond");
    ruv::Csr a;
    int rc = take_csr(vals, cols, row_ptr, nrows, nnz, &a);
    if (rc) return rc;
    ruv::PlanImpl* impl = nullptr;
    rc = ruv::plan_create(a, rhs_per_row, precond, mode, &impl);
    if (rc) return rc;
    *out = new ruv_plan{impl};
    return RUV_OK;
}

int ruv_plan_solve_dev(ruv_plan* plan, const double* d_b, double* d_x, uint64_t max_iter, double tol,
                       ruv_solve_info* info) {
    if (!plan || !d_b || !d_x) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::plan_solve_dev(plan->impl, d_b, d_x, max_iter, tol, info);
}

int ruv_plan_solve(ruv_plan* plan, const double* b, double* x, uint64_t max_iter, double tol, ruv_solve_info* info) {
    if (!plan || !b || !x) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::plan_solve_host(plan->impl, b, x, max_iter, tol, info);
}

int ruv_plan_spmv_time(ruv_plan* plan, int32_t reps, double* seconds_per_launch, double* bytes_per_launch) {
    if (!plan || !seconds_per_launch || !bytes_per_launch) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::plan_spmv_time(plan->impl, reps, seconds_per_launch, bytes_per_launch);
}

void ruv_plan_destroy(ruv_plan* plan) {
    if (!plan) return;
    ruv::plan_destroy(plan->impl);
    delete plan;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------------
// charting and the per-chart driver
// ------------------------------------------------------------------------------------------------------
#include "host_lscm.hpp"

struct ruv_charts {
    ruv::Charts c;
};

extern "C" {

int ruv_weld(const double* pos, int64_t nv, const uint32_t* faces, int64_t nf, double tol, double* pos_out,
             uint32_t* faces_out, int64_t* nv_out, int64_t* nf_out) {
    if ((nv && !pos) || (nf && !faces) || !pos_out || !faces_out || !nv_out || !nf_out)
        return ruv::fail(RUV_ERR_INVALID, "null argument");
    for (int64_t k = 0; k < 3 * nf; ++k)
        if ((int64_t)faces[k] >= nv) return ruv::fail(RUV_ERR_BAD_DIMS, "face index out of range");
    std::vector<double> p;
    std::vector<uint32_t> f;
    ruv::weld(pos, nv, faces, nf, tol, &p, &f);
    std::copy(p.begin(), p.end(), pos_out);
    std::copy(f.begin(), f.end(), faces_out);
    *nv_out = (int64_t)p.size() / 3;
    *nf_out = (int64_t)f.size() / 3;
    return RUV_OK;
}

int ruv_segment(const double* pos, int64_t nv, const uint32_t* faces, int64_t nf, double angle_min_deg,
                ruv_charts** out) {
    if (!out || (nv && !pos) || (nf && !faces)) return ruv::fail(RUV_ERR_INVALID, "null argument");
    for (int64_t k = 0; k < 3 * nf; ++k)
        if ((int64_t)faces[k] >= nv) return ruv::fail(RUV_ERR_BAD_DIMS, "face index out of range");
    auto* h = new ruv_charts();
    ruv::segment(pos, nv, faces, nf, angle_min_deg, &h->c);
    *out = h;
    return RUV_OK;
}

int64_t ruv_charts_count(const ruv_charts* charts) { return charts ? (int64_t)charts->c.charts.size() : 0; }

int ruv_chart_sizes(const ruv_charts* charts, int64_t i, int64_t* sizes, double* area3d) {
    if (!charts || !sizes || i < 0 || i >= (int64_t)charts->c.charts.size())
        return ruv::fail(RUV_ERR_INVALID, "bad chart handle or index");
    const ruv::Chart& ch = charts->c.charts[i];
    int64_t tot = 0;
    for (const auto& l : ch.loops) tot += (int64_t)l.size();
    sizes[0] = (int64_t)ch.face_ids.size();
    sizes[1] = (int64_t)ch.vertex_ids.size();
    sizes[2] = (int64_t)ch.border.size();
    sizes[3] = (int64_t)ch.loops.size();
    sizes[4] = tot;
    sizes[5] = (int64_t)charts->c.cut.size();
    if (area3d) *area3d = ch.area3d;
    return RUV_OK;
}

int ruv_chart_arrays(const ruv_charts* charts, int64_t i, uint32_t* face_ids, uint32_t* vertex_ids, uint32_t* tris,
                     uint32_t* border, uint32_t* loops, int64_t* loop_off) {
    if (!charts || i < 0 || i >= (int64_t)charts->c.charts.size())
        return ruv::fail(RUV_ERR_INVALID, "bad chart handle or index");
    const ruv::Chart& ch = charts->c.charts[i];
    if (face_ids) std::copy(ch.face_ids.begin(), ch.face_ids.end(), face_ids);
    if (vertex_ids) std::copy(ch.vertex_ids.begin(), ch.vertex_ids.end(), vertex_ids);
    if (tris) std::copy(ch.tris.begin(), ch.tris.end(), tris);
    if (border) std::copy(ch.border.begin(), ch.border.end(), border);
    int64_t o = 0;
    if (loop_off) loop_off[0] = 0;
    for (size_t l = 0; l < ch.loops.size(); ++l) {
        if (loops) std::copy(ch.loops[l].begin(), ch.loops[l].end(), loops + o);
        o += (int64_t)ch.loops[l].size();
        if (loop_off) loop_off[l + 1] = o;
    }
    return RUV_OK;
}

void ruv_charts_destroy(ruv_charts* charts) { delete charts; }

int ruv_lscm_system(const double* mesh_pos, const ruv_charts* charts, int64_t i, double angle_mix, int64_t* sizes,
                    double* p2, int32_t* pin_ids, double* pin_uv, int32_t* free_ids, double* l_vals, int32_t* l_cols,
                    int32_t* l_ptr, double* rhs) {
    if (!mesh_pos || !charts || !sizes || i < 0 || i >= (int64_t)charts->c.charts.size())
        return ruv::fail(RUV_ERR_INVALID, "bad argument");
    ruv::ChartSystem S;
    ruv::build_system(mesh_pos, charts->c.charts[i], angle_mix, &S);
    sizes[0] = S.nv;
    sizes[1] = S.nfree;
    sizes[2] = (int64_t)S.pin_ids.size();
    sizes[3] = S.l.nnz();
    sizes[4] = (int64_t)S.edges;
    if (p2) std::copy(S.p2.begin(), S.p2.end(), p2);
    if (pin_ids) std::copy(S.pin_ids.begin(), S.pin_ids.end(), pin_ids);
    if (pin_uv) std::copy(S.pin_uv.begin(), S.pin_uv.end(), pin_uv);
    if (free_ids) std::copy(S.free_ids.begin(), S.free_ids.end(), free_ids);
    if (l_vals) std::copy(S.l.vals.begin(), S.l.vals.end(), l_vals);
    if (l_cols) std::copy(S.l.cols.begin(), S.l.cols.end(), l_cols);
    if (l_ptr) std::copy(S.l.ptr.begin(), S.l.ptr.end(), l_ptr);
    if (rhs) std::copy(S.rhs.begin(), S.rhs.end(), rhs);
    return RUV_OK;
}

int ruv_unfold_chart(const double* mesh_pos, const ruv_charts* charts, int64_t i, const ruv_unfold_opts* opts,
                     double* uv_out, double* stats, ruv_solve_info* info) {
    if (!mesh_pos || !charts || !opts || !uv_out || i < 0 || i >= (int64_t)charts->c.charts.size())
        return ruv::fail(RUV_ERR_INVALID, "bad argument");
    if (opts->precond < RUV_PC_NONE || opts->precond > RUV_PC_AMGSA) return ruv::fail(RUV_ERR_INVALID, "unknown precond");
    return ruv::unfold_chart(mesh_pos, charts->c.charts[i], *opts, uv_out, stats, info);
}

}  // extern "C"
