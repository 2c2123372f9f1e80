// api.cpp -- the extern "C" entry points declared in include/ruv_lscm.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <string>
#include <vector>
#include <algorithm>

#include "host_lscm.hpp"
#include "ruv_internal.hpp"

namespace ruv {
struct PlanImpl;
int plan_create(const Csr& a, int R, int precond, int mode, PlanImpl** out);
void plan_destroy(PlanImpl* p);
int plan_solve_dev(PlanImpl* P, const double* d_b, double* d_x, uint64_t max_iter, double tol, ruv_solve_info* info);
int plan_solve_host(PlanImpl* P, const double* b, double* x, uint64_t max_iter, double tol, ruv_solve_info* info);
int plan_spmv_time(PlanImpl* P, int reps, double* sec, double* bytes);
int plan_update_values(PlanImpl* P, const double* vals);
int ic0_factor_device(const Csr& a, std::vector<int32_t>* ptr, std::vector<int32_t>* cols, std::vector<double>* vals);
struct BatchImpl;
int batch_create(int64_t nrows, const int64_t* ptr, const int32_t* cols, const double* vals, int nchart,
                 const int64_t* crow, int R, int precond, int mode, BatchImpl** out);
void batch_destroy(BatchImpl* q);
int batch_solve_dev(BatchImpl* Q, const double* d_b, double* d_x, uint64_t max_iter, double tol,
                    ruv_solve_info* info, uint64_t* chart_iters, double* chart_resid);
int batch_spmv_time(BatchImpl* Q, int reps, double* sec, double* bytes, double* stored);
int batch_schedule(BatchImpl* Q, int32_t* rows_per_tile, int32_t* col_bytes, int64_t* ell_entries,
                   int64_t* col_entries, int32_t* stencil_slices, int32_t* slices, int32_t* nontemporal);
int batch_kernel_timing(BatchImpl* Q, int enable);
int batch_precond_info(BatchImpl* Q, int32_t* levels, double* setup_seconds, int64_t* coarse_rows);
int batch_setup_precond(BatchImpl* Q);
int batch_precond_levels(BatchImpl* Q, int64_t* rows, int32_t cap);
int batch_precond_fallbacks(BatchImpl* Q, int64_t* not_spd, int64_t* coupled, int64_t* too_large);
int batch_band_info(const BatchImpl* Q, int32_t* band_charts, int64_t* band_values);
hipStream_t batch_stream(const BatchImpl* Q);
int batch_coarsest_block(BatchImpl* Q, int c, int64_t* m_out, double* block, double* inverse);
int batch_amg_level(BatchImpl* Q, int l, int64_t* sizes, int64_t* row_ptr, int32_t* cols, double* vals, int32_t* agg,
                    int32_t* chart);
int batch_kernel_stats(BatchImpl* Q, int64_t* launches, double* seconds, double* bytes, double* stored);
int batch_step_profile(BatchImpl* Q, int reps, int cap, int32_t* kind, int32_t* level, double* sec, double* bytes,
                       double* stored, int32_t* count);
int lscm_batch_build(int32_t nmesh, const double* pos, const int64_t* pos_off, const uint32_t* faces,
                     const int64_t* face_off, double tol, double angle, double angle_mix, int32_t threads,
                     LscmBatch** out);
void lscm_batch_destroy(LscmBatch* b);
}  // namespace ruv

struct ruv_batch {
    ruv::BatchImpl* impl;
    int64_t n;  // rows * rhs_per_row
};

struct ruv_plan {
    ruv::PlanImpl* impl;
};

namespace ruv {
struct UnfoldBatch;
int unfold_batch_run(int32_t nmesh, const double* pos, const int64_t* pos_off, const uint32_t* faces,
                     const int64_t* face_off, double weld_tol, double angle_min_deg, const ruv_unfold_opts& o,
                     int32_t threads, UnfoldBatch** out);
void unfold_batch_destroy(UnfoldBatch* u);
int unfold_batch_sizes(const UnfoldBatch* u, int64_t* sizes);
int unfold_batch_results(const UnfoldBatch* u, int32_t* chart_mesh, int64_t* v_off, uint32_t* vertex_ids, double* uv,
                         double* rectified, double* ext, double* metrics, int64_t* counts, double* stats,
                         ruv_solve_info* info);
}  // namespace ruv

struct ruv_unfold_batch {
    ruv::UnfoldBatch* impl;
};

struct ruv_amg_host {
    std::vector<ruv::AmgHostLevel> levels;
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

int ruv_amg_aggregate(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                      int32_t max_members, int32_t* agg, int32_t* n_coarse) {
    if (!agg || !n_coarse || max_members < 1) return ruv::fail(RUV_ERR_INVALID, "null argument or max_members < 1");
    ruv::Csr a;
    int rc = take_csr(vals, cols, row_ptr, n, nnz, &a);
    if (rc) return rc;
    std::vector<int32_t> g;
    ruv::amg_aggregate(a, max_members, &g, n_coarse);
    std::copy(g.begin(), g.end(), agg);
    return RUV_OK;
}

// out CSR of a host matrix into caller arrays (sizes[0] rows, sizes[1] entries; arrays null: sizes only)
static int put_csr(const ruv::Csr& m, int64_t* sizes, double* vals, int32_t* cols, int32_t* ptr) {
    sizes[0] = m.n;
    sizes[1] = m.nnz();
    if (vals) std::copy(m.vals.begin(), m.vals.end(), vals);
    if (cols) std::copy(m.cols.begin(), m.cols.end(), cols);
    if (ptr) std::copy(m.ptr.begin(), m.ptr.end(), ptr);
    return RUV_OK;
}

int ruv_amg_prolongator(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                        const int32_t* agg, int32_t n_coarse, int32_t smoothed, int64_t* sizes, double* p_vals,
                        int32_t* p_cols, int32_t* p_ptr) {
    if (!agg || !sizes) return ruv::fail(RUV_ERR_INVALID, "null argument");
    ruv::Csr a;
    int rc = take_csr(vals, cols, row_ptr, n, nnz, &a);
    if (rc) return rc;
    for (int32_t i = 0; i < n; ++i)
        if (agg[i] < 0 || agg[i] >= n_coarse) return ruv::fail(RUV_ERR_BAD_DIMS, "aggregate out of range");
    const ruv::Csr p = ruv::amg_prolongator(a, std::vector<int32_t>(agg, agg + n), smoothed != 0);
    return put_csr(p, sizes, p_vals, p_cols, p_ptr);
}

int ruv_amg_galerkin(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                     const double* p_vals, const int32_t* p_cols, const int32_t* p_ptr, int32_t n_coarse,
                     int64_t* sizes, double* c_vals, int32_t* c_cols, int32_t* c_ptr) {
    if (!p_vals || !p_cols || !p_ptr || !sizes) return ruv::fail(RUV_ERR_INVALID, "null argument");
    ruv::Csr a;
    int rc = take_csr(vals, cols, row_ptr, n, nnz, &a);
    if (rc) return rc;
    ruv::Csr p;
    p.n = n;
    p.ptr.assign(p_ptr, p_ptr + n + 1);
    p.cols.assign(p_cols, p_cols + p_ptr[n]);
    p.vals.assign(p_vals, p_vals + p_ptr[n]);
    for (int32_t c : p.cols)
        if (c < 0 || c >= n_coarse) return ruv::fail(RUV_ERR_BAD_DIMS, "P column out of range");
    const ruv::Csr ac = ruv::amg_galerkin(a, p, n_coarse);
    return put_csr(ac, sizes, c_vals, c_cols, c_ptr);
}

int ruv_amg_host_build(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                       int32_t smoothed, ruv_amg_host** out) {
    if (!out) return ruv::fail(RUV_ERR_INVALID, "null argument");
    ruv::Csr a;
    int rc = take_csr(vals, cols, row_ptr, n, nnz, &a);
    if (rc) return rc;
    std::unique_ptr<ruv_amg_host> h(new ruv_amg_host());
    std::string why;
    if (!ruv::amg_hierarchy(a, smoothed != 0, &h->levels, &why)) return ruv::fail(RUV_ERR_PRECOND, "amg: " + why);
    *out = h.release();
    return RUV_OK;
}

int32_t ruv_amg_host_levels(const ruv_amg_host* h) { return h ? (int32_t)h->levels.size() : 0; }

int ruv_amg_host_level(const ruv_amg_host* h, int32_t level, int64_t* sizes, double* op_vals, int32_t* op_cols,
                       int32_t* op_ptr, double* p_vals, int32_t* p_cols, int32_t* p_ptr, int32_t* agg) {
    if (!h || !sizes) return ruv::fail(RUV_ERR_INVALID, "null argument");
    if (level < 0 || level >= (int32_t)h->levels.size()) return ruv::fail(RUV_ERR_INVALID, "level out of range");
    const ruv::AmgHostLevel& L = h->levels[(size_t)level];
    sizes[0] = L.op.n;
    sizes[1] = L.op.nnz();
    sizes[2] = L.p.nnz();
    sizes[3] = L.n_coarse;
    sizes[4] = L.has_factor ? 1 : 0;
    if (op_vals) std::copy(L.op.vals.begin(), L.op.vals.end(), op_vals);
    if (op_cols) std::copy(L.op.cols.begin(), L.op.cols.end(), op_cols);
    if (op_ptr) std::copy(L.op.ptr.begin(), L.op.ptr.end(), op_ptr);
    if (p_vals) std::copy(L.p.vals.begin(), L.p.vals.end(), p_vals);
    if (p_cols) std::copy(L.p.cols.begin(), L.p.cols.end(), p_cols);
    if (p_ptr && !L.p.ptr.empty()) std::copy(L.p.ptr.begin(), L.p.ptr.end(), p_ptr);
    if (agg) std::copy(L.agg.begin(), L.agg.end(), agg);
    return RUV_OK;
}

void ruv_amg_host_destroy(ruv_amg_host* h) { delete h; }

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

int ruv_plan_update_values(ruv_plan* plan, const double* vals) {
    if (!plan || !vals) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::plan_update_values(plan->impl, vals);
}

int ruv_ic0_factor(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                   int64_t* l_nnz, int32_t* l_ptr, int32_t* l_cols, double* l_vals) {
    ruv::Csr a;
    int rc = take_csr(vals, cols, row_ptr, n, nnz, &a);
    if (rc) return rc;
    std::vector<int32_t> p, c;
    std::vector<double> v;
    rc = ruv::ic0_factor_device(a, &p, &c, &v);
    if (rc) return rc;
    if (l_nnz) *l_nnz = (int64_t)v.size();
    if (l_ptr) std::copy(p.begin(), p.end(), l_ptr);
    if (l_cols) std::copy(c.begin(), c.end(), l_cols);
    if (l_vals) std::copy(v.begin(), v.end(), l_vals);
    return RUV_OK;
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

// ------------------------------------------------------------------------------------------------------
// batched solves
// ------------------------------------------------------------------------------------------------------
extern "C" {

int ruv_batch_create(int64_t nrows, const int64_t* row_ptr, const int32_t* cols, const double* vals, int32_t nchart,
                     const int64_t* chart_row_off, int32_t rhs_per_row, int32_t precond, int32_t mode,
                     ruv_batch** out) {
    if (!out) return ruv::fail(RUV_ERR_INVALID, "null batch pointer");
    *out = nullptr;
    if (nrows <= 0 || nchart <= 0) return ruv::fail(RUV_ERR_BAD_DIMS, "empty batch");
    if (!row_ptr || !cols || !vals || !chart_row_off) return ruv::fail(RUV_ERR_INVALID, "null argument");
    if (nrows >= INT32_MAX) return ruv::fail(RUV_ERR_BAD_DIMS, "too many rows for 32-bit column indices");
    if (row_ptr[0] != 0 || chart_row_off[0] != 0 || chart_row_off[nchart] != nrows)
        return ruv::fail(RUV_ERR_BAD_DIMS, "row_ptr / chart_row_off do not start at 0 or end at nrows");
    for (int32_t c = 0; c < nchart; ++c) {
        if (chart_row_off[c + 1] < chart_row_off[c]) return ruv::fail(RUV_ERR_BAD_DIMS, "chart_row_off not monotone");
        for (int64_t i = chart_row_off[c]; i < chart_row_off[c + 1]; ++i) {
            if (row_ptr[i + 1] < row_ptr[i]) return ruv::fail(RUV_ERR_BAD_DIMS, "row_ptr not monotone");
            for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
                if (cols[k] < chart_row_off[c] || cols[k] >= chart_row_off[c + 1])
                    return ruv::fail(RUV_ERR_BAD_DIMS, "column outside its chart's block (chart " + std::to_string(c) + ")");
        }
    }
    ruv::BatchImpl* impl = nullptr;
    int rc = ruv::batch_create(nrows, row_ptr, cols, vals, nchart, chart_row_off, rhs_per_row, precond, mode, &impl);
    if (rc) return rc;
    *out = new ruv_batch{impl, nrows * rhs_per_row};
    return RUV_OK;
}

int ruv_batch_solve_dev(ruv_batch* batch, const double* d_b, double* d_x, uint64_t max_iter, double tol,
                        ruv_solve_info* info, uint64_t* chart_iters, double* chart_true_resid) {
    if (!batch || !d_b || !d_x) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::batch_solve_dev(batch->impl, d_b, d_x, max_iter, tol, info, chart_iters, chart_true_resid);
}

int ruv_batch_solve(ruv_batch* batch, const double* b, double* x, uint64_t max_iter, double tol,
                    ruv_solve_info* info, uint64_t* chart_iters, double* chart_true_resid) {
    if (!batch || !b || !x) return ruv::fail(RUV_ERR_INVALID, "null argument");
    const size_t N = (size_t)batch->n;
    const hipStream_t s = ruv::batch_stream(batch->impl);  // the copies on the batch's own stream
    double *db = nullptr, *dx = nullptr;
    hipError_t e = hipMalloc((void**)&db, N * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&dx, N * sizeof(double));
    if (e == hipSuccess) e = hipMemcpyAsync(db, b, N * sizeof(double), hipMemcpyHostToDevice, s);
    int rc = (e == hipSuccess) ? ruv_batch_solve_dev(batch, db, dx, max_iter, tol, info, chart_iters, chart_true_resid)
                               : ruv::fail(RUV_ERR_DEVICE, hipGetErrorString(e));
    if (rc == RUV_OK) {
        e = hipMemcpyAsync(x, dx, N * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = ruv::fail(RUV_ERR_DEVICE, hipGetErrorString(e));
    }
    if (db) (void)hipFree(db);
    if (dx) (void)hipFree(dx);
    if (rc == RUV_OK)
        for (size_t i = 0; i < N; ++i)
            if (!std::isfinite(x[i])) return ruv::fail(RUV_ERR_NONFINITE, "non-finite solution");
    return rc;
}

int ruv_batch_spmv_time(ruv_batch* batch, int32_t reps, double* seconds_per_launch, double* bytes_per_launch,
                        double* stored_bytes_per_launch) {
    if (!batch || !seconds_per_launch || !bytes_per_launch) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::batch_spmv_time(batch->impl, reps, seconds_per_launch, bytes_per_launch, stored_bytes_per_launch);
}

int ruv_batch_schedule(const ruv_batch* batch, int32_t* rows_per_tile, int32_t* col_bytes, int64_t* ell_entries,
                       int64_t* col_entries, int32_t* stencil_slices, int32_t* slices, int32_t* nontemporal) {
    if (!batch) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::batch_schedule(batch->impl, rows_per_tile, col_bytes, ell_entries, col_entries, stencil_slices,
                               slices, nontemporal);
}

int ruv_batch_band_info(const ruv_batch* batch, int32_t* band_charts, int64_t* band_values) {
    if (!batch) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::batch_band_info(batch->impl, band_charts, band_values);
}

int ruv_batch_precond_fallbacks(ruv_batch* batch, int64_t* not_spd_charts, int64_t* coupled_charts,
                                int64_t* too_large_charts) {
    if (!batch) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::batch_precond_fallbacks(batch->impl, not_spd_charts, coupled_charts, too_large_charts);
}

int ruv_batch_coarsest_block(ruv_batch* batch, int32_t chart, int64_t* rows, double* block, double* inverse) {
    if (!batch) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::batch_coarsest_block(batch->impl, chart, rows, block, inverse);
}

int ruv_batch_amg_level(ruv_batch* batch, int32_t level, int64_t* sizes, int64_t* row_ptr, int32_t* cols,
                        double* vals, int32_t* agg, int32_t* chart) {
    if (!batch) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::batch_amg_level(batch->impl, level, sizes, row_ptr, cols, vals, agg, chart);
}

int ruv_batch_precond_levels(ruv_batch* batch, int64_t* rows, int32_t cap) {
    if (!batch || (cap > 0 && !rows)) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::batch_precond_levels(batch->impl, rows, cap);
}

int ruv_batch_setup_precond(ruv_batch* batch) {
    if (!batch) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::batch_setup_precond(batch->impl);
}

int ruv_batch_precond_info(ruv_batch* batch, int32_t* levels, double* setup_seconds, int64_t* coarse_rows) {
    if (!batch) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::batch_precond_info(batch->impl, levels, setup_seconds, coarse_rows);
}

int ruv_batch_kernel_timing(ruv_batch* batch, int32_t enable) {
    if (!batch) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::batch_kernel_timing(batch->impl, enable);
}

int ruv_batch_kernel_stats(ruv_batch* batch, int64_t* launches, double* seconds, double* bytes,
                           double* stored_bytes) {
    if (!batch) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::batch_kernel_stats(batch->impl, launches, seconds, bytes, stored_bytes);
}

int ruv_batch_step_profile(ruv_batch* batch, int32_t reps, int32_t cap, int32_t* kind, int32_t* level,
                           double* seconds, double* bytes, double* stored_bytes, int32_t* count) {
    if (!batch) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::batch_step_profile(batch->impl, reps, cap, kind, level, seconds, bytes, stored_bytes, count);
}

void ruv_batch_destroy(ruv_batch* batch) {
    if (!batch) return;
    ruv::batch_destroy(batch->impl);
    delete batch;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------------
// host batch building
// ------------------------------------------------------------------------------------------------------
struct ruv_lscm_batch {
    ruv::LscmBatch* b;
};

extern "C" {

int ruv_lscm_batch_build(int32_t nmesh, const double* pos, const int64_t* pos_off, const uint32_t* faces,
                         const int64_t* face_off, double weld_tol, double angle_min_deg, double angle_mix,
                         int32_t threads, ruv_lscm_batch** out) {
    if (!out || nmesh <= 0 || !pos || !pos_off || !faces || !face_off) return ruv::fail(RUV_ERR_INVALID, "bad argument");
    for (int32_t m = 0; m < nmesh; ++m) {
        if (pos_off[m + 1] < pos_off[m] || face_off[m + 1] < face_off[m]) return ruv::fail(RUV_ERR_BAD_DIMS, "offsets not monotone");
        const int64_t nv = pos_off[m + 1] - pos_off[m];
        for (int64_t k = 3 * face_off[m]; k < 3 * face_off[m + 1]; ++k)
            if ((int64_t)faces[k] >= nv) return ruv::fail(RUV_ERR_BAD_DIMS, "face index outside its mesh");
    }
    ruv::LscmBatch* b = nullptr;
    int rc = ruv::lscm_batch_build(nmesh, pos, pos_off, faces, face_off, weld_tol, angle_min_deg, angle_mix, threads, &b);
    if (rc) return rc;
    *out = new ruv_lscm_batch{b};
    return RUV_OK;
}

int ruv_lscm_batch_sizes(const ruv_lscm_batch* lb, int64_t* sizes) {
    if (!lb || !sizes) return ruv::fail(RUV_ERR_INVALID, "null argument");
    sizes[0] = lb->b->nchart;
    sizes[1] = lb->b->nrows;
    sizes[2] = (int64_t)lb->b->vals.size();
    sizes[3] = lb->b->nv_total;
    return RUV_OK;
}

int ruv_lscm_batch_arrays(const ruv_lscm_batch* lb, int64_t* row_ptr, int32_t* cols, double* vals, double* rhs,
                          int64_t* chart_row_off, int32_t* chart_mesh, int64_t* chart_nv, int64_t* chart_npins) {
    if (!lb) return ruv::fail(RUV_ERR_INVALID, "null argument");
    const ruv::LscmBatch& b = *lb->b;
    if (row_ptr) std::copy(b.row_ptr.begin(), b.row_ptr.end(), row_ptr);
    if (cols) std::copy(b.cols.begin(), b.cols.end(), cols);
    if (vals) std::copy(b.vals.begin(), b.vals.end(), vals);
    if (rhs) std::copy(b.rhs.begin(), b.rhs.end(), rhs);
    if (chart_row_off) std::copy(b.chart_row_off.begin(), b.chart_row_off.end(), chart_row_off);
    if (chart_mesh) std::copy(b.chart_mesh.begin(), b.chart_mesh.end(), chart_mesh);
    if (chart_nv) std::copy(b.chart_nv.begin(), b.chart_nv.end(), chart_nv);
    if (chart_npins) std::copy(b.chart_npins.begin(), b.chart_npins.end(), chart_npins);
    return RUV_OK;
}

int ruv_batch_create_lscm(const ruv_lscm_batch* lb, int32_t precond, int32_t mode, ruv_batch** out) {
    if (!lb) return ruv::fail(RUV_ERR_INVALID, "null argument");
    const ruv::LscmBatch& b = *lb->b;
    return ruv_batch_create(b.nrows, b.row_ptr.data(), b.cols.data(), b.vals.data(), (int32_t)b.nchart,
                            b.chart_row_off.data(), 2, precond, mode, out);
}

void ruv_lscm_batch_destroy(ruv_lscm_batch* lb) {
    if (!lb) return;
    ruv::lscm_batch_destroy(lb->b);
    delete lb;
}

}  // extern "C"

extern "C" int ruv_lscm_system_device(const double* mesh_pos, const ruv_charts* charts, int64_t i, double angle_mix,
                                      int64_t* sizes, double* l_vals, int32_t* l_cols, int32_t* l_ptr, double* rhs) {
    if (!mesh_pos || !charts || !sizes || i < 0 || i >= (int64_t)charts->c.charts.size())
        return ruv::fail(RUV_ERR_INVALID, "bad argument");
    ruv::ChartSystem S;
    int rc = ruv::device_assemble(mesh_pos, charts->c.charts[i], angle_mix, &S);
    if (rc) return rc;
    sizes[0] = S.nv;
    sizes[1] = S.nfree;
    sizes[2] = (int64_t)S.pin_ids.size();
    sizes[3] = S.l.nnz();
    sizes[4] = (int64_t)S.edges;
    if (l_vals) std::copy(S.l.vals.begin(), S.l.vals.end(), l_vals);
    if (l_cols) std::copy(S.l.cols.begin(), S.l.cols.end(), l_cols);
    if (l_ptr) std::copy(S.l.ptr.begin(), S.l.ptr.end(), l_ptr);
    if (rhs) std::copy(S.rhs.begin(), S.rhs.end(), rhs);
    return RUV_OK;
}

namespace ruv {
int lscm_forward(const double* p3, int64_t nv, const uint32_t* tris, int64_t ntri, const int32_t* pin_ids,
                 const double* pin_uv, int64_t npins, int precond, int mode, double tol, double* uv_out,
                 double* energy_out, ruv_solve_info* info);
int lscm_vjp(const double* p3, int64_t nv, const uint32_t* tris, int64_t ntri, const int32_t* pin_ids,
             const double* pin_uv, int64_t npins, const double* uv, const double* g_uv, double a_energy, int precond,
             int mode, double tol, double* grad_p3, ruv_solve_info* info);
}  // namespace ruv

extern "C" int ruv_lscm_forward(const double* p3, int64_t nv, const uint32_t* tris, int64_t ntri,
                                const int32_t* pin_ids, const double* pin_uv, int64_t npins, int32_t precond,
                                int32_t mode, double tol, double* uv_out, double* energy_out, ruv_solve_info* info) {
    if (!p3 || nv <= 0 || !tris || ntri < 0 || (npins && (!pin_ids || !pin_uv)) || !uv_out)
        return ruv::fail(RUV_ERR_INVALID, "bad argument");
    if (precond < RUV_PC_NONE || precond > RUV_PC_AMGSA) return ruv::fail(RUV_ERR_INVALID, "unknown precond");
    return ruv::lscm_forward(p3, nv, tris, ntri, pin_ids, pin_uv, npins, precond, mode, tol, uv_out, energy_out, info);
}

extern "C" int ruv_lscm_vjp(const double* p3, int64_t nv, const uint32_t* tris, int64_t ntri, const int32_t* pin_ids,
                            const double* pin_uv, int64_t npins, const double* uv, const double* g_uv,
                            double energy_weight, int32_t precond, int32_t mode, double tol, double* grad_p3,
                            ruv_solve_info* info) {
    if (!p3 || nv <= 0 || !tris || ntri < 0 || (npins && (!pin_ids || !pin_uv)) || !uv || !g_uv || !grad_p3)
        return ruv::fail(RUV_ERR_INVALID, "bad argument");
    if (precond < RUV_PC_NONE || precond > RUV_PC_AMGSA) return ruv::fail(RUV_ERR_INVALID, "unknown precond");
    return ruv::lscm_vjp(p3, nv, tris, ntri, pin_ids, pin_uv, npins, uv, g_uv, energy_weight, precond, mode, tol,
                         grad_p3, info);
}

// ------------------------------------------------------------------------------------------------------
// the differentiable batch on the device
// ------------------------------------------------------------------------------------------------------
namespace ruv {
struct DiffBatchImpl;
int diff_create(const double* p3, int64_t nv, const uint32_t* tris, int64_t ntri, const int32_t* pin_ids,
                const double* pin_uv, int64_t npins, const int64_t* voff, int32_t nchart, int precond, int mode,
                double tol, DiffBatchImpl** out);
void diff_destroy(DiffBatchImpl* d);
int diff_forward(DiffBatchImpl* D, const double* p3, double* uv_out, double* energy_out, ruv_solve_info* info);
int diff_vjp(DiffBatchImpl* D, const double* g_uv, double a_energy, double* grad_out, ruv_solve_info* info);
}  // namespace ruv

struct ruv_lscm_diff {
    ruv::DiffBatchImpl* impl;
};

extern "C" {

int ruv_lscm_diff_create(const double* p3, int64_t nv, const uint32_t* tris, int64_t ntri, const int32_t* pin_ids,
                         const double* pin_uv, int64_t npins, const int64_t* chart_vertex_off, int32_t nchart,
                         int32_t precond, int32_t mode, double tol, ruv_lscm_diff** out) {
    if (!out) return ruv::fail(RUV_ERR_INVALID, "null handle pointer");
    *out = nullptr;
    if (!p3 || !tris || (npins > 0 && (!pin_ids || !pin_uv)) || !chart_vertex_off)
        return ruv::fail(RUV_ERR_INVALID, "null argument");
    if (nv <= 0 || ntri <= 0 || npins < 0) return ruv::fail(RUV_ERR_BAD_DIMS, "empty mesh");
    if (precond != RUV_PC_NONE && precond != RUV_PC_JACOBI && precond != RUV_PC_AMG)
        return ruv::fail(RUV_ERR_INVALID, "the differentiable batch supports none, jacobi and amg");
    ruv::DiffBatchImpl* d = nullptr;
    int rc = ruv::diff_create(p3, nv, tris, ntri, pin_ids, pin_uv, npins, chart_vertex_off, nchart, precond, mode,
                              tol, &d);
    if (rc) return rc;
    *out = new ruv_lscm_diff{d};
    return RUV_OK;
}

int ruv_lscm_diff_forward(ruv_lscm_diff* h, const double* p3, double* uv_out, double* energy_out,
                          ruv_solve_info* info) {
    if (!h || !p3 || !uv_out) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::diff_forward(h->impl, p3, uv_out, energy_out, info);
}

int ruv_lscm_diff_vjp(ruv_lscm_diff* h, const double* g_uv, double energy_weight, double* grad_p3,
                      ruv_solve_info* info) {
    if (!h || !g_uv || !grad_p3) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::diff_vjp(h->impl, g_uv, energy_weight, grad_p3, info);
}

void ruv_lscm_diff_destroy(ruv_lscm_diff* h) {
    if (!h) return;
    ruv::diff_destroy(h->impl);
    delete h;
}

int ruv_chart_metrics(int32_t nchart, const int64_t* v_off, const double* p3, const double* uv,
                      const int64_t* t_off, const uint32_t* tris, double* metrics, int64_t* counts) {
    return ruv::chart_metrics(nchart, v_off, p3, uv, t_off, tris, metrics, counts);
}

int ruv_rectangularize(int32_t nchart, const int64_t* v_off, const double* uv, const int64_t* b_off,
                       const int64_t* bnd, double* q, double* ext) {
    return ruv::rectangularize(nchart, v_off, uv, b_off, bnd, q, ext);
}

int ruv_chart_metrics_dev(int32_t nchart, const int64_t* v_off, const double* p3, const double* uv,
                          const int64_t* t_off, const uint32_t* tris, double* metrics, int64_t* counts) {
    return ruv::chart_metrics_dev(nchart, v_off, p3, uv, t_off, tris, metrics, counts);
}

int ruv_chart_metrics_dev_noalloc(int32_t nchart, const int64_t* v_off, const double* p3, const double* uv,
                                  const int64_t* t_off, const uint32_t* tris, double* metrics, int64_t* counts) {
    return ruv::chart_metrics_dev_noalloc(nchart, v_off, p3, uv, t_off, tris, metrics, counts);
}

int ruv_rectangularize_dev(int32_t nchart, const int64_t* v_off, const double* uv, const int64_t* b_off,
                           const int64_t* bnd, double* q, double* ext) {
    return ruv::rectangularize_dev(nchart, v_off, uv, b_off, bnd, q, ext);
}

}  // extern "C"

extern "C" {

int ruv_unfold_batch_run(int32_t nmesh, const double* pos, const int64_t* pos_off, const uint32_t* faces,
                         const int64_t* face_off, double weld_tol, double angle_min_deg, const ruv_unfold_opts* opts,
                         int32_t threads, ruv_unfold_batch** out) {
    if (!pos_off || !face_off || !opts || !out || (nmesh > 0 && (!pos || !faces)))
        return ruv::fail(RUV_ERR_INVALID, "null argument");
    for (int32_t m = 0; m < nmesh; ++m) {
        if (pos_off[m + 1] < pos_off[m] || face_off[m + 1] < face_off[m])
            return ruv::fail(RUV_ERR_BAD_DIMS, "unfold batch: offsets not monotone");
        const int64_t nv = pos_off[m + 1] - pos_off[m];
        for (int64_t k = 3 * face_off[m]; k < 3 * face_off[m + 1]; ++k)
            if ((int64_t)faces[k] >= nv) return ruv::fail(RUV_ERR_BAD_DIMS, "unfold batch: face index outside its mesh");
    }
    ruv::UnfoldBatch* u = nullptr;
    int rc = ruv::unfold_batch_run(nmesh, pos, pos_off, faces, face_off, weld_tol, angle_min_deg, *opts, threads, &u);
    if (rc) return rc;
    *out = new ruv_unfold_batch{u};
    return RUV_OK;
}

int ruv_unfold_batch_sizes(const ruv_unfold_batch* u, int64_t* sizes) {
    if (!u || !sizes) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::unfold_batch_sizes(u->impl, sizes);
}

int ruv_unfold_batch_results(const ruv_unfold_batch* u, int32_t* chart_mesh, int64_t* v_off, uint32_t* vertex_ids,
                             double* uv, double* rectified, double* ext, double* metrics, int64_t* counts,
                             double* stats, ruv_solve_info* info) {
    if (!u) return ruv::fail(RUV_ERR_INVALID, "null argument");
    return ruv::unfold_batch_results(u->impl, chart_mesh, v_off, vertex_ids, uv, rectified, ext, metrics, counts,
                                     stats, info);
}

void ruv_unfold_batch_destroy(ruv_unfold_batch* u) {
    if (!u) return;
    ruv::unfold_batch_destroy(u->impl);
    delete u;
}

}  // extern "C"
