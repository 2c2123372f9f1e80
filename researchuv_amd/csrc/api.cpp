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

void ruv_plan_destroy(ruv_plan* plan) {
    if (!plan) return;
    ruv::plan_destroy(plan->impl);
    delete plan;
}

}  // extern "C"
