// Internal declarations shared by the host (.cpp) and device (.hip) translation units of libruv_lscm.so.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ruv_lscm.h"

namespace ruv {

// Thread-local last-error message behind ruv_last_error().
void set_error(const std::string& msg);
const char* last_error();
int fail(int status, const std::string& msg);

// Host CSR (the CsrMatrix hand-off type, solver.rs:14-44): rows with ascending columns.
struct Csr {
    int32_t n = 0;
    std::vector<double> vals;
    std::vector<int32_t> cols;
    std::vector<int32_t> ptr;  // n + 1
    int64_t nnz() const { return (int64_t)vals.size(); }
    double get(int32_t i, int32_t j) const;
};

// Recognise A = L (x) I2 in interleaved order (rows 2i / 2i+1 are row i of L on even / odd columns).
bool detect_kron2(const Csr& a, Csr* l);

// A triangular factor pair with diagonal-LAST rows and the dependency levels of both sweeps
// (Ic0Factor, precond.rs:127-140).
struct TriFactor {
    Csr l, lt;
    std::vector<int32_t> fwd_rows, fwd_off;  // rows grouped by level, ascending
    std::vector<int32_t> bwd_rows, bwd_off;  // over L^T, descending
};

// factor_ic0's structure (precond.rs:145-219, the values are computed on the device): the L and L^T patterns and
// levels, each strict-lower L entry's and each diagonal's position in the values of A (through val_of), and the
// L entry of every L^T entry.
void ic0_pattern(const Csr& a, const std::vector<int64_t>* val_of, TriFactor* f, std::vector<int64_t>* src,
                 std::vector<int64_t>* dsrc, std::vector<int32_t>* lt_src);
// tri_levels (precond.rs:227-249) flattened as (rows, offsets)
void tri_levels(const Csr& l, bool descending, std::vector<int32_t>* rows, std::vector<int32_t>* off);
// transpose_lower (precond.rs:253-290)
Csr transpose_lower(const Csr& l);
// graph_coloring / color_permutation / permute_csr (precond.rs:52-123)
std::vector<uint32_t> graph_coloring(const Csr& a);
void color_permutation(const std::vector<uint32_t>& color, std::vector<int32_t>* to_perm, std::vector<int32_t>* to_orig);
Csr permute_csr(const Csr& a, const std::vector<int32_t>& to_perm, const std::vector<int32_t>& to_orig);
// the position in a's values of every entry of permute_csr(a, to_perm, to_orig)
std::vector<int64_t> permuted_positions(const Csr& a, const std::vector<int32_t>& to_perm,
                                        const std::vector<int32_t>& to_orig);

// AMG hierarchy (amg.rs:127-455), host side.
struct AmgHostLevel {
    Csr op;        // level operator
    Csr p;         // prolongator (n x n_coarse), empty on the last level
    Csr r;         // restriction R = P^T
    int32_t n_coarse = 0;
    std::vector<int32_t> agg;  // row -> aggregate (the next level's row), empty on the last level
    bool diag_only = false;
    bool has_factor = false;
    // coarsest level (n <= 512): the dense inverse A^-1 (row-major n x n) from a dense Cholesky, in place of
    // the reference's sparse exact factor under a nested-dissection order (precond.rs:300-498)
    std::vector<double> dense_inv;
};
bool amg_hierarchy(const Csr& a, bool smooth, std::vector<AmgHostLevel>* levels, std::string* why);
void amg_aggregate(const Csr& a, int max_members, std::vector<int32_t>* agg, int32_t* n_coarse);
Csr amg_prolongator(const Csr& a, const std::vector<int32_t>& agg, bool smooth);
Csr amg_galerkin(const Csr& a, const Csr& p, int32_t n_coarse);
Csr transpose_rect(const Csr& p, int32_t n_coarse);

// Device solve entry used by the C ABI and the unfold driver (pcg.hip).
// a: the operator rows (R right-hand sides interleaved per row).  Plain FP64 (mode 0) or mixed (mode 1).
int device_solve(const Csr& a, int R, const double* b, uint64_t max_iter, double tol, int precond, int mode,
                 double* x, ruv_solve_info* info);
int device_apply_ic0(const Csr& a, int R, const double* r, double* z);

// post-solve chart metrics on the device (metrics.hip)
int chart_metrics(int32_t nchart, const int64_t* v_off, const double* p3, const double* uv, const int64_t* t_off,
                  const uint32_t* tris, double* metrics, int64_t* counts);
int rectangularize(int32_t nchart, const int64_t* v_off, const double* uv, const int64_t* b_off, const int64_t* bnd,
                   double* q, double* ext);
int chart_metrics_dev(int32_t nchart, const int64_t* v_off, const double* p3, const double* uv, const int64_t* t_off,
                      const uint32_t* tris, double* metrics, int64_t* counts);
int chart_metrics_dev_noalloc(int32_t nchart, const int64_t* v_off, const double* p3, const double* uv,
                              const int64_t* t_off, const uint32_t* tris, double* metrics, int64_t* counts);
int rectangularize_dev(int32_t nchart, const int64_t* v_off, const double* uv, const int64_t* b_off,
                       const int64_t* bnd, double* q, double* ext);

}  // namespace ruv
