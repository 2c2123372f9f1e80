/*
 * ruv_lscm.h -- C ABI of the MI355X LSCM solve engine (libruv_lscm.so).
 *
 * The library replaces the researchuv-gpu CUDA backend behind ResearchUV's
 * solver switch (UnfoldOptions.solver = SolverBackend::Gpu,
 * crates/researchuv-unwrap/src/lscm.rs:57-66,677-681).  Every entry point takes
 * plain host pointers and sizes; the *_dev entry points take device pointers
 * (memory from hipMalloc or any allocator on the same device).  No pointer is
 * retained after a call returns.
 *
 * Threading: every entry point is thread-safe.  Each calling thread gets its
 * own HIP stream (default priority), so concurrent callers do not serialise on
 * a shared stream and never wait on each other's work.
 *
 * Status codes: 0 on success, a negative RUV_ERR_* otherwise; the reference
 * returned Option::None in every failure case and the caller fell back to the
 * CPU LU (lscm.rs:678).  ruv_last_error() gives the calling thread's message
 * for its last failure.
 */
#ifndef RUV_LSCM_H
#define RUV_LSCM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum ruv_status {
    RUV_OK = 0,
    RUV_ERR_UNAVAILABLE = -1,   /* no HIP device (GpuSolver::global() == None, solver.rs:104-126)        */
    RUV_ERR_BAD_DIMS = -2,      /* n == 0 or length mismatch (solver.rs:296-299)                           */
    RUV_ERR_BREAKDOWN = -3,     /* |p.Ap| <= 1e-300 or non-finite (solver.rs:505-508)                      */
    RUV_ERR_NOT_CONVERGED = -4, /* max_iter reached (solver.rs:551-553)                                    */
    RUV_ERR_NONFINITE = -5,     /* non-finite solution (solver.rs:561-563)                                 */
    RUV_ERR_DEVICE = -6,        /* a HIP runtime call failed                                               */
    RUV_ERR_PRECOND = -7,       /* preconditioner setup failed: zero diagonal (precond.rs:596-608,
                                   amg.rs:140-142), IC(0) pivot (precond.rs:179-188), AMG depth
                                   (amg.rs:159-161)                                                         */
    RUV_ERR_INVALID = -8        /* null pointer or unknown enum value                                     */
};

/* Precond (crates/researchuv-gpu/src/precond.rs:20-47); same names as RESEARCHUV_GPU_PRECOND
 * (lscm.rs:492-499): none, jacobi, ic0, ic0color, amg, amgsa. */
enum ruv_precond {
    RUV_PC_NONE = 0,
    RUV_PC_JACOBI = 1,
    RUV_PC_IC0 = 2,
    RUV_PC_IC0COLOR = 3,
    RUV_PC_AMG = 4,
    RUV_PC_AMGSA = 5
};

/* Solve statistics (no reference counterpart: the reference discarded its iteration count,
 * lscm.rs:500-502). */
typedef struct ruv_solve_info {
    uint64_t iters;          /* PCG iterations (inner FP32 iterations summed for the mixed form)      */
    uint32_t refinements;    /* outer FP64 refinement steps (0 for the plain FP64 form)               */
    int32_t precond;         /* preconditioner actually applied                                        */
    double true_rel_resid;   /* ||b - A x||_2 / ||b||_2 computed in FP64 after the solve              */
    double seconds;          /* device time of the solve (HIP events), excluding setup and copies     */
} ruv_solve_info;

/* GpuSolver::global().is_some() (solver.rs:104-113): 1 when a HIP device is usable. */
int ruv_available(void);

/* Message for the calling thread's last non-zero status ("" if none). */
const char* ruv_last_error(void);

/* GpuSolver::cg_solve_precond (solver.rs:284-565): PCG in FP64, x0 = 0, stop when the recursive residual
 * satisfies ||r|| <= tol*||b|| (tol absolute when ||b|| <= 1e-300), at most max_iter iterations.  The
 * reference ran flexible CG for its AMG preconditioners (solver.rs:474-481) because its K-cycle varies from
 * one application to the next; here AMG is a fixed symmetric V(1,1) cycle (damped Jacobi smoothing, exact
 * coarsest solve), an SPD operator, so every preconditioner runs under plain PCG (beta = rz_new / rz).
 * A is SPD CSR (vals/cols of length nnz, row_ptr of length n+1), b of length b_len (must equal n).
 * x_out (n) and iters_out are written on success.  An interleaved
 * A = L (x) I2 (the LSCM free block of lscm.rs:605-640 as produced by Sparse::to_csr) is detected and
 * streamed as L against two right-hand sides; the iterates are those of CG on A. */
int ruv_cg_solve_precond(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                         const double* b, int64_t b_len, uint64_t max_iter, double tol, int32_t precond,
                         double* x_out, uint64_t* iters_out);

/* GpuSolver::cg_solve (solver.rs:274-277): ruv_cg_solve_precond with RUV_PC_NONE. */
int ruv_cg_solve(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                 const double* b, int64_t b_len, uint64_t max_iter, double tol, double* x_out);

/* GpuSolver::apply_ic0_for_test (solver.rs:205-270): z = (L L^T)^-1 r with the IC(0) factor of A
 * (precond.rs:145-219) applied by the device triangular sweeps. */
int ruv_apply_ic0(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                  const double* r, double* z_out);
/* Inspection hook: the IC(0) factor of A (factor_ic0, precond.rs:145-219) as the device factorisation computes it
 * (structure on the host, values by level-scheduled device kernels in the reference's operation order): L's rows,
 * strict-lower entries with ascending columns and the diagonal last (l_ptr n + 1, l_cols and l_vals *l_nnz).  Call
 * with null arrays for *l_nnz.  RUV_ERR_PRECOND where factor_ic0 returns None (a vanishing pivot or diagonal). */
int ruv_ic0_factor(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                   int64_t* l_nnz, int32_t* l_ptr, int32_t* l_cols, double* l_vals);

/* Mixed-precision form (no reference counterpart; the reference GPU path is FP64 throughout,
 * solver.rs:284-565): the SpMV, preconditioner and Krylov vectors run in FP32 inside an FP64
 * iterative-refinement loop (r = b - A x, the x update and the CG alpha/beta reductions in FP64) until the
 * true FP64 residual satisfies ||b - A x|| <= tol*||b|| and the forward-error estimate does not exceed
 * 1e-7 max|x|: each refinement round solves A d = r_start, so gamma = max|d| / ||r_start|| (the largest of
 * the rounds) measures how far a residual reaches in x, and gamma ||r|| estimates the error left.  A residual
 * test alone does not bound the UV error on large charts (the 512^2 height field at ||r|| 2.5e-9 ||b|| was
 * 2e-6 from the LU).  The environment variable RUV_UV_TOL replaces 1e-7 (0: residual test only).  max_iter
 * bounds the summed inner iterations. */
int ruv_cg_solve_mixed(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                       const double* b, int64_t b_len, uint64_t max_iter, double tol, int32_t precond,
                       double* x_out, ruv_solve_info* info);

/* ---------------------------------------------------------------------------------------------------------
 * AMG setup inspection (test hooks; the reference checks its setup in amg.rs:799-942).  The plan path builds
 * the reference's hierarchy on the host (amg.rs:127-455: matched aggregation, tentative or smoothed
 * prolongator, Galerkin operators, exactly factored coarsest level); these return it level by level.
 * ------------------------------------------------------------------------------------------------------- */
typedef struct ruv_amg_host ruv_amg_host;
/* aggregate (amg.rs:254-291): agg (n) row -> aggregate, *n_coarse aggregates */
int ruv_amg_aggregate(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                      int32_t max_members, int32_t* agg, int32_t* n_coarse);
/* prolongator (amg.rs:308-362): the tentative (smoothed = 0) or smoothed P = (I - 2/3 D^-1 A) B, truncated to 32
 * entries a row; sizes (2): rows, entries; arrays null: sizes only. */
int ruv_amg_prolongator(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                        const int32_t* agg, int32_t n_coarse, int32_t smoothed, int64_t* sizes, double* p_vals,
                        int32_t* p_cols, int32_t* p_ptr);
/* galerkin_p (amg.rs:384-431): A_c = P^T A P; sizes (2) as above. */
int ruv_amg_galerkin(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                     const double* p_vals, const int32_t* p_cols, const int32_t* p_ptr, int32_t n_coarse,
                     int64_t* sizes, double* c_vals, int32_t* c_cols, int32_t* c_ptr);
/* hierarchy (amg.rs:127-164); smoothed = 0: Amg (pairs, tentative P), 1: AmgSa (SA_AGG members, smoothed P) */
int ruv_amg_host_build(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t n, int64_t nnz,
                       int32_t smoothed, ruv_amg_host** out);
int32_t ruv_amg_host_levels(const ruv_amg_host* h);
/* sizes (5): rows, nnz of the level operator, nnz of P (0 on the last level), coarse rows, 1 when the level is
 * factored exactly (the last).  Any array may be null: op (CSR, rows + 1 pointers), P (rows x coarse rows, CSR),
 * agg (rows; absent on the last level). */
int ruv_amg_host_level(const ruv_amg_host* h, int32_t level, int64_t* sizes, double* op_vals, int32_t* op_cols,
                       int32_t* op_ptr, double* p_vals, int32_t* p_cols, int32_t* p_ptr, int32_t* agg);
void ruv_amg_host_destroy(ruv_amg_host* h);

/* ---------------------------------------------------------------------------------------------------------
 * Device-resident plans: upload and set up once, solve many times on device buffers.  A plan holds the
 * system L (x) I_R for R = 1 (generic CSR) or R = 2 (the LSCM free block: L is the cotan Laplacian on the
 * free vertices, vectors interleaved (u0,v0,u1,v1,...) as in lscm.rs:616-629).  mode 0 = plain FP64 PCG,
 * mode 1 = FP32 PCG inside FP64 refinement.  No reference counterpart (the reference re-uploaded and re-set
 * up every solve, solver.rs:301-314).
 * ------------------------------------------------------------------------------------------------------- */
typedef struct ruv_plan ruv_plan;

enum ruv_mode { RUV_MODE_FP64 = 0, RUV_MODE_MIXED = 1 };

int ruv_plan_create(const double* vals, const int32_t* cols, const int32_t* row_ptr, int32_t nrows, int64_t nnz,
                    int32_t rhs_per_row, int32_t precond, int32_t mode, ruv_plan** out);
/* d_b, d_x: device arrays of nrows*rhs_per_row doubles (interleaved).  tol as in ruv_cg_solve_precond for
 * mode 0 (recursive residual) and ruv_cg_solve_mixed for mode 1 (true residual). */
int ruv_plan_solve_dev(ruv_plan* plan, const double* d_b, double* d_x, uint64_t max_iter, double tol,
                       ruv_solve_info* info);
/* Host-buffer convenience around ruv_plan_solve_dev (copies b in, x out). */
int ruv_plan_solve(ruv_plan* plan, const double* b, double* x, uint64_t max_iter, double tol, ruv_solve_info* info);
void ruv_plan_destroy(ruv_plan* plan);
/* New values for the plan's operator (nnz doubles in the CSR order it was created with; same pattern): the
 * operator and its preconditioner are refreshed on the device -- the Jacobi diagonal, or the IC(0) factor
 * recomputed by the device factorisation from its kept structure (factor_ic0, precond.rs:145-219; the reference
 * builds a new factor per solve, solver.rs:284-360) -- so refinement rounds and driver iterations reuse the plan.
 * RUV_ERR_PRECOND on a vanishing pivot or diagonal (the plan then must not be solved with); RUV_ERR_INVALID for
 * the amg preconditioners, whose hierarchy is built on the host (create a new plan).  Not concurrent with a solve
 * of the same plan. */
int ruv_plan_update_values(ruv_plan* plan, const double* vals);
/* Measurement hook (no reference counterpart): run the plan's inner-precision SpMV kernel (the one its PCG
 * iterations launch) reps times back to back on the plan's stream and return the mean launch duration from
 * HIP events, plus the kernel's algorithmic bytes per launch: nnz*(s_v + 4) + (nrows + 1)*4 +
 * 2*nrows*rhs_per_row*s_x (every matrix entry, index, x element and y element counted once). */
int ruv_plan_spmv_time(ruv_plan* plan, int32_t reps, double* seconds_per_launch, double* bytes_per_launch);

/* ---------------------------------------------------------------------------------------------------------
 * Batched solves (no reference counterpart: the reference solved one chart per worker-thread call,
 * pipeline.rs:171-194 -> lscm.rs:471-506).  A batch holds nchart independent systems L_c (x) I_R as one
 * block-diagonal CSR: rows of chart c are [chart_row_off[c], chart_row_off[c+1]) and its columns stay inside
 * that range (row_ptr is 64-bit; columns are global row indices).  Every chart runs its own PCG (own
 * alpha/beta/stopping test) inside shared launches; precond none, jacobi or amg (an aggregation hierarchy
 * built on the device at create time, block-diagonal by chart, applied as a V(1,1) cycle).  mode 0: FP64 PCG per chart with
 * the recursive-residual test tol*||b_c||; mode 1: FP32 PCG inside per-chart FP64 refinement until
 * ||b_c - A_c x_c|| <= tol*||b_c|| and the chart's forward-error estimate is at most 1e-7 max|x_c| (the test
 * of ruv_cg_solve_mixed, per chart).  info->iters is the largest per-chart iteration count and
 * info->true_rel_resid the worst chart; chart_iters / chart_true_resid (nchart, optional) give every chart.
 * ------------------------------------------------------------------------------------------------------- */
typedef struct ruv_batch ruv_batch;
int ruv_batch_create(int64_t nrows, const int64_t* row_ptr, const int32_t* cols, const double* vals, int32_t nchart,
                     const int64_t* chart_row_off, int32_t rhs_per_row, int32_t precond, int32_t mode,
                     ruv_batch** out);
int ruv_batch_solve_dev(ruv_batch* batch, const double* d_b, double* d_x, uint64_t max_iter, double tol,
                        ruv_solve_info* info, uint64_t* chart_iters, double* chart_true_resid);
int ruv_batch_solve(ruv_batch* batch, const double* b, double* x, uint64_t max_iter, double tol,
                    ruv_solve_info* info, uint64_t* chart_iters, double* chart_true_resid);
/* Measurement hook: mean duration of the batch's inner-precision SpMV kernel (q = A p' with the p'.q partials)
 * over every tile, and its bytes per launch: algorithmic (every nonzero's value and, in a general slice, its
 * stored column index; the slice tables; p' read and q written once) and as stored (the zero slots that pad a
 * sliced-ELL slice to its width included; null to skip). */
int ruv_batch_spmv_time(ruv_batch* batch, int32_t reps, double* seconds_per_launch, double* bytes_per_launch,
                        double* stored_bytes_per_launch);
/* The batch's storage and the schedule its autotuner chose: rows per workgroup tile; bytes per stored column
 * index of the general slices (2: 16-bit offsets from the row, 4: 32-bit rows); stored values (padding
 * included) and stored column indices; slices in stencil form (column offsets per slice, none per entry) and
 * all slices; 1 when the step kernel loads the values with the non-temporal hint.  Any pointer may be null. */
int ruv_batch_schedule(const ruv_batch* batch, int32_t* rows_per_tile, int32_t* col_bytes, int64_t* ell_entries,
                       int64_t* col_entries, int32_t* stencil_slices, int32_t* slices, int32_t* nontemporal);
/* Charts the batch also keeps in band form -- free rows forming a grid of width W in row-major order (column
 * offsets 0, +-1, +-(W - 1), +-W, +-(W + 1) only) with an exactly symmetric operator -- and the values that form
 * stores (the diagonal and the upper offsets present, one slot per row each, zeros included).  The step's SpMV
 * and the AMG's finest smoothing passes stream that form on those charts (RUV_BATCH_BAND=0 at creation turns it
 * off).  Either pointer may be null.  No counterpart in the reference (its CSR SpMV stores every nonzero). */
int ruv_batch_band_info(const ruv_batch* batch, int32_t* band_charts, int64_t* band_values);
/* Step-kernel accounting over the solves that follow (measurement hook): with enable != 0 every launch of the
 * SpMV kernel is bracketed by HIP events on the batch's stream.  ruv_batch_kernel_stats returns, and resets,
 * the launch count, the summed device time of those launches, their algorithmic bytes and their bytes as
 * stored (per chart as in ruv_batch_spmv_time, times the steps the chart ran; any pointer may be null). */
int ruv_batch_kernel_timing(ruv_batch* batch, int32_t enable);
/* precond amg: levels of the device-built aggregation hierarchy (finest included), the device time its setup
 * took inside ruv_batch_create, and the rows of all coarse levels together; zeros for the other
 * preconditioners. */
int ruv_batch_precond_info(ruv_batch* batch, int32_t* levels, double* setup_seconds, int64_t* coarse_rows);
/* precond amg: rebuild the hierarchy on the device from the batch's operator (device work only; create builds
 * it once); a no-op for the other preconditioners. */
int ruv_batch_setup_precond(ruv_batch* batch);
/* precond amg: rows of each level, finest first, into rows[0..cap); returns the number of levels (0 for the
 * other preconditioners). */
int ruv_batch_precond_levels(ruv_batch* batch, int64_t* rows, int32_t cap);
/* precond amg: charts of the last build whose coarsest block is applied through its diagonal instead of its exact
 * inverse (the coarsest blocks are factored on the device in FP64, Cholesky then the inverse of the factor,
 * amg.rs:685-773 and precond.rs:439-498 factor theirs exactly on the host): blocks that are not numerically
 * positive definite, blocks coupled to another chart, and blocks of more than 88 rows.  Zeros for the other
 * preconditioners.  Any pointer may be null. */
int ruv_batch_precond_fallbacks(ruv_batch* batch, int64_t* not_spd_charts, int64_t* coupled_charts,
                                int64_t* too_large_charts);
/* precond amg, inspection hook: chart's block of the coarsest operator (dense, row-major, FP64 from the stored
 * values) and the FP64 inverse the device factorisation produced for it; *rows receives the block size (call with
 * null block and inverse to size the buffers).  RUV_ERR_PRECOND for a chart kept on its diagonal. */
int ruv_batch_coarsest_block(ruv_batch* batch, int32_t chart, int64_t* rows, double* block, double* inverse);
/* precond amg, inspection hook: level `level` of the device hierarchy in its stored order.  sizes (3): rows, stored
 * operator entries (padding dropped; 0 for level 0, whose operator is the batch's own), rows of the next level
 * (0 on the coarsest).  Any array may be null: the operator as CSR (row_ptr rows + 1, cols, vals: the values the
 * cycle applies, in FP64; the levels between the finest and the coarsest are stored symmetrically scaled,
 * a_ij / sqrt(d_i d_j), and returned times the stored sqrt(d_i) sqrt(d_j)), agg (rows: each row's aggregate, a row of the next level; absent on the coarsest) and chart
 * (rows: each row's chart). */
int ruv_batch_amg_level(ruv_batch* batch, int32_t level, int64_t* sizes, int64_t* row_ptr, int32_t* cols,
                        double* vals, int32_t* agg, int32_t* chart);
int ruv_batch_kernel_stats(ruv_batch* batch, int64_t* launches, double* seconds, double* bytes,
                           double* stored_bytes);
/* precond amg, measurement hook (no counterpart in the reference): reps steps of the AMG-preconditioned inner
 * iteration over the whole batch (every chart running, every level whole), each kernel launch bracketed by HIP
 * events on the batch's stream.  One entry per (kind, level): kind 0 direction update, 1 SpMV, 2 x/r update,
 * 3 finest pre-smoothing + residual, 4 finest restriction, 5 coarse pre-smoothing + residual + restriction,
 * 6 coarsest solve, 7 coarse prolongation, 8 coarse post-smoothing, 9 finest prolongation, 10 finest
 * post-smoothing; the mean seconds per launch and the bytes of one launch, algorithmic (every nonzero's value and
 * stored column, every vector element read or written once, the index arrays) and as stored (padding slots
 * included).  Entries go to [0, cap) in launch order; *count receives their number.  The solver's state is
 * overwritten (a later solve starts afresh).  RUV_ERR_INVALID without a built AMG hierarchy. */
int ruv_batch_step_profile(ruv_batch* batch, int32_t reps, int32_t cap, int32_t* kind, int32_t* level,
                           double* seconds, double* bytes, double* stored_bytes, int32_t* count);
void ruv_batch_destroy(ruv_batch* batch);

/* Build the LSCM systems of nmesh meshes on host worker threads (the worker-thread unfold stage,
 * pipeline.rs:171-194 over exec::par_map, core/src/exec.rs:14-60; threads 0 = all cores): each mesh m
 * (positions pos[pos_off[m]..pos_off[m+1]) as xyz triples, faces[face_off[m]..face_off[m+1]) as index
 * triples local to the mesh) is welded, segmented and every chart assembled; charts are concatenated in
 * (mesh, chart) order as the diagonal blocks of one batch (rhs_per_row = 2). */
typedef struct ruv_lscm_batch ruv_lscm_batch;
int ruv_lscm_batch_build(int32_t nmesh, const double* pos, const int64_t* pos_off, const uint32_t* faces,
                         const int64_t* face_off, double weld_tol, double angle_min_deg, double angle_mix,
                         int32_t threads, ruv_lscm_batch** out);
/* sizes (4): nchart, nrows (free vertices), nnz(L), total chart vertices */
int ruv_lscm_batch_sizes(const ruv_lscm_batch* lb, int64_t* sizes);
/* any pointer may be null: row_ptr (nrows+1), cols / vals (nnz), rhs (2*nrows), chart_row_off (nchart+1),
 * chart_mesh / chart_nv / chart_npins (nchart) */
int ruv_lscm_batch_arrays(const ruv_lscm_batch* lb, int64_t* row_ptr, int32_t* cols, double* vals, double* rhs,
                          int64_t* chart_row_off, int32_t* chart_mesh, int64_t* chart_nv, int64_t* chart_npins);
int ruv_batch_create_lscm(const ruv_lscm_batch* lb, int32_t precond, int32_t mode, ruv_batch** out);
void ruv_lscm_batch_destroy(ruv_lscm_batch* lb);

/* ---------------------------------------------------------------------------------------------------------
 * Charting and the per-chart LSCM driver (host side of the unfold path; the solves run on the device).
 * ------------------------------------------------------------------------------------------------------- */

/* weld::weld (crates/researchuv-unwrap/src/weld.rs:19-57).  pos_out (nv*3) and faces_out (nf*3) are caller
 * buffers of the input sizes; *nv_out / *nf_out receive the welded counts. */
int ruv_weld(const double* pos, int64_t nv, const uint32_t* faces, int64_t nf, double tol, double* pos_out,
             uint32_t* faces_out, int64_t* nv_out, int64_t* nf_out);

/* segment::segment (segment.rs:130-140) with build_charts / chart_border_loops (:143-331): the charts of a
 * welded mesh.  The handle keeps the charts; the mesh arrays are not retained. */
typedef struct ruv_charts ruv_charts;
int ruv_segment(const double* pos, int64_t nv, const uint32_t* faces, int64_t nf, double angle_min_deg,
                ruv_charts** out);
int64_t ruv_charts_count(const ruv_charts* charts);
/* sizes (6): nfaces, nverts, nborder, nloops, total loop length, cut edges of the whole segmentation */
int ruv_chart_sizes(const ruv_charts* charts, int64_t i, int64_t* sizes, double* area3d);
/* segment::Chart fields (segment.rs:32-46); loops flattened with loop_off (nloops + 1).  Any pointer may
 * be null to skip that field. */
int ruv_chart_arrays(const ruv_charts* charts, int64_t i, uint32_t* face_ids, uint32_t* vertex_ids,
                     uint32_t* tris, uint32_t* border, uint32_t* loops, int64_t* loop_off);
void ruv_charts_destroy(ruv_charts* charts);

/* The free-vertex system of chart i (lscm.rs:512-666): tangent frame p2 (nv*2), pins (local ids ascending
 * by source id, pin_uv 2 per pin), free_ids (nfree), and A = L (x) I2 as L in CSR (nfree rows; columns
 * ascending) with the interleaved right-hand side rhs (2*nfree).  sizes (5): nv, nfree, npins, nnz(L),
 * weighted edges.  Call with null arrays to get the sizes. */
int ruv_lscm_system(const double* mesh_pos, const ruv_charts* charts, int64_t i, double angle_mix, int64_t* sizes,
                    double* p2, int32_t* pin_ids, double* pin_uv, int32_t* free_ids, double* l_vals,
                    int32_t* l_cols, int32_t* l_ptr, double* rhs);

typedef struct ruv_unfold_opts {
    int32_t max_iter;   /* UnfoldOptions.max_iter (0 -> 100, lscm.rs:46,574)                            */
    double angle_mix;   /* UnfoldOptions.angle_mix (1.0, lscm.rs:50)                                    */
    double mix_w;       /* UnfoldOptions.mix_w (1.0, lscm.rs:54); <= 0 is rejected like lscm.rs:479     */
    int32_t precond;    /* ruv_precond; RESEARCHUV_GPU_PRECOND default is jacobi (lscm.rs:492-499)       */
    int32_t mode;       /* ruv_mode                                                                     */
    int32_t min_vars;   /* GPU size gate on 2*nfree: lscm.rs:462 uses 1024; 0 disables it              */
    double tol;         /* 1e-12 reproduces lscm.rs:501 (mode 0); the mixed form targets the true residual */
    int32_t device_assembly; /* 1: cotan weights, L and rhs computed on the device (bitwise the host ones) */
} ruv_unfold_opts;

/* ---------------------------------------------------------------------------------------------------------
 * Differentiable form (no reference counterpart).  One chart, chart-local: p3 (nv*3) positions, tris
 * (ntri*3), frozen pins (pin_ids, pin_uv 2 per pin); default path (angle_mix = 1).  Forward: uv (nv*2) and the
 * solver's quadratic energy E = sum over assembled edges of c_e |uv_i - uv_j|^2 (c_e the cotan weight of
 * lscm.rs:252-274,610).  VJP: grad_p3 (nv*3) = (d uv / d p)^T g_uv + energy_weight * (dE/dp at fixed uv),
 * computed with one extra PCG on the same SPD matrix (L^T = L) and a device gather through the assembly.
 * ------------------------------------------------------------------------------------------------------- */
int ruv_lscm_forward(const double* p3, int64_t nv, const uint32_t* tris, int64_t ntri, const int32_t* pin_ids,
                     const double* pin_uv, int64_t npins, int32_t precond, int32_t mode, double tol, double* uv_out,
                     double* energy_out, ruv_solve_info* info);
int ruv_lscm_vjp(const double* p3, int64_t nv, const uint32_t* tris, int64_t ntri, const int32_t* pin_ids,
                 const double* pin_uv, int64_t npins, const double* uv, const double* g_uv, double energy_weight,
                 int32_t precond, int32_t mode, double tol, double* grad_p3, ruv_solve_info* info);

/* The same system as ruv_lscm_system with the cotan weights (lscm.rs:252-274) and the free-block assembly
 * (lscm.rs:594-666) computed on the device: one thread per edge, one per free row.  sizes (5) as in
 * ruv_lscm_system; call with null arrays for the sizes.  The result is bitwise the host assembly. */
int ruv_lscm_system_device(const double* mesh_pos, const ruv_charts* charts, int64_t i, double angle_mix,
                           int64_t* sizes, double* l_vals, int32_t* l_cols, int32_t* l_ptr, double* rhs);

/* lscm::unfold_chart (lscm.rs:512-727) with SolverBackend::Gpu.  uv_out (nv*2) in chart-local order; stats
 * (optional, 4): iters, err_accum, target, area3d (UnfoldResult, lscm.rs:104-116); info optional.  A chart
 * below the size gate returns RUV_ERR_BAD_DIMS and a failed solve its status: the caller then runs its CPU
 * LU, as lscm.rs:677-681 does. */
int ruv_unfold_chart(const double* mesh_pos, const ruv_charts* charts, int64_t i, const ruv_unfold_opts* opts,
                     double* uv_out, double* stats, ruv_solve_info* info);

/* The unfold stage of pipeline::run (pipeline.rs:171-194) over nmesh meshes in one call: weld (weld_tol) and segment
 * (angle_min_deg) each mesh and assemble its charts on host worker threads (threads 0 = all cores, input order as
 * exec::par_map), solve every chart's system in one batched device solve (opts: precond none / jacobi / amg, mode,
 * tol as ruv_batch_solve, max_iter, angle_mix, mix_w; the size gate min_vars does not apply: every chart goes to
 * the device), run lscm::unfold_chart's driver iterations per chart (lscm.rs:602-718), and compute
 * rectangularize and chart_distortion of every chart on the device (pipeline.rs:176-177).  The results per chart
 * are those of the reference's stage: its uv (pins included), UnfoldResult (lscm.rs:104-116), rectified uv and
 * extent, and ChartMetrics, whose fold count is recut::chart_folds' measure (recut.rs:71-80). */
typedef struct ruv_unfold_batch ruv_unfold_batch;
int ruv_unfold_batch_run(int32_t nmesh, const double* pos, const int64_t* pos_off, const uint32_t* faces,
                         const int64_t* face_off, double weld_tol, double angle_min_deg, const ruv_unfold_opts* opts,
                         int32_t threads, ruv_unfold_batch** out);
/* sizes (3): charts, chart vertices in all, chart triangles in all */
int ruv_unfold_batch_sizes(const ruv_unfold_batch* u, int64_t* sizes);
/* Any pointer may be null.  Charts in (mesh, chart) order: chart_mesh (nchart); v_off (nchart + 1) into the per-
 * vertex arrays, chart c's vertices in ascending welded-vertex order; vertex_ids (the welded mesh's vertex of each);
 * uv and rectified (2 per vertex); ext (2 per chart); metrics (3 per chart: conformal_mean, conformal_max,
 * area_ratio_mean); counts (2 per chart: flips, folds); stats (4 per chart: iters, err_accum, target, area3d);
 * info: the batched solve. */
int ruv_unfold_batch_results(const ruv_unfold_batch* u, int32_t* chart_mesh, int64_t* v_off, uint32_t* vertex_ids,
                             double* uv, double* rectified, double* ext, double* metrics, int64_t* counts,
                             double* stats, ruv_solve_info* info);
void ruv_unfold_batch_destroy(ruv_unfold_batch* u);

/* ---------------------------------------------------------------------------------------------------------
 * The differentiable batch (no reference counterpart; the forward is lscm.rs:594-666 + the GPU solve, with pins
 * frozen as in ruv_lscm_forward).  nchart charts as one disjoint union: chart c owns vertices
 * [chart_vertex_off[c], chart_vertex_off[c+1]) of p3 (xyz triples); tris index the union and never span two
 * charts; pin_ids index the union.  The topology, the operator's structure (every edge with a free end, zero
 * weights held as explicit zeros) and the batched solver (precond none / jacobi / amg, mode as ruv_batch_create,
 * tol its tolerance) are built once.  forward: positions -> uv of every vertex (free: solved; pins: targets)
 * and the solver's energy sum_e c_e |uv_i - uv_j|^2.  vjp: with the state of the last forward, the gradient
 * w.r.t. the positions of energy_weight * E + <g_uv, uv> (one adjoint solve with the same SPD operator and a
 * gather through the assembly).
 * ------------------------------------------------------------------------------------------------------- */
typedef struct ruv_lscm_diff ruv_lscm_diff;
int ruv_lscm_diff_create(const double* p3, int64_t nv, const uint32_t* tris, int64_t ntri, const int32_t* pin_ids,
                         const double* pin_uv, int64_t npins, const int64_t* chart_vertex_off, int32_t nchart,
                         int32_t precond, int32_t mode, double tol, ruv_lscm_diff** out);
int ruv_lscm_diff_forward(ruv_lscm_diff* h, const double* p3, double* uv_out, double* energy_out,
                          ruv_solve_info* info);
int ruv_lscm_diff_vjp(ruv_lscm_diff* h, const double* g_uv, double energy_weight, double* grad_p3,
                      ruv_solve_info* info);
void ruv_lscm_diff_destroy(ruv_lscm_diff* h);

/* ---------------------------------------------------------------------------------------------------------
 * Post-solve chart metrics, many charts per call on the device (the unfold stage runs both on every chart
 * after its solve, pipeline.rs:171-182).  Chart c owns vertices [v_off[c], v_off[c+1]) of p3 (xyz: the mesh
 * positions of the chart's vertex_ids, chart-local order) and uv (u, v); v_off[0] = 0.  Host pointers.
 * ------------------------------------------------------------------------------------------------------- */
/* chart_distortion (crates/researchuv-unwrap/src/metrics.rs:73-141; svd2 :59-71, winding_flips :39-56): chart
 * c's triangles are [t_off[c], t_off[c+1]) of tris, chart-local corner triples in mesh face order.
 * metrics (nchart*3): conformal_mean, conformal_max, area_ratio_mean; counts (nchart*2): flips, folds. */
int ruv_chart_metrics(int32_t nchart, const int64_t* v_off, const double* p3, const double* uv,
                      const int64_t* t_off, const uint32_t* tris, double* metrics, int64_t* counts);
/* rectangularize (metrics.rs:147-227): chart c's border vertices are [b_off[c], b_off[c+1]) of bnd (distinct
 * chart-local ids).  q (2 per vertex): the rectified uv in [0,1]^2; ext (nchart*2): the extent before
 * normalisation (the pack rectangle, pipeline.rs:178). */
int ruv_rectangularize(int32_t nchart, const int64_t* v_off, const double* uv, const int64_t* b_off,
                       const int64_t* bnd, double* q, double* ext);
/* The same two with every array a device pointer (inputs resident in HBM, results written there).  Each call
 * runs on a stream of its own (leased for the call; concurrent calls lease different streams and scratch
 * buffers) that first waits for the work issued so far on the null stream, so inputs written there or on any
 * blocking stream are read complete; inputs written on another non-blocking stream must be complete (that
 * stream synchronised) before the call.  The call returns when its results are written.  Offsets are trusted;
 * a triangle corner or border id outside its chart is skipped on the device and reported as RUV_ERR_BAD_DIMS;
 * repeated border ids give an unspecified snap. */
int ruv_chart_metrics_dev(int32_t nchart, const int64_t* v_off, const double* p3, const double* uv,
                          const int64_t* t_off, const uint32_t* tris, double* metrics, int64_t* counts);
/* ruv_chart_metrics_dev without any allocation (no counterpart in the reference; for callers that cannot
 * allocate device memory: tight serving loops, harnesses that forbid hipMalloc).  The same kernels and results
 * (bitwise) as ruv_chart_metrics_dev; its scratch (error flag, tile table, per-tile partials) is a fixed 8 MiB
 * device array held by each GPU the library runs on until the process ends, enough for about 116 000 tiles of
 * 1024 triangles (nchart + ceil(triangles / 1024) tiles); larger inputs are rejected with RUV_ERR_BAD_DIMS
 * before anything is launched.  Runs on the null stream and returns when its results are written.  Every call on
 * one GPU shares that GPU's scratch: calls that overlap on one GPU race. */
int ruv_chart_metrics_dev_noalloc(int32_t nchart, const int64_t* v_off, const double* p3, const double* uv,
                                  const int64_t* t_off, const uint32_t* tris, double* metrics, int64_t* counts);
int ruv_rectangularize_dev(int32_t nchart, const int64_t* v_off, const double* uv, const int64_t* b_off,
                           const int64_t* bnd, double* q, double* ext);

#ifdef __cplusplus
}
#endif
#endif /* RUV_LSCM_H */
