// host_unfold_batch.cpp -- the pipeline's unfold stage for many meshes in one call (pipeline.rs:171-194).
//
// The reference runs, on exec::par_map workers, one lscm::unfold_chart per chart (each a GPU or CPU solve of its
// own), then rectangularize and chart_distortion on the chart's uv.  Here the host work stays on worker threads
// (weld, segment and assembly per mesh; the driver iterations per chart) but every chart's system goes into one
// batched device solve, and both metrics run for all charts at once on the device.  Each chart's results are
// those of unfold_chart + rectangularize + chart_distortion: uv with the pins at their targets, the UnfoldResult
// fields (lscm.rs:104-116), the rectified uv and its extent, and ChartMetrics (whose fold count is
// recut::chart_folds' measure, recut.rs:71-80).  Charts and vertices keep the reference's order: charts in
// (mesh, chart) order, each chart's vertices in its ascending welded-vertex order.
#include <algorithm>
#include <memory>
#include <thread>
#include <vector>

#include "host_lscm.hpp"

extern "C" {
int ruv_batch_create(int64_t nrows, const int64_t* row_ptr, const int32_t* cols, const double* vals, int32_t nchart,
                     const int64_t* chart_row_off, int32_t rhs_per_row, int32_t precond, int32_t mode,
                     ruv_batch** out);
int ruv_batch_solve(ruv_batch* batch, const double* b, double* x, uint64_t max_iter, double tol, ruv_solve_info* info,
                    uint64_t* chart_iters, double* chart_true_resid);
void ruv_batch_destroy(ruv_batch* batch);
}

namespace ruv {

struct UnfoldBatch {
    int64_t nchart = 0, nv = 0, nt = 0;
    std::vector<int32_t> chart_mesh;
    std::vector<int64_t> v_off{0};
    std::vector<uint32_t> vertex_ids;
    std::vector<double> uv, rectified, ext, metrics, stats;
    std::vector<int64_t> counts;
    ruv_solve_info info{};
};

namespace {

struct ChartWork {
    Chart ch;
    ChartSystem S;
};

// contiguous chunks over [0, n) on `workers` threads, in input order (exec::par_map, core/src/exec.rs:28-60)
template <typename F>
void par_chunks(int64_t n, unsigned workers, F f) {
    workers = (unsigned)std::max<int64_t>(1, std::min<int64_t>(workers, n));
    if (workers <= 1) {
        f((int64_t)0, n);
        return;
    }
    const int64_t chunk = (n + workers - 1) / workers;
    std::vector<std::thread> th;
    for (int64_t lo = 0; lo < n; lo += chunk) th.emplace_back(f, lo, std::min(n, lo + chunk));
    for (auto& t : th) t.join();
}

}  // namespace

int unfold_batch_run(int32_t nmesh, const double* pos, const int64_t* pos_off, const uint32_t* faces,
                     const int64_t* face_off, double weld_tol, double angle_min_deg, const ruv_unfold_opts& o,
                     int32_t threads, UnfoldBatch** out) {
    if (nmesh < 0) return fail(RUV_ERR_INVALID, "unfold batch: negative mesh count");
    if (o.mix_w <= 0.0) return fail(RUV_ERR_INVALID, "unfold batch: mix_w must be positive (lscm.rs:479)");
    if (o.precond != RUV_PC_NONE && o.precond != RUV_PC_JACOBI && o.precond != RUV_PC_AMG)
        return fail(RUV_ERR_INVALID, "unfold batch: the batched solver takes the none, jacobi and amg preconditioners");
    const unsigned workers = threads > 0 ? (unsigned)threads : std::max(1u, std::thread::hardware_concurrency());
    // 1. weld, segment and assemble every mesh's charts (worker threads, input order)
    std::vector<std::vector<double>> welded(nmesh);
    std::vector<std::vector<ChartWork>> per(nmesh);
    par_chunks(nmesh, workers, [&](int64_t lo, int64_t hi) {
        for (int64_t m = lo; m < hi; ++m) {
            std::vector<uint32_t> fw;
            weld(pos + 3 * pos_off[m], pos_off[m + 1] - pos_off[m], faces + 3 * face_off[m],
                 face_off[m + 1] - face_off[m], weld_tol, &welded[m], &fw);
            Charts cs;
            segment(welded[m].data(), (int64_t)welded[m].size() / 3, fw.data(), (int64_t)fw.size() / 3, angle_min_deg,
                    &cs);
            per[m].resize(cs.charts.size());
            for (size_t c = 0; c < cs.charts.size(); ++c) {
                per[m][c].ch = std::move(cs.charts[c]);
                build_system(welded[m].data(), per[m][c].ch, o.angle_mix, &per[m][c].S);
            }
        }
    });
    std::vector<ChartWork*> charts;
    std::unique_ptr<UnfoldBatch> U(new UnfoldBatch());
    for (int32_t m = 0; m < nmesh; ++m)
        for (auto& w : per[m]) {
            charts.push_back(&w);
            U->chart_mesh.push_back(m);
        }
    const int64_t C = (int64_t)charts.size();
    U->nchart = C;
    // 2. one block-diagonal system, one batched solve
    std::vector<int64_t> row_ptr{0}, crow{0};
    std::vector<int32_t> cols;
    std::vector<double> vals, rhs;
    for (ChartWork* w : charts) {
        const Csr& l = w->S.l;
        const int64_t base = crow.back();
        for (int32_t i = 0; i < l.n; ++i) {
            for (int32_t k = l.ptr[i]; k < l.ptr[i + 1]; ++k) {
                cols.push_back((int32_t)(base + l.cols[k]));
                vals.push_back(l.vals[k]);
            }
            row_ptr.push_back((int64_t)vals.size());
        }
        rhs.insert(rhs.end(), w->S.rhs.begin(), w->S.rhs.end());
        crow.push_back(base + w->S.nfree);
    }
    const int64_t nrows = crow.back();
    std::vector<double> x(2 * (size_t)nrows, 0.0);
    if (nrows > 0) {
        ruv_batch* bt = nullptr;
        int rc = ruv_batch_create(nrows, row_ptr.data(), cols.data(), vals.data(), (int32_t)C, crow.data(), 2,
                                  o.precond, o.mode, &bt);
        if (rc) return rc;
        const int max_iter = o.max_iter > 0 ? o.max_iter : 100;
        const uint64_t mi = (uint64_t)std::max(max_iter, 1024) * 16;  // lscm.rs:501, per chart
        rc = ruv_batch_solve(bt, rhs.data(), x.data(), mi, o.tol, &U->info, nullptr, nullptr);
        ruv_batch_destroy(bt);
        if (rc) return rc;
    }
    // 3. the driver iterations per chart (worker threads), uv laid out by chart
    for (ChartWork* w : charts) U->v_off.push_back(U->v_off.back() + w->S.nv);
    U->nv = U->v_off.back();
    U->uv.assign(2 * (size_t)U->nv, 0.0);
    U->stats.assign(4 * (size_t)C, 0.0);
    U->vertex_ids.resize((size_t)U->nv);
    par_chunks(C, workers, [&](int64_t lo, int64_t hi) {
        for (int64_t c = lo; c < hi; ++c) {
            ChartWork* w = charts[(size_t)c];
            driver_iterations(w->S, w->ch, x.data() + 2 * crow[c], o.max_iter, U->uv.data() + 2 * U->v_off[c],
                              U->stats.data() + 4 * c);
            std::copy(w->ch.vertex_ids.begin(), w->ch.vertex_ids.end(), U->vertex_ids.begin() + U->v_off[c]);
        }
    });
    // 4. chart_distortion and rectangularize of every chart on the device (pipeline.rs:176-177)
    std::vector<int64_t> t_off{0}, b_off{0};
    std::vector<double> p3;
    std::vector<uint32_t> tris;
    std::vector<int64_t> bnd;
    p3.reserve(3 * (size_t)U->nv);
    for (ChartWork* w : charts) {
        p3.insert(p3.end(), w->S.p3.begin(), w->S.p3.end());
        tris.insert(tris.end(), w->ch.tris.begin(), w->ch.tris.end());
        t_off.push_back(t_off.back() + (int64_t)w->ch.tris.size() / 3);
        for (uint32_t v : w->ch.border) {  // border vertices in chart-local ids (pipeline.rs:174-175)
            const int64_t li = w->ch.local_of(v);
            if (li >= 0) bnd.push_back(li);
        }
        b_off.push_back((int64_t)bnd.size());
    }
    U->nt = t_off.back();
    U->metrics.assign(3 * (size_t)C, 0.0);
    U->counts.assign(2 * (size_t)C, 0);
    U->rectified.assign(2 * (size_t)U->nv, 0.0);
    U->ext.assign(2 * (size_t)C, 0.0);
    if (C > 0) {
        int rc = chart_metrics((int32_t)C, U->v_off.data(), p3.data(), U->uv.data(), t_off.data(), tris.data(),
                               U->metrics.data(), U->counts.data());
        if (!rc) rc = rectangularize((int32_t)C, U->v_off.data(), U->uv.data(), b_off.data(), bnd.data(),
                                     U->rectified.data(), U->ext.data());
        if (rc) return rc;
    }
    *out = U.release();
    return RUV_OK;
}

void unfold_batch_destroy(UnfoldBatch* u) { delete u; }

int unfold_batch_sizes(const UnfoldBatch* u, int64_t* sizes) {
    sizes[0] = u->nchart;
    sizes[1] = u->nv;
    sizes[2] = u->nt;
    return RUV_OK;
}

int unfold_batch_results(const UnfoldBatch* u, int32_t* chart_mesh, int64_t* v_off, uint32_t* vertex_ids, double* uv,
                         double* rectified, double* ext, double* metrics, int64_t* counts, double* stats,
                         ruv_solve_info* info) {
    auto put = [](auto* dst, const auto& v) {
        if (dst) std::copy(v.begin(), v.end(), dst);
    };
    put(chart_mesh, u->chart_mesh);
    put(v_off, u->v_off);
    put(vertex_ids, u->vertex_ids);
    put(uv, u->uv);
    put(rectified, u->rectified);
    put(ext, u->ext);
    put(metrics, u->metrics);
    put(counts, u->counts);
    put(stats, u->stats);
    if (info) *info = u->info;
    return RUV_OK;
}

}  // namespace ruv
