// host_batch.cpp -- build the LSCM systems of many meshes on host worker threads, laid out as one batch.
//
// The worker-thread path of the reference (pipeline.rs:171-194 over exec::par_map, core/src/exec.rs:14-60):
// items are split into contiguous chunks, one per worker, and the results are joined in input order, so
// the output is identical for every worker count.  Each mesh is welded (weld.rs:19-57), segmented
// (segment.rs:130-223) and every chart's free block assembled (lscm.rs:512-666); the charts are then
// concatenated in (mesh, chart) order as the diagonal blocks of one CSR.
#include <algorithm>
#include <thread>
#include <vector>

#include "host_lscm.hpp"

namespace ruv {

namespace {

struct MeshOut {
    std::vector<ChartSystem> sys;
};

void build_one(const double* pos, int64_t nv, const uint32_t* faces, int64_t nf, double tol, double angle,
               double angle_mix, MeshOut* out) {
    std::vector<double> pw;
    std::vector<uint32_t> fw;
    weld(pos, nv, faces, nf, tol, &pw, &fw);
    Charts ch;
    segment(pw.data(), (int64_t)pw.size() / 3, fw.data(), (int64_t)fw.size() / 3, angle, &ch);
    out->sys.resize(ch.charts.size());
    for (size_t c = 0; c < ch.charts.size(); ++c) {
        build_system(pw.data(), ch.charts[c], angle_mix, &out->sys[c]);
        // only the solve inputs are kept
        std::vector<double>().swap(out->sys[c].p3);
    }
}

}  // namespace

int lscm_batch_build(int32_t nmesh, const double* pos, const int64_t* pos_off, const uint32_t* faces,
                     const int64_t* face_off, double tol, double angle, double angle_mix, int32_t threads,
                     LscmBatch** out) {
    std::vector<MeshOut> res(nmesh);
    unsigned workers = threads > 0 ? (unsigned)threads : std::max(1u, std::thread::hardware_concurrency());
    workers = std::min<unsigned>(workers, (unsigned)std::max(nmesh, 1));
    auto run = [&](int32_t lo, int32_t hi) {
        for (int32_t m = lo; m < hi; ++m)
            build_one(pos + 3 * pos_off[m], pos_off[m + 1] - pos_off[m], faces + 3 * face_off[m],
                      face_off[m + 1] - face_off[m], tol, angle, angle_mix, &res[m]);
    };
    if (workers <= 1) {
        run(0, nmesh);
    } else {
        const int32_t chunk = (nmesh + (int32_t)workers - 1) / (int32_t)workers;
        std::vector<std::thread> th;
        for (int32_t lo = 0; lo < nmesh; lo += chunk) th.emplace_back(run, lo, std::min(nmesh, lo + chunk));
        for (auto& t : th) t.join();
    }
    auto* B = new LscmBatch();
    int64_t nnz = 0, rows = 0, nch = 0;
    for (const auto& m : res)
        for (const auto& s : m.sys) {
            nnz += s.l.nnz();
            rows += s.nfree;
            ++nch;
        }
    B->cols.reserve(nnz);
    B->vals.reserve(nnz);
    B->row_ptr.reserve(rows + 1);
    B->rhs.reserve(2 * rows);
    for (int32_t m = 0; m < nmesh; ++m)
        for (auto& s : res[m].sys) {
            const int64_t base = B->nrows;
            for (int32_t i = 0; i < s.l.n; ++i) {
                for (int32_t k = s.l.ptr[i]; k < s.l.ptr[i + 1]; ++k) {
                    B->cols.push_back((int32_t)(base + s.l.cols[k]));
                    B->vals.push_back(s.l.vals[k]);
                }
                B->row_ptr.push_back((int64_t)B->vals.size());
            }
            B->rhs.insert(B->rhs.end(), s.rhs.begin(), s.rhs.end());
            B->nrows += s.nfree;
            B->chart_row_off.push_back(B->nrows);
            B->chart_mesh.push_back(m);
            B->chart_nv.push_back(s.nv);
            B->chart_npins.push_back((int64_t)s.pin_ids.size());
            B->nv_total += s.nv;
            ++B->nchart;
            std::vector<double>().swap(s.l.vals);  // release as the batch grows
            std::vector<int32_t>().swap(s.l.cols);
        }
    *out = B;
    return RUV_OK;
}

void lscm_batch_destroy(LscmBatch* b) { delete b; }

}  // namespace ruv
