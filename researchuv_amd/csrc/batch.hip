// batch.hip -- batched LSCM solves: thousands of independent charts, three launches per PCG step.
//
// The reference solved charts one at a time from worker threads (crates/researchuv-unwrap/src/pipeline.rs:
// 171-194 -> lscm.rs:471-506 -> solver.rs:284-565), each solve paying its own uploads, host-synchronised dot
// products and launches.  Here every chart of a batch is one diagonal block of a single block-diagonal matrix
// in sliced-ELL form (batch_kernels.hpp); each chart keeps its own CG scalars, so each chart runs exactly an
// independent PCG while all of them share the launches.  Converged charts drop out: between chunks of steps the
// host compacts the list of live tiles, so late steps launch only the tiles of the charts still running.
//
// Mixed precision: the inner PCG runs in FP32 (matrix, Jacobi diagonal, Krylov vectors); the outer loop keeps
// x and r = b - A x in FP64 per chart and stops a chart when its true residual ||b - A x|| <= tol ||b||.  An
// inner solve after the first goes on from the previous one's last direction (the residual replaced by the
// refined one) when the two residuals agree closely, else starts afresh.
//
// Schedule: the number of slices per tile (rows per workgroup) is picked per batch by timing the step kernel
// over the whole batch for each candidate (RUV_BATCH_SLICES_PER_TILE=1|2|4|8 forces one).  The storage form of
// each slice (stencil or general, see batch_kernels.hpp) is decided from its column offsets at setup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "batch_impl.hpp"
#include "batch_kernels.hpp"
#include "ruv_internal.hpp"

namespace ruv {

int bamg_upload_fine(BatchImpl& Q, const int64_t* ptr, const int32_t* cols, const double* vals);
int bamg_build(BatchImpl& Q);
int bamg_set_fine_values(BatchImpl& Q, const double* d_vals);
template <typename T, int R, int CI>
int bamg_apply(BatchImpl& Q, int na, int par, const T* r, T* z);
int bamg_set_live(BatchImpl& Q, const int32_t* done);
int bamg_info(const BatchImpl& Q, int* levels, double* setup_seconds, int64_t* coarse_rows);
int bamg_level_rows(const BatchImpl& Q, int64_t* rows, int cap);
int bamg_fallbacks(BatchImpl& Q, int64_t* not_spd, int64_t* coupled, int64_t* too_large);
int bamg_coarsest_block(BatchImpl& Q, int c, int64_t* m_out, double* block, double* inverse);
int bamg_level(BatchImpl& Q, int l, int64_t* sizes, int64_t* row_ptr, int32_t* cols, double* vals, int32_t* agg,
               int32_t* chart);

namespace {

// run f(lo, hi) over [0, n) on up to 16 host threads
template <typename F>
void parallel_for(int64_t n, F f) {
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(16, n / 4096));
    if (nt == 1) {
        f((int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int k = 0; k < nt; ++k) th.emplace_back(f, n * k / nt, n * (k + 1) / nt);
    for (auto& t : th) t.join();
}

// (Re)build the tiles for `spt` slices per tile and their per-tile buffers.
int build_tiles(BatchImpl& Q, int spt) {
    std::vector<int32_t> tile_s0, tile_chart;
    std::vector<int64_t> tr0, tr1;
    Q.chart_t0.assign(Q.nchart + 1, 0);
    for (int c = 0; c < Q.nchart; ++c) {
        Q.chart_t0[c] = (int)tile_chart.size();
        const int s0 = Q.chart_s0[c], s1 = Q.chart_s0[c + 1];
        if (s0 == s1) {  // an empty chart still owns one empty tile (slices [s0, s0)) so its scalars exist
            tile_s0.push_back(s0);
            tile_chart.push_back(c);
            tr0.push_back(Q.crow[c]);
            tr1.push_back(Q.crow[c]);
            continue;
        }
        for (int s = s0; s < s1; s += spt) {
            const int e = std::min(s + spt, s1);
            tile_s0.push_back(s);
            tile_chart.push_back(c);
            tr0.push_back(Q.h_sl_row0[s]);
            tr1.push_back(Q.h_sl_row0[e - 1] + Q.h_sl_n[e - 1]);
        }
    }
    Q.chart_t0[Q.nchart] = (int)tile_chart.size();
    const int T = (int)tile_chart.size();
    int most = 0;  // tiles of the largest chart
    for (int c = 0; c < Q.nchart; ++c) most = std::max(most, Q.chart_t0[c + 1] - Q.chart_t0[c]);
    int big_tiles = 1024;  // from this many tiles on, a chart's sums are formed once (kb_chart_sums)
    if (const char* e = std::getenv("RUV_BATCH_BIG_TILES")) big_tiles = std::max(1, std::atoi(e));
    Q.B.big = most >= big_tiles ? 1 : 0;
    std::vector<int32_t> ts0(tile_s0);
    ts0.push_back(Q.nslice);  // tile t spans slices [ts0[t], ts0[t + 1])
    Q.free_tiles();
    Q.ntile = T;
    Q.spt = spt;
    Q.h_tile_r0 = tr0;
    Q.h_tile_r1 = tr1;
    auto up = [&](auto** dst, const auto* src, size_t n) -> hipError_t {
        using E = std::remove_pointer_t<std::remove_reference_t<decltype(*dst)>>;
        hipError_t e = hipMalloc((void**)dst, std::max<size_t>(n, 1) * sizeof(E));
        if (e == hipSuccess && n) e = hipMemcpyAsync(*dst, src, n * sizeof(E), hipMemcpyHostToDevice, Q.stream);
        return e;
    };
    HIPCHK(up(&Q.d_tile_s0, ts0.data(), ts0.size()));
    HIPCHK(up(&Q.d_tile_chart, tile_chart.data(), tile_chart.size()));
    HIPCHK(up(&Q.d_chart_t0, Q.chart_t0.data(), Q.chart_t0.size()));
    HIPCHK(up(&Q.d_tile_r0, tr0.data(), tr0.size()));
    HIPCHK(up(&Q.d_tile_r1, tr1.data(), tr1.size()));
    HIPCHK(hipMalloc((void**)&Q.d_part_a, sizeof(double) * (size_t)T));
    HIPCHK(hipMalloc((void**)&Q.d_part_b, sizeof(double) * 2 * (size_t)T));
    HIPCHK(hipMalloc((void**)&Q.d_part_c, sizeof(double) * 2 * (size_t)T));
    HIPCHK(hipMalloc((void**)&Q.d_part_o, sizeof(double) * (size_t)T));
    HIPCHK(hipMalloc((void**)&Q.d_part_d, sizeof(double) * 2 * (size_t)T));
    HIPCHK(hipMalloc((void**)&Q.act, sizeof(int) * (size_t)T));
    HIPCHK(hipHostMalloc((void**)&Q.h_act, sizeof(int) * 2 * (size_t)T, hipHostMallocDefault));  // two halves
    HIPCHK(hipStreamSynchronize(Q.stream));  // host staging vectors die at return
    BDev& B = Q.B;
    B.tile_s0 = Q.d_tile_s0;
    B.tile_chart = Q.d_tile_chart;
    B.chart_t0 = Q.d_chart_t0;
    B.tile_r0 = Q.d_tile_r0;
    B.tile_r1 = Q.d_tile_r1;
    B.part_a = Q.d_part_a;
    B.part_b = Q.d_part_b;
    B.part_c = Q.d_part_c;
    B.part_o = Q.d_part_o;
    B.part_d = Q.d_part_d;
    B.ntile = T;
    return RUV_OK;
}

// Band form of a chart (batch_kernels.hpp): the chart's positive column offsets, at most four, fit
// {1, W - 1, W, W + 1} for one W >= 3, and the operator is exactly symmetric inside it.  Returns W (0: no) and the
// offset mask.
int band_of(int64_t lo, int64_t hi, const int64_t* ptr, const int32_t* cols, const double* vals, int* mask) {
    int64_t ds[4];
    int nd = 0;
    int64_t upper = 0, lower = 0;
    for (int64_t i = lo; i < hi; ++i)
        for (int64_t a = ptr[i]; a < ptr[i + 1]; ++a) {
            const int64_t d = (int64_t)cols[a] - i;
            if (cols[a] < lo || cols[a] >= hi) return 0;
            if (d < 0) ++lower;
            if (d <= 0) continue;
            ++upper;
            int k = 0;
            while (k < nd && ds[k] != d) ++k;
            if (k == nd) {
                if (nd == 4) return 0;
                ds[nd++] = d;
            }
        }
    if (nd == 0 || upper != lower) return 0;
    const int64_t dmax = *std::max_element(ds, ds + nd);
    // the width: the first candidate that fits all offsets, a candidate whose mask the kernels compile (the
    // offsets of triangulated and plain grids) before any other
    int W = 0, mk = 0;
    bool compiled = false;
    for (const int64_t w : {dmax - 1, dmax, dmax + 1}) {
        if (w < 3 || w > (int64_t)INT32_MAX / 16) continue;
        int m = 0;
        bool fits = true;
        for (int k = 0; k < nd; ++k) {
            const int64_t d = ds[k];
            if (d == 1) m |= 1;
            else if (d == w - 1) m |= 2;
            else if (d == w) m |= 4;
            else if (d == w + 1) m |= 8;
            else fits = false;
        }
        const bool cm = band_mask_compiled(m);
        if (fits && (!W || (cm && !compiled))) {
            W = (int)w;
            mk = m;
            compiled = cm;
        }
    }
    if (!W) return 0;
    // every upper entry (i, j) has its mirror (j, i) with the same value; with as many lower entries as upper ones,
    // the lower triangle is exactly the mirror of the upper
    for (int64_t i = lo; i < hi; ++i)
        for (int64_t a = ptr[i]; a < ptr[i + 1]; ++a) {
            const int64_t j = cols[a];
            if (j <= i) continue;
            int64_t b = ptr[j];
            while (b < ptr[j + 1] && cols[b] != (int32_t)i) ++b;
            if (b == ptr[j + 1] || !(vals[b] == vals[a])) return 0;
        }
    *mask = mk;
    return W;
}

// slot of offset d >= 0 in a band chart of width w and offset mask m (0: the diagonal)
inline int band_slot(int64_t d, int w, int m) {
    if (d == 0) return 0;
    const int bit = d == 1 ? 0 : d == w - 1 ? 1 : d == w ? 2 : 3;
    return 1 + __builtin_popcount((unsigned)(m & ((1 << bit) - 1)));
}

// Band form of every band chart: widths, masks, value offsets, and the values (the mode's precision) uploaded.
// RUV_BATCH_BAND=0 keeps every chart in the sliced-ELL form.
int band_setup(BatchImpl& Q, const int64_t* ptr, const int32_t* cols, const double* vals) {
    const int nchart = Q.nchart;
    Q.h_bw.assign(nchart, 0);
    Q.h_bmask.assign(nchart, 0);
    Q.h_bbase.assign(nchart + 1, 0);
    const char* env = std::getenv("RUV_BATCH_BAND");
    if (env && std::atoi(env) == 0) return RUV_OK;
    parallel_for(nchart, [&](int64_t c0, int64_t c1) {
        for (int64_t c = c0; c < c1; ++c) {
            int m = 0;
            const int w = band_of(Q.crow[c], Q.crow[c + 1], ptr, cols, vals, &m);
            Q.h_bw[c] = w;
            Q.h_bmask[c] = w ? m : 0;
        }
    });
    Q.nband = 0;
    for (int c = 0; c < nchart; ++c) {
        const int64_t n = Q.crow[c + 1] - Q.crow[c];
        Q.h_bbase[c + 1] = Q.h_bbase[c] + (Q.h_bw[c] ? (1 + __builtin_popcount((unsigned)Q.h_bmask[c])) * n : 0);
        Q.nband += Q.h_bw[c] ? 1 : 0;
    }
    Q.nbv = Q.h_bbase[nchart];
    if (Q.nband == 0) return RUV_OK;
    std::vector<double> bv((size_t)Q.nbv, 0.0);
    parallel_for(nchart, [&](int64_t c0, int64_t c1) {
        for (int64_t c = c0; c < c1; ++c) {
            const int w = Q.h_bw[c], m = Q.h_bmask[c];
            if (!w) continue;
            const int64_t lo = Q.crow[c], n = Q.crow[c + 1] - lo;
            for (int64_t i = lo; i < lo + n; ++i)
                for (int64_t a = ptr[i]; a < ptr[i + 1]; ++a) {
                    const int64_t d = (int64_t)cols[a] - i;
                    if (d >= 0) bv[Q.h_bbase[c] + (int64_t)band_slot(d, w, m) * n + (i - lo)] += vals[a];
                }
        }
    });
    // FP64 for the refinement residual (and the plain solve), FP32 for the mixed inner solve
    HIPCHK(Q.upload(&Q.bv64, bv.data(), bv.size()));
    if (Q.mode == RUV_MODE_MIXED) {
        std::vector<float> f((size_t)Q.nbv);
        parallel_for(Q.nbv, [&](int64_t a, int64_t b) {
            for (int64_t k = a; k < b; ++k) f[k] = (float)bv[k];
        });
        HIPCHK(Q.upload(&Q.bv32, f.data(), f.size()));
    }
    HIPCHK(hipStreamSynchronize(Q.stream));
    int32_t *d_bw = nullptr, *d_bm = nullptr;
    int64_t* d_bb = nullptr;
    HIPCHK(Q.upload(&d_bw, Q.h_bw.data(), Q.h_bw.size()));
    HIPCHK(Q.upload(&d_bm, Q.h_bmask.data(), Q.h_bmask.size()));
    HIPCHK(Q.upload(&d_bb, Q.h_bbase.data(), Q.h_bbase.size()));
    Q.B.c_bw = d_bw;
    Q.B.c_bmask = d_bm;
    Q.B.c_bbase = d_bb;
    return RUV_OK;
}

int batch_setup(BatchImpl& Q, int64_t nrows, const int64_t* ptr, const int32_t* cols, const double* vals,
                int nchart, const int64_t* crow, int R, int precond, int mode) {
    Q.R = R;
    Q.mode = mode;
    Q.precond = precond;
    Q.nchart = nchart;
    Q.nrows = nrows;
    Q.nnz = ptr[nrows];
    Q.crow.assign(crow, crow + nchart + 1);
    HIPCHK(hipStreamCreateWithFlags(&Q.stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&Q.ev0));
    HIPCHK(hipEventCreate(&Q.ev1));
    // slices: each chart's rows in runs of at most SLICE
    Q.chart_s0.assign(nchart + 1, 0);
    for (int c = 0; c < nchart; ++c) {
        Q.chart_s0[c] = (int)Q.h_sl_row0.size();
        for (int64_t r = crow[c]; r < crow[c + 1]; r += SLICE) {
            Q.h_sl_row0.push_back(r);
            Q.h_sl_n.push_back((int)std::min<int64_t>(SLICE, crow[c + 1] - r));
        }
    }
    Q.chart_s0[nchart] = (int)Q.h_sl_row0.size();
    const int nsl = Q.nslice = (int)Q.h_sl_row0.size();
    int rc = band_setup(Q, ptr, cols, vals);
    if (rc) return rc;
    // classify: a stencil slice has at most SLOTS distinct column offsets (col - row); its slot k holds the
    // entry at offset dtab[k] of every row (or 0), so it stores no column indices and its gathers coalesce.
    // Any other slice is padded to its longest row and stores a column (offset) per entry.
    std::vector<int32_t> sl_w(nsl, 0), dtab((size_t)nsl * SLOTS, 0);
    std::vector<char> stencil(nsl, 0);
    std::vector<int64_t> wide(1, 0);  // general entries whose offset does not fit 16 bits
    parallel_for(nsl, [&](int64_t s0, int64_t s1) {
        std::vector<int64_t> ds;
        int64_t wide_local = 0;
        for (int64_t s = s0; s < s1; ++s) {
            const int64_t r0 = Q.h_sl_row0[s];
            const int n = Q.h_sl_n[s];
            ds.clear();
            int wmax = 0;
            for (int64_t i = r0; i < r0 + n; ++i) {
                wmax = std::max<int>(wmax, (int)(ptr[i + 1] - ptr[i]));
                for (int64_t a = ptr[i]; a < ptr[i + 1]; ++a) ds.push_back((int64_t)cols[a] - i);
            }
            std::sort(ds.begin(), ds.end());
            ds.erase(std::unique(ds.begin(), ds.end()), ds.end());
            if ((int)ds.size() <= SLOTS && (ds.empty() || (ds.front() >= INT32_MIN && ds.back() <= INT32_MAX))) {
                stencil[s] = 1;
                sl_w[s] = (int)ds.size();
                for (size_t k = 0; k < ds.size(); ++k) dtab[(size_t)s * SLOTS + k] = (int32_t)ds[k];
            } else {
                sl_w[s] = wmax;
                for (int64_t i = r0; i < r0 + n; ++i)
                    for (int64_t a = ptr[i]; a < ptr[i + 1]; ++a)
                        if (cols[a] - i > 32767 || cols[a] - i < -32767) ++wide_local;
            }
        }
        if (wide_local) __atomic_fetch_add(&wide[0], wide_local, __ATOMIC_RELAXED);
    });
    const bool narrow = wide[0] == 0;
    Q.ci = narrow ? 0 : 1;
    std::vector<int64_t> sl_off(nsl), sl_coff(nsl);
    int64_t nell = 0, ncol = 0;
    Q.nstencil = 0;
    for (int s = 0; s < nsl; ++s) {
        sl_off[s] = nell;
        nell += (int64_t)sl_w[s] * Q.h_sl_n[s];
        if (stencil[s]) {
            sl_coff[s] = -1;
            ++Q.nstencil;
        } else {
            sl_coff[s] = ncol;
            ncol += (int64_t)sl_w[s] * Q.h_sl_n[s];
        }
    }
    Q.nell = nell;
    Q.ncol = ncol;
    {
        // algorithmic bytes of one step's SpMV per chart: every nonzero's value (and, in a general slice, its
        // stored column), the slice table (and a stencil slice's offsets), p' read and q written once.  The
        // stored form also streams the zero slots of its padding; that figure is kept beside it.
        const double sv = mode == RUV_MODE_MIXED ? 4.0 : 8.0, sc = narrow ? 2.0 : 4.0;
        Q.chart_step_bytes.assign(nchart, 0.0);
        Q.chart_step_stored.assign(nchart, 0.0);
        Q.fin_colb = Q.fin_colb_st = Q.fin_meta = Q.fin_slots = 0.0;
        for (int c = 0; c < nchart; ++c) {
            const double vec = 2.0 * (double)(crow[c + 1] - crow[c]) * R * sv;  // p' read, q written
            double b = vec, st = vec;
            if (Q.h_bw[c]) {  // band form: the nonzeros' values, as stored its slots (mirrored entries once)
                const double slots = (double)(Q.h_bbase[c + 1] - Q.h_bbase[c]);
                Q.chart_step_bytes[c] = b + (double)(ptr[crow[c + 1]] - ptr[crow[c]]) * sv;
                Q.chart_step_stored[c] = st + slots * sv;
                Q.fin_slots += slots;
                continue;
            }
            for (int s = Q.chart_s0[c]; s < Q.chart_s0[c + 1]; ++s) {
                const int64_t r0 = Q.h_sl_row0[s];
                const double nz = (double)(ptr[r0 + Q.h_sl_n[s]] - ptr[r0]);
                const double meta = 32.0 + (stencil[s] ? 4.0 * SLOTS : 0.0);
                b += nz * (stencil[s] ? sv : sv + sc) + meta;
                st += (double)sl_w[s] * Q.h_sl_n[s] * (stencil[s] ? sv : sv + sc) + meta;
                Q.fin_slots += (double)sl_w[s] * Q.h_sl_n[s];
                Q.fin_meta += meta;
                if (!stencil[s]) {
                    Q.fin_colb += nz * sc;
                    Q.fin_colb_st += (double)sl_w[s] * Q.h_sl_n[s] * sc;
                }
            }
            Q.chart_step_bytes[c] = b;
            Q.chart_step_stored[c] = st;
        }
    }
    // fill the values (and the general slices' columns, and the Jacobi diagonal) on the host threads
    std::vector<double> e64((size_t)nell, 0.0);
    std::vector<int16_t> c16(narrow ? (size_t)ncol : 0, 0);
    std::vector<int32_t> c32(narrow ? 0 : (size_t)ncol, 0);
    std::vector<double> dinv(precond != RUV_PC_NONE ? (size_t)nrows : 0, 0.0);
    std::vector<int64_t> bad(1, -1);
    parallel_for(nsl, [&](int64_t s0, int64_t s1) {
        for (int64_t s = s0; s < s1; ++s) {
            const int64_t r0 = Q.h_sl_row0[s];
            const int n = Q.h_sl_n[s];
            const int32_t* dt = dtab.data() + (size_t)s * SLOTS;
            for (int l = 0; l < n; ++l) {
                const int64_t i = r0 + l;
                int k = 0;
                double d = 0.0;
                for (int64_t a = ptr[i]; a < ptr[i + 1]; ++a, ++k) {
                    if (cols[a] == i) d = vals[a];
                    if (stencil[s]) {
                        const int slot = (int)(std::lower_bound(dt, dt + sl_w[s], (int32_t)(cols[a] - i)) - dt);
                        e64[sl_off[s] + (int64_t)slot * n + l] += vals[a];
                    } else {
                        e64[sl_off[s] + (int64_t)k * n + l] = vals[a];
                        const int64_t e = sl_coff[s] + (int64_t)k * n + l;
                        if (narrow) c16[e] = (int16_t)(cols[a] - i);
                        else c32[e] = cols[a];
                    }
                }
                if (!stencil[s])
                    for (; k < sl_w[s]; ++k) {  // padding: (0, the row itself)
                        const int64_t e = sl_coff[s] + (int64_t)k * n + l;
                        if (!narrow) c32[e] = (int32_t)i;
                    }
                if (!dinv.empty()) {
                    if (std::fabs(d) < 1e-300) bad[0] = i;
                    else dinv[i] = 1.0 / d;
                }
            }
        }
    });
    if (bad[0] >= 0) return fail(RUV_ERR_PRECOND, "zero diagonal at row " + std::to_string(bad[0]));
    Q.h_sl_off = sl_off;
    {  // the sliced-ELL entries of the charts not in band form, in ranges of at most 1 << 16 (for their copies)
        std::vector<int64_t> rng;
        for (int c = 0; c < nchart; ++c) {
            if (Q.h_bw[c] || Q.chart_s0[c] == Q.chart_s0[c + 1]) continue;
            const int64_t a = sl_off[Q.chart_s0[c]];
            const int64_t e = Q.chart_s0[c + 1] < nsl ? sl_off[Q.chart_s0[c + 1]] : nell;
            for (int64_t k = a; k < e; k += (1 << 16)) {
                rng.push_back(k);
                rng.push_back(std::min<int64_t>(1 << 16, e - k));
            }
        }
        Q.n_ell_ranges = (int64_t)rng.size() / 2;
        if (Q.n_ell_ranges) HIPCHK(Q.upload(&Q.ell_ranges, rng.data(), rng.size()));
    }
    Q.h_sl_w = sl_w;
    Q.h_dtab = dtab;
    Q.h_stencil = stencil;
    int64_t *d_row0 = nullptr, *d_off = nullptr, *d_coff = nullptr, *d_crow = nullptr;
    int32_t *d_w = nullptr, *d_n = nullptr, *d_dtab = nullptr;
    void* d_col = nullptr;
    HIPCHK(Q.upload(&d_row0, Q.h_sl_row0.data(), Q.h_sl_row0.size()));
    HIPCHK(Q.upload(&d_off, sl_off.data(), sl_off.size()));
    HIPCHK(Q.upload(&d_coff, sl_coff.data(), sl_coff.size()));
    HIPCHK(Q.upload(&d_w, sl_w.data(), sl_w.size()));
    HIPCHK(Q.upload(&d_n, Q.h_sl_n.data(), Q.h_sl_n.size()));
    HIPCHK(Q.upload(&d_dtab, dtab.data(), dtab.size()));
    HIPCHK(Q.upload(&d_crow, crow, (size_t)nchart + 1));
    if (narrow) {
        int16_t* p = nullptr;
        HIPCHK(Q.upload(&p, c16.data(), c16.size()));
        d_col = p;
    } else {
        int32_t* p = nullptr;
        HIPCHK(Q.upload(&p, c32.data(), c32.size()));
        d_col = p;
    }
    HIPCHK(Q.upload(&Q.ev64, e64.data(), e64.size()));
    std::vector<float> f32;
    if (mode == RUV_MODE_MIXED) {
        f32.resize(e64.size());
        parallel_for((int64_t)e64.size(), [&](int64_t a, int64_t b) {
            for (int64_t k = a; k < b; ++k) f32[k] = (float)e64[k];
        });
        HIPCHK(Q.upload(&Q.ev32, f32.data(), f32.size()));
    }
    std::vector<float> dinv32;
    if (!dinv.empty()) {
        HIPCHK(Q.upload(&Q.dinv64, dinv.data(), dinv.size()));
        if (mode == RUV_MODE_MIXED) {
            dinv32.assign(dinv.begin(), dinv.end());
            HIPCHK(Q.upload(&Q.dinv32, dinv32.data(), dinv32.size()));
        }
    }
    const size_t N = (size_t)nrows * R;
    const size_t sv = mode == RUV_MODE_MIXED ? sizeof(float) : sizeof(double);
    HIPCHK(Q.alloc(&Q.r64, N));
    for (void** v : {&Q.xv, &Q.rv, &Q.zv, &Q.pv0, &Q.pv1, &Q.qv}) {
        char* p = nullptr;
        HIPCHK(Q.alloc(&p, N * sv));
        *v = p;
    }
    BDev& B = Q.B;
    B.sl_row0 = d_row0;
    B.sl_off = d_off;
    B.sl_coff = d_coff;
    B.sl_w = d_w;
    B.sl_n = d_n;
    B.sl_dtab = d_dtab;
    B.crow = d_crow;
    B.ecol = d_col;
    B.nchart = nchart;
    HIPCHK(Q.alloc(&B.c_rz, 2 * (size_t)nchart));
    HIPCHK(Q.alloc(&B.c_alpha, nchart));
    HIPCHK(hipMemsetAsync(B.c_alpha, 0, sizeof(double) * (size_t)std::max(nchart, 1), Q.stream));
    HIPCHK(Q.alloc(&B.c_tol, nchart));
    HIPCHK(Q.alloc(&B.c_bnorm, nchart));
    HIPCHK(Q.alloc(&B.c_true, nchart));
    HIPCHK(Q.alloc(&B.c_rrec, nchart));
    HIPCHK(Q.alloc(&B.c_rin, nchart));
    HIPCHK(Q.alloc(&B.c_gamma, nchart));
    HIPCHK(Q.alloc(&B.c_tgt, nchart));
    HIPCHK(Q.alloc(&B.c_sa, nchart));
    HIPCHK(Q.alloc(&B.c_sb, 2 * (size_t)nchart));
    HIPCHK(Q.alloc(&B.c_sc, 2 * (size_t)nchart));
    HIPCHK(Q.alloc(&B.c_done, nchart));
    HIPCHK(Q.alloc(&B.c_final, nchart));
    HIPCHK(Q.alloc(&B.c_iters, nchart));
    HIPCHK(hipHostMalloc((void**)&Q.h_done, sizeof(int) * (size_t)std::max(nchart, 1), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&Q.h_done_b, sizeof(int) * (size_t)std::max(nchart, 1), hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&Q.fev[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&Q.fev[1], hipEventDisableTiming));
    HIPCHK(hipHostMalloc((void**)&Q.h_final, sizeof(int) * (size_t)std::max(nchart, 1), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&Q.h_iters, sizeof(unsigned) * (size_t)std::max(nchart, 1), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&Q.h_true, sizeof(double) * (size_t)std::max(nchart, 1), hipHostMallocDefault));
    HIPCHK(hipStreamSynchronize(Q.stream));  // staging vectors above must outlive their copies
    return build_tiles(Q, 2);
}

// upload the tiles of the charts whose flag[c] satisfies pred; returns the count.  The pinned staging alternates
// between two halves, so a list can be written while the copy of the previous one is still queued (the copy of the
// one before that has completed: every caller waits on the stream, or on a look recorded after that copy, at
// least once between two lists).
// The tiles of band charts come first (Q.na_band of them), so a pass whose two paths differ in register use can
// launch one kernel per storage form.
template <typename Pred>
int set_active(BatchImpl& Q, Pred pred, int* count) {
    Q.act_half ^= 1;
    int* h = Q.h_act + (size_t)Q.act_half * (size_t)Q.ntile;
    int n = 0;
    const bool hb = !Q.h_bw.empty();
    for (int pass = 0; pass < 2; ++pass) {
        for (int c = 0; c < Q.nchart; ++c)
            if ((hb && Q.h_bw[c] > 0) == (pass == 0) && pred(c))
                for (int t = Q.chart_t0[c]; t < Q.chart_t0[c + 1]; ++t) h[n++] = t;
        if (pass == 0) Q.na_band = n;
    }
    *count = n;
    if (n) HIPCHK(hipMemcpyAsync(Q.act, h, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, Q.stream));
    return RUV_OK;
}

inline dim3 grid_of(int na) { return dim3((unsigned)(((na + XCDS - 1) / XCDS) * XCDS)); }

template <typename T, int R, int CI, int PROBE = 0>
void launch_spmv(BatchImpl& Q, int na, const T* p, T* q) {
    const T* vals = (sizeof(T) == 4) ? (const T*)(const void*)Q.ev32 : (const T*)(const void*)Q.ev64;
    const T* bvals = (sizeof(T) == 4) ? (const T*)(const void*)Q.bv32 : (const T*)(const void*)Q.bv64;
    if (Q.nt)
        hipLaunchKernelGGL((kb_spmv<T, R, CI, 1, PROBE>), grid_of(na), dim3(BLK), 0, Q.stream, Q.B,
                           (const int*)Q.act, na, vals, bvals, p, q);
    else
        hipLaunchKernelGGL((kb_spmv<T, R, CI, 0, PROBE>), grid_of(na), dim3(BLK), 0, Q.stream, Q.B,
                           (const int*)Q.act, na, vals, bvals, p, q);
}

// One inner PCG over the charts not yet final.  cont: go on from the previous inner solve's last direction (the
// residual replaced by the refined FP64 one) instead of starting afresh.
template <typename T, int R, int PC, int CI>
int inner_pcg(BatchImpl& Q, double tol, uint64_t max_iter, int cont, double cut = 0.0) {
    hipStream_t s = Q.stream;
    BDev& B = Q.B;
    B.max_iter = (unsigned)std::min<uint64_t>(max_iter, 0xffffffffu);
    const T* dinv = (sizeof(T) == 4) ? (const T*)(const void*)Q.dinv32 : (const T*)(const void*)Q.dinv64;
    T* x = (T*)Q.xv;
    T* r = (T*)Q.rv;
    T* z = (PC >= 1) ? (T*)Q.zv : r;
    T* pb[2] = {(T*)Q.pv0, (T*)Q.pv1};
    T* q = (T*)Q.qv;
    int na = 0;
    int rc = set_active(Q, [&](int c) { return Q.h_final[c] == 0; }, &na);
    if (rc) return rc;
    if (na == 0) return RUV_OK;
    // step 0 reads its old direction from pb[1]
    hipLaunchKernelGGL((kb_prep<T, R, PC>), grid_of(na), dim3(BLK), 0, s, B, (const int*)Q.act, na,
                       (const double*)Q.r64, dinv, x, r, z, (const T*)pb[0], pb[1], cont);
    hipLaunchKernelGGL(kb_init, dim3((unsigned)Q.nchart), dim3(BLK), 0, s, B, tol, Q.inner_tol,
                       Q.mode == RUV_MODE_MIXED ? 1 : 0, cont, Q.cont_gap, Q.final_aim);
    HIPCHK(hipGetLastError());
    if (PC == 2) {  // z0 = M^-1 r0 and its r.z partials (parity 0); kb_init has marked the running charts
        rc = bamg_set_live(Q, nullptr);
        if (!rc) rc = bamg_apply<T, R, CI>(Q, na, 0, (const T*)r, z);
        if (rc) return rc;
    }
    // charts of very many tiles: the chart-wide sums once per producer, not in every consuming workgroup
    auto chart_sums = [&](const double* pa, const double* pb, double* sa, double* sb) {
        if (B.big)
            hipLaunchKernelGGL(kb_chart_sums, dim3((unsigned)Q.nchart), dim3(BLK), 0, s, B, pa, pb, sa, sb);
    };
    chart_sums(B.part_b, B.part_c, B.c_sb, B.c_sc);  // parity 0
    // Which charts still run is looked at after every chunk of steps: the flags after chunk j are copied to a pinned
    // buffer, marked by an event, and chunk j + 1 launches the tiles of the charts still running.  With look_lag the
    // look is one chunk late while most tiles run: chunk j + 1 is queued with the charts that ran after chunk j - 1
    // while the host waits for chunk j's flags, so the device never idles for a look (the lists only shrink and
    // every kernel leaves at once for a stopped chart, so an older list is a superset that costs early exits).
    // Flags alternate between two pinned buffers, each copy with its own event.
    int* flags[2] = {Q.h_done, Q.h_done_b};
    HIPCHK(hipMemcpyAsync(flags[0], B.c_done, sizeof(int) * Q.nchart, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // steps between looks: growing up to 64 for the cheap Jacobi/plain steps; the AMG step costs a V-cycle, so a
    // look every few steps keeps the lists tight
    int chunk = (PC == 2) ? Q.amg_chunk : 8;
    uint32_t iter = 0;
    int par = 0;
    auto run_chunk = [&](int steps) -> int {
        for (int it = 0; it < steps; ++it) {
            hipLaunchKernelGGL((kb_dir<T, R>), grid_of(na), dim3(BLK), 0, s, B, (const int*)Q.act, na, par, iter, z,
                               (const T*)pb[par ^ 1], pb[par], x);
            if (Q.ktime) {
                if (Q.kev_used + 2 > Q.kev.size()) {
                    hipEvent_t e0, e1;
                    HIPCHK(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
                    HIPCHK(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
                    Q.kev.push_back(e0);
                    Q.kev.push_back(e1);
                }
                HIPCHK(hipEventRecord(Q.kev[Q.kev_used], s));
            }
            launch_spmv<T, R, CI>(Q, na, pb[par], q);
            if (Q.ktime) {
                HIPCHK(hipEventRecord(Q.kev[Q.kev_used + 1], s));
                Q.kev_used += 2;
                Q.k_launches += 1;
            }
            chart_sums(B.part_a, nullptr, B.c_sa, nullptr);
            hipLaunchKernelGGL((kb_update<T, R, PC>), grid_of(na), dim3(BLK), 0, s, B, (const int*)Q.act, na, par,
                               dinv, (const T*)q, r, z);
            if (PC == 2) {
                int rc2 = bamg_apply<T, R, CI>(Q, na, par ^ 1, (const T*)r, z);
                if (rc2) return rc2;
            }
            chart_sums(B.part_b + (int64_t)(par ^ 1) * Q.ntile, B.part_c + (int64_t)(par ^ 1) * Q.ntile,
                       B.c_sb + (int64_t)(par ^ 1) * Q.nchart, B.c_sc + (int64_t)(par ^ 1) * Q.nchart);
            par ^= 1;
            ++iter;
        }
        HIPCHK(hipGetLastError());
        return RUV_OK;
    };
    // flags[j % 2] receives the flags after chunk j (chunk 0: the state kb_init left, in flags[0]); `known` is the
    // buffer of the latest flags waited for
    int known = 0;
    const int na0 = na;
    for (int64_t j = 1;; ++j) {
        const int* kf = flags[known];
        rc = set_active(Q, [&](int c) { return kf[c] == B_RUNNING; }, &na);
        if (rc) return rc;
        if (na == 0) break;
        if (cut > 0.0 && (double)na < cut * (double)na0) {  // the few charts still running stop here
            hipLaunchKernelGGL((kb_dir<T, R>), grid_of(na), dim3(BLK), 0, s, B, (const int*)Q.act, na, par, iter, z,
                               (const T*)pb[par ^ 1], pb[par], x, 1);
            HIPCHK(hipGetLastError());
            break;
        }
        // once most charts have stopped, the coarse levels launch only the running charts' slices and blocks
        if (PC == 2) rc = bamg_set_live(Q, (double)na < Q.list_below * (double)Q.ntile ? kf : nullptr);
        if (rc) return rc;
        rc = run_chunk(chunk);
        if (rc) return rc;
        // the flags after chunk j go to the buffer read above (the list is built); read one iteration later
        HIPCHK(hipMemcpyAsync(flags[j & 1], B.c_done, sizeof(int) * Q.nchart, hipMemcpyDeviceToHost, s));
        HIPCHK(hipEventRecord(Q.fev[j & 1], s));
        // the look one chunk late pays while most tiles run (a chunk then takes milliseconds, and the lists barely
        // change); late in a solve it would queue a chunk of early exits after the last chart stops
        if (!Q.look_lag || 2 * (int64_t)na < (int64_t)Q.ntile) {  // this chunk's flags (the device waits)
            HIPCHK(hipEventSynchronize(Q.fev[j & 1]));
            known = (int)(j & 1);
        } else if (j > 1) {  // chunk j - 1's flags, while chunk j runs
            HIPCHK(hipEventSynchronize(Q.fev[(j - 1) & 1]));
            known = (int)((j - 1) & 1);
        }
        if (PC != 2) chunk = std::min(chunk * 2, 64);
    }
    // every queued chunk done, and the final states
    HIPCHK(hipMemcpyAsync(Q.h_done, B.c_done, sizeof(int) * Q.nchart, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpyAsync(Q.h_iters, B.c_iters, sizeof(unsigned) * Q.nchart, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int c = 0; c < Q.nchart; ++c) {
        if (Q.h_final[c]) continue;
        Q.iters_total[c] += Q.h_iters[c];
        if (Q.h_done[c] == B_BREAKDOWN || Q.h_done[c] == B_BREAKDOWN_PQ)
            return fail(RUV_ERR_BREAKDOWN, "batch: chart " + std::to_string(c) + " broke down");
        if (Q.h_done[c] == B_MAXITER)
            return fail(RUV_ERR_NOT_CONVERGED, "batch: chart " + std::to_string(c) + " hit max_iter");
    }
    return RUV_OK;
}

template <typename T, int R, int PC, int CI>
int batch_run(BatchImpl& Q, const double* d_b, double* d_x, uint64_t max_iter, double tol) {
    hipStream_t s = Q.stream;
    BDev& B = Q.B;
    const bool mixed = Q.mode == RUV_MODE_MIXED;
    Q.iters_total.assign(Q.nchart, 0);
    hipLaunchKernelGGL((kb_start<R>), dim3(Q.ntile), dim3(BLK), 0, s, B, d_b, d_x, Q.r64);
    hipLaunchKernelGGL(kb_bnorm, dim3((unsigned)Q.nchart), dim3(BLK), 0, s, B);
    HIPCHK(hipGetLastError());
    std::fill(Q.h_final, Q.h_final + Q.nchart, 0);
    int na = 0;
    for (int round = 0;; ++round) {
        if (round >= Q.max_rounds) return fail(RUV_ERR_NOT_CONVERGED, "batch: refinement did not converge");
        int rc = inner_pcg<T, R, PC, CI>(Q, tol, max_iter, (round > 0 && Q.continue_rounds) ? 1 : 0,
                                         (mixed && round == 0) ? Q.first_cut : 0.0);
        if (rc) return rc;
        rc = set_active(Q, [&](int c) { return Q.h_final[c] == 0; }, &na);
        if (rc) return rc;
        if (na == 0) break;
        // x64 += x_inner over the live charts, then the FP64 residual and the per-chart outer test
        hipLaunchKernelGGL((kb_axpy<T, R>), grid_of(na), dim3(BLK), 0, s, B, (const int*)Q.act, na,
                           (const T*)Q.xv, d_x);
        hipLaunchKernelGGL((kb_residual<R, CI>), grid_of(na), dim3(BLK), 0, s, B, (const int*)Q.act, na,
                           (const double*)Q.ev64, (const double*)Q.bv64, d_b, (const double*)d_x, Q.r64);
        hipLaunchKernelGGL(kb_outer, dim3((unsigned)Q.nchart), dim3(BLK), 0, s, B, tol, mixed ? Q.uv_tol : 0.0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(Q.h_final, B.c_final, sizeof(int) * Q.nchart, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (!mixed) break;  // plain FP64: one PCG per chart, the true residual is only reported
        bool all = true;
        for (int c = 0; c < Q.nchart; ++c) all = all && Q.h_final[c];
        if (all) break;
    }
    HIPCHK(hipMemcpyAsync(Q.h_true, B.c_true, sizeof(double) * Q.nchart, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return RUV_OK;
}

template <typename T, int R, int CI>
int run_pc(BatchImpl& Q, const double* d_b, double* d_x, uint64_t max_iter, double tol) {
    if (Q.precond == RUV_PC_AMG) return batch_run<T, R, 2, CI>(Q, d_b, d_x, max_iter, tol);
    return Q.precond == RUV_PC_JACOBI ? batch_run<T, R, 1, CI>(Q, d_b, d_x, max_iter, tol)
                                      : batch_run<T, R, 0, CI>(Q, d_b, d_x, max_iter, tol);
}
template <typename T, int R>
int run_ci(BatchImpl& Q, const double* d_b, double* d_x, uint64_t max_iter, double tol) {
    return Q.ci == 0 ? run_pc<T, R, 0>(Q, d_b, d_x, max_iter, tol) : run_pc<T, R, 1>(Q, d_b, d_x, max_iter, tol);
}

// The step's SpMV kernel over every tile with never-converging state: mean seconds per launch.
template <typename T, int R, int CI>
int time_step_kernel(BatchImpl& Q, int reps, double* sec) {
    hipStream_t s = Q.stream;
    BDev& B = Q.B;
    B.max_iter = 0xffffffffu;
    const int C = Q.nchart;
    std::vector<int> zeros(C, 0);
    std::vector<double> ones(2 * (size_t)C, 1.0), never(C, -1.0);
    HIPCHK(hipMemcpyAsync(B.c_done, zeros.data(), sizeof(int) * C, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B.c_rz, ones.data(), sizeof(double) * 2 * C, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B.c_tol, never.data(), sizeof(double) * C, hipMemcpyHostToDevice, s));
    std::vector<double> pones(2 * (size_t)Q.ntile, 1.0);
    HIPCHK(hipMemcpyAsync(B.part_b, pones.data(), sizeof(double) * pones.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B.part_c, pones.data(), sizeof(double) * pones.size(), hipMemcpyHostToDevice, s));
    const size_t N = (size_t)Q.nrows * Q.R * sizeof(T);
    for (void* v : {Q.zv, Q.pv0, Q.pv1, Q.qv}) HIPCHK(hipMemsetAsync(v, 0, N, s));
    int na = 0;
    int rc = set_active(Q, [](int) { return true; }, &na);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s));
    launch_spmv<T, R, CI, 1>(Q, na, (const T*)Q.pv0, (T*)Q.qv);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(Q.ev0, s));
    for (int i = 0; i < reps; ++i) launch_spmv<T, R, CI, 1>(Q, na, (const T*)Q.pv0, (T*)Q.qv);
    HIPCHK(hipEventRecord(Q.ev1, s));
    HIPCHK(hipEventSynchronize(Q.ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, Q.ev0, Q.ev1));
    *sec = ms * 1e-3 / reps;
    return RUV_OK;
}

int time_step(BatchImpl& Q, int reps, double* sec) {
    const bool f = Q.mode == RUV_MODE_MIXED;
    if (Q.R == 2) {
        if (f) return Q.ci == 0 ? time_step_kernel<float, 2, 0>(Q, reps, sec) : time_step_kernel<float, 2, 1>(Q, reps, sec);
        return Q.ci == 0 ? time_step_kernel<double, 2, 0>(Q, reps, sec) : time_step_kernel<double, 2, 1>(Q, reps, sec);
    }
    if (f) return Q.ci == 0 ? time_step_kernel<float, 1, 0>(Q, reps, sec) : time_step_kernel<float, 1, 1>(Q, reps, sec);
    return Q.ci == 0 ? time_step_kernel<double, 1, 0>(Q, reps, sec) : time_step_kernel<double, 1, 1>(Q, reps, sec);
}

// Whole-batch profile of the AMG-preconditioned step (every chart running, whole levels), reps times, each launch
// marked (kind, level, bytes).  The scalars the step's kernels test are reset before each of them (outside the
// marked intervals), so no chart stops; the traffic does not depend on the vectors' values.
template <typename T, int R, int CI>
int step_profile_kernel(BatchImpl& Q, int reps) {
    hipStream_t s = Q.stream;
    BDev& B = Q.B;
    B.max_iter = 0xffffffffu;
    const int C = Q.nchart;
    const size_t nt = (size_t)Q.ntile;
    std::vector<int> zeros(C, 0);
    std::vector<double> ones(2 * std::max<size_t>((size_t)C, nt), 1.0), never(C, -1.0);
    auto reset = [&]() -> int {
        HIPCHK(hipMemcpyAsync(B.c_done, zeros.data(), sizeof(int) * C, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(B.c_tol, never.data(), sizeof(double) * C, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(B.c_rz, ones.data(), sizeof(double) * 2 * C, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(B.c_sa, ones.data(), sizeof(double) * C, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(B.c_sb, ones.data(), sizeof(double) * 2 * C, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(B.c_sc, ones.data(), sizeof(double) * 2 * C, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(B.part_a, ones.data(), sizeof(double) * nt, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(B.part_b, ones.data(), sizeof(double) * 2 * nt, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(B.part_c, ones.data(), sizeof(double) * 2 * nt, hipMemcpyHostToDevice, s));
        return Q.mark(-1, 0, 0.0, 0.0) == hipSuccess ? RUV_OK : fail(RUV_ERR_DEVICE, "profile event");
    };
    const size_t N = (size_t)Q.nrows * R * sizeof(T);
    for (void* v : {Q.xv, Q.rv, Q.zv, Q.pv0, Q.pv1, Q.qv}) HIPCHK(hipMemsetAsync(v, 0, N, s));
    int na = 0;
    int rc = set_active(Q, [](int) { return true; }, &na);
    if (!rc) rc = bamg_set_live(Q, nullptr);
    if (rc) return rc;
    const T* dinv = (sizeof(T) == 4) ? (const T*)(const void*)Q.dinv32 : (const T*)(const void*)Q.dinv64;
    T *x = (T*)Q.xv, *r = (T*)Q.rv, *z = (T*)Q.zv, *q = (T*)Q.qv;
    T* pb[2] = {(T*)Q.pv0, (T*)Q.pv1};
    const double vb = (double)R * sizeof(T), n = (double)Q.nrows;
    double spmv = 0.0, spmv_st = 0.0;
    for (int c = 0; c < C; ++c) {
        spmv += Q.chart_step_bytes[c];
        spmv_st += Q.chart_step_stored[c];
    }
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipEventRecord(Q.mev[0], s));
    int par = 0;
    for (int it = 0; it < reps; ++it) {
        if ((rc = reset())) return rc;
        hipLaunchKernelGGL((kb_dir<T, R>), grid_of(na), dim3(BLK), 0, s, B, (const int*)Q.act, na, par, 1u, z,
                           (const T*)pb[par ^ 1], pb[par], x);
        HIPCHK(Q.mark(BK_DIR, 0, 5.0 * n * vb, 5.0 * n * vb));  // z, old p, x read; p, x written
        launch_spmv<T, R, CI>(Q, na, pb[par], q);
        HIPCHK(Q.mark(BK_SPMV, 0, spmv, spmv_st));
        if ((rc = reset())) return rc;
        hipLaunchKernelGGL((kb_update<T, R, 2>), grid_of(na), dim3(BLK), 0, s, B, (const int*)Q.act, na, par, dinv,
                           (const T*)q, r, z);
        HIPCHK(Q.mark(BK_UPDATE, 0, 3.0 * n * vb, 3.0 * n * vb));  // q, r read; r written
        if ((rc = reset())) return rc;
        if ((rc = bamg_apply<T, R, CI>(Q, na, par ^ 1, (const T*)r, z))) return rc;
        par ^= 1;
    }
    HIPCHK(hipGetLastError());
    return RUV_OK;
}

// Pick slices-per-tile and the value-load hint by timing the step's SpMV kernel over the whole batch.  Both can
// be pinned instead (RUV_BATCH_SLICES_PER_TILE = 1, 2, 4 or 8, RUV_BATCH_NONTEMPORAL = 0 or 1): a profiler that
// serialises launches and reads counters distorts the timings, so its runs pin the choice of the timed run.
int autotune(BatchImpl& Q) {
    const char* env = std::getenv("RUV_BATCH_SLICES_PER_TILE");
    if (env) {
        const int v = std::atoi(env);
        if (v == 1 || v == 2 || v == 4 || v == 8) {
            const char* nt = std::getenv("RUV_BATCH_NONTEMPORAL");
            Q.nt = (nt && std::atoi(nt) != 0) ? 1 : 0;
            return build_tiles(Q, v);
        }
    }
    int best = 2, best_nt = 0;
    double best_t = 1e300;
    for (int spt : {1, 2, 4, 8}) {
        int rc = build_tiles(Q, spt);
        if (rc) return rc;
        for (int nt : {0, 1}) {
            Q.nt = nt;
            double t = 0.0;
            rc = time_step(Q, 8, &t);
            if (rc) return rc;
            if (t < best_t) {
                best_t = t;
                best = spt;
                best_nt = nt;
            }
        }
    }
    Q.nt = best_nt;
    return Q.spt == best ? RUV_OK : build_tiles(Q, best);
}

}  // namespace

int batch_create(int64_t nrows, const int64_t* ptr, const int32_t* cols, const double* vals, int nchart,
                 const int64_t* crow, int R, int precond, int mode, BatchImpl** out) {
    if (R != 1 && R != 2) return fail(RUV_ERR_INVALID, "rhs_per_row must be 1 or 2");
    if (precond != RUV_PC_NONE && precond != RUV_PC_JACOBI && precond != RUV_PC_AMG)
        return fail(RUV_ERR_INVALID, "the batched solver supports the none, jacobi and amg preconditioners");
    if (mode != RUV_MODE_FP64 && mode != RUV_MODE_MIXED) return fail(RUV_ERR_INVALID, "unknown mode");
    if (nrows < 0 || nrows >= (int64_t)INT32_MAX)  // the kernels index rows in 32 bits
        return fail(RUV_ERR_INVALID, "batch: the row count must be below 2^31 - 1");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(RUV_ERR_UNAVAILABLE, "no HIP device");
    std::unique_ptr<BatchImpl> Q(new BatchImpl());
    // relative reduction each FP32 inner solve aims for before the FP64 residual is recomputed
    if (const char* e = std::getenv("RUV_BATCH_INNER_TOL")) {
        const double v = std::atof(e);
        if (v > 0.0 && v < 1.0) Q->inner_tol = v;
    }
    if (const char* e = std::getenv("RUV_BATCH_AMG_CHUNK")) Q->amg_chunk = std::min(64, std::max(1, std::atoi(e)));
    if (const char* e = std::getenv("RUV_BATCH_RESTART")) Q->continue_rounds = std::atoi(e) ? 0 : 1;
    if (const char* e = std::getenv("RUV_BATCH_CONT_GAP")) Q->cont_gap = std::atof(e);
    if (const char* e = std::getenv("RUV_BATCH_LIST_BELOW")) Q->list_below = std::atof(e);
    if (const char* e = std::getenv("RUV_UV_TOL")) Q->uv_tol = std::atof(e);
    if (const char* e = std::getenv("RUV_BATCH_FINAL_AIM")) Q->final_aim = std::atof(e);
    if (const char* e = std::getenv("RUV_BATCH_FIRST_CUT")) Q->first_cut = std::atof(e);
    if (const char* e = std::getenv("RUV_BATCH_LOOK_LAG")) Q->look_lag = std::atoi(e) ? 1 : 0;
    int rc = batch_setup(*Q, nrows, ptr, cols, vals, nchart, crow, R, precond, mode);
    if (rc) return rc;
    rc = autotune(*Q);
    if (rc) return rc;
    if (precond == RUV_PC_AMG) {  // the finest aggregates follow the tiles, so after the autotuner
        rc = bamg_upload_fine(*Q, ptr, cols, vals);
        if (!rc) rc = bamg_build(*Q);
        if (rc) return rc;
    }
    *out = Q.release();
    return RUV_OK;
}

void batch_destroy(BatchImpl* q) { delete q; }

int batch_solve_dev(BatchImpl* Q, const double* d_b, double* d_x, uint64_t max_iter, double tol,
                    ruv_solve_info* info, uint64_t* chart_iters, double* chart_resid) {
    const uintptr_t align = sizeof(double) * (uintptr_t)Q->R;  // rows are loaded as R-wide vectors
    if (((uintptr_t)d_b % align) != 0 || ((uintptr_t)d_x % align) != 0)
        return fail(RUV_ERR_INVALID, "device vectors must be aligned to rhs_per_row * 8 bytes");
    HIPCHK(hipEventRecord(Q->ev0, Q->stream));
    int rc;
    if (Q->mode == RUV_MODE_MIXED) rc = Q->R == 2 ? run_ci<float, 2>(*Q, d_b, d_x, max_iter, tol) : run_ci<float, 1>(*Q, d_b, d_x, max_iter, tol);
    else rc = Q->R == 2 ? run_ci<double, 2>(*Q, d_b, d_x, max_iter, tol) : run_ci<double, 1>(*Q, d_b, d_x, max_iter, tol);
    if (rc) return rc;
    HIPCHK(hipEventRecord(Q->ev1, Q->stream));
    HIPCHK(hipEventSynchronize(Q->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, Q->ev0, Q->ev1));
    if (Q->ktime) {
        for (size_t k = 0; k + 1 < Q->kev_used; k += 2) {
            float kms = 0;
            HIPCHK(hipEventElapsedTime(&kms, Q->kev[k], Q->kev[k + 1]));
            Q->k_seconds += kms * 1e-3;
        }
        Q->kev_used = 0;
        for (int c = 0; c < Q->nchart; ++c) {
            Q->k_bytes += Q->chart_step_bytes[c] * (double)Q->iters_total[c];
            Q->k_stored += Q->chart_step_stored[c] * (double)Q->iters_total[c];
        }
    }
    if (info) {
        uint64_t mx = 0;
        double worst = 0.0;
        for (int c = 0; c < Q->nchart; ++c) {
            mx = std::max<uint64_t>(mx, Q->iters_total[c]);
            worst = std::max(worst, Q->h_true[c]);
        }
        info->iters = mx;
        info->refinements = 0;
        info->precond = Q->precond;
        info->true_rel_resid = worst;
        info->seconds = ms * 1e-3;
    }
    if (chart_iters)
        for (int c = 0; c < Q->nchart; ++c) chart_iters[c] = Q->iters_total[c];
    if (chart_resid)
        for (int c = 0; c < Q->nchart; ++c) chart_resid[c] = Q->h_true[c];
    return RUV_OK;
}

// Mean duration and bytes of the step's SpMV kernel (kb_spmv) over the whole batch: algorithmic (every nonzero's
// value and stored column, the slice table, two vectors streamed once: p read, q written; neighbour gathers are
// served by the caches) and as stored (the padding slots of the sliced ELL included).
int batch_spmv_time(BatchImpl* Q, int reps, double* sec, double* bytes, double* stored) {
    if (reps <= 0) return fail(RUV_ERR_INVALID, "reps must be positive");
    int rc = time_step(*Q, reps, sec);
    if (rc) return rc;
    double b = 0.0, st = 0.0;
    for (int c = 0; c < Q->nchart; ++c) {
        b += Q->chart_step_bytes[c];
        st += Q->chart_step_stored[c];
    }
    if (bytes) *bytes = b;
    if (stored) *stored = st;
    return RUV_OK;
}

int batch_step_profile(BatchImpl* Q, int reps, int cap, int32_t* kind, int32_t* level, double* sec, double* bytes,
                       double* stored, int32_t* count) {
    if (reps <= 0 || cap < 0) return fail(RUV_ERR_INVALID, "reps must be positive and cap non-negative");
    if (Q->precond != RUV_PC_AMG || !Q->amg) return fail(RUV_ERR_INVALID, "the step profile needs a built AMG");
    std::vector<BatchImpl::LaunchMark> marks;
    if (Q->mev.empty()) {
        hipEvent_t e = nullptr;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        Q->mev.push_back(e);
    }
    Q->marks = &marks;
    int rc;
    const bool f = Q->mode == RUV_MODE_MIXED;
    if (Q->R == 2) {
        if (f) rc = Q->ci == 0 ? step_profile_kernel<float, 2, 0>(*Q, reps) : step_profile_kernel<float, 2, 1>(*Q, reps);
        else rc = Q->ci == 0 ? step_profile_kernel<double, 2, 0>(*Q, reps) : step_profile_kernel<double, 2, 1>(*Q, reps);
    } else {
        if (f) rc = Q->ci == 0 ? step_profile_kernel<float, 1, 0>(*Q, reps) : step_profile_kernel<float, 1, 1>(*Q, reps);
        else rc = Q->ci == 0 ? step_profile_kernel<double, 1, 0>(*Q, reps) : step_profile_kernel<double, 1, 1>(*Q, reps);
    }
    Q->marks = nullptr;
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(Q->stream));
    // per (kind, level): mean interval from the previous event to the launch's own
    struct Acc {
        int kind, level;
        double sec, bytes, stored;
        int n;
    };
    std::vector<Acc> acc;
    for (size_t k = 0; k < marks.size(); ++k) {
        const BatchImpl::LaunchMark& m = marks[k];
        if (m.kind < 0) continue;
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, Q->mev[k], Q->mev[k + 1]));
        size_t j = 0;
        while (j < acc.size() && !(acc[j].kind == m.kind && acc[j].level == m.level)) ++j;
        if (j == acc.size()) acc.push_back({m.kind, m.level, 0.0, m.bytes, m.stored, 0});
        acc[j].sec += ms * 1e-3;
        acc[j].n += 1;
    }
    const int nout = (int)std::min<size_t>(acc.size(), (size_t)cap);
    for (int j = 0; j < nout; ++j) {
        if (kind) kind[j] = acc[j].kind;
        if (level) level[j] = acc[j].level;
        if (sec) sec[j] = acc[j].sec / acc[j].n;
        if (bytes) bytes[j] = acc[j].bytes;
        if (stored) stored[j] = acc[j].stored;
    }
    if (count) *count = (int32_t)acc.size();
    return RUV_OK;
}

int batch_precond_levels(BatchImpl* Q, int64_t* rows, int32_t cap) { return bamg_level_rows(*Q, rows, cap); }

int batch_precond_fallbacks(BatchImpl* Q, int64_t* not_spd, int64_t* coupled, int64_t* too_large) {
    return bamg_fallbacks(*Q, not_spd, coupled, too_large);
}

int batch_coarsest_block(BatchImpl* Q, int c, int64_t* m_out, double* block, double* inverse) {
    return bamg_coarsest_block(*Q, c, m_out, block, inverse);
}

int batch_amg_level(BatchImpl* Q, int l, int64_t* sizes, int64_t* row_ptr, int32_t* cols, double* vals, int32_t* agg,
                    int32_t* chart) {
    return bamg_level(*Q, l, sizes, row_ptr, cols, vals, agg, chart);
}

int batch_setup_precond(BatchImpl* Q) { return Q->precond == RUV_PC_AMG ? bamg_build(*Q) : RUV_OK; }

int batch_precond_info(BatchImpl* Q, int32_t* levels, double* setup_seconds, int64_t* coarse_rows) {
    int lv = 0;
    int rc = bamg_info(*Q, &lv, setup_seconds, coarse_rows);
    if (levels) *levels = lv;
    return rc;
}

// Position of every CSR entry of the batch's operator in its value arrays (the storage form of batch_setup): the
// sliced-ELL position at map[a], and at map[nnz + a] the band position (-1 outside band charts and for the lower
// entries, which the band form keeps as their mirrors).
int batch_value_map(const BatchImpl& Q, const int64_t* ptr, const int32_t* cols, std::vector<int64_t>* map) {
    map->assign(2 * (size_t)Q.nnz, -1);
    for (int c = 0; c < Q.nchart; ++c) {
        const int w = Q.h_bw[c];
        if (!w) continue;
        const int64_t lo = Q.crow[c], n = Q.crow[c + 1] - lo;
        for (int64_t i = lo; i < lo + n; ++i)
            for (int64_t a = ptr[i]; a < ptr[i + 1]; ++a) {
                const int64_t d = (int64_t)cols[a] - i;
                if (d >= 0) (*map)[Q.nnz + a] = Q.h_bbase[c] + (int64_t)band_slot(d, w, Q.h_bmask[c]) * n + (i - lo);
            }
    }
    for (int s = 0; s < Q.nslice; ++s) {
        const int64_t r0 = Q.h_sl_row0[s];
        const int n = Q.h_sl_n[s];
        const int32_t* dt = Q.h_dtab.data() + (size_t)s * SLOTS;
        for (int l = 0; l < n; ++l) {
            const int64_t i = r0 + l;
            int k = 0;
            for (int64_t a = ptr[i]; a < ptr[i + 1]; ++a, ++k) {
                if (Q.h_stencil[s]) {
                    const int slot = (int)(std::lower_bound(dt, dt + Q.h_sl_w[s], (int32_t)(cols[a] - i)) - dt);
                    (*map)[a] = Q.h_sl_off[s] + (int64_t)slot * n + l;
                } else {
                    (*map)[a] = Q.h_sl_off[s] + (int64_t)k * n + l;
                }
            }
        }
    }
    return RUV_OK;
}

namespace {
// new values (CSR order) into the value arrays (entries of one slot sum, as at setup: each slot is written by
// one CSR entry except with duplicate columns, which the assembly never produces) and the inverse diagonal
__global__ __launch_bounds__(BLK) void kb_scatter_values(int64_t nnz, const double* __restrict__ v,
                                                         const int64_t* __restrict__ map, double* __restrict__ e64,
                                                         float* __restrict__ e32, double* __restrict__ b64,
                                                         float* __restrict__ b32) {
    const int64_t a = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (a >= nnz) return;
    const int64_t e = map[a];
    e64[e] = v[a];
    if (e32) e32[e] = (float)v[a];
    const int64_t b = map[nnz + a];  // band form: upper entries of band charts
    if (b >= 0) {
        if (b64) b64[b] = v[a];
        if (b32) b32[b] = (float)v[a];
    }
}
__global__ __launch_bounds__(BLK) void kb_inv_diag(int64_t n, const double* __restrict__ v,
                                                   const int64_t* __restrict__ diagpos, double* __restrict__ d64,
                                                   float* __restrict__ d32) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const double d = 1.0 / v[diagpos[i]];
    d64[i] = d;
    if (d32) d32[i] = (float)d;
}
}  // namespace

// Replace the operator's values (same structure, CSR order, device) and rebuild what depends on them.  The band
// form keeps the upper entries of its charts, so the new values must stay symmetric (the batch's operator is SPD).
int batch_update_values(BatchImpl* Q, const double* d_vals, const int64_t* d_map, const int64_t* d_diagpos) {
    hipStream_t s = Q->stream;
    const dim3 ge((unsigned)std::max<int64_t>(1, (Q->nnz + BLK - 1) / BLK));
    const dim3 gr((unsigned)std::max<int64_t>(1, (Q->nrows + BLK - 1) / BLK));
    hipLaunchKernelGGL(kb_scatter_values, ge, dim3(BLK), 0, s, Q->nnz, d_vals, d_map, Q->ev64, Q->ev32, Q->bv64,
                       Q->bv32);
    if (Q->dinv64)
        hipLaunchKernelGGL(kb_inv_diag, gr, dim3(BLK), 0, s, Q->nrows, d_vals, d_diagpos, Q->dinv64, Q->dinv32);
    HIPCHK(hipGetLastError());
    if (Q->precond == RUV_PC_AMG) {
        int rc = bamg_set_fine_values(*Q, d_vals);
        return rc ? rc : bamg_build(*Q);
    }
    return RUV_OK;
}

int batch_kernel_timing(BatchImpl* Q, int enable) {
    Q->ktime = enable != 0;
    Q->kev_used = 0;
    Q->k_launches = 0;
    Q->k_seconds = 0.0;
    Q->k_bytes = 0.0;
    Q->k_stored = 0.0;
    return RUV_OK;
}

int batch_kernel_stats(BatchImpl* Q, int64_t* launches, double* seconds, double* bytes, double* stored) {
    if (launches) *launches = Q->k_launches;
    if (seconds) *seconds = Q->k_seconds;
    if (bytes) *bytes = Q->k_bytes;
    if (stored) *stored = Q->k_stored;
    Q->k_launches = 0;
    Q->k_seconds = 0.0;
    Q->k_bytes = 0.0;
    Q->k_stored = 0.0;
    return RUV_OK;
}

hipStream_t batch_stream(const BatchImpl* Q) { return Q->stream; }

int batch_band_info(const BatchImpl* Q, int32_t* band_charts, int64_t* band_values) {
    if (band_charts) *band_charts = Q->nband;
    if (band_values) *band_values = Q->nbv;
    return RUV_OK;
}

int batch_schedule(BatchImpl* Q, int32_t* rows_per_tile, int32_t* col_bytes, int64_t* ell_entries,
                   int64_t* col_entries, int32_t* stencil_slices, int32_t* slices, int32_t* nontemporal) {
    if (rows_per_tile) *rows_per_tile = Q->spt * SLICE;
    if (nontemporal) *nontemporal = Q->nt;
    if (col_bytes) *col_bytes = Q->ci == 0 ? 2 : 4;
    if (ell_entries) *ell_entries = Q->nell;
    if (col_entries) *col_entries = Q->ncol;
    if (stencil_slices) *stencil_slices = Q->nstencil;
    if (slices) *slices = Q->nslice;
    return RUV_OK;
}

}  // namespace ruv
