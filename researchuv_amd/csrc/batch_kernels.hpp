// batch_kernels.hpp -- device side of the batched solver: sliced-ELL storage and a three-kernel PCG iteration.
//
// Storage.  Every chart is one diagonal block; its rows are cut into slices of at most SLICE (= BLK) rows, one
// row per thread.  Slice s holds rows [sl_row0, sl_row0 + n_s) in w_s value columns; value k of row lane sits at
// sl_off + k * n_s + lane, so every load of a slice is one coalesced sweep.  A slice is stored in one of two
// forms:
//   stencil  the slice's rows use at most SLOTS distinct column offsets (col - row), listed in dtab; value
//            column k holds every row's entry at offset dtab[k] (0 where a row has none).  No column indices are
//            stored, and the gathers of one value column read consecutive rows.  Offsets that land outside the
//            chart (only ever with value 0) are clamped into it.  Mesh charts in a locality-preserving vertex
//            order are stencil slices almost everywhere.
//   general  rows padded to the slice's longest row; a column per entry at sl_coff + k * n_s + lane, as a
//            16-bit offset from the row or, when some offset in the batch does not fit, a 32-bit row.  Padding is
//            (value 0, the row itself).
// A tile is a run of slices of one chart; one workgroup owns one tile.
//
// Band charts.  A chart whose free-vertex rows form a grid of width W in row-major order -- every column offset
// is one of 0, +-1, +-(W - 1), +-W, +-(W + 1), and the operator is exactly symmetric -- is also stored in band
// form for the kernels that stream the finest operator once per step (the SpMV and the two finest smoothing
// passes): only the diagonal and the upper offsets present in the chart (1, W - 1, W, W + 1, in that order;
// bit b of the chart's mask says whether offset b is present), each as one slot-major column of the chart's n
// rows, value 0 where a row has no entry.  Row i's entry at offset -d is row (i - d)'s entry at +d.  A workgroup
// thread owns the rows r0, r0 + W, ..., r0 + (H - 1) W of a band of H = BAND_H grid rows and
// walks down them with a three-row window of the gathered vector (rows r - W, r, r + W, each at columns -1, 0,
// +1), so each row gathers three new vector elements that lie in the cache lines of its own row, and the lower
// entry at -W is the upper entry at +W of the thread's previous row.  The chart's band columns are dealt to its
// tiles in chunks of 256, chunk k to tile k, k + ntile, ...
//
// Iteration.  With z = M r (Jacobi) the reference recurrence (solver.rs:405-552)
//     alpha = rz / pAp;  x += alpha p;  r -= alpha q;  z = M r;  beta = rz' / rz;  p = z + beta p
// needs two chart-wide reductions per step, so it runs as three kernels:
//   kb_dir     opens the step (convergence on ||r||, beta from the r.z partials), p' = z + beta p (a stream);
//   kb_spmv    q = A p', partials p'.q;
//   kb_update  alpha = rz / p'.q; x, r, z updated; partials r.r and r.z.
// (Forming p' for every gathered neighbour inside the SpMV instead, from z and the old p, saves the direction
// pass but doubles the gathers; measured slower.)
// The old and new directions live in two buffers swapped every step, and the per-step scalars (r.r and r.z
// partials, chart r.z) are double-buffered by step parity, so no workgroup reads a value another workgroup of
// the same launch writes.  All workgroups of a chart sum the same partials in the same order, so they take the
// same decisions and the recurrence of each chart is exactly an independent PCG.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dev_common.hpp"

namespace ruv {

constexpr int SLICE = BLK;  // rows per slice (one per thread)
constexpr int SLOTS = 8;    // column offsets of a stencil slice
constexpr int BAND_H = 4;   // band charts: grid rows per thread
constexpr int XCDS = 8;     // MI355X: workgroups are dispatched round-robin over 8 XCDs, each with its own L2

// chart states; B_BREAKDOWN_PQ: p.q broke down in the update pass (its step's x update is not made), told apart from
// the direction pass's own stops
enum : int32_t { B_RUNNING = 0, B_CONVERGED = 1, B_BREAKDOWN = 2, B_MAXITER = 3, B_SKIP = 4, B_BREAKDOWN_PQ = 5 };

struct BDev {
    // sliced ELL (per slice)
    const int64_t* sl_row0;
    const int64_t* sl_off;   // first value
    const int64_t* sl_coff;  // first column of a general slice, -1 for a stencil slice
    const int32_t* sl_w;
    const int32_t* sl_n;
    const int32_t* sl_dtab;  // [nslice][SLOTS] offsets of a stencil slice
    const int64_t* crow;     // nchart + 1 chart row offsets
    const void* ecol;        // general slices: int16 offsets (CI = 0) or int32 rows (CI = 1)
    // tiles: tile t = slices [tile_s0[t], tile_s0[t + 1]) = rows [tile_r0[t], tile_r1[t]) of chart tile_chart[t]
    const int32_t* tile_s0;
    const int64_t* tile_r0;
    const int64_t* tile_r1;
    const int32_t* tile_chart;
    const int32_t* chart_t0;  // nchart + 1
    // per chart
    double* c_rz;  // [2][nchart] by step parity
    double* c_alpha;  // alpha of the chart's last update pass (its x update is made by the next direction pass)
    double* c_tol;
    double* c_bnorm;
    double* c_true;
    double* c_rrec;  // recursive ||r|| of the step an inner solve converged at
    // forward-error test of the mixed refinement: c_rin the FP64 ||r|| each round starts from, c_gamma the largest
    // ratio max|correction| / ||r_start|| of a round so far, c_tgt the residual norm the chart must reach
    double* c_rin;
    double* c_gamma;
    double* c_tgt;
    // big: charts with very many tiles; the step's chart-wide sums come from kb_chart_sums (p.q in c_sa, r.r and
    // r.z by parity in c_sb, c_sc) instead of every workgroup summing the chart's partials
    int32_t big;
    double* c_sa;
    double* c_sb;
    double* c_sc;
    int32_t* c_done;
    int32_t* c_final;
    uint32_t* c_iters;
    // per tile partials
    double* part_a;  // p.q
    double* part_b;  // [2][ntile] r.r by parity
    double* part_c;  // [2][ntile] r.z by parity
    double* part_o;  // outer FP64 residual / ||b||
    double* part_d;  // [2][ntile] largest |correction| and |x| of the tile after a refinement round
    uint32_t max_iter;
    int32_t nchart;
    int32_t ntile;
    // band charts (null when the batch has none): grid width W (0: the chart is not in band form), offset mask,
    // first value of the chart in the band arrays
    const int32_t* c_bw;
    const int32_t* c_bmask;
    const int64_t* c_bbase;
};

template <typename T, int R>
struct alignas(sizeof(T) * R) Vr {
    T c[R];
};

template <typename T, int R>
__device__ __forceinline__ Vr<T, R> ld(const T* __restrict__ p, int64_t row) {
    return *reinterpret_cast<const Vr<T, R>*>(p + row * R);
}
template <typename T, int R>
__device__ __forceinline__ void st(T* __restrict__ p, int64_t row, const Vr<T, R>& v) {
    *reinterpret_cast<Vr<T, R>*>(p + row * R) = v;
}

template <int CI>
__device__ __forceinline__ int64_t col_of(const void* __restrict__ ec, int64_t e, int64_t row) {
    if constexpr (CI == 0) return row + (int64_t)static_cast<const int16_t*>(ec)[e];
    else return (int64_t) static_cast<const int32_t*>(ec)[e];
}

// Workgroup -> slot of the active-tile list such that each XCD receives one contiguous run of slots (tiles of a
// chart are contiguous, so a chart's rows, neighbours and partials stay in one L2).  Returns -1 for the few
// padding workgroups of the last run.
__device__ __forceinline__ int xcd_slot(int na) {
    const int per = (na + XCDS - 1) / XCDS;
    const int b = (int)blockIdx.x;
    const int slot = (b % XCDS) * per + b / XCDS;
    return slot < na ? slot : -1;
}
// the same slots, each XCD walking its range from the end: a pass that reads what the pass before it wrote in the
// other direction meets the most recently written rows first, while they are still in the memory-side cache
__device__ __forceinline__ int xcd_slot_rev(int na) {
    const int per = (na + XCDS - 1) / XCDS;
    const int b = (int)blockIdx.x;
    const int slot = (b % XCDS) * per + (per - 1 - b / XCDS);
    return slot < na ? slot : -1;
}

// sums of two partial arrays over [a, b), fixed order, identical in every workgroup
__device__ __forceinline__ void sum_range2(const double* __restrict__ pa, const double* __restrict__ pb, int64_t a,
                                           int64_t b, double* sh, double* sa, double* sb) {
    double va = 0.0, vb = 0.0;
    for (int64_t i = a + threadIdx.x; i < b; i += BLK) {
        va += pa[i];
        vb += pb[i];
    }
    va = wave_sum(va);
    vb = wave_sum(vb);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        sh[threadIdx.x >> 6] = va;
        sh[WAVES + (threadIdx.x >> 6)] = vb;
    }
    __syncthreads();
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int i = 0; i < WAVES; ++i) {
        s0 += sh[i];
        s1 += sh[WAVES + i];
    }
    *sa = s0;
    *sb = s1;
}

// NT = 1 loads the matrix values with the non-temporal hint (they are read once per step; the gathered vectors
// are the data worth keeping in L2); the autotuner picks it or not.
template <typename T>
__device__ __forceinline__ T ld_val(const T* __restrict__ p, int nt) {
    return nt ? __builtin_nontemporal_load(p) : *p;
}

// bfloat16 matrix values (the AMG cycle's copies of the operators under FP32): the upper 16 bits of an FP32
struct Bf16 {
    uint16_t bits;
};
__device__ __forceinline__ float cval(Bf16 v) { return __uint_as_float((uint32_t)v.bits << 16); }
__device__ __forceinline__ double cval(double v) { return v; }
__device__ __forceinline__ float cval(float v) { return v; }
__device__ __forceinline__ float ld_val(const Bf16* __restrict__ p, int) { return cval(*p); }

// Row i of slice s (lane threadIdx.x): acc[c] += sum over the row's slots of a_ij * g(j)[c].  The slot loop is
// unrolled for the slice's width through a switch (the width is uniform over the workgroup): a loop of run-time
// length issues its loads a few at a time, and the kernels built on it measured 20-35 % slower.
// Stencil slots in 32-bit row arithmetic (a batch has fewer than 2^31 rows); CLAMP = 0 for a slice whose every
// offset stays inside its chart, which skips the clamp.
template <int W, int CLAMP, typename T, int R, int NT, typename V, typename G>
__device__ __forceinline__ void stencil_slots(const V* __restrict__ v, int n, const int32_t* __restrict__ dt,
                                              int32_t i, int32_t lo, int32_t hi, G g, T (&acc)[R]) {
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const T a = (T)ld_val(v + k * n, NT);
        const int32_t j = CLAMP ? min(max(i + dt[k], lo), hi) : i + dt[k];
        const Vr<T, R> x = g((int64_t)j);
#pragma unroll
        for (int cc = 0; cc < R; ++cc) acc[cc] = fma(a, x.c[cc], acc[cc]);
    }
}
template <int W, typename T, int R, int CI, int NT, typename V, typename G>
__device__ __forceinline__ void general_slots(const V* __restrict__ v, int n, const void* __restrict__ ec, int64_t f0,
                                              int64_t i, G g, T (&acc)[R]) {
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const T a = (T)ld_val(v + k * n, NT);
        const Vr<T, R> x = g(col_of<CI>(ec, f0 + k * n, i));
#pragma unroll
        for (int cc = 0; cc < R; ++cc) acc[cc] = fma(a, x.c[cc], acc[cc]);
    }
}
constexpr int UNROLL_W = 16;  // general slices wider than this take a plain loop

template <typename T, int R, int CI, int NT, typename V, typename G>
__device__ __forceinline__ void row_slots(const BDev& B, int s, int n, int64_t i, int64_t lo, int64_t hi,
                                          const V* __restrict__ vals, G g, T (&acc)[R]) {
    const V* v = vals + B.sl_off[s] + threadIdx.x;
    const int64_t c0 = B.sl_coff[s];
    const int w = B.sl_w[s];
    if (c0 < 0) {
        if (w <= 0) return;
        const int32_t* dt = B.sl_dtab + (int64_t)s * SLOTS;
        const int64_t r0 = B.sl_row0[s];
        const bool inner = r0 + dt[0] >= lo && r0 + n - 1 + dt[w - 1] <= hi;  // offsets ascend
        const int32_t i32 = (int32_t)i, lo32 = (int32_t)lo, hi32 = (int32_t)hi;
        if (inner) {
            switch (w) {
#define RUV_ST(W) \
    case W: stencil_slots<W, 0, T, R, NT>(v, n, dt, i32, lo32, hi32, g, acc); break;
                RUV_ST(1) RUV_ST(2) RUV_ST(3) RUV_ST(4) RUV_ST(5) RUV_ST(6) RUV_ST(7) RUV_ST(8)
#undef RUV_ST
                default: break;
            }
        } else {
            switch (w) {
#define RUV_ST(W) \
    case W: stencil_slots<W, 1, T, R, NT>(v, n, dt, i32, lo32, hi32, g, acc); break;
                RUV_ST(1) RUV_ST(2) RUV_ST(3) RUV_ST(4) RUV_ST(5) RUV_ST(6) RUV_ST(7) RUV_ST(8)
#undef RUV_ST
                default: break;
            }
        }
    } else {
        const int64_t f0 = c0 + threadIdx.x;
        switch (w) {
#define RUV_GS(W) \
    case W: general_slots<W, T, R, CI, NT>(v, n, B.ecol, f0, i, g, acc); break;
            RUV_GS(1) RUV_GS(2) RUV_GS(3) RUV_GS(4) RUV_GS(5) RUV_GS(6) RUV_GS(7) RUV_GS(8)
            RUV_GS(9) RUV_GS(10) RUV_GS(11) RUV_GS(12) RUV_GS(13) RUV_GS(14) RUV_GS(15) RUV_GS(16)
#undef RUV_GS
            default:
                for (int k = 0; k < w; ++k) {
                    const T a = (T)ld_val(v + k * n, NT);
                    const Vr<T, R> x = g(col_of<CI>(B.ecol, f0 + (int64_t)k * n, i));
#pragma unroll
                    for (int cc = 0; cc < R; ++cc) acc[cc] = fma(a, x.c[cc], acc[cc]);
                }
        }
    }
}

// P full inner stencil slices of one width at once (row threadIdx.x of each): P times the loads in flight per
// thread before the first one is waited for.  v[p], d[p], i[p]: values, offsets and row of slice p.
template <int P, int W, typename T, int R, int NT, typename V, typename G>
__device__ __forceinline__ void stencil_group(const V* const (&v)[P], const int32_t* const (&d)[P],
                                              const int32_t (&i)[P], G g, T (&acc)[P][R]) {
#pragma unroll
    for (int k = 0; k < W; ++k) {
        T u[P];
        Vr<T, R> x[P];
#pragma unroll
        for (int q = 0; q < P; ++q) {
            u[q] = (T)ld_val(v[q] + k * SLICE, NT);
            x[q] = g((int64_t)(i[q] + d[q][k]));
        }
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
            for (int cc = 0; cc < R; ++cc) acc[q][cc] = fma(u[q], x[q].c[cc], acc[q][cc]);
    }
}

// Width of slices s .. s + P - 1 when all are full stencil slices of that width whose every offset stays inside
// the chart [lo, hi], else 0.
template <int P>
__device__ __forceinline__ int group_width(const BDev& B, int s, int64_t lo, int64_t hi) {
    const int w = B.sl_w[s];
    if (w <= 0) return 0;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        if (B.sl_coff[s + q] >= 0 || B.sl_n[s + q] != SLICE || B.sl_w[s + q] != w) return 0;
        const int32_t* d = B.sl_dtab + (int64_t)(s + q) * SLOTS;
        const int64_t r0 = B.sl_row0[s + q];
        if (r0 + d[0] < lo || r0 + SLICE - 1 + d[w - 1] > hi) return 0;
    }
    return w;
}

// Band form (see the header): the chart's band columns of chunk k, k + ntl, ... of this workgroup; for each of
// its rows, acc = sum over the row's entries of a_ij * g(j), then fin(row, acc).  NT: non-temporal upper values.
// M: the chart's offset mask when known at compile time, so a row issues all its loads before any is waited for
// (a load inside a branch is waited for inside it); M = 0 takes the mask at run time and loads every slot, an
// absent one from the diagonal's slot, weighted 0.
template <int H, int M, typename T, int R, int NT, typename V, typename G, typename F>
__device__ __forceinline__ void band_rows_hm(int64_t lo, int n, int W, int m, const V* __restrict__ v0, int k,
                                             int ntl, G g, F fin) {
    const int mk = M ? M : m;
    int slot = 1;
    const V* v1 = (mk & 1) ? v0 + (int64_t)(slot++) * n : v0;   // +1
    const V* vwm = (mk & 2) ? v0 + (int64_t)(slot++) * n : v0;  // +(W - 1)
    const V* vw = (mk & 4) ? v0 + (int64_t)(slot++) * n : v0;   // +W
    const V* vwp = (mk & 8) ? v0 + (int64_t)(slot++) * n : v0;  // +(W + 1)
    // present offsets: compile-time with M, else weights 1 / 0 on every slot
    constexpr bool ALL = M == 0;
    const T f1 = (mk & 1) ? T(1) : T(0), fwm = (mk & 2) ? T(1) : T(0), fw = (mk & 4) ? T(1) : T(0),
            fwp = (mk & 8) ? T(1) : T(0);
    const auto gx = [&](int r) { return g(lo + (int64_t)min(max(r, 0), n - 1)); };
    const auto low = [&](const V* a, int r) { return r >= 0 ? (T)cval(a[r]) : T(0); };  // r < n always
    const auto up = [&](const V* a, int r) { return (T)ld_val(a + r, NT); };
    const int hw = H * W;
    const int64_t ncol = (int64_t)((n + hw - 1) / hw) * W;
    const int nchunk = (int)((ncol + BLK - 1) / BLK);
    for (int ch = k; ch < nchunk; ch += ntl) {
        const int gi = ch * BLK + (int)threadIdx.x;
        const int band = gi / W, j = gi - band * W;
        const int r0 = band * hw + j;  // >= n for the columns past the chart's last band
        if (r0 >= n) continue;
        Vr<T, R> xm0 = gx(r0 - W - 1), xm1 = gx(r0 - W), xm2 = gx(r0 - W + 1);
        Vr<T, R> xc0 = gx(r0 - 1), xc1 = gx(r0), xc2 = gx(r0 + 1);
        T aw_prev = (ALL || (M & 4)) ? fw * low(vw, r0 - W) : T(0);
        // row by row: each row's loads issue together; its store keeps the next row's loads behind it (the
        // register use of hoisting them costs more occupancy than their latency)
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int r = r0 + h * W;
            if (r >= n) break;
            T acc[R];
            const Vr<T, R> xn0 = gx(r + W - 1), xn1 = gx(r + W), xn2 = gx(r + W + 1);
            const T a0 = up(v0, r);
#pragma unroll
            for (int cc = 0; cc < R; ++cc) acc[cc] = a0 * xc1.c[cc];
            if (ALL || (M & 1)) {
                const T u = f1 * up(v1, r), l = f1 * low(v1, r - 1);
#pragma unroll
                for (int cc = 0; cc < R; ++cc) acc[cc] = fma(l, xc0.c[cc], fma(u, xc2.c[cc], acc[cc]));
            }
            if (ALL || (M & 2)) {
                const T u = fwm * up(vwm, r), l = fwm * low(vwm, r - W + 1);
#pragma unroll
                for (int cc = 0; cc < R; ++cc) acc[cc] = fma(l, xm2.c[cc], fma(u, xn0.c[cc], acc[cc]));
            }
            if (ALL || (M & 4)) {
                const T u = fw * up(vw, r);
#pragma unroll
                for (int cc = 0; cc < R; ++cc) acc[cc] = fma(aw_prev, xm1.c[cc], fma(u, xn1.c[cc], acc[cc]));
                aw_prev = u;
            }
            if (ALL || (M & 8)) {
                const T u = fwp * up(vwp, r), l = fwp * low(vwp, r - W - 1);
#pragma unroll
                for (int cc = 0; cc < R; ++cc) acc[cc] = fma(l, xm0.c[cc], fma(u, xn2.c[cc], acc[cc]));
            }
            xm0 = xc0;
            xm1 = xc1;
            xm2 = xc2;
            xc0 = xn0;
            xc1 = xn1;
            xc2 = xn2;
            fin(lo + (int64_t)r, acc);
        }
    }
}
// band_rows_hm with BAND_H grid rows per thread, the masks of triangulated grids (diagonals one way: 1, W, W + 1
// or the other: 1, W - 1, W; both: all four) and of plain grids (1, W) at compile time
constexpr bool band_mask_compiled(int m) { return m == 13 || m == 7 || m == 15 || m == 5; }
template <typename T, int R, int NT, typename V, typename G, typename F>
__device__ __forceinline__ void band_rows(const BDev& B, int c, int k, int ntl, const V* __restrict__ bvals, G g,
                                          F fin) {
    const int64_t lo = B.crow[c];
    const int n = (int)(B.crow[c + 1] - lo);
    const int W = B.c_bw[c], m = B.c_bmask[c];
    const V* v0 = bvals + B.c_bbase[c];
    switch (m) {
        case 13: band_rows_hm<BAND_H, 13, T, R, NT>(lo, n, W, m, v0, k, ntl, g, fin); break;
        case 7: band_rows_hm<BAND_H, 7, T, R, NT>(lo, n, W, m, v0, k, ntl, g, fin); break;
        case 15: band_rows_hm<BAND_H, 15, T, R, NT>(lo, n, W, m, v0, k, ntl, g, fin); break;
        case 5: band_rows_hm<BAND_H, 5, T, R, NT>(lo, n, W, m, v0, k, ntl, g, fin); break;
        default: band_rows_hm<BAND_H, 0, T, R, NT>(lo, n, W, m, v0, k, ntl, g, fin); break;
    }
}

// f(i, i2, two) over the rows r0 + threadIdx.x + k BLK < r1 of a tile, two at a time (i and i2 = i + BLK), so a
// body can issue both rows' loads before it stores either; past the tile's last row i2 repeats i and two is false.
// A body uses the second row's loads unconditionally (where two is false it stores row i's results to row i once
// more, and adds nothing to its sums through a select): loads used only under a branch are moved into it by the
// compiler, and then wait for the first row's stores.
template <typename F>
__device__ __forceinline__ void tile_rows2(int64_t r0, int64_t r1, F f) {
    for (int64_t i = r0 + threadIdx.x; i < r1; i += 2 * BLK) {
        const bool two = i + BLK < r1;
        f(i, two ? i + BLK : i, two);
    }
}

// Opens step `iter` (parity par) of every live chart -- convergence on ||r||, max_iter, breakdown, beta from the
// r.z partials -- then p' = z + beta p over the tile's rows, and the previous step's x update x += alpha p (alpha
// from the update pass, p the old direction this pass reads anyway): also when the chart stops here, so an inner
// solve ends with x complete.  The pass's workgroups of one chart each make the same decision from the same sums;
// the first writes the chart's state, so another one of the same launch may already see the new state: a
// workgroup leaves at once only for a chart stopped before this step (state not running and c_iters below iter --
// a stop here leaves c_iters at iter) or broken down in the last update pass.
template <typename T, int R>
__global__ __launch_bounds__(BLK) void kb_dir(BDev B, const int* __restrict__ act, int na, int par, uint32_t iter,
                                              const T* __restrict__ z, const T* __restrict__ pold,
                                              T* __restrict__ pnew, T* __restrict__ x, int stop = 0) {
    __shared__ double sh[2 * WAVES];
    const int slot = xcd_slot_rev(na);
    if (slot < 0) return;
    const int t = act[slot];
    const int c = B.tile_chart[t];
    const int32_t state = B.c_done[c];
    if (state != B_RUNNING && (state == B_BREAKDOWN_PQ || B.c_iters[c] < iter)) return;
    const int t0 = B.chart_t0[c], t1 = B.chart_t0[c + 1];
    double rr, rz;
    if (B.big) {  // summed per chart by kb_chart_sums
        rr = B.c_sb[(int64_t)par * B.nchart + c];
        rz = B.c_sc[(int64_t)par * B.nchart + c];
    } else {
        sum_range2(B.part_b + (int64_t)par * B.ntile, B.part_c + (int64_t)par * B.ntile, t0, t1, sh, &rr, &rz);
    }
    const bool lead = (t == t0) && threadIdx.x == 0;
    const T a = iter > 0 ? (T)B.c_alpha[c] : T(0);
    // the previous step's x update alone, for a chart that stops here
    const auto x_only = [&]() {
        if (iter == 0) return;
        tile_rows2(B.tile_r0[t], B.tile_r1[t], [&](int64_t i, int64_t i2, bool two) {
            const Vr<T, R> pi = ld<T, R>(pold, i), xi = ld<T, R>(x, i);
            const Vr<T, R> p2 = ld<T, R>(pold, i2), x2 = ld<T, R>(x, i2);
            Vr<T, R> xn, xn2;
#pragma unroll
            for (int cc = 0; cc < R; ++cc) {
                xn.c[cc] = fma(a, pi.c[cc], xi.c[cc]);
                xn2.c[cc] = fma(a, p2.c[cc], x2.c[cc]);
            }
            st<T, R>(x, i, xn);
            st<T, R>(x, i2, two ? xn2 : xn);
        });
    };
    if (stop || sqrt(rr) <= B.c_tol[c]) {  // stop: the inner solve ends here for every chart still running
        if (lead) {
            B.c_done[c] = B_CONVERGED;
            B.c_rrec[c] = sqrt(rr);
        }
        x_only();
        return;
    }
    if (iter >= B.max_iter) {
        if (lead) B.c_done[c] = B_MAXITER;
        x_only();
        return;
    }
    const double rzold = B.c_rz[(int64_t)(par ^ 1) * B.nchart + c];
    if (!isfinite(rz) || !(fabs(rzold) > 0.0)) {
        if (lead) B.c_done[c] = B_BREAKDOWN;
        x_only();
        return;
    }
    if (lead) {
        B.c_rz[(int64_t)par * B.nchart + c] = rz;
        B.c_iters[c] = iter + 1;
    }
    const T beta = (T)(rz / rzold);
    if (iter == 0) {  // no update before the first step
        tile_rows2(B.tile_r0[t], B.tile_r1[t], [&](int64_t i, int64_t i2, bool two) {
            const Vr<T, R> zi = ld<T, R>(z, i), pi = ld<T, R>(pold, i);
            const Vr<T, R> z2 = ld<T, R>(z, i2), p2 = ld<T, R>(pold, i2);
            Vr<T, R> pn, pn2;
#pragma unroll
            for (int cc = 0; cc < R; ++cc) {
                pn.c[cc] = fma(beta, pi.c[cc], zi.c[cc]);
                pn2.c[cc] = fma(beta, p2.c[cc], z2.c[cc]);
            }
            st<T, R>(pnew, i, pn);
            st<T, R>(pnew, i2, two ? pn2 : pn);
        });
        return;
    }
    tile_rows2(B.tile_r0[t], B.tile_r1[t], [&](int64_t i, int64_t i2, bool two) {
        const Vr<T, R> zi = ld<T, R>(z, i), pi = ld<T, R>(pold, i), xi = ld<T, R>(x, i);
        const Vr<T, R> z2 = ld<T, R>(z, i2), p2 = ld<T, R>(pold, i2), x2 = ld<T, R>(x, i2);
        Vr<T, R> pn, pn2, xn, xn2;
#pragma unroll
        for (int cc = 0; cc < R; ++cc) {
            pn.c[cc] = fma(beta, pi.c[cc], zi.c[cc]);
            pn2.c[cc] = fma(beta, p2.c[cc], z2.c[cc]);
            xn.c[cc] = fma(a, pi.c[cc], xi.c[cc]);
            xn2.c[cc] = fma(a, p2.c[cc], x2.c[cc]);
        }
        st<T, R>(pnew, i, pn);
        st<T, R>(pnew, i2, two ? pn2 : pn);
        st<T, R>(x, i, xn);
        st<T, R>(x, i2, two ? xn2 : xn);
    });
}

// Rows threadIdx.x of slices s .. s + P - 1 through stencil_group when group_width allows, each finished by fin;
// false (nothing done) otherwise.
template <int P, typename T, int R, int NT, typename V, typename G, typename F>
__device__ __forceinline__ bool spmv_group(const BDev& B, int s, int64_t lo, int64_t hi, const V* __restrict__ vals,
                                           G g, F fin) {
    const int w = group_width<P>(B, s, lo, hi);
    if (w == 0) return false;
    const V* v[P];
    const int32_t* d[P];
    int32_t i[P];
    T acc[P][R];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        v[q] = vals + B.sl_off[s + q] + threadIdx.x;
        d[q] = B.sl_dtab + (int64_t)(s + q) * SLOTS;
        i[q] = (int32_t)(B.sl_row0[s + q] + threadIdx.x);
#pragma unroll
        for (int cc = 0; cc < R; ++cc) acc[q][cc] = T(0);
    }
    switch (w) {
#define RUV_SG(W) \
    case W: stencil_group<P, W, T, R, NT>(v, d, i, g, acc); break;
        RUV_SG(1) RUV_SG(2) RUV_SG(3) RUV_SG(4) RUV_SG(5) RUV_SG(6) RUV_SG(7) RUV_SG(8)
#undef RUV_SG
        default: break;
    }
#pragma unroll
    for (int q = 0; q < P; ++q) fin((int64_t)i[q], acc[q]);
    return true;
}

// q = A p' over one tile of a running chart, partial p'.q of the tile.  PROBE = 1 is the same code under its own
// symbol for the schedule autotuner and the measurement hook, so a kernel trace of the solves is not mixed with
// whole-batch timing launches.
template <typename T, int R, int CI, int NT, int PROBE = 0>
__global__ __launch_bounds__(BLK) void kb_spmv(BDev B, const int* __restrict__ act, int na,
                                               const T* __restrict__ vals, const T* __restrict__ bvals,
                                               const T* __restrict__ p, T* __restrict__ q) {
    __shared__ double sh[WAVES];
    const int slot = xcd_slot(na);
    if (slot < 0) return;
    const int t = act[slot];
    const int c = B.tile_chart[t];
    if (B.c_done[c] != B_RUNNING) return;
    const int64_t lo = B.crow[c], hi = B.crow[c + 1] - 1;
    double acc = 0.0;
    const auto gp = [&](int64_t j) { return ld<T, R>(p, j); };
    const auto finish = [&](int64_t i, const T (&aq)[R]) {
        const Vr<T, R> pi = ld<T, R>(p, i);
        Vr<T, R> qv;
#pragma unroll
        for (int cc = 0; cc < R; ++cc) {
            qv.c[cc] = aq[cc];
            acc += (double)pi.c[cc] * (double)aq[cc];
        }
        st<T, R>(q, i, qv);
    };
    if (B.c_bw && B.c_bw[c] > 0) {
        const int t0 = B.chart_t0[c];
        band_rows<T, R, NT>(B, c, t - t0, B.chart_t0[c + 1] - t0, bvals, gp, finish);
        const double sum = block_sum_all(acc, sh);
        if (threadIdx.x == 0) B.part_a[t] = sum;
        return;
    }
    const int s1 = B.tile_s0[t + 1];
    for (int s = B.tile_s0[t]; s < s1; ++s) {
        if (s + 1 < s1 && spmv_group<2, T, R, NT>(B, s, lo, hi, vals, gp, finish)) {
            ++s;
            continue;
        }
        const int n = B.sl_n[s];
        if ((int)threadIdx.x >= n) continue;
        const int64_t i = B.sl_row0[s] + threadIdx.x;
        T aq[R];
#pragma unroll
        for (int cc = 0; cc < R; ++cc) aq[cc] = T(0);
        row_slots<T, R, CI, NT>(B, s, n, i, lo, hi, vals, gp, aq);
        finish(i, aq);
    }
    const double sum = block_sum_all(acc, sh);
    if (threadIdx.x == 0) B.part_a[t] = sum;
}

// alpha = rz / p.q (kept per chart: the next direction pass makes x += alpha p);  r -= alpha q;  z = D^-1 r
// (PC 1; z aliases r for PC 0; PC 2 leaves z and the r.z partials to the AMG cycle); partials r.r and r.z for step
// parity par ^ 1.
template <typename T, int R, int PC>
__global__ __launch_bounds__(BLK) void kb_update(BDev B, const int* __restrict__ act, int na, int par,
                                                 const T* __restrict__ dinv, const T* __restrict__ q,
                                                 T* __restrict__ r, T* __restrict__ z) {
    __shared__ double sh[2 * WAVES];
    const int slot = xcd_slot_rev(na);
    if (slot < 0) return;
    const int t = act[slot];
    const int c = B.tile_chart[t];
    if (B.c_done[c] != B_RUNNING) return;
    const int t0 = B.chart_t0[c];
    const double pq = B.big ? B.c_sa[c] : sum_range(B.part_a, t0, B.chart_t0[c + 1], sh);
    if (!(fabs(pq) > 1e-300) || !isfinite(pq)) {
        if (t == t0 && threadIdx.x == 0) B.c_done[c] = B_BREAKDOWN_PQ;
        return;
    }
    const T a = (T)(B.c_rz[(int64_t)par * B.nchart + c] / pq);
    if (t == t0 && threadIdx.x == 0) B.c_alpha[c] = (double)a;
    double arr = 0.0, arz = 0.0;
    tile_rows2(B.tile_r0[t], B.tile_r1[t], [&](int64_t i0, int64_t i2, bool two) {
        const int64_t ix[2] = {i0, i2};
        T d[2];
        Vr<T, R> qi[2], ri[2], zi[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            d[k] = (PC == 1) ? dinv[ix[k]] : T(1);
            qi[k] = ld<T, R>(q, ix[k]);
            ri[k] = ld<T, R>(r, ix[k]);
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const bool on = k == 0 || two;
#pragma unroll
            for (int cc = 0; cc < R; ++cc) {
                ri[k].c[cc] = fma(-a, qi[k].c[cc], ri[k].c[cc]);
                zi[k].c[cc] = d[k] * ri[k].c[cc];
                const double vrr = (double)ri[k].c[cc] * (double)ri[k].c[cc];
                const double vrz = (double)ri[k].c[cc] * (double)zi[k].c[cc];
                arr += on ? vrr : 0.0;
                arz += on ? vrz : 0.0;
            }
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int q = (k == 0 || two) ? k : 0;  // where two is false row i's results go to row i again
            st<T, R>(r, ix[k], ri[q]);
            if (PC == 1) st<T, R>(z, ix[k], zi[q]);
        }
    });
    // PC 2 (AMG): the cycle that follows writes z and the r.z partials
    double srr, srz;
    {
        double va = wave_sum(arr), vb = wave_sum(arz);
        __syncthreads();
        if ((threadIdx.x & 63) == 0) {
            sh[threadIdx.x >> 6] = va;
            sh[WAVES + (threadIdx.x >> 6)] = vb;
        }
        __syncthreads();
        srr = 0.0;
        srz = 0.0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            srr += sh[w];
            srz += sh[WAVES + w];
        }
    }
    if (threadIdx.x == 0) {
        B.part_b[(int64_t)(par ^ 1) * B.ntile + t] = srr;
        if (PC != 2) B.part_c[(int64_t)(par ^ 1) * B.ntile + t] = (PC == 1) ? srz : srr;
    }
}

// inner start on the live charts: x = 0, r = T(r64), z = D^-1 r; partials r.r / r.z for parity 0 (PC 2: z
// and r.z come from the AMG cycle).  The first step's old direction p (= p1, read by the step of parity 0): 0
// for a fresh start; with cont, the chart's last direction of the previous inner solve (the residual is
// replaced by the refined one and the iteration goes on, as in reliable-update CG), which ended in p0 or p1 by
// the parity of its last step.
template <typename T, int R, int PC>
__global__ __launch_bounds__(BLK) void kb_prep(BDev B, const int* __restrict__ act, int na,
                                               const double* __restrict__ r64, const T* __restrict__ dinv,
                                               T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                               const T* __restrict__ p0, T* __restrict__ p, int cont) {
    __shared__ double sh[WAVES];
    const int slot = xcd_slot(na);
    if (slot < 0) return;
    const int t = act[slot];
    const uint32_t last = B.c_iters[B.tile_chart[t]];  // steps of the previous inner solve (read before kb_init)
    // 0: zero direction; 1: keep p (the last step had parity 1); 2: copy p0 (parity 0)
    const int keep = (cont && last > 0) ? (((last - 1) & 1) ? 1 : 2) : 0;
    double arr = 0.0, arz = 0.0, agap = 0.0;
    const int64_t r1 = B.tile_r1[t];
    for (int64_t i = B.tile_r0[t] + threadIdx.x; i < r1; i += BLK) {
        const T d = (PC == 1) ? dinv[i] : T(1);
        Vr<T, R> xi, ri, zi;
        Vr<T, R> rold;  // the recursive residual the previous inner solve ended with
        if (cont) rold = ld<T, R>(r, i);
#pragma unroll
        for (int cc = 0; cc < R; ++cc) {
            ri.c[cc] = (T)r64[i * R + cc];
            if (cont) agap += ((double)ri.c[cc] - (double)rold.c[cc]) * ((double)ri.c[cc] - (double)rold.c[cc]);
            zi.c[cc] = d * ri.c[cc];
            xi.c[cc] = T(0);
            arr += (double)ri.c[cc] * (double)ri.c[cc];
            arz += (double)ri.c[cc] * (double)zi.c[cc];
        }
        st<T, R>(x, i, xi);
        st<T, R>(r, i, ri);
        if (PC == 1) st<T, R>(z, i, zi);
        if (keep == 0) st<T, R>(p, i, xi);  // xi is zero
        else if (keep == 2) st<T, R>(p, i, ld<T, R>(p0, i));
    }
    const double srr = block_sum_all(arr, sh);
    const double srz = block_sum_all(arz, sh);
    const double sgap = block_sum_all(agap, sh);
    if (threadIdx.x == 0) {
        B.part_b[t] = srr;
        if (PC != 2) B.part_c[t] = srz;
        B.part_o[t] = sgap;  // ||refined - recursive residual||^2 of the tile (part_o is free until the next outer test)
    }
}

// per chart (not final): inner tolerance, step state.  plain: tol * ||b|| on the recursive residual
// (solver.rs:465-468); mixed: inner_tol * ||r64|| but no tighter than a quarter of the outer target, and, when that
// reduction would bring the chart within twice its outer target (the round can end the chart), aim times the
// target (aim > 0): stopping the recursive residual right at the target left the refined one above it on most
// charts, which then paid a whole further round.  One workgroup per chart (the chart's tile partials summed in a
// fixed order).
static __global__ __launch_bounds__(BLK) void kb_init(BDev B, double tol, double inner_tol, int mixed, int cont,
                                                      double gap, double aim) {
    __shared__ double sh[WAVES];
    const int c = blockIdx.x;
    if (B.c_final[c]) {
        if (threadIdx.x == 0) B.c_done[c] = B_SKIP;
        return;
    }
    const int t0 = B.chart_t0[c], t1 = B.chart_t0[c + 1];
    const double rn = sqrt(sum_range(B.part_b, t0, t1, sh)), bn = B.c_bnorm[c];
    // the "old" r.z of step 0.  Fresh start: 1 (the direction is zero, only finiteness matters).  cont: that of
    // the previous inner solve's last step, whose direction kb_prep kept, when the refined residual differs from
    // the recursive one the solve stopped at by at most gap times the latter's norm; a larger difference (FP32
    // drift on a badly conditioned chart) makes the kept direction a poor one, so the step restarts instead:
    // infinity makes beta 0 and p' = z exactly.
    const uint32_t last = B.c_iters[c];
    double rz_old = 1.0;
    if (cont && last > 0) {
        const double g = sum_range(B.part_o, t0, t1, sh);
        rz_old = (sqrt(g) <= gap * B.c_rrec[c]) ? B.c_rz[(int64_t)((last - 1) & 1) * B.nchart + c] : INFINITY;
    }
    if (threadIdx.x != 0) return;
    double ta;
    const double tres = (bn <= 1e-300) ? tol : tol * bn;
    if (mixed) {
        const double tg = B.c_tgt[c] > 0.0 ? B.c_tgt[c] : tres;
        ta = fmax(inner_tol * rn, 0.25 * tg);
        if (aim > 0.0 && inner_tol * rn <= 2.0 * tg) ta = aim * tg;
    } else {
        ta = tres;
    }
    B.c_tol[c] = ta;
    B.c_iters[c] = 0;
    B.c_rz[B.nchart + c] = rz_old;
    B.c_done[c] = (rn <= ta) ? B_CONVERGED : B_RUNNING;
}

// start of a solve: x64 = 0, r64 = b, partials of b.b (every tile)
template <int R>
__global__ __launch_bounds__(BLK) void kb_start(BDev B, const double* __restrict__ b, double* __restrict__ x64,
                                                double* __restrict__ r64) {
    __shared__ double sh[WAVES];
    const int t = blockIdx.x;
    double acc = 0.0;
    const int64_t r1 = B.tile_r1[t] * R;
    for (int64_t k = B.tile_r0[t] * R + threadIdx.x; k < r1; k += BLK) {
        const double v = b[k];
        x64[k] = 0.0;
        r64[k] = v;
        acc += v * v;
    }
    const double s = block_sum_all(acc, sh);
    if (threadIdx.x == 0) B.part_o[t] = s;
}

// per chart: ||b||, not final
static __global__ __launch_bounds__(BLK) void kb_bnorm(BDev B) {  // one workgroup per chart
    __shared__ double sh[WAVES];
    const int c = blockIdx.x;
    const double s = sum_range(B.part_o, B.chart_t0[c], B.chart_t0[c + 1], sh);
    if (threadIdx.x != 0) return;
    B.c_bnorm[c] = sqrt(s);
    B.c_final[c] = 0;
    B.c_true[c] = 1.0;
    B.c_rin[c] = sqrt(s);
    B.c_gamma[c] = 0.0;
    B.c_tgt[c] = -1.0;  // before the first round: the residual test alone
}

// x64 += x (the inner correction) on the live charts
template <typename T, int R>
__global__ __launch_bounds__(BLK) void kb_axpy(BDev B, const int* __restrict__ act, int na,
                                               const T* __restrict__ xin, double* __restrict__ x64) {
    __shared__ double sh[WAVES];
    const int slot = xcd_slot(na);
    if (slot < 0) return;
    const int t = act[slot];
    if (B.c_final[B.tile_chart[t]]) return;
    double dmax = 0.0, xmax = 0.0;
    const int64_t r1 = B.tile_r1[t] * R;
    for (int64_t k = B.tile_r0[t] * R + threadIdx.x; k < r1; k += BLK) {
        const double d = (double)xin[k];
        const double v = x64[k] + d;
        x64[k] = v;
        dmax = fmax(dmax, fabs(d));
        xmax = fmax(xmax, fabs(v));
    }
    dmax = block_max_all(dmax, sh);
    xmax = block_max_all(xmax, sh);
    if (threadIdx.x == 0) {
        B.part_d[t] = dmax;
        B.part_d[B.ntile + t] = xmax;
    }
}

// r64 = b - A64 x64 in FP64 on the live charts, partials of r64.r64 (after kb_axpy: a tile's rows read x64 of
// rows other tiles update)
template <int R, int CI>
__global__ __launch_bounds__(BLK) void kb_residual(BDev B, const int* __restrict__ act, int na,
                                                   const double* __restrict__ vals64,
                                                   const double* __restrict__ bvals64, const double* __restrict__ b,
                                                   const double* __restrict__ x64, double* __restrict__ r64) {
    __shared__ double sh[WAVES];
    const int slot = xcd_slot(na);
    if (slot < 0) return;
    const int t = act[slot];
    const int c = B.tile_chart[t];
    if (B.c_final[c]) return;
    double acc = 0.0;
    if (B.c_bw && B.c_bw[c] > 0) {  // band chart: its FP64 band form
        const int t0 = B.chart_t0[c];
        band_rows<double, R, 0>(B, c, t - t0, B.chart_t0[c + 1] - t0, bvals64,
                                [&](int64_t j) { return ld<double, R>(x64, j); },
                                [&](int64_t i, const double (&sum)[R]) {
                                    const Vr<double, R> bi = ld<double, R>(b, i);
                                    Vr<double, R> ri;
#pragma unroll
                                    for (int cc = 0; cc < R; ++cc) {
                                        ri.c[cc] = bi.c[cc] - sum[cc];
                                        acc += ri.c[cc] * ri.c[cc];
                                    }
                                    st<double, R>(r64, i, ri);
                                });
        const double ss = block_sum_all(acc, sh);
        if (threadIdx.x == 0) B.part_o[t] = ss;
        return;
    }
    const int64_t lo = B.crow[c], hi = B.crow[c + 1] - 1;
    const int s1 = B.tile_s0[t + 1];
    for (int s = B.tile_s0[t]; s < s1; ++s) {
        const int n = B.sl_n[s];
        if ((int)threadIdx.x >= n) continue;
        const int64_t i = B.sl_row0[s] + threadIdx.x;
        double sum[R];
#pragma unroll
        for (int cc = 0; cc < R; ++cc) sum[cc] = 0.0;
        row_slots<double, R, CI, 0>(B, s, n, i, lo, hi, vals64, [&](int64_t j) { return ld<double, R>(x64, j); },
                                    sum);
        const Vr<double, R> bi = ld<double, R>(b, i);
        Vr<double, R> ri;
#pragma unroll
        for (int cc = 0; cc < R; ++cc) {
            ri.c[cc] = bi.c[cc] - sum[cc];
            acc += ri.c[cc] * ri.c[cc];
        }
        st<double, R>(r64, i, ri);
    }
    const double ss = block_sum_all(acc, sh);
    if (threadIdx.x == 0) B.part_o[t] = ss;
}

// per chart: true residual; final when ||b - A x|| <= tol ||b||
// one workgroup per chart
// The per-chart outer test after a refinement round.  The round's correction d solved A d = r_start, so
// gamma = max|d| / ||r_start|| is how far that residual's solution reaches per unit of residual; the error left,
// A^-1 r, is estimated as gamma ||r|| with the largest gamma of the rounds so far.  The chart is final when
// ||r|| <= tol ||b|| and that estimate is at most uv_tol max|x| (uv_tol <= 0: the residual test alone), i.e. when
// ||r|| <= c_tgt = min(tol ||b||, uv_tol max|x| / gamma), never asked below 1e-4 tol ||b||; a round that fails to
// halve a residual already below tol ||b|| also ends the chart (the FP64 residual has stagnated).
static __global__ __launch_bounds__(BLK) void kb_outer(BDev B, double tol, double uv_tol) {
    __shared__ double sh[WAVES];
    const int c = blockIdx.x;
    if (B.c_final[c]) return;
    const int t0 = B.chart_t0[c], t1 = B.chart_t0[c + 1];
    const double rn = sqrt(sum_range(B.part_o, t0, t1, sh)), bn = B.c_bnorm[c];
    double dmax = 0.0, xmax = 0.0;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += BLK) {
        dmax = fmax(dmax, B.part_d[t]);
        xmax = fmax(xmax, B.part_d[B.ntile + t]);
    }
    dmax = block_max_all(dmax, sh);
    xmax = block_max_all(xmax, sh);
    if (threadIdx.x != 0) return;
    B.c_true[c] = (bn > 0.0) ? rn / bn : rn;
    const double tres = (bn <= 1e-300) ? tol : tol * bn;
    const double rin = B.c_rin[c];
    double tgt = tres;
    if (uv_tol > 0.0) {
        const double g = fmax(B.c_gamma[c], rin > 0.0 ? dmax / rin : 0.0);
        B.c_gamma[c] = g;
        if (g > 0.0) tgt = fmax(fmin(tres, uv_tol * xmax / g), 1e-4 * tres);
    }
    B.c_tgt[c] = tgt;
    B.c_rin[c] = rn;
    if (rn <= tgt || (rn <= tres && rn > 0.5 * rin)) B.c_final[c] = 1;
}

// Per running chart (one workgroup each), the fixed-order sums of the tile partials pa (and pb) into sa[c] (sb[c]):
// for batches with charts of very many tiles, whose every workgroup would otherwise sum them all again.
static __global__ __launch_bounds__(BLK) void kb_chart_sums(BDev B, const double* __restrict__ pa,
                                                            const double* __restrict__ pb, double* __restrict__ sa,
                                                            double* __restrict__ sb) {
    __shared__ double sh[2 * WAVES];
    const int c = blockIdx.x;
    if (B.c_done[c] != B_RUNNING) return;
    const int t0 = B.chart_t0[c], t1 = B.chart_t0[c + 1];
    if (pb) {
        double a, b;
        sum_range2(pa, pb, t0, t1, sh, &a, &b);
        if (threadIdx.x == 0) {
            sa[c] = a;
            sb[c] = b;
        }
    } else {
        const double a = sum_range(pa, t0, t1, sh);
        if (threadIdx.x == 0) sa[c] = a;
    }
}

}  // namespace ruv
