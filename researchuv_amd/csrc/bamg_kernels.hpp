// bamg_kernels.hpp -- device-built aggregation AMG for the batched solver: setup (pairwise matching, Galerkin
// products) and the V(1,1) cycle.
//
// The hierarchy follows the reference's aggregation AMG (crates/researchuv-gpu/src/amg.rs:127-455): rows are
// matched with their strongest unmatched neighbour, two matchings make one level (aggregates of up to four
// rows, amg.rs SA_AGG = 4), the prolongator is the piecewise-constant aggregate map, the coarse operator is
// the Galerkin product P^T A P, and the cycle smooths with damped Jacobi (omega = 2/3).  What differs: the
// matching runs on the device by handshakes (a row proposes its strongest unmatched neighbour, mutual
// proposals pair up, three rounds), every aggregate stays inside one chart -- and on the finest level inside
// one workgroup tile -- so the hierarchy of a batch is block-diagonal by chart, and the coarsest level is one
// Jacobi step (exact once a chart has coarsened to a single row).  No inner products: the cycle is a fixed
// linear SPD operator, so each chart's PCG stays an independent PCG.
//
// Setup works on FP64 CSR levels (64-bit row pointers, 32-bit level-global columns).  Ties in the matching are
// broken by the lower row index, and every Galerkin row is summed in a fixed order, so the hierarchy is
// deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "batch_kernels.hpp"
#include "dev_common.hpp"

namespace ruv {

constexpr int AMG_MEMBERS = 4;  // fine rows per aggregate after two matchings
constexpr int SCAN_BLK = 1024;  // elements per workgroup of the scan

// ---------------------------------------------------------------- exclusive scan of int64 counts
__global__ __launch_bounds__(BLK) void k_scan_blocks(int64_t n, const int64_t* __restrict__ in,
                                                     int64_t* __restrict__ out, int64_t* __restrict__ totals) {
    __shared__ int64_t sh[SCAN_BLK];
    const int64_t base = (int64_t)blockIdx.x * SCAN_BLK;
    for (int k = threadIdx.x; k < SCAN_BLK; k += BLK) sh[k] = (base + k < n) ? in[base + k] : 0;
    __syncthreads();
    // Hillis-Steele inclusive scan over SCAN_BLK entries with BLK threads (each owns 4 consecutive)
    int64_t loc[SCAN_BLK / BLK];
    int64_t run = 0;
#pragma unroll
    for (int q = 0; q < SCAN_BLK / BLK; ++q) {
        run += sh[threadIdx.x * (SCAN_BLK / BLK) + q];
        loc[q] = run;
    }
    __shared__ int64_t part[BLK];
    part[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < BLK; off <<= 1) {
        const int64_t v = (threadIdx.x >= (unsigned)off) ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    const int64_t before = (threadIdx.x == 0) ? 0 : part[threadIdx.x - 1];
#pragma unroll
    for (int q = 0; q < SCAN_BLK / BLK; ++q) {
        const int64_t k = base + threadIdx.x * (SCAN_BLK / BLK) + q;
        const int64_t incl = before + loc[q];
        if (k < n) out[k] = incl - ((k < n) ? in[k] : 0);  // exclusive
    }
    if (threadIdx.x == BLK - 1) totals[blockIdx.x] = part[BLK - 1];
}

// one workgroup: exclusive scan of the block totals in place, grand total in totals[nb]
__global__ __launch_bounds__(BLK) void k_scan_totals(int64_t nb, int64_t* __restrict__ totals) {
    __shared__ int64_t part[BLK];
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += BLK) {
        const int64_t k = b0 + threadIdx.x;
        const int64_t v = (k < nb) ? totals[k] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < BLK; off <<= 1) {
            const int64_t u = (threadIdx.x >= (unsigned)off) ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += u;
            __syncthreads();
        }
        if (k < nb) totals[k] = carry + part[threadIdx.x] - v;
        carry += part[BLK - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[nb] = carry;
}

__global__ __launch_bounds__(BLK) void k_scan_add(int64_t n, int64_t* __restrict__ out,
                                                  const int64_t* __restrict__ totals) {
    const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (k < n) out[k] += totals[k / SCAN_BLK];
}

// ---------------------------------------------------------------- matching
struct SetupCsr {
    int64_t n;
    const int64_t* ptr;
    const int32_t* cols;
    const double* vals;
};

// strongest coupling of each row: max over off-diagonal a_ij < 0 of -a_ij
__global__ __launch_bounds__(BLK) void k_smax(SetupCsr A, double* __restrict__ smax) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= A.n) return;
    double m = 0.0;
    for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
        if (A.cols[k] != i && A.vals[k] < 0.0) m = fmax(m, -A.vals[k]);
    smax[i] = m;
}

// an unmatched row proposes its strongest unmatched neighbour of the same group (coupling >= theta * its
// strongest coupling); ties go to the lower index
__global__ __launch_bounds__(BLK) void k_propose(SetupCsr A, const double* __restrict__ smax,
                                                 const int32_t* __restrict__ group, const int64_t* __restrict__ mate,
                                                 int64_t* __restrict__ prop, double theta) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= A.n) return;
    int64_t best = -1;
    if (mate[i] < 0) {
        const double thr = theta * smax[i];
        const int32_t g = group[i];
        double bs = 0.0;
        for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
            const int64_t j = A.cols[k];
            if (j == i || !(A.vals[k] < 0.0) || mate[j] >= 0 || group[j] != g) continue;
            const double s = -A.vals[k];
            if (s < thr) continue;
            if (best < 0 || s > bs || (s == bs && j < best)) {
                bs = s;
                best = j;
            }
        }
    }
    prop[i] = best;
}

__global__ __launch_bounds__(BLK) void k_accept(int64_t n, const int64_t* __restrict__ prop,
                                                int64_t* __restrict__ mate) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n || mate[i] >= 0) return;
    const int64_t j = prop[i];
    if (j >= 0 && prop[j] == i) mate[i] = j;
}

// unmatched rows become singletons; leader flag = the lower row of its pair
__global__ __launch_bounds__(BLK) void k_leaders(int64_t n, int64_t* __restrict__ mate, int64_t* __restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    if (mate[i] < 0) mate[i] = i;
    flag[i] = (mate[i] >= i) ? 1 : 0;
}

// aggregate index of every row; leader list of every aggregate; group of every aggregate
__global__ __launch_bounds__(BLK) void k_aggregate(int64_t n, const int64_t* __restrict__ mate,
                                                   const int64_t* __restrict__ cid, const int32_t* __restrict__ group,
                                                   int32_t* __restrict__ agg, int64_t* __restrict__ lead_of,
                                                   int32_t* __restrict__ group_c) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const int64_t l = min(i, mate[i]);
    agg[i] = (int32_t)cid[l];
    if (i == l) {
        lead_of[cid[i]] = i;
        group_c[cid[i]] = group[i];
    }
}

// Galerkin row I = rows {l, mate[l]} with columns mapped through agg, duplicates summed.  The distinct coarse
// columns are emitted in ascending order by repeated selection, each summed over the entries in row order.
template <bool FILL>
__global__ __launch_bounds__(BLK) void k_galerkin(int64_t nc, SetupCsr A, const int64_t* __restrict__ lead_of,
                                                  const int64_t* __restrict__ mate, const int32_t* __restrict__ agg,
                                                  int64_t* __restrict__ cnt, const int64_t* __restrict__ cptr,
                                                  int32_t* __restrict__ ccols, double* __restrict__ cvals,
                                                  double* __restrict__ cdiag) {
    const int64_t I = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (I >= nc) return;
    const int64_t r0 = lead_of[I], r1 = mate[r0];
    const int nr = (r1 != r0) ? 2 : 1;
    int64_t last = -1;
    int64_t out = FILL ? cptr[I] : 0;
    int64_t count = 0;
    double diag = 0.0;
    for (;;) {
        // smallest coarse column > last
        int64_t key = INT64_MAX;
        for (int q = 0; q < nr; ++q) {
            const int64_t r = q ? r1 : r0;
            for (int64_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k) {
                const int64_t c = agg[A.cols[k]];
                if (c > last && c < key) key = c;
            }
        }
        if (key == INT64_MAX) break;
        if (FILL) {
            double v = 0.0;
            for (int q = 0; q < nr; ++q) {
                const int64_t r = q ? r1 : r0;
                for (int64_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k)
                    if (agg[A.cols[k]] == key) v += A.vals[k];
            }
            ccols[out] = (int32_t)key;
            cvals[out] = v;
            ++out;
            if (key == I) diag = v;
        }
        ++count;
        last = key;
    }
    if (!FILL) cnt[I] = count;
    else cdiag[I] = diag;
}

// composed map of two matchings: fine row -> level aggregate, and each aggregate's (up to 4) fine members
__global__ __launch_bounds__(BLK) void k_compose(int64_t nc2, const int64_t* __restrict__ lead2,
                                                 const int64_t* __restrict__ mate2, const int64_t* __restrict__ lead1,
                                                 const int64_t* __restrict__ mate1, int64_t* __restrict__ members) {
    const int64_t I = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (I >= nc2) return;
    int64_t m[AMG_MEMBERS] = {-1, -1, -1, -1};
    int k = 0;
    const int64_t j0 = lead2[I], j1 = mate2[j0];
    for (int q = 0; q < 2; ++q) {
        const int64_t j = q ? j1 : j0;
        if (q && j1 == j0) break;
        const int64_t i0 = lead1[j], i1 = mate1[i0];
        m[k++] = i0;
        if (i1 != i0) m[k++] = i1;
    }
#pragma unroll
    for (int q = 0; q < AMG_MEMBERS; ++q) members[I * AMG_MEMBERS + q] = m[q];
}

__global__ __launch_bounds__(BLK) void k_compose_agg(int64_t n, const int32_t* __restrict__ agg1,
                                                     const int32_t* __restrict__ agg2, int32_t* __restrict__ agg) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) agg[i] = agg2[agg1[i]];
}

// group (chart) of every row from the chart row offsets
__global__ __launch_bounds__(BLK) void k_chart_groups(int64_t n, const int64_t* __restrict__ crow, int nchart,
                                                      int32_t* __restrict__ group) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nchart;  // largest c with crow[c] <= i
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (crow[mid] <= i) lo = mid;
        else hi = mid;
    }
    group[i] = lo;
}

template <typename T>
__global__ __launch_bounds__(BLK) void k_cast_inv(int64_t n, const double* __restrict__ d, T* __restrict__ dinv) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) dinv[i] = (T)(1.0 / d[i]);
}

template <typename T>
__global__ __launch_bounds__(BLK) void k_cast(int64_t n, const double* __restrict__ a, T* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) b[i] = (T)a[i];
}

// ---------------------------------------------------------------- cycle
template <typename T>
struct CycLevel {
    int64_t n;
    const int64_t* ptr;      // level operator (levels >= 1)
    const int32_t* cols;
    const T* vals;
    const T* dinv;
    const int32_t* agg;      // row -> aggregate on the next level
    const int64_t* members;  // next level's aggregates -> AMG_MEMBERS rows of this level (-1 padded)
    T* b;                    // right-hand side on this level
    T* z;                    // correction on this level
};

// level >= 1, per aggregate I of the next level: b_next[I] = sum over members m of (b - A (w D^-1 b))_m
template <typename T, int R>
__global__ __launch_bounds__(BLK) void kc_down(CycLevel<T> L, int64_t nc, T w, T* __restrict__ bnext) {
    const int64_t I = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (I >= nc) return;
    T acc[R];
#pragma unroll
    for (int c = 0; c < R; ++c) acc[c] = T(0);
    for (int q = 0; q < AMG_MEMBERS; ++q) {
        const int64_t m = L.members[I * AMG_MEMBERS + q];
        if (m < 0) break;
        const Vr<T, R> bm = ld<T, R>(L.b, m);
        T s[R];
#pragma unroll
        for (int c = 0; c < R; ++c) s[c] = bm.c[c];
        for (int64_t k = L.ptr[m]; k < L.ptr[m + 1]; ++k) {
            const int64_t j = L.cols[k];
            const T v = L.vals[k] * (w * L.dinv[j]);
            const Vr<T, R> bj = ld<T, R>(L.b, j);
#pragma unroll
            for (int c = 0; c < R; ++c) s[c] = fma(-v, bj.c[c], s[c]);
        }
#pragma unroll
        for (int c = 0; c < R; ++c) acc[c] += s[c];
    }
    Vr<T, R> o;
#pragma unroll
    for (int c = 0; c < R; ++c) o.c[c] = acc[c];
    st<T, R>(bnext, I, o);
}

// level >= 1, per row: e1 = w D^-1 b + P z_next, then z = e1 + w D^-1 (b - A e1)
template <typename T, int R>
__global__ __launch_bounds__(BLK) void kc_up(CycLevel<T> L, T w, const T* __restrict__ znext) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= L.n) return;
    T s[R];
#pragma unroll
    for (int c = 0; c < R; ++c) s[c] = T(0);
    for (int64_t k = L.ptr[i]; k < L.ptr[i + 1]; ++k) {
        const int64_t j = L.cols[k];
        const T dj = w * L.dinv[j];
        const Vr<T, R> bj = ld<T, R>(L.b, j);
        const Vr<T, R> zj = ld<T, R>(znext, L.agg[j]);
        const T v = L.vals[k];
#pragma unroll
        for (int c = 0; c < R; ++c) s[c] = fma(v, fma(dj, bj.c[c], zj.c[c]), s[c]);
    }
    const T di = w * L.dinv[i];
    const Vr<T, R> bi = ld<T, R>(L.b, i);
    const Vr<T, R> zc = ld<T, R>(znext, L.agg[i]);
    Vr<T, R> o;
#pragma unroll
    for (int c = 0; c < R; ++c) {
        const T e1 = fma(di, bi.c[c], zc.c[c]);
        o.c[c] = fma(di, bi.c[c] - s[c], e1);
    }
    st<T, R>(L.z, i, o);
}

// coarsest level: z = D^-1 b (exact for a chart coarsened to one row)
template <typename T, int R>
__global__ __launch_bounds__(BLK) void kc_coarsest(CycLevel<T> L) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= L.n) return;
    const T d = L.dinv[i];
    const Vr<T, R> bi = ld<T, R>(L.b, i);
    Vr<T, R> o;
#pragma unroll
    for (int c = 0; c < R; ++c) o.c[c] = d * bi.c[c];
    st<T, R>(L.z, i, o);
}

// Finest level (the batch's sliced-ELL operator), per live tile.  b is the PCG residual r.
//   down: t = r - A (w D^-1 r) for the tile's rows into LDS, then every aggregate of the tile (aggregates of
//         the finest level never leave their tile) sums its members: b1[I] = sum t_m.
template <typename T, int R, int CI>
__global__ __launch_bounds__(BLK) void kc_down0(BDev B, const int* __restrict__ act, int na, const T* __restrict__ vals,
                                                const T* __restrict__ dinv, T w, const T* __restrict__ r,
                                                const int64_t* __restrict__ tile_c0, const int64_t* __restrict__ members,
                                                T* __restrict__ b1, int max_tile_rows) {
    extern __shared__ char smem[];
    T* tl = reinterpret_cast<T*>(smem);  // max_tile_rows * R
    const int slot = xcd_slot(na);
    if (slot < 0) return;
    const int t = act[slot];
    const int c = B.tile_chart[t];
    if (B.c_done[c] != B_RUNNING) return;
    const int64_t lo = B.crow[c], hi = B.crow[c + 1] - 1;
    const int64_t r0 = B.tile_r0[t];
    const int s1 = B.tile_s0[t + 1];
    for (int s = B.tile_s0[t]; s < s1; ++s) {
        const int n = B.sl_n[s];
        if ((int)threadIdx.x >= n) continue;
        const int64_t i = B.sl_row0[s] + threadIdx.x;
        const int64_t e0 = B.sl_off[s] + threadIdx.x;
        const int64_t c0 = B.sl_coff[s];
        const int wd = B.sl_w[s];
        const Vr<T, R> ri = ld<T, R>(r, i);
        T acc[R];
#pragma unroll
        for (int cc = 0; cc < R; ++cc) acc[cc] = ri.c[cc];
        for (int k = 0; k < wd; ++k) {
            const T v = vals[e0 + (int64_t)k * n];
            const int64_t j = (c0 < 0) ? min(max(i + B.sl_dtab[(int64_t)s * SLOTS + k], lo), hi)
                                       : col_of<CI>(B.ecol, c0 + threadIdx.x + (int64_t)k * n, i);
            const T dj = w * dinv[j];
            const Vr<T, R> rj = ld<T, R>(r, j);
#pragma unroll
            for (int cc = 0; cc < R; ++cc) acc[cc] = fma(-v * dj, rj.c[cc], acc[cc]);
        }
#pragma unroll
        for (int cc = 0; cc < R; ++cc) tl[(i - r0) * R + cc] = acc[cc];
    }
    __syncthreads();
    const int64_t a0 = tile_c0[t], a1 = tile_c0[t + 1];
    for (int64_t I = a0 + threadIdx.x; I < a1; I += BLK) {
        T acc[R];
#pragma unroll
        for (int cc = 0; cc < R; ++cc) acc[cc] = T(0);
        for (int q = 0; q < AMG_MEMBERS; ++q) {
            const int64_t m = members[I * AMG_MEMBERS + q];
            if (m < 0) break;
#pragma unroll
            for (int cc = 0; cc < R; ++cc) acc[cc] += tl[(m - r0) * R + cc];
        }
        Vr<T, R> o;
#pragma unroll
        for (int cc = 0; cc < R; ++cc) o.c[cc] = acc[cc];
        st<T, R>(b1, I, o);
    }
}

//   up: e1 = w D^-1 r + P z1; z = e1 + w D^-1 (r - A e1) for the tile's rows; partial r.z of the tile into
//       part_c[par] (the r.z the next step opens with).
template <typename T, int R, int CI>
__global__ __launch_bounds__(BLK) void kc_up0(BDev B, const int* __restrict__ act, int na, int par,
                                              const T* __restrict__ vals, const T* __restrict__ dinv, T w,
                                              const T* __restrict__ r, const int32_t* __restrict__ agg,
                                              const T* __restrict__ z1, T* __restrict__ z) {
    __shared__ double sh[WAVES];
    const int slot = xcd_slot(na);
    if (slot < 0) return;
    const int t = act[slot];
    const int c = B.tile_chart[t];
    if (B.c_done[c] != B_RUNNING) return;
    const int64_t lo = B.crow[c], hi = B.crow[c + 1] - 1;
    double arz = 0.0;
    const int s1 = B.tile_s0[t + 1];
    for (int s = B.tile_s0[t]; s < s1; ++s) {
        const int n = B.sl_n[s];
        if ((int)threadIdx.x >= n) continue;
        const int64_t i = B.sl_row0[s] + threadIdx.x;
        const int64_t e0 = B.sl_off[s] + threadIdx.x;
        const int64_t c0 = B.sl_coff[s];
        const int wd = B.sl_w[s];
        T acc[R];
#pragma unroll
        for (int cc = 0; cc < R; ++cc) acc[cc] = T(0);
        for (int k = 0; k < wd; ++k) {
            const T v = vals[e0 + (int64_t)k * n];
            const int64_t j = (c0 < 0) ? min(max(i + B.sl_dtab[(int64_t)s * SLOTS + k], lo), hi)
                                       : col_of<CI>(B.ecol, c0 + threadIdx.x + (int64_t)k * n, i);
            const T dj = w * dinv[j];
            const Vr<T, R> rj = ld<T, R>(r, j);
            const Vr<T, R> zj = ld<T, R>(z1, agg[j]);
#pragma unroll
            for (int cc = 0; cc < R; ++cc) acc[cc] = fma(v, fma(dj, rj.c[cc], zj.c[cc]), acc[cc]);
        }
        const T di = w * dinv[i];
        const Vr<T, R> ri = ld<T, R>(r, i);
        const Vr<T, R> zc = ld<T, R>(z1, agg[i]);
        Vr<T, R> o;
#pragma unroll
        for (int cc = 0; cc < R; ++cc) {
            const T e1 = fma(di, ri.c[cc], zc.c[cc]);
            o.c[cc] = fma(di, ri.c[cc] - acc[cc], e1);
            arz += (double)ri.c[cc] * (double)o.c[cc];
        }
        st<T, R>(z, i, o);
    }
    const double srz = block_sum_all(arz, sh);
    if (threadIdx.x == 0) B.part_c[(int64_t)par * B.ntile + t] = srz;
}

}  // namespace ruv
