// metrics.hip -- post-solve chart metrics on the device, one workgroup per chart:
//   chart_distortion (crates/researchuv-unwrap/src/metrics.rs:73-141, with svd2 :59-71 and winding_flips
//   :39-56) and rectangularize (:147-227), which the reference's unfold stage runs on every chart after its
//   solve (pipeline.rs:171-182).
// Every operation follows the reference's order with multiply-add contraction disabled in this file, so each
// triangle's terms, every min / max and the rectangularized coordinates are the reference's bit for bit; the
// per-chart means sum the triangles in a fixed order (tiles of 1024 triangles, each a tree) instead of the
// reference's left to right.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <vector>

#include "dev_common.hpp"
#include "ruv_internal.hpp"

#pragma clang fp contract(off)

namespace ruv {

namespace {

#define HIPCHK(call)                                                                                           \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess) return fail(RUV_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int LADDER = 72;  // metrics.rs:176-179

struct V3 {
    double x, y, z;
};
__device__ __forceinline__ V3 sub3(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 div3(V3 a, double s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }  // vec.rs:104-106
__device__ __forceinline__ V3 cross3(V3 a, V3 o) {                                                  // vec.rs:108-114
    return V3{a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x};
}
__device__ __forceinline__ double len3(V3 a) { return sqrt(dot3(a, a)); }
__device__ __forceinline__ V3 ld3(const double* p, int64_t i) { return V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// tri_signed (metrics.rs:31-34)
__device__ __forceinline__ double tri_signed(double uu, double uv, double vu, double vv, double wu, double wv) {
    return (vu - uu) * (wv - uv) - (wu - uu) * (vv - uv);
}

// block reductions in a fixed order
__device__ __forceinline__ double block_max(double v, double* sh) { return block_max_all(v, sh); }
__device__ __forceinline__ long long block_count(long long v, long long* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    long long s = 0;
#pragma unroll
    for (int i = 0; i < WAVES; ++i) s += sh[i];
    return s;
}

// Per tile of at most TILE_TRIS triangles of one chart (a big chart spans many tiles, so its triangles spread
// over the whole device): the tile's sums of conformal ratio, area ratio and signed area, its largest conformal
// ratio and its counts (measured triangles, flips, signed areas <= 0 and >= 0).
constexpr int TILE_TRIS = 1024;
struct TilePart {
    double conf_sum, conf_max, area_sum, signed_sum, signed_abs;
    long long nconf, flips, le0, ge0;
};

// chart_tile0[c] (c = 0..nchart) = the first tile of chart c, tiles of TILE_TRIS triangles from each chart's first
// triangle; one workgroup scans the charts in runs of BLK.  The kernel first clears the call's error flag; bit 4
// reports triangle offsets that do not start at 0 or decrease (such a chart gets no tiles).
__global__ __launch_bounds__(BLK) void k_tile_prefix(int32_t nchart, const int64_t* __restrict__ t_off,
                                                     int32_t* __restrict__ chart_tile0, int* __restrict__ err) {
    __shared__ int32_t sh[WAVES];
    if (threadIdx.x == 0) {
        *err = (t_off[0] != 0) ? 4 : 0;
        chart_tile0[0] = 0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int32_t base = 0;
    for (int32_t c0 = 0; c0 < nchart; c0 += BLK) {
        const int32_t c = c0 + (int32_t)threadIdx.x;
        int32_t n = 0;
        if (c < nchart) {
            const int64_t a = t_off[c], b = t_off[c + 1];
            if (b < a) atomicOr(err, 4);
            else n = (int32_t)((b - a + TILE_TRIS - 1) / TILE_TRIS);
        }
        int32_t inc = n;  // inclusive scan over the wave, then over the waves
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        if (lane == 63) sh[wv] = inc;
        __syncthreads();
        int32_t before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            before += (w < wv) ? sh[w] : 0;
            total += sh[w];
        }
        if (c < nchart) chart_tile0[c + 1] = base + before + inc;
        base += total;
        __syncthreads();
    }
}

// One workgroup per tile (grid: an upper bound on the tiles; the ones past chart_tile0[nchart] leave at once).
__global__ __launch_bounds__(BLK) void k_distortion_tiles(int32_t nchart, const int64_t* __restrict__ v_off,
                                                          const double* __restrict__ p3,
                                                          const double* __restrict__ uv,
                                                          const int64_t* __restrict__ t_off,
                                                          const int32_t* __restrict__ chart_tile0,
                                                          const uint32_t* __restrict__ tris,
                                                          TilePart* __restrict__ part, int* __restrict__ err) {
    __shared__ double sh[WAVES];
    __shared__ long long shc[WAVES];
    const int32_t tile = (int32_t)blockIdx.x;
    if (tile >= chart_tile0[nchart]) return;
    int32_t lo = 0, hi = nchart;  // the chart of the tile: the last c with chart_tile0[c] <= tile
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (chart_tile0[mid] <= tile) lo = mid;
        else hi = mid;
    }
    const int c = lo;
    const int64_t tt0 = t_off[c] + (int64_t)(tile - chart_tile0[c]) * TILE_TRIS;
    const int64_t tt1 = min(t_off[c + 1], tt0 + TILE_TRIS);
    const int64_t v0 = v_off[c], nvc = v_off[c + 1] - v0;
    const double* P = p3 + 3 * v0;
    const double* U = uv + 2 * v0;
    double conf_sum = 0.0, conf_max = 0.0, area_sum = 0.0, signed_sum = 0.0, signed_abs = 0.0;
    long long nconf = 0, flips = 0, le0 = 0, ge0 = 0;
    for (int64_t t = tt0 + threadIdx.x; t < tt1; t += BLK) {
        const uint32_t a = tris[3 * t], b = tris[3 * t + 1], cc = tris[3 * t + 2];
        if ((int64_t)a >= nvc || (int64_t)b >= nvc || (int64_t)cc >= nvc) {  // reported, never read
            atomicOr(err, 1);
            continue;
        }
        const double u0u = U[2 * a], u0v = U[2 * a + 1], u1u = U[2 * b], u1v = U[2 * b + 1];
        const double u2u = U[2 * cc], u2v = U[2 * cc + 1];
        // winding_flips' signed area, for every triangle
        const double sa = tri_signed(u0u, u0v, u1u, u1v, u2u, u2v);
        signed_sum += sa;
        signed_abs += fabs(sa);
        le0 += (sa <= 0.0) ? 1 : 0;
        ge0 += (sa >= 0.0) ? 1 : 0;
        const V3 a3 = ld3(P, a), b3 = ld3(P, b), c3 = ld3(P, cc);
        const V3 e1 = sub3(b3, a3), e2 = sub3(c3, a3);
        const V3 nrm = cross3(e1, e2);
        const double ln = len3(nrm);
        if (ln < 1e-30) continue;
        const V3 n = div3(nrm, ln);
        const V3 t1 = div3(e1, fmax(len3(e1), 1e-30));
        const V3 t2 = cross3(n, t1);
        const double e1t1 = dot3(e1, t1), e2t1 = dot3(e2, t1), e1t2 = dot3(e1, t2), e2t2 = dot3(e2, t2);
        const double det = e1t1 * e2t2 - e2t1 * e1t2;
        if (fabs(det) < 1e-30) {
            ++flips;
            continue;
        }
        const double ip00 = e2t2 / det, ip01 = -e2t1 / det, ip10 = -e1t2 / det, ip11 = e1t1 / det;
        const double d1u = u1u - u0u, d1v = u1v - u0v, d2u = u2u - u0u, d2v = u2v - u0v;
        const double f00 = d1u * ip00 + d2u * ip10, f01 = d1u * ip01 + d2u * ip11;
        const double f10 = d1v * ip00 + d2v * ip10, f11 = d1v * ip01 + d2v * ip11;
        // svd2 (metrics.rs:59-71)
        const double ga = f00 * f00 + f10 * f10;
        const double gb = f00 * f01 + f10 * f11;
        const double gc = f01 * f01 + f11 * f11;
        const double tr = ga + gc;
        const double d = 0.5 * (ga - gc);
        const double disc = sqrt(fmax(d * d + gb * gb, 0.0));
        const double l1 = 0.5 * tr + disc;
        const double l2 = fmax(0.5 * tr - disc, 0.0);
        const double s1 = sqrt(fmax(l1, 0.0)), s2 = sqrt(l2);
        const double conf = s1 / fmax(s2, 1e-30);
        conf_sum += conf;
        conf_max = fmax(conf_max, conf);
        ++nconf;
        const double uv_area = 0.5 * fabs(d1u * d2v - d2u * d1v);
        area_sum += uv_area / (0.5 * ln);
    }
    TilePart r;
    r.conf_sum = block_sum_all(conf_sum, sh);
    r.conf_max = block_max(conf_max, sh);
    r.area_sum = block_sum_all(area_sum, sh);
    r.signed_sum = block_sum_all(signed_sum, sh);
    r.signed_abs = block_sum_all(signed_abs, sh);
    r.nconf = block_count(nconf, shc);
    r.flips = block_count(flips, shc);
    r.le0 = block_count(le0, shc);
    r.ge0 = block_count(ge0, shc);
    if (threadIdx.x == 0) part[tile] = r;
}

// per chart: its tiles' parts summed in tile order.  metrics (3 per chart): conformal mean, conformal max,
// area-ratio mean; counts (2 per chart): flips, folds.  Tiles at or past ntile_cap (the launched tiles, which
// offsets that decrease can leave short of the table; the call then reports an error) are not read.
// The fold count needs the sign of the signed-area sum as winding_flips forms it, left to right from +0.0
// (metrics.rs:39-56).  The tiles' sum has another rounding, but any order of summing n terms lands within
// (n - 1) u sum|s| of the exact sum (u = 2^-53), so when the tiles' sum is farther than twice that from 0 both
// sums have its sign; otherwise (terms that nearly cancel, as on a chart folded onto itself) the chart's thread
// sums its triangles again in the reference's order.
__global__ __launch_bounds__(BLK) void k_distortion_finish(int32_t nchart, const int32_t* __restrict__ chart_tile0,
                                                           int32_t ntile_cap, const TilePart* __restrict__ part,
                                                           const int64_t* __restrict__ v_off,
                                                           const double* __restrict__ uv,
                                                           const int64_t* __restrict__ t_off,
                                                           const uint32_t* __restrict__ tris,
                                                           double* __restrict__ metrics,
                                                           long long* __restrict__ counts) {
    const int c = blockIdx.x * BLK + threadIdx.x;
    if (c >= nchart) return;
    double cs = 0.0, cm = 0.0, as = 0.0, ss = 0.0, sa = 0.0;
    long long nc = 0, fl = 0, nle = 0, nge = 0;
    const int t1 = min(chart_tile0[c + 1], ntile_cap);
    for (int t = chart_tile0[c]; t < t1; ++t) {
        const TilePart r = part[t];
        cs += r.conf_sum;
        cm = fmax(cm, r.conf_max);
        as += r.area_sum;
        ss += r.signed_sum;
        sa += r.signed_abs;
        nc += r.nconf;
        fl += r.flips;
        nle += r.le0;
        nge += r.ge0;
    }
    const int64_t ntri = t_off[c + 1] - t_off[c];
    if (ntri > 0 && !(fabs(ss) > 2.0 * (double)ntri * 0x1p-53 * sa)) {
        const int64_t v0 = v_off[c], nvc = v_off[c + 1] - v0;
        const double* U = uv + 2 * v0;
        double seq = 0.0;
        for (int64_t t = t_off[c]; t < t_off[c + 1]; ++t) {
            const uint32_t a = tris[3 * t], b = tris[3 * t + 1], d = tris[3 * t + 2];
            if ((int64_t)a >= nvc || (int64_t)b >= nvc || (int64_t)d >= nvc) continue;  // reported by the tiles
            seq += tri_signed(U[2 * a], U[2 * a + 1], U[2 * b], U[2 * b + 1], U[2 * d], U[2 * d + 1]);
        }
        ss = seq;
    }
    const double den = (double)(nc > 0 ? nc : 1);
    metrics[3 * c] = cs / den;
    metrics[3 * c + 1] = cm;
    metrics[3 * c + 2] = as / den;
    counts[2 * c] = fl;
    // dominant winding = sign of the signed-area sum; folds disagree with it (zero areas count)
    counts[2 * c + 1] = (chart_tile0[c + 1] == chart_tile0[c]) ? 0 : ((ss < 0.0) ? nge : nle);
}

// rectangularize (metrics.rs:147-227), chart c: q (2 per vertex) and ext (2 per chart); ladder[2k], [2k+1]
// the cosine and sine of pi k / 71 as the host's libm gives them
// N maxima over the workgroup at once (one pair of barriers); every thread receives them in v
template <int N>
__device__ __forceinline__ void block_max_vec(double (&v)[N], double* sh /* [WAVES * N] */) {
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[k] = fmax(v[k], __shfl_xor(v[k], o, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) sh[(threadIdx.x >> 6) * N + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double m = sh[k];
#pragma unroll
        for (int w = 1; w < WAVES; ++w) m = fmax(m, sh[w * N + k]);
        v[k] = m;
    }
}

constexpr int LADDER_GROUP = 8;  // divides LADDER
static_assert(LADDER % LADDER_GROUP == 0, "ladder groups");

// rectangularize in six passes so a big chart spreads over the device:
//   k_rect_scores  one wave per (chart, 8 ladder angles): the border's bounding-box score of each angle
//   k_rect_pick    per chart: the chosen rotation (or none)
//   k_rect_box     per vertex tile: the rotated coordinates' bounding box
//   k_rect_finish  per chart: the box of its tiles -> minimum and extent (ext)
//   k_rect_write   per vertex tile: q = (rotated - minimum) / max(extent, 1e-30)
//   k_rect_snap    per border tile: border points onto the nearest side
// Min / max are exact in any order, so every pass gives the reference's values bit for bit.
struct RectChart {
    double c, s;     // rotation (cos, sin); rot = 0: none
    double mnu, mnv, eu, ev;
    int rot, pad;
};

// the ladder's scores: one wave per (chart c, group g) takes angles g*LADDER_GROUP.. over c's whole border
constexpr int LGROUPS = LADDER / LADDER_GROUP;
__global__ __launch_bounds__(BLK) void k_rect_scores(int32_t nchart, const int64_t* __restrict__ v_off,
                                                     const double* __restrict__ uv, const int64_t* __restrict__ b_off,
                                                     const int64_t* __restrict__ bnd,
                                                     const double* __restrict__ ladder, double* __restrict__ score,
                                                     int* __restrict__ bad_chart) {
    const int64_t pair = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (pair >= (int64_t)nchart * LGROUPS) return;  // whole waves; no workgroup barrier below
    const int c = (int)(pair / LGROUPS), k0 = (int)(pair % LGROUPS) * LADDER_GROUP;
    const int64_t v0 = v_off[c], nv = v_off[c + 1] - v0;
    const int64_t b0 = b_off[c], nb = b_off[c + 1] - b0;
    const double* U = uv + 2 * v0;
    double cs[LADDER_GROUP], sn[LADDER_GROUP];
#pragma unroll
    for (int g = 0; g < LADDER_GROUP; ++g) {
        cs[g] = ladder[2 * (k0 + g)];
        sn[g] = ladder[2 * (k0 + g) + 1];
    }
    double m[4 * LADDER_GROUP];  // per angle: -min u, max u, -min v, max v
#pragma unroll
    for (int g = 0; g < 4 * LADDER_GROUP; ++g) m[g] = -INFINITY;
    int bad = 0;
    for (int64_t j = lane; j < nb; j += 64) {
        const int64_t bi = bnd[b0 + j];
        if (bi < 0 || bi >= nv) {  // reported, never read
            bad = 1;
            continue;
        }
        const double pu = U[2 * bi], pv = U[2 * bi + 1];
#pragma unroll
        for (int g = 0; g < LADDER_GROUP; ++g) {
            const double qu = cs[g] * pu + sn[g] * pv;
            const double qv = -sn[g] * pu + cs[g] * pv;
            m[4 * g] = fmax(m[4 * g], -qu);
            m[4 * g + 1] = fmax(m[4 * g + 1], qu);
            m[4 * g + 2] = fmax(m[4 * g + 2], -qv);
            m[4 * g + 3] = fmax(m[4 * g + 3], qv);
        }
    }
    if (bad) bad_chart[c] = 1;
#pragma unroll
    for (int g = 0; g < 4 * LADDER_GROUP; ++g)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m[g] = fmax(m[g], __shfl_xor(m[g], o, 64));
    if (lane < LADDER_GROUP) {
        double e0 = 0.0, e1 = 0.0;
#pragma unroll
        for (int g = 0; g < LADDER_GROUP; ++g)
            if (g == lane) {
                e0 = m[4 * g + 1] - (-m[4 * g]);
                e1 = m[4 * g + 3] - (-m[4 * g + 2]);
            }
        score[(int64_t)c * LADDER + k0 + lane] = sqrt(e0 * (e0 + 1e-30) + e1 * (e1 + 1e-30));
    }
}

// per chart: the first angle of smallest score, a later one replacing it only when strictly smaller
__global__ __launch_bounds__(BLK) void k_rect_pick(int32_t nchart, const int64_t* __restrict__ b_off,
                                                   const double* __restrict__ ladder,
                                                   const double* __restrict__ score,
                                                   const int* __restrict__ bad_chart, RectChart* __restrict__ rc,
                                                   int* __restrict__ err) {
    const int c = blockIdx.x * BLK + threadIdx.x;
    if (c >= nchart) return;
    const bool border = b_off[c + 1] > b_off[c];
    const int bad = border ? bad_chart[c] : 0;
    if (bad) atomicOr(err, 2);
    RectChart r{};
    r.c = 1.0;
    r.s = 0.0;
    if (border && !bad) {
        double bscore = INFINITY;
        bool have = false;
        for (int k = 0; k < LADDER; ++k) {
            const double sc = score[(int64_t)c * LADDER + k];
            if (!(have && sc >= bscore)) {
                have = true;
                r.c = ladder[2 * k];
                r.s = ladder[2 * k + 1];
                bscore = sc;
            }
        }
    }
    r.rot = (border && !bad) ? 1 : 0;
    r.pad = bad;
    rc[c] = r;
}

// tiles: tile_chart, [tile_i0, tile_i1) of the chart's local vertex (or border-entry) indices
__device__ __forceinline__ void rect_rotated(const RectChart& r, double pu, double pv, double* qu, double* qv) {
    if (r.rot) {
        *qu = r.c * pu + r.s * pv;
        *qv = -r.s * pu + r.c * pv;
    } else {  // without border the chart is only normalised (metrics.rs:152-172)
        *qu = pu;
        *qv = pv;
    }
}

struct Box {
    double nu, xu, nv, xv;  // -min u, max u, -min v, max v
};

__global__ __launch_bounds__(BLK) void k_rect_box(const int32_t* __restrict__ tile_chart,
                                                  const int64_t* __restrict__ tile_i0,
                                                  const int64_t* __restrict__ tile_i1,
                                                  const int64_t* __restrict__ v_off, const double* __restrict__ uv,
                                                  const RectChart* __restrict__ rc, Box* __restrict__ box) {
    __shared__ double shv[WAVES * 4];
    const int tile = blockIdx.x;
    const int c = tile_chart[tile];
    const RectChart r = rc[c];
    const double* U = uv + 2 * v_off[c];
    double m[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
    if (!r.pad)
        for (int64_t i = tile_i0[tile] + threadIdx.x; i < tile_i1[tile]; i += BLK) {
            double qu, qv;
            rect_rotated(r, U[2 * i], U[2 * i + 1], &qu, &qv);
            m[0] = fmax(m[0], -qu);
            m[1] = fmax(m[1], qu);
            m[2] = fmax(m[2], -qv);
            m[3] = fmax(m[3], qv);
        }
    block_max_vec<4>(m, shv);
    if (threadIdx.x == 0) box[tile] = Box{m[0], m[1], m[2], m[3]};
}

__global__ __launch_bounds__(BLK) void k_rect_finish(int32_t nchart, const int32_t* __restrict__ chart_tile0,
                                                     const Box* __restrict__ box, RectChart* __restrict__ rc,
                                                     double* __restrict__ ext) {
    const int c = blockIdx.x * BLK + threadIdx.x;
    if (c >= nchart) return;
    double nu = -INFINITY, xu = -INFINITY, nv = -INFINITY, xv = -INFINITY;
    for (int t = chart_tile0[c]; t < chart_tile0[c + 1]; ++t) {
        nu = fmax(nu, box[t].nu);
        xu = fmax(xu, box[t].xu);
        nv = fmax(nv, box[t].nv);
        xv = fmax(xv, box[t].xv);
    }
    RectChart r = rc[c];
    r.mnu = -nu;
    r.mnv = -nv;
    r.eu = xu - r.mnu;
    r.ev = xv - r.mnv;
    if (chart_tile0[c + 1] == chart_tile0[c] || r.pad) r.eu = r.ev = 0.0;  // empty chart (or a reported error)
    rc[c] = r;
    ext[2 * c] = r.eu;
    ext[2 * c + 1] = r.ev;
}

__global__ __launch_bounds__(BLK) void k_rect_write(const int32_t* __restrict__ tile_chart,
                                                    const int64_t* __restrict__ tile_i0,
                                                    const int64_t* __restrict__ tile_i1,
                                                    const int64_t* __restrict__ v_off, const double* __restrict__ uv,
                                                    const RectChart* __restrict__ rc, double* __restrict__ q) {
    const int tile = blockIdx.x;
    const int c = tile_chart[tile];
    const RectChart r = rc[c];
    if (r.pad) return;
    const double* U = uv + 2 * v_off[c];
    double* Q = q + 2 * v_off[c];
    for (int64_t i = tile_i0[tile] + threadIdx.x; i < tile_i1[tile]; i += BLK) {
        double qu, qv;
        rect_rotated(r, U[2 * i], U[2 * i + 1], &qu, &qv);
        Q[2 * i] = (qu - r.mnu) / fmax(r.eu, 1e-30);
        Q[2 * i + 1] = (qv - r.mnv) / fmax(r.ev, 1e-30);
    }
}

__global__ __launch_bounds__(BLK) void k_rect_snap(const int32_t* __restrict__ tile_chart,
                                                   const int64_t* __restrict__ tile_i0,
                                                   const int64_t* __restrict__ tile_i1,
                                                   const int64_t* __restrict__ v_off,
                                                   const int64_t* __restrict__ b_off,
                                                   const int64_t* __restrict__ bnd,
                                                   const RectChart* __restrict__ rc, double* __restrict__ q) {
    const int tile = blockIdx.x;
    const int c = tile_chart[tile];
    if (!rc[c].rot) return;
    double* Q = q + 2 * v_off[c];
    const int64_t* B = bnd + b_off[c];
    for (int64_t j = tile_i0[tile] + threadIdx.x; j < tile_i1[tile]; j += BLK) {
        const int64_t bi = B[j];
        const double pu = Q[2 * bi], pv = Q[2 * bi + 1];
        const double dx = fmin(pu, 1.0 - pu), dy = fmin(pv, 1.0 - pv);
        if (dx <= dy) Q[2 * bi] = (pu < 0.5) ? 0.0 : 1.0;
        else Q[2 * bi + 1] = (pv < 0.5) ? 0.0 : 1.0;
    }
}

// device copies of host arrays, freed together
struct DevSet {
    std::vector<void*> mem;
    ~DevSet() {
        for (void* p : mem) (void)hipFree(p);
    }
    template <typename E>
    hipError_t put(E** d, const E* h, size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(E));
        if (e != hipSuccess) return e;
        mem.push_back(q);
        *d = (E*)q;
        return h ? hipMemcpy(q, h, n * sizeof(E), hipMemcpyHostToDevice) : hipSuccess;
    }
};

int check_charts(int32_t nchart, const int64_t* v_off, const int64_t* x_off, int64_t* nv, int64_t* nx) {
    if (nchart < 0) return fail(RUV_ERR_INVALID, "metrics: negative chart count");
    if (v_off[0] != 0 || x_off[0] != 0) return fail(RUV_ERR_BAD_DIMS, "metrics: offsets must start at 0");
    for (int32_t c = 0; c < nchart; ++c)
        if (v_off[c + 1] < v_off[c] || x_off[c + 1] < x_off[c])
            return fail(RUV_ERR_BAD_DIMS, "metrics: offsets not monotone");
    *nv = v_off[nchart];
    *nx = x_off[nchart];
    return RUV_OK;
}

}  // namespace

namespace {

// the ladder's angles, pi k / 71 (metrics.rs:176-179), through the host's libm like the reference
void ladder_table(double* lad) {
    for (int k = 0; k < LADDER; ++k) {
        const double ang = 3.14159265358979323846 * (double)k / (double)(LADDER - 1);
        lad[2 * k] = std::cos(ang);
        lad[2 * k + 1] = std::sin(ang);
    }
}

// The ladder table of each device: uploaded on the first call on that device, read-only afterwards.
int ladder_for_device(int dev, const double** out) {
    static std::mutex mu;
    static std::vector<double*> tab;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)tab.size() <= dev) tab.resize((size_t)dev + 1, nullptr);
    if (!tab[(size_t)dev]) {
        double lad[2 * LADDER];
        ladder_table(lad);
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, sizeof(lad)));
        HIPCHK(hipMemcpy(p, lad, sizeof(lad), hipMemcpyHostToDevice));
        tab[(size_t)dev] = (double*)p;
    }
    *out = tab[(size_t)dev];
    return RUV_OK;
}

// The resources of one call of a device-pointer entry point: a non-blocking stream of default priority, an event,
// pinned host words for the results read back, and a scratch block grown on demand.  A call leases a context of
// the current device for its duration; concurrent calls lease different ones, so they neither share a stream nor
// a buffer nor wait on each other.  Released contexts are kept (per device) for later calls until the process
// ends, so a call allocates nothing once its device has as many contexts as calls have overlapped.
struct CallCtx {
    int dev = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr;
    int64_t* h_word = nullptr;  // pinned: [0] the error flag, [1] a count read back
    void* buf = nullptr;
    size_t cap = 0;
};

struct CtxPool {
    std::mutex mu;
    std::vector<std::vector<CallCtx*>> free_by_dev;
};
CtxPool& ctx_pool() {
    static CtxPool* p = new CtxPool();  // never destroyed: contexts live until the process ends
    return *p;
}

struct CtxLease {
    CallCtx* c = nullptr;
    ~CtxLease() {
        if (!c) return;
        CtxPool& P = ctx_pool();
        std::lock_guard<std::mutex> lock(P.mu);
        P.free_by_dev[(size_t)c->dev].push_back(c);
    }
};

int lease_ctx(CtxLease* L) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    CtxPool& P = ctx_pool();
    {
        std::lock_guard<std::mutex> lock(P.mu);
        if ((int)P.free_by_dev.size() <= dev) P.free_by_dev.resize((size_t)dev + 1);
        auto& fl = P.free_by_dev[(size_t)dev];
        if (!fl.empty()) {
            L->c = fl.back();
            fl.pop_back();
            return RUV_OK;
        }
    }
    std::unique_ptr<CallCtx> c(new CallCtx());
    c->dev = dev;
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ready, hipEventDisableTiming));
    HIPCHK(hipHostMalloc((void**)&c->h_word, 2 * sizeof(int64_t), hipHostMallocDefault));
    L->c = c.release();
    return RUV_OK;
}

// the context's scratch block, at least `need` bytes (the stream has finished with the old block: every call
// ends with a wait on its stream)
int ctx_scratch(CallCtx* c, size_t need) {
    if (c->cap >= need) return RUV_OK;
    if (c->buf) HIPCHK(hipFree(c->buf));
    c->buf = nullptr;
    c->cap = 0;
    HIPCHK(hipMalloc(&c->buf, need));
    c->cap = need;
    return RUV_OK;
}

// The call's stream waits for the work issued so far on the null stream (which itself follows the blocking
// streams): inputs a caller wrote there, or on any blocking stream, are complete before the call reads them.
int ctx_order_after_caller(CallCtx* c) {
    HIPCHK(hipEventRecord(c->ready, nullptr));
    HIPCHK(hipStreamWaitEvent(c->stream, c->ready, 0));
    return RUV_OK;
}

// the error flag at `derr`, read back on `s` after the call's kernels (host word `h`, pinned or not)
int finish_call(hipStream_t s, const int* derr, int64_t* h, const char* what) {
    *h = 0;
    HIPCHK(hipMemcpyAsync(h, derr, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if ((int)*h & 4) return fail(RUV_ERR_BAD_DIMS, std::string(what) + ": chart offsets not monotone from 0");
    if (*h) return fail(RUV_ERR_BAD_DIMS, std::string(what) + ": an index lies outside its chart");
    return RUV_OK;
}

inline size_t al256(size_t b) { return (std::max<size_t>(b, 1) + 255) & ~(size_t)255; }

// Scratch of chart_distortion for nchart charts with `ntri` triangles in all: [error flag | chart_tile0 | one
// TilePart per tile], tiles bounded by nchart + ntri / TILE_TRIS.
struct DistortionScratch {
    int64_t tiles_bound = 0;
    size_t off_ct = 0, off_part = 0, bytes = 0;
};
DistortionScratch distortion_scratch(int32_t nchart, int64_t ntri) {
    DistortionScratch d;
    d.tiles_bound = (int64_t)nchart + (std::max<int64_t>(ntri, 0) + TILE_TRIS - 1) / TILE_TRIS;
    d.off_ct = 256;
    d.off_part = d.off_ct + al256(sizeof(int32_t) * ((size_t)nchart + 1));
    d.bytes = d.off_part + al256(sizeof(TilePart) * (size_t)std::max<int64_t>(d.tiles_bound, 1));
    return d;
}

// The chart_distortion launches on stream s with scratch at `base` (laid out by distortion_scratch): tile table
// on the device, one workgroup per tile, per-chart finish.  Shared by the leased-context and static-scratch entry
// points.
int distortion_launch(hipStream_t s, char* base, const DistortionScratch& d, int32_t nchart, const int64_t* v_off,
                      const double* p3, const double* uv, const int64_t* t_off, const uint32_t* tris,
                      double* metrics, int64_t* counts, int64_t* h_word) {
    int* d_err = (int*)base;
    int32_t* d_ct = (int32_t*)(base + d.off_ct);
    TilePart* d_p = (TilePart*)(base + d.off_part);
    hipLaunchKernelGGL(k_tile_prefix, dim3(1), dim3(BLK), 0, s, nchart, t_off, d_ct, d_err);
    if (d.tiles_bound > 0)
        hipLaunchKernelGGL(k_distortion_tiles, dim3((unsigned)d.tiles_bound), dim3(BLK), 0, s, nchart, v_off, p3, uv,
                           t_off, (const int32_t*)d_ct, tris, d_p, d_err);
    hipLaunchKernelGGL(k_distortion_finish, dim3((unsigned)((nchart + BLK - 1) / BLK)), dim3(BLK), 0, s, nchart,
                       (const int32_t*)d_ct, (int32_t)d.tiles_bound, (const TilePart*)d_p, v_off, uv, t_off, tris,
                       metrics, (long long*)counts);
    HIPCHK(hipGetLastError());
    return finish_call(s, d_err, h_word, "metrics");
}

// tiles of at most `size` items of each chart from its offsets (host copy): chart_tile0 (nchart + 1), per tile
// its chart and local range
void make_tiles(int32_t nchart, const std::vector<int64_t>& off, int64_t size, std::vector<int32_t>* chart_tile0,
                std::vector<int32_t>* tile_chart, std::vector<int64_t>* i0, std::vector<int64_t>* i1) {
    chart_tile0->assign((size_t)nchart + 1, 0);
    for (int32_t c = 0; c < nchart; ++c) {
        for (int64_t t = 0; t < off[c + 1] - off[c]; t += size) {
            tile_chart->push_back(c);
            i0->push_back(t);
            i1->push_back(std::min<int64_t>(off[c + 1] - off[c], t + size));
        }
        (*chart_tile0)[(size_t)c + 1] = (int32_t)tile_chart->size();
    }
}

}  // namespace

int chart_metrics_dev(int32_t nchart, const int64_t* v_off, const double* p3, const double* uv, const int64_t* t_off,
                      const uint32_t* tris, double* metrics, int64_t* counts) {
    if (nchart < 0) return fail(RUV_ERR_INVALID, "metrics: negative chart count");
    if (nchart == 0) return RUV_OK;
    if (!v_off || !t_off || !metrics || !counts) return fail(RUV_ERR_INVALID, "metrics: null argument");
    CtxLease L;
    int rc = lease_ctx(&L);
    if (rc) return rc;
    CallCtx* C = L.c;
    rc = ctx_order_after_caller(C);
    if (rc) return rc;
    // the triangle count bounds the tiles (one word read back; the tile table itself is built on the device)
    HIPCHK(hipMemcpyAsync(&C->h_word[1], t_off + nchart, sizeof(int64_t), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    const DistortionScratch d = distortion_scratch(nchart, C->h_word[1]);
    if (d.tiles_bound > (int64_t)INT32_MAX) return fail(RUV_ERR_BAD_DIMS, "metrics: too many triangles");
    rc = ctx_scratch(C, d.bytes);
    if (rc) return rc;
    return distortion_launch(C->stream, (char*)C->buf, d, nchart, v_off, p3, uv, t_off, tris, metrics, counts,
                             &C->h_word[0]);
}

namespace {

// The fixed scratch of the allocation-free chart_distortion: the error flag, the tile table and one TilePart per
// tile (distortion_scratch's layout).  Each device holds its own copy until the process ends.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// the device address of `scratch` on the current device, looked up once per device (hipGetSymbolAddress with the
// device current) and kept: host code never passes the array's name itself to a kernel or a copy
int noalloc_scratch(char** out) {
    static std::mutex mu;
    static std::vector<char*> addr;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)addr.size() <= dev) addr.resize((size_t)dev + 1, nullptr);
    if (!addr[(size_t)dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(RUV_ERR_DEVICE, std::string("metrics: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        addr[(size_t)dev] = (char*)ptr;
    }
    *out = addr[(size_t)dev];
    return RUV_OK;
}

}  // namespace

int chart_metrics_dev_noalloc(int32_t nchart, const int64_t* v_off, const double* p3, const double* uv,
                              const int64_t* t_off, const uint32_t* tris, double* metrics, int64_t* counts) {
    if (nchart < 0) return fail(RUV_ERR_INVALID, "metrics: negative chart count");
    if (nchart == 0) return RUV_OK;
    if (!v_off || !t_off || !metrics || !counts) return fail(RUV_ERR_INVALID, "metrics: null argument");
    char* base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    // the null stream: ordered after the caller's null-stream (and blocking-stream) work; nothing is created
    int64_t ntri = 0;
    HIPCHK(hipMemcpyAsync(&ntri, t_off + nchart, sizeof(int64_t), hipMemcpyDeviceToHost, nullptr));
    HIPCHK(hipStreamSynchronize(nullptr));
    const DistortionScratch d = distortion_scratch(nchart, ntri);
    if (d.bytes > NOALLOC_SCRATCH_BYTES)
        return fail(RUV_ERR_BAD_DIMS, "metrics (no-alloc): " + std::to_string(nchart) + " charts of " +
                                          std::to_string(ntri) + " triangles need " + std::to_string(d.bytes) +
                                          " scratch bytes, above the fixed " + std::to_string(NOALLOC_SCRATCH_BYTES));
    int64_t herr = 0;
    return distortion_launch(nullptr, base, d, nchart, v_off, p3, uv, t_off, tris, metrics, counts, &herr);
}

int rectangularize_dev(int32_t nchart, const int64_t* v_off, const double* uv, const int64_t* b_off,
                       const int64_t* bnd, double* q, double* ext) {
    if (nchart < 0) return fail(RUV_ERR_INVALID, "rectangularize: negative chart count");
    if (nchart == 0) return RUV_OK;
    if (!v_off || !b_off || !ext) return fail(RUV_ERR_INVALID, "rectangularize: null argument");
    CtxLease L;
    int rc = lease_ctx(&L);
    if (rc) return rc;
    CallCtx* C = L.c;
    const double* ladder = nullptr;
    rc = ladder_for_device(C->dev, &ladder);
    if (rc) return rc;
    const hipStream_t s = C->stream;
    rc = ctx_order_after_caller(C);
    if (rc) return rc;
    std::vector<int64_t> voff((size_t)nchart + 1), boff((size_t)nchart + 1);
    HIPCHK(hipMemcpyAsync(voff.data(), v_off, sizeof(int64_t) * voff.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(boff.data(), b_off, sizeof(int64_t) * boff.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int32_t c = 0; c < nchart; ++c)
        if (voff[c + 1] < voff[c] || boff[c + 1] < boff[c]) return fail(RUV_ERR_BAD_DIMS, "rectangularize: offsets not monotone");
    constexpr int64_t VT = 4096, BT = 4096;
    std::vector<int32_t> vct0, vtc, bct0, btc;
    std::vector<int64_t> vi0, vi1, bi0, bi1;
    make_tiles(nchart, voff, VT, &vct0, &vtc, &vi0, &vi1);
    make_tiles(nchart, boff, BT, &bct0, &btc, &bi0, &bi1);
    const size_t nvt = vtc.size(), nbt = btc.size();
    const std::vector<size_t> bytes = {sizeof(int), sizeof(RectChart) * (size_t)nchart, sizeof(int32_t) * vct0.size(),
                                       sizeof(int32_t) * nvt, sizeof(int64_t) * nvt, sizeof(int64_t) * nvt,
                                       sizeof(Box) * nvt, sizeof(int32_t) * nbt, sizeof(int64_t) * nbt,
                                       sizeof(int64_t) * nbt, sizeof(double) * LADDER * (size_t)nchart,
                                       sizeof(int) * (size_t)nchart};
    size_t need = 0;
    for (size_t b : bytes) need += al256(b);
    rc = ctx_scratch(C, need);
    if (rc) return rc;
    std::vector<char*> P;
    char* p = (char*)C->buf;
    for (size_t b : bytes) {
        P.push_back(p);
        p += al256(b);
    }
    int* d_err = (int*)P[0];
    RectChart* d_rc = (RectChart*)P[1];
    int32_t *d_vct0 = (int32_t*)P[2], *d_vtc = (int32_t*)P[3], *d_btc = (int32_t*)P[7];
    int64_t *d_vi0 = (int64_t*)P[4], *d_vi1 = (int64_t*)P[5], *d_bi0 = (int64_t*)P[8], *d_bi1 = (int64_t*)P[9];
    Box* d_box = (Box*)P[6];
    double* d_score = (double*)P[10];
    int* d_bad = (int*)P[11];
    auto h2d = [s](void* d, const void* h, size_t b) {
        return b ? hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, s) : hipSuccess;
    };
    HIPCHK(h2d(d_vct0, vct0.data(), sizeof(int32_t) * vct0.size()));
    HIPCHK(h2d(d_vtc, vtc.data(), sizeof(int32_t) * nvt));
    HIPCHK(h2d(d_vi0, vi0.data(), sizeof(int64_t) * nvt));
    HIPCHK(h2d(d_vi1, vi1.data(), sizeof(int64_t) * nvt));
    HIPCHK(h2d(d_btc, btc.data(), sizeof(int32_t) * nbt));
    HIPCHK(h2d(d_bi0, bi0.data(), sizeof(int64_t) * nbt));
    HIPCHK(h2d(d_bi1, bi1.data(), sizeof(int64_t) * nbt));
    HIPCHK(hipMemsetAsync(d_err, 0, sizeof(int), s));
    HIPCHK(hipMemsetAsync(d_bad, 0, sizeof(int) * (size_t)nchart, s));
    hipLaunchKernelGGL(k_rect_scores, dim3((unsigned)(((int64_t)nchart * LGROUPS + WAVES - 1) / WAVES)), dim3(BLK), 0, s,
                       nchart, v_off, uv, b_off, bnd, ladder, d_score, d_bad);
    hipLaunchKernelGGL(k_rect_pick, dim3((unsigned)((nchart + BLK - 1) / BLK)), dim3(BLK), 0, s, nchart, b_off, ladder,
                       (const double*)d_score, (const int*)d_bad, d_rc, d_err);
    if (nvt)
        hipLaunchKernelGGL(k_rect_box, dim3((unsigned)nvt), dim3(BLK), 0, s, (const int32_t*)d_vtc,
                           (const int64_t*)d_vi0, (const int64_t*)d_vi1, v_off, uv, (const RectChart*)d_rc, d_box);
    hipLaunchKernelGGL(k_rect_finish, dim3((unsigned)((nchart + BLK - 1) / BLK)), dim3(BLK), 0, s, nchart,
                       (const int32_t*)d_vct0, (const Box*)d_box, d_rc, ext);
    if (nvt)
        hipLaunchKernelGGL(k_rect_write, dim3((unsigned)nvt), dim3(BLK), 0, s, (const int32_t*)d_vtc,
                           (const int64_t*)d_vi0, (const int64_t*)d_vi1, v_off, uv, (const RectChart*)d_rc, q);
    if (nbt)
        hipLaunchKernelGGL(k_rect_snap, dim3((unsigned)nbt), dim3(BLK), 0, s, (const int32_t*)d_btc,
                           (const int64_t*)d_bi0, (const int64_t*)d_bi1, v_off, b_off, bnd, (const RectChart*)d_rc, q);
    HIPCHK(hipGetLastError());
    return finish_call(s, d_err, &C->h_word[0], "rectangularize");
}

int chart_metrics(int32_t nchart, const int64_t* v_off, const double* p3, const double* uv, const int64_t* t_off,
                  const uint32_t* tris, double* metrics, int64_t* counts) {
    if (!v_off || !t_off || (nchart > 0 && (!metrics || !counts))) return fail(RUV_ERR_INVALID, "metrics: null argument");
    int64_t nv = 0, nt = 0;
    int rc = check_charts(nchart, v_off, t_off, &nv, &nt);
    if (rc) return rc;
    if (nchart == 0) return RUV_OK;
    if ((nv > 0 && (!p3 || !uv)) || (nt > 0 && !tris)) return fail(RUV_ERR_INVALID, "metrics: null argument");
    for (int32_t c = 0; c < nchart; ++c) {
        const int64_t n = v_off[c + 1] - v_off[c];
        for (int64_t k = 3 * t_off[c]; k < 3 * t_off[c + 1]; ++k)
            if ((int64_t)tris[k] >= n) return fail(RUV_ERR_BAD_DIMS, "metrics: triangle corner outside its chart");
    }
    DevSet D;
    int64_t *dv = nullptr, *dt = nullptr, *dc = nullptr;
    double *dp = nullptr, *du = nullptr, *dm = nullptr;
    uint32_t* dtri = nullptr;
    HIPCHK(D.put(&dv, v_off, (size_t)nchart + 1));
    HIPCHK(D.put(&dt, t_off, (size_t)nchart + 1));
    HIPCHK(D.put(&dp, p3, 3 * (size_t)nv));
    HIPCHK(D.put(&du, uv, 2 * (size_t)nv));
    HIPCHK(D.put(&dtri, tris, 3 * (size_t)nt));
    HIPCHK(D.put(&dm, (const double*)nullptr, 3 * (size_t)nchart));
    HIPCHK(D.put(&dc, (const int64_t*)nullptr, 2 * (size_t)nchart));
    rc = chart_metrics_dev(nchart, dv, dp, du, dt, dtri, dm, dc);
    if (rc) return rc;
    HIPCHK(hipMemcpy(metrics, dm, sizeof(double) * 3 * (size_t)nchart, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(counts, dc, sizeof(int64_t) * 2 * (size_t)nchart, hipMemcpyDeviceToHost));
    return RUV_OK;
}

int rectangularize(int32_t nchart, const int64_t* v_off, const double* uv, const int64_t* b_off, const int64_t* bnd,
                   double* q, double* ext) {
    if (!v_off || !b_off || (nchart > 0 && !ext)) return fail(RUV_ERR_INVALID, "rectangularize: null argument");
    int64_t nv = 0, nb = 0;
    int rc = check_charts(nchart, v_off, b_off, &nv, &nb);
    if (rc) return rc;
    if (nchart == 0) return RUV_OK;
    if ((nv > 0 && (!uv || !q)) || (nb > 0 && !bnd)) return fail(RUV_ERR_INVALID, "rectangularize: null argument");
    std::vector<char> seen;
    for (int32_t c = 0; c < nchart; ++c) {
        const int64_t n = v_off[c + 1] - v_off[c];
        seen.assign((size_t)n, 0);
        for (int64_t j = b_off[c]; j < b_off[c + 1]; ++j) {
            if (bnd[j] < 0 || bnd[j] >= n) return fail(RUV_ERR_BAD_DIMS, "rectangularize: border vertex outside its chart");
            if (seen[(size_t)bnd[j]]) return fail(RUV_ERR_INVALID, "rectangularize: repeated border vertex");
            seen[(size_t)bnd[j]] = 1;
        }
    }
    DevSet D;
    int64_t *dv = nullptr, *db = nullptr, *dbn = nullptr;
    double *du = nullptr, *dq = nullptr, *de = nullptr;
    HIPCHK(D.put(&dv, v_off, (size_t)nchart + 1));
    HIPCHK(D.put(&db, b_off, (size_t)nchart + 1));
    HIPCHK(D.put(&dbn, bnd, (size_t)nb));
    HIPCHK(D.put(&du, uv, 2 * (size_t)nv));
    HIPCHK(D.put(&dq, (const double*)nullptr, 2 * (size_t)nv));
    HIPCHK(D.put(&de, (const double*)nullptr, 2 * (size_t)nchart));
    rc = rectangularize_dev(nchart, dv, du, db, dbn, dq, de);
    if (rc) return rc;
    HIPCHK(hipMemcpy(q, dq, sizeof(double) * 2 * (size_t)nv, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ext, de, sizeof(double) * 2 * (size_t)nchart, hipMemcpyDeviceToHost));
    return RUV_OK;
}

}  // namespace ruv
