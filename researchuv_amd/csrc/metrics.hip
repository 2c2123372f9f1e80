// metrics.hip -- post-solve chart metrics on the device, one workgroup per chart:
//   chart_distortion (crates/researchuv-unwrap/src/metrics.rs:73-141, with svd2 :59-71 and winding_flips
//   :39-56) and rectangularize (:147-227), which the reference's unfold stage runs on every chart after its
//   solve (pipeline.rs:171-182).
// Every operation follows the reference's order with multiply-add contraction disabled in this file, so each
// triangle's terms, every min / max and the rectangularized coordinates are the reference's bit for bit; the
// per-chart means sum the triangles in a fixed tree order instead of the reference's left-to-right order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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
__device__ __forceinline__ double block_min(double v, double* sh) { return -block_max_all(-v, sh); }
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

// metrics (3 per chart): conformal mean, conformal max, area-ratio mean; counts (2 per chart): flips, folds
__global__ __launch_bounds__(BLK) void k_chart_distortion(const int64_t* __restrict__ v_off,
                                                          const double* __restrict__ p3,
                                                          const double* __restrict__ uv,
                                                          const int64_t* __restrict__ t_off,
                                                          const uint32_t* __restrict__ tris,
                                                          double* __restrict__ metrics,
                                                          long long* __restrict__ counts) {
    __shared__ double sh[WAVES];
    __shared__ long long shc[WAVES];
    const int c = blockIdx.x;
    const int64_t v0 = v_off[c];
    const double* P = p3 + 3 * v0;
    const double* U = uv + 2 * v0;
    double conf_sum = 0.0, conf_max = 0.0, area_sum = 0.0, signed_sum = 0.0;
    long long nconf = 0, flips = 0, le0 = 0, ge0 = 0;
    for (int64_t t = t_off[c] + threadIdx.x; t < t_off[c + 1]; t += BLK) {
        const uint32_t a = tris[3 * t], b = tris[3 * t + 1], cc = tris[3 * t + 2];
        const double u0u = U[2 * a], u0v = U[2 * a + 1], u1u = U[2 * b], u1v = U[2 * b + 1];
        const double u2u = U[2 * cc], u2v = U[2 * cc + 1];
        // winding_flips' signed area, for every triangle
        const double sa = tri_signed(u0u, u0v, u1u, u1v, u2u, u2v);
        signed_sum += sa;
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
    const double cs = block_sum_all(conf_sum, sh);
    const double cm = block_max(conf_max, sh);
    const double as = block_sum_all(area_sum, sh);
    const double ss = block_sum_all(signed_sum, sh);
    const long long nc = block_count(nconf, shc);
    const long long fl = block_count(flips, shc);
    const long long nle = block_count(le0, shc);
    const long long nge = block_count(ge0, shc);
    if (threadIdx.x == 0) {
        const double den = (double)(nc > 0 ? nc : 1);
        metrics[3 * c] = cs / den;
        metrics[3 * c + 1] = cm;
        metrics[3 * c + 2] = as / den;
        counts[2 * c] = fl;
        // dominant winding = sign of the signed-area sum; folds disagree with it (zero areas count)
        counts[2 * c + 1] = (t_off[c + 1] == t_off[c]) ? 0 : ((ss < 0.0) ? nge : nle);
    }
}

// rectangularize (metrics.rs:147-227), chart c: q (2 per vertex) and ext (2 per chart); ladder[2k], [2k+1]
// the cosine and sine of pi k / 71 as the host's libm gives them
__global__ __launch_bounds__(BLK) void k_rectangularize(const int64_t* __restrict__ v_off,
                                                        const double* __restrict__ uv,
                                                        const int64_t* __restrict__ b_off,
                                                        const int64_t* __restrict__ bnd,
                                                        const double* __restrict__ ladder,
                                                        double* __restrict__ q, double* __restrict__ ext) {
    __shared__ double sh[WAVES];
    const int c = blockIdx.x;
    const int64_t v0 = v_off[c], nv = v_off[c + 1] - v0;
    const int64_t b0 = b_off[c], nb = b_off[c + 1] - b0;
    const double* U = uv + 2 * v0;
    double* Q = q + 2 * v0;
    if (nv == 0) {
        if (threadIdx.x == 0) ext[2 * c] = ext[2 * c + 1] = 0.0;
        return;
    }
    if (nb == 0) {
        double mnu = INFINITY, mnv = INFINITY, mxu = -INFINITY, mxv = -INFINITY;
        for (int64_t i = threadIdx.x; i < nv; i += BLK) {
            mnu = fmin(mnu, U[2 * i]); mnv = fmin(mnv, U[2 * i + 1]);
            mxu = fmax(mxu, U[2 * i]); mxv = fmax(mxv, U[2 * i + 1]);
        }
        mnu = block_min(mnu, sh); mnv = block_min(mnv, sh);
        mxu = block_max(mxu, sh); mxv = block_max(mxv, sh);
        const double eu = mxu - mnu, ev = mxv - mnv;
        for (int64_t i = threadIdx.x; i < nv; i += BLK) {
            Q[2 * i] = (U[2 * i] - mnu) / fmax(eu, 1e-30);
            Q[2 * i + 1] = (U[2 * i + 1] - mnv) / fmax(ev, 1e-30);
        }
        if (threadIdx.x == 0) {
            ext[2 * c] = eu;
            ext[2 * c + 1] = ev;
        }
        return;
    }
    // the ladder: the first angle of smallest score (a later one replaces it only when strictly smaller)
    double bc = 1.0, bs = 0.0, bscore = INFINITY;
    bool have = false;
    for (int k = 0; k < LADDER; ++k) {
        const double cs = ladder[2 * k], sn = ladder[2 * k + 1];
        double mn = INFINITY, mx = -INFINITY, mnv = INFINITY, mxv = -INFINITY;
        for (int64_t j = threadIdx.x; j < nb; j += BLK) {
            const int64_t bi = bnd[b0 + j];
            const double pu = U[2 * bi], pv = U[2 * bi + 1];
            const double qu = cs * pu + sn * pv;
            const double qv = -sn * pu + cs * pv;
            mn = fmin(mn, qu); mx = fmax(mx, qu);
            mnv = fmin(mnv, qv); mxv = fmax(mxv, qv);
        }
        mn = block_min(mn, sh); mx = block_max(mx, sh);
        mnv = block_min(mnv, sh); mxv = block_max(mxv, sh);
        const double e0 = mx - mn, e1 = mxv - mnv;
        const double score = sqrt(e0 * (e0 + 1e-30) + e1 * (e1 + 1e-30));
        if (!(have && score >= bscore)) {  // every thread holds the same values
            have = true;
            bc = cs;
            bs = sn;
            bscore = score;
        }
    }
    double mnu = INFINITY, mnv = INFINITY, mxu = -INFINITY, mxv = -INFINITY;
    for (int64_t i = threadIdx.x; i < nv; i += BLK) {
        const double pu = U[2 * i], pv = U[2 * i + 1];
        const double qu = bc * pu + bs * pv, qv = -bs * pu + bc * pv;
        mnu = fmin(mnu, qu); mnv = fmin(mnv, qv);
        mxu = fmax(mxu, qu); mxv = fmax(mxv, qv);
    }
    mnu = block_min(mnu, sh); mnv = block_min(mnv, sh);
    mxu = block_max(mxu, sh); mxv = block_max(mxv, sh);
    const double eu = mxu - mnu, ev = mxv - mnv;
    for (int64_t i = threadIdx.x; i < nv; i += BLK) {
        const double pu = U[2 * i], pv = U[2 * i + 1];
        const double qu = bc * pu + bs * pv, qv = -bs * pu + bc * pv;
        Q[2 * i] = (qu - mnu) / fmax(eu, 1e-30);
        Q[2 * i + 1] = (qv - mnv) / fmax(ev, 1e-30);
    }
    __syncthreads();  // the border snap reads what other threads wrote (border entries are distinct)
    for (int64_t j = threadIdx.x; j < nb; j += BLK) {
        const int64_t bi = bnd[b0 + j];
        const double pu = Q[2 * bi], pv = Q[2 * bi + 1];
        const double dx = fmin(pu, 1.0 - pu), dy = fmin(pv, 1.0 - pv);
        if (dx <= dy) Q[2 * bi] = (pu < 0.5) ? 0.0 : 1.0;
        else Q[2 * bi + 1] = (pv < 0.5) ? 0.0 : 1.0;
    }
    if (threadIdx.x == 0) {
        ext[2 * c] = eu;
        ext[2 * c + 1] = ev;
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
    int64_t *dv = nullptr, *dt = nullptr;
    double *dp = nullptr, *du = nullptr, *dm = nullptr;
    uint32_t* dtri = nullptr;
    long long* dc = nullptr;
    HIPCHK(D.put(&dv, v_off, (size_t)nchart + 1));
    HIPCHK(D.put(&dt, t_off, (size_t)nchart + 1));
    HIPCHK(D.put(&dp, p3, 3 * (size_t)nv));
    HIPCHK(D.put(&du, uv, 2 * (size_t)nv));
    HIPCHK(D.put(&dtri, tris, 3 * (size_t)nt));
    HIPCHK(D.put(&dm, (const double*)nullptr, 3 * (size_t)nchart));
    HIPCHK(D.put(&dc, (const long long*)nullptr, 2 * (size_t)nchart));
    hipLaunchKernelGGL(k_chart_distortion, dim3((unsigned)nchart), dim3(BLK), 0, 0, (const int64_t*)dv,
                       (const double*)dp, (const double*)du, (const int64_t*)dt, (const uint32_t*)dtri, dm, dc);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(metrics, dm, sizeof(double) * 3 * (size_t)nchart, hipMemcpyDeviceToHost));
    std::vector<long long> hc(2 * (size_t)nchart);
    HIPCHK(hipMemcpy(hc.data(), dc, sizeof(long long) * hc.size(), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < hc.size(); ++k) counts[k] = (int64_t)hc[k];
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
    // the ladder's angles, pi k / 71 (metrics.rs:176-179), through the host's libm like the reference
    double lad[2 * LADDER];
    for (int k = 0; k < LADDER; ++k) {
        const double ang = 3.14159265358979323846 * (double)k / (double)(LADDER - 1);
        lad[2 * k] = std::cos(ang);
        lad[2 * k + 1] = std::sin(ang);
    }
    DevSet D;
    int64_t *dv = nullptr, *db = nullptr, *dbn = nullptr;
    double *du = nullptr, *dq = nullptr, *de = nullptr, *dl = nullptr;
    HIPCHK(D.put(&dv, v_off, (size_t)nchart + 1));
    HIPCHK(D.put(&db, b_off, (size_t)nchart + 1));
    HIPCHK(D.put(&dbn, bnd, (size_t)nb));
    HIPCHK(D.put(&du, uv, 2 * (size_t)nv));
    HIPCHK(D.put(&dl, (const double*)lad, 2 * (size_t)LADDER));
    HIPCHK(D.put(&dq, (const double*)nullptr, 2 * (size_t)nv));
    HIPCHK(D.put(&de, (const double*)nullptr, 2 * (size_t)nchart));
    hipLaunchKernelGGL(k_rectangularize, dim3((unsigned)nchart), dim3(BLK), 0, 0, (const int64_t*)dv,
                       (const double*)du, (const int64_t*)db, (const int64_t*)dbn, (const double*)dl, dq, de);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(q, dq, sizeof(double) * 2 * (size_t)nv, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ext, de, sizeof(double) * 2 * (size_t)nchart, hipMemcpyDeviceToHost));
    return RUV_OK;
}

}  // namespace ruv
