// bamg_kernels.hpp -- device-built aggregation AMG for the batched solver: setup (pairwise matching, Galerkin
// products) and the V(1,1) cycle.
//
// The hierarchy follows the reference's aggregation AMG (crates/researchuv-gpu/src/amg.rs:127-455) in kind:
// aggregates of rows strongly coupled to each other, the piecewise-constant aggregate map as the prolongator, the
// Galerkin product P^T A P as the coarse operator, damped Jacobi smoothing, a direct solve on the coarsest level.
// What differs, for the device and for batches of charts:
//  * one pairwise matching per level, by handshakes on the device: each row ranks its neighbours coupled at least
//    theta (0.25) times as strongly as its strongest -- strength in quarter-octave bins, then a pair hash -- keeps
//    the best MCAND, and in three rounds proposes the first still unmatched; mutual proposals pair up, and a row
//    left unmatched joins the pair of its first matched candidate.  Aggregates never leave their chart, so the
//    hierarchy of a batch is block-diagonal by chart;
//  * the coarse levels are stored with each next-level aggregate's rows consecutive (the restriction is a sum
//    over runs of rows, fused into the residual pass) as sliced ELL whose slices start at aggregate starts
//    (16-bit offsets from the row, or 32-bit columns when some offset does not fit; bfloat16 values under the
//    FP32 cycle), so every coarse pass streams coalesced;
//  * V(1,1) with Jacobi weight min(0.9, 1.9 / g) per level (g the level's Gershgorin bound on D^-1 A), coarse
//    corrections over-corrected by a fixed factor, and each chart's coarsest block inverted (dense) when small.
// No inner products: the cycle is a fixed linear SPD operator, so each chart's PCG stays an independent PCG.
//
// Setup works on CSR levels with FP32 values (64-bit row pointers, 32-bit level-global columns; sums in FP64).
// Ties in the matching are broken by a pair hash and the lower row, and every Galerkin row is summed in a fixed
// order, so the hierarchy is deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "batch_kernels.hpp"
#include "dev_common.hpp"

namespace ruv {

constexpr int SCAN_BLK = 1024;  // elements per workgroup of the scan

// ---------------------------------------------------------------- exclusive scan of int64 counts
// Reduce, then scan: k_scan_reduce sums each block of SCAN_BLK entries, k_scan_totals scans the block sums,
// k_scan_down scans each block again with its offset and writes the result (two reads and one write of the
// entries).  A thread owns SCAN_BLK / BLK consecutive entries; the waves scan their threads' sums by shuffles.
constexpr int SCAN_PER = SCAN_BLK / BLK;

// the thread's entries (zeros past n) and their sum
__device__ __forceinline__ int64_t scan_load(int64_t n, const int64_t* __restrict__ in, int64_t base,
                                             int64_t (&v)[SCAN_PER]) {
    int64_t sum = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q) {
        const int64_t k = base + threadIdx.x * SCAN_PER + q;
        v[q] = k < n ? in[k] : 0;
        sum += v[q];
    }
    return sum;
}

__global__ __launch_bounds__(BLK) void k_scan_reduce(int64_t n, const int64_t* __restrict__ in,
                                                     int64_t* __restrict__ totals) {
    __shared__ int64_t sh[WAVES];
    int64_t v[SCAN_PER];
    int64_t sum = scan_load(n, in, (int64_t)blockIdx.x * SCAN_BLK, v);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) t += sh[w];
        totals[blockIdx.x] = t;
    }
}

// out[k] = the exclusive scan within the block plus the block's offset; the last block also writes out[n] = the
// grand total.  scanned: totals holds the scanned block sums and the grand total at nb (k_scan_totals); else it
// holds the block sums, and each block adds up those before it (few blocks: no launch for the totals).
__global__ __launch_bounds__(BLK) void k_scan_down(int64_t n, int64_t nb, const int64_t* __restrict__ in,
                                                   const int64_t* __restrict__ totals, int64_t* __restrict__ out,
                                                   int scanned) {
    __shared__ int64_t sh[WAVES];
    __shared__ int64_t sb[WAVES];
    const int64_t base = (int64_t)blockIdx.x * SCAN_BLK;
    int64_t v[SCAN_PER];
    const int64_t sum = scan_load(n, in, base, v);
    int64_t offset;
    if (scanned) {
        offset = totals[blockIdx.x];
    } else {
        int64_t a = 0;
        for (int64_t k = threadIdx.x; k < (int64_t)blockIdx.x; k += BLK) a += totals[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
        if ((threadIdx.x & 63) == 0) sb[threadIdx.x >> 6] = a;
        __syncthreads();
        offset = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) offset += sb[w];
    }
    const int lane = threadIdx.x & 63;
    int64_t incl = sum;  // inclusive scan of the threads' sums within the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
    }
    if (lane == 63) sh[threadIdx.x >> 6] = incl;
    __syncthreads();
    int64_t run = offset + incl - sum;  // the block's offset, earlier waves, earlier threads
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) run += sh[w];
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q) {
        const int64_t k = base + threadIdx.x * SCAN_PER + q;
        if (k < n) out[k] = run;
        run += v[q];
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == BLK - 1) out[n] = run;  // the last thread's run: the grand total
}

// one workgroup of TOT_BLK threads: exclusive scan of the block totals in place, grand total in totals[nb]; each
// thread sums a contiguous run of the totals, the runs' sums are scanned in LDS, then each run is written out
constexpr int TOT_BLK = 1024;
__global__ __launch_bounds__(TOT_BLK) void k_scan_totals(int64_t nb, int64_t* __restrict__ totals) {
    __shared__ int64_t part[TOT_BLK];
    const int64_t per = (nb + TOT_BLK - 1) / TOT_BLK;
    const int64_t a = (int64_t)threadIdx.x * per, b = min(nb, a + per);
    int64_t sum = 0;
    for (int64_t k = a; k < b; ++k) sum += totals[k];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < TOT_BLK; off <<= 1) {
        const int64_t u = (threadIdx.x >= (unsigned)off) ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += u;
        __syncthreads();
    }
    int64_t run = (threadIdx.x == 0) ? 0 : part[threadIdx.x - 1];
    for (int64_t k = a; k < b; ++k) {
        const int64_t v = totals[k];
        totals[k] = run;
        run += v;
    }
    if (threadIdx.x == TOT_BLK - 1) totals[nb] = part[TOT_BLK - 1];
}


// ---------------------------------------------------------------- matching
// A level's operator during the setup (CSR).  Values in FP32: the cycle stores them in bfloat16 (FP64 under the
// FP64 cycle, where FP32 setup values still make a fixed SPD preconditioner), and the setup passes read them
// once or twice each, so half the bytes of FP64 shorten every pass; sums (Galerkin entries, strengths, diagonals)
// are formed in FP64.
struct SetupCsr {
    int64_t n;
    const int64_t* ptr;
    const int32_t* cols;
    const float* vals;
};

// The cycle's operator values (coarse levels, and the finest level's copy for the smoothing legs) are stored as
// bfloat16 under an FP32 cycle (half the bytes of the level sweeps, 8 significant bits: the operators stay
// symmetric, since a_ij and a_ji round alike, and the cycle a fixed SPD preconditioner) and as double under an
// FP64 cycle.
template <typename T>
struct CoarseVal {
    using type = T;
};
template <>
struct CoarseVal<float> {
    using type = Bf16;
};
// double -> stored value: round to nearest, ties to even, through FP32 for bfloat16
__device__ __forceinline__ void cput(Bf16* p, double v) {
    const uint32_t u = __float_as_uint((float)v);
    p->bits = (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ void cput(double* p, double v) { *p = v; }


// The finest operator's band form (batch_kernels.hpp) as a source of the setup's level-0 rows: null bw for the other
// levels or a batch without band charts.  Values in FP32 (f32, the mixed batch) or FP64 (f64, cast to FP32 as the
// setup CSR holds them: the same values).
struct BandSrc {
    const int32_t* bw;     // per chart: grid width, 0 when not in band form
    const int32_t* bmask;
    const int64_t* bbase;
    const int64_t* crow;
    const int32_t* chart;  // row -> chart
    const float* f32;
    const double* f64;
};

// Row i's entries from the band form in ascending column order -- the order of the setup CSR -- as col[q], val[q]
// for the offsets -(W+1), -W, -(W-1), -1, 0, 1, W-1, W, W+1 (an absent entry reads as 0 and is to be skipped: the
// CSR holds no zero entries).  Returns false (nothing read) when row i's chart is not in band form.
__device__ __forceinline__ bool band_row(const BandSrc& S, int64_t i, int32_t (&col)[9], float (&val)[9]) {
    if (!S.bw) return false;
    const int c = S.chart[i];
    const int W = S.bw[c];
    if (W <= 0) return false;
    const int m = S.bmask[c];
    const int64_t lo = S.crow[c];
    const int n = (int)(S.crow[c + 1] - lo);
    const int l = (int)(i - lo);
    // slot of each offset in the chart's band arrays (an absent offset reads the diagonal's slot, zeroed below)
    int sl[4], k = 1;
#pragma unroll
    for (int b = 0; b < 4; ++b) sl[b] = ((m >> b) & 1) ? k++ : 0;
    const int64_t base = S.bbase[c];
    const int d[4] = {1, W - 1, W, W + 1};
    const int off[9] = {-d[3], -d[2], -d[1], -d[0], 0, d[0], d[1], d[2], d[3]};
    const int slots[9] = {sl[3], sl[2], sl[1], sl[0], 0, sl[0], sl[1], sl[2], sl[3]};
    const int bits[9] = {3, 2, 1, 0, -1, 0, 1, 2, 3};
    int64_t at[9];
    bool ok[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int r = q < 4 ? l + off[q] : l;  // a lower entry is the upper entry of row l - d
        ok[q] = r >= 0 && (bits[q] < 0 || ((m >> bits[q]) & 1));
        at[q] = base + (int64_t)slots[q] * n + (ok[q] ? r : l);
        col[q] = (int32_t)(i + off[q]);
    }
    // the nine loads issue together (one branch on the precision, outside the loop)
    if (S.f32) {
#pragma unroll
        for (int q = 0; q < 9; ++q) val[q] = S.f32[at[q]];
    } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) val[q] = (float)S.f64[at[q]];
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) val[q] = ok[q] ? val[q] : 0.0f;
    return true;
}

// Preference of the coupling (i, j): its strength in bins of a quarter octave, then a hash of the unordered
// pair.  Couplings that vary smoothly along a line (uniform or geometric spacing) fall into one bin, where the
// hash -- the same seen from either end -- acts as a random priority; ordering by exact strength or by index
// would make every row propose to the same side and leave chains instead of pairs.
__device__ __forceinline__ uint64_t pair_pref(double s, int64_t i, int64_t j) {
    const int64_t a = min(i, j), b = max(i, j);
    uint32_t h = (uint32_t)a * 0x9E3779B1u ^ ((uint32_t)b + 0x7F4A7C15u) * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    // bin = floor(4 log2 s) for s > 0, from the exponent and the mantissa against 2^(1/4), 2^(1/2), 2^(3/4)
    // (the rounded constants; no logarithm); subnormal s fall into the lowest bins
    const uint64_t u = (uint64_t)__double_as_longlong(s);
    const int ex = (int)((u >> 52) & 0x7FF) - 1023;
    const double mant = __longlong_as_double((long long)((u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull));
    const int q = (mant >= 1.6817928305074290) + (mant >= 1.4142135623730951) + (mant >= 1.1892071150027210);
    const int bin = 4 * ex + q + 4096;
    return ((uint64_t)(uint32_t)max(bin, 0) << 32) | h;
}

// Handshake matching.  Each row ranks its eligible neighbours once -- same group, coupling a_ij < 0 with
// -a_ij >= theta * the row's strongest -- by pair_pref (ties to the lower row), keeping the best MCAND; every round
// an unmatched row proposes the first of them still unmatched, and mutual proposals become pairs.
constexpr int MCAND = 4;
constexpr int CAND_B = 4;  // CSR entries whose loads issue together (8 costs occupancy)

__device__ __forceinline__ bool pref_before(uint64_t p, int32_t j, uint64_t q, int32_t k) {
    return p > q || (p == q && j < k);
}

// ranking and the first round's proposals (every row is unmatched then)
// Also the level's Gershgorin bound on D^-1 A, max over rows of sum_j |a_ij| / a_ii, into *gmax (the bits of a
// non-negative double order like the double).
// The level-0 rows of band charts are read from the band form (the same entries in the same order, coalesced), the
// others from the setup CSR.
__global__ __launch_bounds__(BLK) void k_candidates(SetupCsr A, BandSrc band, const int32_t* __restrict__ group,
                                                    double theta, int32_t* __restrict__ cand,
                                                    int32_t* __restrict__ ncand, int64_t* __restrict__ prop,
                                                    unsigned long long* __restrict__ gmax) {
    __shared__ double sh[WAVES];
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    const bool in = i < A.n;
    int32_t bc[9];
    float bv[9];
    const bool fromband = in && band_row(band, i, bc, bv);
    const int64_t e0 = (in && !fromband) ? A.ptr[i] : 0, e1 = (in && !fromband) ? A.ptr[i + 1] : 0;
    // f(j, a_ij, group of j) over the row's entries in ascending column order (the group read when G).  CSR rows
    // in batches of CAND_B entries whose loads issue together with no branch: past the row's end the index repeats
    // its last entry, passed as a_ij = 0, which both passes skip (it adds 0 to the sums: the diagonal and |a| sums
    // start at +0.0 and add nonzero entries, so adding +0.0 changes nothing)
    const auto each = [&](auto&& f, auto gflag) {
        constexpr bool G = decltype(gflag)::value;
        if (fromband) {
#pragma unroll
            for (int q = 0; q < 9; ++q)
                if (bv[q] != 0.0f) f(bc[q], bv[q], 0);
        } else {
            for (int64_t e = e0; e < e1; e += CAND_B) {
                int32_t cj[CAND_B], gj[CAND_B];
                float vj[CAND_B];
#pragma unroll
                for (int u = 0; u < CAND_B; ++u) {
                    const int64_t a = min(e + u, e1 - 1);
                    cj[u] = A.cols[a];
                    vj[u] = A.vals[a];
                }
#pragma unroll
                for (int u = 0; u < CAND_B; ++u) gj[u] = G ? group[cj[u]] : 0;
#pragma unroll
                for (int u = 0; u < CAND_B; ++u) f(cj[u], e + u < e1 ? vj[u] : 0.0f, gj[u]);
            }
        }
    };
    double smax = 0.0, asum = 0.0, d = 0.0;
    each(
        [&](int32_t j, float a, int32_t) {
            const double v = a;
            asum += fabs(v);
            if (j == i) d += v;
            else if (v < 0.0) smax = fmax(smax, -v);
        },
        std::false_type{});
    const double gb = block_max_all((d > 0.0) ? asum / d : 0.0, sh);
    // one atomic on one word from every workgroup serialises (about 12 ns each: 1.7 ms for the 142 000 workgroups
    // of 36 M rows); the bound only grows, so a workgroup whose maximum the word already covers skips it
    if (threadIdx.x == 0) {
        const unsigned long long v = (unsigned long long)__double_as_longlong(gb);
        if (v > __hip_atomic_load(gmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(gmax, v);
    }
    if (!in) return;
    const double thr = theta * smax;
    const int32_t g = group[i];
    uint64_t kp[MCAND];
    int32_t kj[MCAND];
#pragma unroll
    for (int k = 0; k < MCAND; ++k) {
        kp[k] = 0;
        kj[k] = INT32_MAX;
    }
    int cnt = 0;
    each([&](int32_t j, float a, int32_t gj) {
        // a band row's neighbours lie in its chart, the level-0 group
        if (j == i || !(a < 0.0f) || (!fromband && gj != g)) return;
        const double sv = -(double)a;
        if (sv < thr) return;
        // carry (p, j) down the descending list: it takes the first slot it precedes, or the first free one
        uint64_t cp = pair_pref(sv, i, j);
        int32_t cj = j;
        bool carry = true;
#pragma unroll
        for (int k = 0; k < MCAND; ++k) {
            if (carry) {
                if (k >= cnt) {
                    kp[k] = cp;
                    kj[k] = cj;
                    carry = false;
                } else if (pref_before(cp, cj, kp[k], kj[k])) {
                    const uint64_t tp = kp[k];
                    const int32_t tj = kj[k];
                    kp[k] = cp;
                    kj[k] = cj;
                    cp = tp;
                    cj = tj;
                }
            }
        }
        if (cnt < MCAND) ++cnt;
    }, std::true_type{});
#pragma unroll
    for (int k = 0; k < MCAND; ++k)
        if (k < cnt) cand[i * MCAND + k] = kj[k];
    ncand[i] = cnt;
    prop[i] = cnt ? kj[0] : -1;
}

// later rounds: an unmatched row proposes its first candidate still unmatched
__global__ __launch_bounds__(BLK) void k_propose(int64_t n, const int32_t* __restrict__ cand,
                                                 const int32_t* __restrict__ ncand,
                                                 const int64_t* __restrict__ mate, int64_t* __restrict__ prop) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    int64_t best = -1;
    // the row's state, candidate count and candidates, then the candidates' states, each group of loads issued
    // together (a candidate past the count repeats the first, whose state is loaded anyway)
    const int64_t mi = mate[i];
    const int c = ncand[i];
    int32_t cj[MCAND];
#pragma unroll
    for (int k = 0; k < MCAND; ++k) cj[k] = cand[i * MCAND + (k < c ? k : 0)];
    if (mi < 0 && c > 0) {
        int64_t mj[MCAND];
#pragma unroll
        for (int k = 0; k < MCAND; ++k) mj[k] = mate[cj[k]];
#pragma unroll
        for (int k = MCAND - 1; k >= 0; --k)
            if (k < c && mj[k] < 0) best = cj[k];
    }
    prop[i] = best;
}

__global__ __launch_bounds__(BLK) void k_accept(int64_t n, const int64_t* __restrict__ prop,
                                                int64_t* __restrict__ mate) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const int64_t mi = mate[i], j = prop[i];  // loaded together
    if (mi >= 0) return;
    if (j >= 0 && prop[j] == i) mate[i] = j;
}

// a row still unmatched joins the pair of its first candidate that is matched; join = that pair's lower row, or
// -1 (the row stays alone)
__global__ __launch_bounds__(BLK) void k_join(int64_t n, const int32_t* __restrict__ cand,
                                              const int32_t* __restrict__ ncand, const int64_t* __restrict__ mate,
                                              int64_t* __restrict__ join) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    int64_t best = -1;
    // loads grouped as in k_propose
    const int64_t mi = mate[i];
    const int c = ncand[i];
    int32_t cj[MCAND];
#pragma unroll
    for (int k = 0; k < MCAND; ++k) cj[k] = cand[i * MCAND + (k < c ? k : 0)];
    if (mi < 0 && c > 0) {
        int64_t mj[MCAND];
#pragma unroll
        for (int k = 0; k < MCAND; ++k) mj[k] = mate[cj[k]];
#pragma unroll
        for (int k = MCAND - 1; k >= 0; --k)
            if (k < c && mj[k] >= 0) best = min((int64_t)cj[k], mj[k]);
    }
    join[i] = best;
}

// roots: the lower row of a pair, and rows left alone
__global__ __launch_bounds__(BLK) void k_roots(int64_t n, const int64_t* __restrict__ mate,
                                               const int64_t* __restrict__ join, int64_t* __restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    flag[i] = (mate[i] >= 0) ? (i < mate[i] ? 1 : 0) : (join[i] < 0 ? 1 : 0);
}

// agg[i] = the aggregate of row i (the scan position cid of its root: the lower row of its pair, the pair it
// joined, or itself), and the aggregates' sizes into cnt (zeroed before): each root adds its pair (1 or 2), each
// row that joined a pair adds 1 -- one atomic per aggregate and per joined row, not per row
__global__ __launch_bounds__(BLK) void k_root_agg(int64_t n, const int64_t* __restrict__ mate,
                                                  const int64_t* __restrict__ join, const int64_t* __restrict__ cid,
                                                  int32_t* __restrict__ agg, unsigned long long* __restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const int64_t m = mate[i], jr = join[i];  // loaded together (join is written for every row)
    const int64_t jn = (m >= 0) ? -1 : jr;
    const int64_t root = (m >= 0) ? min(i, m) : (jn >= 0 ? jn : i);
    const int64_t a = cid[root];
    agg[i] = (int32_t)a;
    if (root == i) atomicAdd(&cnt[a], (m >= 0) ? 2ull : 1ull);
    else if (jn >= 0) atomicAdd(&cnt[a], 1ull);
}

// member lists of a matching's aggregates: a pair's rows at its first two slots (lower row first), rows that
// joined it after them (slot order by an atomic cursor, zeroed before; k_member_sort orders the lists)
__global__ __launch_bounds__(BLK) void k_member_fill(int64_t n, const int64_t* __restrict__ mate,
                                                     const int64_t* __restrict__ join, const int32_t* __restrict__ agg,
                                                     const int64_t* __restrict__ mptr,
                                                     unsigned long long* __restrict__ cursor,
                                                     int64_t* __restrict__ mrows) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const int64_t a = agg[i], m = mate[i], jr = join[i];  // loaded together
    const int64_t base = mptr[a];
    int64_t slot;
    if (m >= 0) slot = (i < m) ? 0 : 1;
    else if (jr >= 0) slot = 2 + (int64_t)atomicAdd(&cursor[a], 1ull);
    else slot = 0;
    mrows[base + slot] = i;
}

// each aggregate's member list sorted ascending; with group, the aggregate's group (that of its rows)
__global__ __launch_bounds__(BLK) void k_member_sort(int64_t nc, const int64_t* __restrict__ mptr,
                                                     int64_t* __restrict__ mrows, const int32_t* __restrict__ group,
                                                     int32_t* __restrict__ group_c) {
    const int64_t I = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (I >= nc) return;
    const int64_t a = mptr[I], b = mptr[I + 1];
    if (b - a <= 4) {  // the usual sizes (pairs, and rows that joined them): sorted in registers, loads together
        int64_t v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = (a + k < b) ? mrows[min(a + k, b - 1)] : INT64_MAX;
        const auto cs = [&](int x, int y) {
            const int64_t lo = min(v[x], v[y]), hi = max(v[x], v[y]);
            v[x] = lo;
            v[y] = hi;
        };
        cs(0, 1);
        cs(2, 3);
        cs(0, 2);
        cs(1, 3);
        cs(1, 2);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (a + k < b) mrows[a + k] = v[k];
        if (group_c && b > a) group_c[I] = group[v[0]];
        return;
    }
    for (int64_t p = a + 1; p < b; ++p) {
        const int64_t v = mrows[p];
        int64_t q = p - 1;
        while (q >= a && mrows[q] > v) {
            mrows[q + 1] = mrows[q];
            --q;
        }
        mrows[q + 1] = v;
    }
    if (group_c && b > a) group_c[I] = group[mrows[a]];
}

// Galerkin row I = the rows of aggregate I with columns mapped through agg, duplicates summed, columns
// ascending, each column summed over its entries in row order.  A row of up to GAL_CAP entries is gathered
// into LDS (slot-major, one column of slots per thread) and stable-sorted there; a longer row falls back to
// repeated selection of the next column.  Both give the same sums in the same order.
constexpr int GAL_BLK = 64;
constexpr int GAL_CAP = 64;

// With `list`, the kernel does the rows list[0 .. nc) (those k_galerkin_reg could not).
template <bool FILL>
__global__ __launch_bounds__(GAL_BLK) void k_galerkin(int64_t nc, SetupCsr A, const int64_t* __restrict__ mptr,
                                                      const int64_t* __restrict__ mrows,
                                                      const int32_t* __restrict__ agg, int64_t* __restrict__ cnt,
                                                      const int64_t* __restrict__ cptr, int32_t* __restrict__ ccols,
                                                      float* __restrict__ cvals, double* __restrict__ cdiag,
                                                      const int64_t* __restrict__ list) {
    __shared__ int32_t skey[GAL_CAP * GAL_BLK];
    __shared__ double sval[GAL_CAP * GAL_BLK];
    const int64_t idx = (int64_t)blockIdx.x * GAL_BLK + threadIdx.x;
    if (idx >= nc) return;
    const int64_t I = list ? list[idx] : idx;
    const int tid = threadIdx.x;
    const int64_t m0 = mptr[I], m1 = mptr[I + 1];
    int m = 0;
    for (int64_t q = m0; q < m1; ++q) {
        const int64_t r = mrows[q];
        m += (int)(A.ptr[r + 1] - A.ptr[r]);
    }
    int64_t out = FILL ? cptr[I] : 0;
    int64_t count = 0;
    double diag = 0.0;
    if (m <= GAL_CAP) {
        int k = 0;
        for (int64_t q = m0; q < m1; ++q) {
            const int64_t r = mrows[q];
            for (int64_t e = A.ptr[r]; e < A.ptr[r + 1]; ++e, ++k) {
                // stable insertion: shift strictly greater keys up
                const int32_t key = agg[A.cols[e]];
                const double v = A.vals[e];
                int p = k - 1;
                while (p >= 0 && skey[p * GAL_BLK + tid] > key) {
                    skey[(p + 1) * GAL_BLK + tid] = skey[p * GAL_BLK + tid];
                    sval[(p + 1) * GAL_BLK + tid] = sval[p * GAL_BLK + tid];
                    --p;
                }
                skey[(p + 1) * GAL_BLK + tid] = key;
                sval[(p + 1) * GAL_BLK + tid] = v;
            }
        }
        for (int p = 0; p < m;) {
            const int32_t key = skey[p * GAL_BLK + tid];
            double v = 0.0;
            while (p < m && skey[p * GAL_BLK + tid] == key) v += sval[(p++) * GAL_BLK + tid];
            if (FILL) {
                ccols[out] = key;
                cvals[out] = (float)v;
                ++out;
                if (key == I) diag = v;
            }
            ++count;
        }
    } else {
        int64_t last = -1;
        for (;;) {
            int64_t key = INT64_MAX;  // smallest coarse column > last
            for (int64_t q = m0; q < m1; ++q) {
                const int64_t r = mrows[q];
                for (int64_t e = A.ptr[r]; e < A.ptr[r + 1]; ++e) {
                    const int64_t c = agg[A.cols[e]];
                    if (c > last && c < key) key = c;
                }
            }
            if (key == INT64_MAX) break;
            if (FILL) {
                double v = 0.0;
                for (int64_t q = m0; q < m1; ++q) {
                    const int64_t r = mrows[q];
                    for (int64_t e = A.ptr[r]; e < A.ptr[r + 1]; ++e)
                        if (agg[A.cols[e]] == key) v += A.vals[e];
                }
                ccols[out] = (int32_t)key;
                cvals[out] = (float)v;
                ++out;
                if (key == I) diag = v;
            }
            ++count;
            last = key;
        }
    }
    if (!FILL) cnt[I] = count;
    else cdiag[I] = diag;
}

// The same Galerkin rows, one thread a row with the distinct columns in registers (up to SL of them): each
// entry adds into the slot of its column, in row order, so the sums equal k_galerkin's.  A row that fits is counted
// and staged in its slots' order (the order in which its columns first appear: the coarse rows are not sorted, and
// nothing that reads them needs it) at tcols / tvals[k * nrow + x], slot-major over the stage's rows, and its
// diagonal goes to cdiag; k_galerkin_unstage writes it out once the row offsets are known.  A row with more
// distinct columns is marked tcols[x] = -1 and listed in `over` (counted by nover) for the next stage: 16 slots
// over all rows, then 32 over the rows listed by the first, then k_galerkin (the search grows with the slots, and
// most coarse rows have at most 16 distinct columns).
constexpr int GR_SLOTS_FIRST = 16;
constexpr int GR_SLOTS = 32;
constexpr int GR_BATCH = 8;  // entries whose loads are issued together

template <int SL>
__global__ __launch_bounds__(BLK) void k_galerkin_stage(int64_t nrow, const int64_t* __restrict__ list, SetupCsr A,
                                                        const int64_t* __restrict__ mptr,
                                                        const int64_t* __restrict__ mrows,
                                                        const int32_t* __restrict__ agg, int64_t* __restrict__ cnt,
                                                        double* __restrict__ cdiag,
                                                        unsigned long long* __restrict__ nover,
                                                        int64_t* __restrict__ over, int32_t* __restrict__ tcols,
                                                        float* __restrict__ tvals) {
    const int64_t x = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (x >= nrow) return;
    const int64_t I = list ? list[x] : x;
    int32_t key[SL];
    double val[SL];
#pragma unroll
    for (int k = 0; k < SL; ++k) {
        key[k] = INT32_MAX;
        val[k] = 0.0;
    }
    int n = 0;
    bool full = false;
    // add a_ij into the slot of column c = agg[j] (a new slot for a new column; full when none is left)
    const auto merge = [&](int32_t c, double v) {
        bool hit = false;
#pragma unroll
        for (int k = 0; k < SL; ++k)
            if (key[k] == c) {
                val[k] += v;
                hit = true;
            }
        if (!hit) {
            if (n < SL) {
#pragma unroll
                for (int k = 0; k < SL; ++k)
                    if (k == n) {
                        key[k] = c;
                        val[k] += v;
                    }
                ++n;
            } else {
                full = true;
            }
        }
    };
    const int64_t m1 = mptr[I + 1];
    for (int64_t q = mptr[I]; q < m1 && !full; ++q) {
        const int64_t r = mrows[q];
        const int64_t e1 = A.ptr[r + 1];
        // entries in batches of GR_BATCH: the column loads, then the aggregate lookups, are issued together with
        // no branch (past the row's end the index repeats its last entry, merged again as + 0.0: its slot exists,
        // and a slot's sum of nonzero entries from +0.0 is never -0.0, so adding +0.0 changes nothing; loads used
        // only under a branch would be moved into it and serialised)
        for (int64_t e0 = A.ptr[r]; e0 < e1; e0 += GR_BATCH) {
            int32_t cb[GR_BATCH];
            float vb[GR_BATCH];
#pragma unroll
            for (int u = 0; u < GR_BATCH; ++u) {
                const int64_t e = min(e0 + u, e1 - 1);
                cb[u] = A.cols[e];
                vb[u] = A.vals[e];
            }
#pragma unroll
            for (int u = 0; u < GR_BATCH; ++u) cb[u] = agg[cb[u]];
#pragma unroll
            for (int u = 0; u < GR_BATCH; ++u) merge(cb[u], e0 + u < e1 ? (double)vb[u] : 0.0);
        }
    }
    if (full) {
        cnt[I] = 0;
        over[atomicAdd(nover, 1ull)] = I;
        tcols[x] = -1;
        return;
    }
    cnt[I] = n;
    double diag = 0.0;
#pragma unroll
    for (int k = 0; k < SL; ++k) {
        if (k < n) {
            tcols[(int64_t)k * nrow + x] = key[k];
            tvals[(int64_t)k * nrow + x] = (float)val[k];
            if (key[k] == I) diag = val[k];
        }
    }
    cdiag[I] = diag;
}

// the staged rows of a k_galerkin_stage pass (rows list[x], or x) out to the coarse CSR at their offsets
__global__ __launch_bounds__(BLK) void k_galerkin_unstage(int64_t nrow, const int64_t* __restrict__ list,
                                                          const int32_t* __restrict__ tcols,
                                                          const float* __restrict__ tvals,
                                                          const int64_t* __restrict__ cptr, int32_t* __restrict__ ccols,
                                                          float* __restrict__ cvals) {
    const int64_t x = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (x >= nrow || tcols[x] < 0) return;  // not staged: a later stage writes the row
    const int64_t I = list ? list[x] : x;
    const int64_t out = cptr[I], m = cptr[I + 1] - out;
    for (int64_t k = 0; k < m; ++k) {
        ccols[out + k] = tcols[k * nrow + x];
        cvals[out + k] = tvals[k * nrow + x];
    }
}

// first and last chart of every BLK rows of a level
__global__ __launch_bounds__(BLK) void k_block_charts(int64_t n, const int32_t* __restrict__ chart,
                                                      int2* __restrict__ blk) {
    const int64_t b = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (b * BLK >= n) return;
    blk[b] = make_int2(chart[b * BLK], chart[min(n, (b + 1) * BLK) - 1]);
}

// the finest operator's values for the cycle's smoothing legs, 8 a thread (two 16-byte loads, one 16-byte
// store; both buffers 16-byte aligned)
constexpr int BF16_VEC = 8;
__global__ __launch_bounds__(BLK) void k_to_bf16(int64_t n, const float* __restrict__ in, Bf16* __restrict__ out) {
    const int64_t i = ((int64_t)blockIdx.x * BLK + threadIdx.x) * BF16_VEC;
    if (i + BF16_VEC <= n) {
        const float4 a = *reinterpret_cast<const float4*>(in + i);
        const float4 b = *reinterpret_cast<const float4*>(in + i + 4);
        const float f[BF16_VEC] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        Bf16 h[BF16_VEC];
#pragma unroll
        for (int k = 0; k < BF16_VEC; ++k) cput(h + k, (double)f[k]);
        uint4 o;
        o.x = (uint32_t)h[0].bits | ((uint32_t)h[1].bits << 16);
        o.y = (uint32_t)h[2].bits | ((uint32_t)h[3].bits << 16);
        o.z = (uint32_t)h[4].bits | ((uint32_t)h[5].bits << 16);
        o.w = (uint32_t)h[6].bits | ((uint32_t)h[7].bits << 16);
        *reinterpret_cast<uint4*>(out + i) = o;
    } else {
        for (int64_t k = i; k < n; ++k) cput(out + k, (double)in[k]);
    }
}

// the same over listed ranges of entries (first, count), one workgroup per range: the sliced-ELL entries of the
// charts not in band form
__global__ __launch_bounds__(BLK) void k_to_bf16_ranges(const int64_t* __restrict__ rng, const float* __restrict__ in,
                                                        Bf16* __restrict__ out) {
    const int64_t a = rng[2 * blockIdx.x], n = rng[2 * blockIdx.x + 1];
    for (int64_t k = threadIdx.x; k < n; k += BLK) cput(out + a + k, (double)in[a + k]);
}

// int64 -> int32 (row indices of one level, below 2^31)
__global__ __launch_bounds__(BLK) void k_narrow(int64_t n, const int64_t* __restrict__ in, int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) out[i] = (int32_t)in[i];
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

// ---------------------------------------------------------------- stored row order of the coarse levels
// The cycle stores every level between the finest and the coarsest with the member rows of each next-level
// aggregate contiguous, the aggregates in the next level's stored order (the coarsest keeps the order of its
// construction; the order is laid down from the coarsest level towards the finest).  Then the restriction of a
// level is a sum over runs of consecutive rows, done in the pass that forms the residual, and the prolongation
// reads the coarse correction in ascending order.  A level's rows are cut into slices that start where an
// aggregate starts: slice b holds rows srow[b] .. srow[b + 1], srow[b] = the first aggregate start at or after
// b * CSLICE_STEP, and sagg[b] is that aggregate.  Permutations map a row of the constructed order to its stored row;
// a null permutation is the identity.  Charts keep their row ranges (aggregates never leave their chart).
// The slice grid steps by CSLICE_STEP = SLICE - 4 rows, so that a slice (its grid step plus the rest of the last
// aggregate it starts, at most 3 rows for aggregates of up to 4 rows) fits one row a thread of a workgroup; a
// slice of more rows has its first threads take a second row.
constexpr int CSLICE_STEP = SLICE - 4;
constexpr int MAX_SLICE_ROWS = 4 * SLICE;  // rows of one slice, bounded by the LDS of the restriction

// sz[J] = member count of aggregate I, J = its stored index
__global__ __launch_bounds__(BLK) void k_perm_sizes(int64_t nI, const int32_t* __restrict__ pos,
                                                    const int64_t* __restrict__ mptr, int64_t* __restrict__ sz) {
    const int64_t I = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (I < nI) sz[pos ? pos[I] : I] = mptr[I + 1] - mptr[I];
}

constexpr int PE_U = 4;
// the k-th member m of aggregate I (stored at J) goes to stored row start[J] + k: rpos[m] = start[J] + k, and
// mout[start[J] + k] = m (each skipped when null)
__global__ __launch_bounds__(BLK) void k_perm_expand(int64_t nI, const int32_t* __restrict__ pos,
                                                     const int64_t* __restrict__ mptr, const int64_t* __restrict__ mrows,
                                                     const int64_t* __restrict__ start, int32_t* __restrict__ rpos,
                                                     int32_t* __restrict__ mout) {
    const int64_t I = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (I >= nI) return;
    const int64_t J = pos ? pos[I] : I, s0 = start[J], m0 = mptr[I], m1 = mptr[I + 1];
    // members PE_U at a time, their loads issued together; past the last member the index repeats it (the same
    // values stored again)
    for (int64_t k0 = 0; k0 < m1 - m0; k0 += PE_U) {
        int64_t m[PE_U];
#pragma unroll
        for (int u = 0; u < PE_U; ++u) m[u] = mrows[m0 + min(k0 + u, m1 - m0 - 1)];
#pragma unroll
        for (int u = 0; u < PE_U; ++u) {
            const int64_t k = min(k0 + u, m1 - m0 - 1);
            if (rpos) rpos[m[u]] = (int32_t)(s0 + k);
            if (mout) mout[s0 + k] = (int32_t)m[u];
        }
    }
}

// out[stored row of i] = stored index of agg[i]
__global__ __launch_bounds__(BLK) void k_perm_agg(int64_t n, const int32_t* __restrict__ agg,
                                                  const int32_t* __restrict__ rpos, const int32_t* __restrict__ npos,
                                                  int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const int32_t a = agg[i];
    out[rpos ? rpos[i] : i] = npos ? npos[a] : a;
}

__global__ __launch_bounds__(BLK) void k_perm_inverse(int64_t n, const int32_t* __restrict__ rpos,
                                                      int32_t* __restrict__ ipos) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) ipos[rpos[i]] = (int32_t)i;
}

// first aggregate start at or after target (start ascends over nI + 1 entries, start[nI] = n)
__device__ __forceinline__ int64_t first_start_at(int64_t target, int64_t nI, const int64_t* __restrict__ start) {
    int64_t lo = 0, hi = nI;  // smallest J with start[J] >= target; start[nI] = n >= target
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (start[mid] >= target) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}
// srow, sagg for b = 0 .. nsl; 1 into *tall when a slice holds more than MAX_SLICE_ROWS rows
__global__ __launch_bounds__(BLK) void k_slice_rows(int64_t nsl, int64_t n, int64_t nI,
                                                    const int64_t* __restrict__ start, int32_t* __restrict__ srow,
                                                    int32_t* __restrict__ sagg, int32_t* __restrict__ tall) {
    const int64_t b = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (b > nsl) return;
    const int64_t J = first_start_at(min(b * CSLICE_STEP, n), nI, start);
    srow[b] = (int32_t)start[J];
    sagg[b] = (int32_t)J;
    if (b < nsl) {
        const int64_t J1 = first_start_at(min((b + 1) * CSLICE_STEP, n), nI, start);
        if (start[J1] - start[J] > MAX_SLICE_ROWS) atomicOr(tall, 1);
    }
}

// slice of stored row p: the largest b with srow[b] <= p (empty slices are passed over).  srow[b] >=
// b * CSLICE_STEP, so it is p / CSLICE_STEP or, where p lies before that slice's first aggregate start, one of the
// slices before it (srow[0] = 0 ends the walk).
__device__ __forceinline__ int64_t slice_of(int64_t p, int64_t nsl, const int32_t* __restrict__ srow) {
    int64_t b = min(p / CSLICE_STEP, nsl - 1);
    while (srow[b] > p) --b;
    return b;
}

// one workgroup per slice: width (longest row) and size of the slice's ELL block; first and last chart
__global__ __launch_bounds__(BLK) void k_pslice_sizes(const int32_t* __restrict__ srow, const int32_t* __restrict__ ipos,
                                                      const int64_t* __restrict__ ptr, const int32_t* __restrict__ chart,
                                                      int32_t* __restrict__ w, int64_t* __restrict__ sz,
                                                      int2* __restrict__ sblk) {
    __shared__ double sh[WAVES];
    const int64_t b = blockIdx.x, r0 = srow[b], r1 = srow[b + 1];
    double m = 0.0;
    for (int64_t p = r0 + threadIdx.x; p < r1; p += BLK) {
        const int64_t i = ipos ? ipos[p] : p;
        m = fmax(m, (double)(ptr[i + 1] - ptr[i]));
    }
    m = block_max_all(m, sh);
    if (threadIdx.x == 0) {
        w[b] = (int32_t)m;
        sz[b] = (int64_t)m * (r1 - r0);
        sblk[b] = (r1 > r0) ? make_int2(chart[r0], chart[r1 - 1]) : make_int2(0, -1);
    }
}

// 1 into *wide when some column is more than 32767 stored rows from its row
__global__ __launch_bounds__(BLK) void k_poffsets_wide(int64_t n, const int64_t* __restrict__ ptr,
                                                       const int32_t* __restrict__ cols, const int32_t* __restrict__ rpos,
                                                       int32_t* __restrict__ wide) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const int64_t p = rpos ? rpos[i] : i;
    const int64_t a0 = ptr[i], a1 = ptr[i + 1];
    // entries 8 at a time: columns, then their stored rows, each group of loads issued together (past the row's
    // end the index repeats its last entry)
    bool far = false;
    for (int64_t ab = a0; ab < a1; ab += 8) {
        int32_t c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) c[u] = cols[min(ab + u, a1 - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int64_t d = (int64_t)(rpos ? rpos[c[u]] : c[u]) - p;
            far = far || d > 32767 || d < -32767;
        }
    }
    if (far) atomicOr(wide, 1);
}

// 1 / sqrt(d) of a diagonal entry, 0 for a non-positive one (a row without couplings keeps zeros)
__device__ __forceinline__ double rsqrt_or0(double d) { return d > 0.0 ? 1.0 / sqrt(d) : 0.0; }

// the operator's rows into the ELL blocks of their stored slices, scaled symmetrically (a_ij / sqrt(d_i d_j), the
// product of the two factors formed first so that (i, j) and (j, i) round alike), columns as stored rows (CI 0:
// 16-bit offsets from the row; CI 1: 32-bit rows); padding: value 0, the row itself.  One thread per stored row p
// (constructed row ipos[p]), so the stores of a slot are contiguous.
// CI 2 (FP32 cycle, 16-bit offsets): value and offset packed in one word, the bfloat16 value in the high half and
// the offset in the low half (ec, uint32), so a slot is one load; ev is unused.
constexpr int PF_U = 8;
template <typename T, int CI>
__global__ __launch_bounds__(BLK) void k_pell_fill(int64_t n, SetupCsr A, const double* __restrict__ diag,
                                                   const int32_t* __restrict__ rpos,
                                                   const int32_t* __restrict__ ipos, int64_t nsl,
                                                   const int32_t* __restrict__ srow, const int64_t* __restrict__ off,
                                                   const int32_t* __restrict__ w,
                                                   typename CoarseVal<T>::type* __restrict__ ev, void* __restrict__ ec) {
    const int64_t p = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (p >= n) return;
    const int64_t i = ipos ? ipos[p] : p;
    const int64_t b = slice_of(p, nsl, srow);
    const int64_t ns = srow[b + 1] - srow[b];
    const int64_t e0 = off[b] + (p - srow[b]);
    const double si = rsqrt_or0(diag[i]);
    const int64_t a0 = A.ptr[i], a1 = A.ptr[i + 1];
    // PF_U entries at a time: their columns and values, then the columns' stored rows and diagonals, each group
    // of loads issued together; past the row's end the index repeats its last entry, whose slot is then written
    // again with the same word (so the stores need no branch)
    const auto batch = [&](int64_t ab, auto perm) {
        constexpr bool P = decltype(perm)::value;
        int32_t cj[PF_U];
        float vj[PF_U];
#pragma unroll
        for (int u = 0; u < PF_U; ++u) {
            const int64_t a = min(ab + u, a1 - 1);
            cj[u] = A.cols[a];
            vj[u] = A.vals[a];
        }
        int64_t cs[PF_U];
        double dj[PF_U];
#pragma unroll
        for (int u = 0; u < PF_U; ++u) {
            cs[u] = P ? rpos[cj[u]] : cj[u];
            dj[u] = diag[cj[u]];
        }
#pragma unroll
        for (int u = 0; u < PF_U; ++u) {
            const int64_t k = min(ab + u, a1 - 1) - a0;
            const int64_t c = cs[u];
            const double v = (double)vj[u] * (si * rsqrt_or0(dj[u]));
            if constexpr (CI == 2) {
                Bf16 h;
                cput(&h, v);
                static_cast<uint32_t*>(ec)[e0 + k * ns] =
                    ((uint32_t)h.bits << 16) | (uint32_t)(uint16_t)(int16_t)(c - p);
            } else {
                cput(ev + e0 + k * ns, v);
                if constexpr (CI == 0) static_cast<int16_t*>(ec)[e0 + k * ns] = (int16_t)(c - p);
                else static_cast<int32_t*>(ec)[e0 + k * ns] = (int32_t)c;
            }
        }
    };
    if (rpos)
        for (int64_t ab = a0; ab < a1; ab += PF_U) batch(ab, std::true_type{});
    else
        for (int64_t ab = a0; ab < a1; ab += PF_U) batch(ab, std::false_type{});
    for (int64_t k = a1 - a0; k < w[b]; ++k) {
        if constexpr (CI == 2) {
            static_cast<uint32_t*>(ec)[e0 + k * ns] = 0u;
            continue;
        } else {
            cput(ev + e0 + k * ns, 0.0);
        }
        if constexpr (CI == 0) static_cast<int16_t*>(ec)[e0 + k * ns] = 0;
        else static_cast<int32_t*>(ec)[e0 + k * ns] = (int32_t)p;
    }
}

// sqrt(d) and 1 / sqrt(d) of a scaled level's diagonal at the stored rows (zeros for a non-positive entry)
template <typename T>
__global__ __launch_bounds__(BLK) void k_pscale(int64_t n, const double* __restrict__ d,
                                                const int32_t* __restrict__ rpos, T* __restrict__ q,
                                                T* __restrict__ rs) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const int64_t p = rpos ? rpos[i] : i;
    const double r = rsqrt_or0(d[i]);
    q[p] = (T)(r > 0.0 ? sqrt(d[i]) : 0.0);
    rs[p] = (T)r;
}

// ---------------------------------------------------------------- cycle
template <typename T>
struct CycLevel {
    int64_t n;
    const int64_t* eoff;     // sliced ELL of the level operator (levels between the finest and the coarsest)
    const int32_t* ew;
    const void* ecols;       // int16 offsets (CI 0), int32 columns (CI 1) or packed bfloat16 value + offset (CI 2)
    const typename CoarseVal<T>::type* evals;
    const int32_t* srow;     // slice row starts (aligned to aggregate starts)
    const int32_t* sagg;     // first next-level aggregate of every slice
    const int2* sblk;        // per slice: first and last chart
    const int32_t* slist;    // slices of the running charts, one per workgroup (null: workgroup b is slice b)
    const int32_t* blist;    // the same for BLK-row blocks
    const T* dinv;
    const int32_t* agg;      // row -> aggregate on the next level
    const int32_t* mptr;     // next level's aggregates -> member rows here (levels >= 1: first stored row)
    const int32_t* mrows;    // (the finest level: the member rows)
    T* b;                    // right-hand side on this level
    T* z;                    // correction on this level
    T* t;                    // work vector on this level
    const int32_t* chart;    // row -> chart
    const int2* blk;         // per BLK rows: first and last chart
    const int32_t* done;     // per chart: B_RUNNING while its solve runs
    const T* q;              // sqrt(d_i) and 1 / sqrt(d_i) of the level's diagonal (scaled levels)
    const T* rs;
    const T* rsn;            // 1 / sqrt(d) of the next level when that one is scaled (else null)
};

// Rows of charts whose solve has stopped are skipped on every level: a chart's rows only couple to rows of the
// same chart, and the running charts never read what a stopped chart's rows hold.
// Decided per workgroup of BLK rows from its first and last chart (charts are contiguous on every level): all
// stopped (the workgroup leaves), all running (no per-row test), or mixed (each row tests its own chart).
// The workgroups of a launch cover either every slice (block) of a level or, late in a solve, only those of the
// running charts, listed: workgroup -> slice (block).
__device__ __forceinline__ int64_t wg_item(const int32_t* __restrict__ list) {
    return list ? (int64_t)list[blockIdx.x] : (int64_t)blockIdx.x;
}
__device__ __forceinline__ int block_state(const int2* __restrict__ blk, int64_t b, const int32_t* __restrict__ done) {
    const int2 fl = blk[b];  // (first, last) chart of the rows of slice or block b; last < first: no rows
    int run = 0;
    for (int c = fl.x; c <= fl.y; ++c) run += (done[c] == B_RUNNING) ? 1 : 0;
    return run == 0 ? 0 : (run == fl.y - fl.x + 1 ? 1 : 2);
}
template <typename T>
__device__ __forceinline__ bool live_row(const CycLevel<T>& L, int bs, int64_t i) {
    return bs == 1 || (bs == 2 && L.done[L.chart[i]] == B_RUNNING);
}

// first entry, slice height and width of a row of a level's sliced ELL
struct EllRow {
    int64_t e0, ns;
    int w;
};

// acc[c] += sum over the ELL row of a level operator of a_ij * g(j)[c], the slot loop unrolled for the slice
// width (a switch, uniform over the workgroup) as in row_slots
// slot e of a level operator: its value and column (CI 2: one packed word, bfloat16 value high, offset low)
template <typename T, int CI>
__device__ __forceinline__ T ell_entry(const CycLevel<T>& L, int64_t e, int64_t i, int64_t* col) {
    if constexpr (CI == 2) {
        const uint32_t u = static_cast<const uint32_t*>(L.ecols)[e];
        *col = i + (int64_t)(int16_t)(u & 0xFFFFu);
        return (T)__uint_as_float(u & 0xFFFF0000u);
    } else {
        *col = col_of<CI>(L.ecols, e, i);
        return (T)cval(L.evals[e]);
    }
}
template <int W, typename T, int R, int CI, typename G>
__device__ __forceinline__ void ell_slots(const CycLevel<T>& L, int64_t e0, int64_t ns, int64_t i, G g,
                                          T (&acc)[R]) {
#pragma unroll
    for (int k = 0; k < W; ++k) {
        int64_t j;
        const T a = ell_entry<T, CI>(L, e0 + k * ns, i, &j);
        const Vr<T, R> x = g(j);
#pragma unroll
        for (int cc = 0; cc < R; ++cc) acc[cc] = fma(a, x.c[cc], acc[cc]);
    }
}
template <typename T, int R, int CI, typename G>
__device__ __forceinline__ void ell_apply(const CycLevel<T>& L, const EllRow& r, int64_t i, G g, T (&acc)[R]) {
    switch (r.w) {
#define RUV_EL(W) \
    case W: ell_slots<W, T, R, CI>(L, r.e0, r.ns, i, g, acc); break;
        RUV_EL(1) RUV_EL(2) RUV_EL(3) RUV_EL(4) RUV_EL(5) RUV_EL(6) RUV_EL(7) RUV_EL(8) RUV_EL(9) RUV_EL(10)
        RUV_EL(11) RUV_EL(12) RUV_EL(13) RUV_EL(14) RUV_EL(15) RUV_EL(16) RUV_EL(17) RUV_EL(18) RUV_EL(19)
        RUV_EL(20)
#undef RUV_EL
        default:
            for (int k = 0; k < r.w; ++k) {
                int64_t j;
                const T a = ell_entry<T, CI>(L, r.e0 + k * r.ns, i, &j);
                const Vr<T, R> x = g(j);
#pragma unroll
                for (int cc = 0; cc < R; ++cc) acc[cc] = fma(a, x.c[cc], acc[cc]);
            }
    }
}

// Levels between the finest and the coarsest are stored symmetrically scaled: the operator as
// A_s = D^-1/2 A D^-1/2 (unit diagonal, the same bfloat16 value for (i, j) and (j, i)), the right-hand side as
// b_s = D^-1/2 b and the work vector of the post-smoothing as e_s = D^1/2 e; the correction z is stored as it is.
// Then the Jacobi step w D^-1 r is the scalar w r_s and every neighbour gather reads one vector (the unscaled form
// gathers d_j as well), while the cycle is the same operator: t_s = b_s - w A_s b_s = D^-1/2 (b - A w D^-1 b),
// the restriction sums sqrt(d_i) t_s,i = t_i, e_s = w b_s + alpha sqrt(d_i) (P z_next)_i, and
// z = D^-1/2 (e_s + w (b_s - A_s e_s)).  The coarsest level keeps the unscaled form (its exact block solve).
//
// One workgroup per slice: t_s = b_s - A_s (w b_s) over the slice's rows (sqrt(d_i) t_s,i into LDS), then b_next[J]
// = the sum over the members of every next-level aggregate J starting in the slice (its members are the
// consecutive rows mptr[J] .. mptr[J + 1], all in the slice), in member order, times 1 / sqrt(d_next,J) when the
// next level is scaled.
template <typename T, int R, int CI>
__global__ __launch_bounds__(BLK) void kc_down(CycLevel<T> L, T w, T* __restrict__ bnext) {
    __shared__ T tl[MAX_SLICE_ROWS * R];
    const int64_t sb = wg_item(L.slist);
    const int bs = block_state(L.sblk, sb, L.done);
    if (bs == 0) return;
    const int64_t r0 = L.srow[sb], h = L.srow[sb + 1] - r0;
    const int wd = L.ew[sb];
    const int64_t e0 = L.eoff[sb];
    for (int64_t r = threadIdx.x; r < h; r += BLK) {
        const int64_t i = r0 + r;
        Vr<T, R> o;
#pragma unroll
        for (int c = 0; c < R; ++c) o.c[c] = T(0);
        if (bs == 1 || L.done[L.chart[i]] == B_RUNNING) {
            T acc[R];
#pragma unroll
            for (int c = 0; c < R; ++c) acc[c] = T(0);
            ell_apply<T, R, CI>(L, EllRow{e0 + r, h, wd}, i,
                                [&](int64_t j) {
                                    Vr<T, R> x = ld<T, R>(L.b, j);
#pragma unroll
                                    for (int c = 0; c < R; ++c) x.c[c] *= w;
                                    return x;
                                },
                                acc);
            const Vr<T, R> bi = ld<T, R>(L.b, i);
            const T qi = L.q[i];
#pragma unroll
            for (int c = 0; c < R; ++c) o.c[c] = qi * (bi.c[c] - acc[c]);
        }
#pragma unroll
        for (int c = 0; c < R; ++c) tl[r * R + c] = o.c[c];
    }
    __syncthreads();
    const int64_t J0 = L.sagg[sb], J1 = L.sagg[sb + 1];
    for (int64_t J = J0 + threadIdx.x; J < J1; J += BLK) {
        const int64_t m0 = L.mptr[J], m1 = L.mptr[J + 1];
        if (bs == 2 && L.done[L.chart[m0]] != B_RUNNING) continue;
        Vr<T, R> o;
#pragma unroll
        for (int c = 0; c < R; ++c) o.c[c] = T(0);
        for (int64_t m = m0; m < m1; ++m)
#pragma unroll
            for (int c = 0; c < R; ++c) o.c[c] += tl[(m - r0) * R + c];
        if (L.rsn) {
            const T sj = L.rsn[J];
#pragma unroll
            for (int c = 0; c < R; ++c) o.c[c] *= sj;
        }
        st<T, R>(bnext, J, o);
    }
}

constexpr int RESTRICT_U = 4;
// per aggregate I of the next level (chart cnext[I]): b_next[I] = sum over its member rows of t (times
// 1 / sqrt(d_next,I) when the next level is scaled)
template <typename T, int R>
__global__ __launch_bounds__(BLK) void kc_restrict(CycLevel<T> L, int64_t nc, const int32_t* __restrict__ cnext,
                                                   const int2* __restrict__ bnext_blk,
                                                   const int32_t* __restrict__ bnext_list, T* __restrict__ bnext) {
    const int64_t nb = wg_item(bnext_list);
    const int64_t I = nb * BLK + threadIdx.x;
    if (I >= nc) return;
    const int bs = block_state(bnext_blk, nb, L.done);
    if (bs == 0 || (bs == 2 && L.done[cnext[I]] != B_RUNNING)) return;
    T acc[R];
#pragma unroll
    for (int c = 0; c < R; ++c) acc[c] = T(0);
    const int32_t q0 = L.mptr[I], q1 = L.mptr[I + 1];
    const T sI = L.rsn ? L.rsn[I] : T(1);
    // the first RESTRICT_U members' row indices, then their rows, each group of loads issued together without
    // branches (past the last member the index repeats the last member's; an aggregate has at least one); the sum
    // keeps member order
    int32_t m[RESTRICT_U];
#pragma unroll
    for (int u = 0; u < RESTRICT_U; ++u) m[u] = L.mrows[min(q0 + u, q1 - 1)];
    Vr<T, R> tm[RESTRICT_U];
#pragma unroll
    for (int u = 0; u < RESTRICT_U; ++u) tm[u] = ld<T, R>(L.t, m[u]);
#pragma unroll
    for (int u = 0; u < RESTRICT_U; ++u) {
        const T f = (u == 0 || q0 + u < q1) ? T(1) : T(0);  // 1 * t + acc rounds as t + acc
#pragma unroll
        for (int c = 0; c < R; ++c) acc[c] = fma(f, tm[u].c[c], acc[c]);
    }
    for (int32_t q = q0 + RESTRICT_U; q < q1; ++q) {
        const Vr<T, R> t1 = ld<T, R>(L.t, L.mrows[q]);
#pragma unroll
        for (int c = 0; c < R; ++c) acc[c] += t1.c[c];
    }
    Vr<T, R> o;
#pragma unroll
    for (int c = 0; c < R; ++c) o.c[c] = L.rsn ? acc[c] * sI : acc[c];
    st<T, R>(bnext, I, o);
}

// scaled level: t = e_s = w b_s + alpha sqrt(d_i) (P z_next)_i
template <typename T, int R>
__global__ __launch_bounds__(BLK) void kc_prolong(CycLevel<T> L, T w, T alpha, const T* __restrict__ znext) {
    const int64_t b = wg_item(L.blist), i = b * BLK + threadIdx.x;
    if (i >= L.n || !live_row(L, block_state(L.blk, b, L.done), i)) return;
    const T aq = alpha * L.q[i];
    const Vr<T, R> bi = ld<T, R>(L.b, i);
    const Vr<T, R> zc = ld<T, R>(znext, L.agg[i]);
    Vr<T, R> o;
#pragma unroll
    for (int c = 0; c < R; ++c) o.c[c] = fma(w, bi.c[c], aq * zc.c[c]);
    st<T, R>(L.t, i, o);
}

// scaled levels, one workgroup per slice: z = D^-1/2 (e_s + w (b_s - A_s e_s)), e_s in t; or, with FUSED (small
// levels, where a launch costs more than the extra gathers), e_s formed for every row read exactly as kc_prolong
// forms it (same operations, so the same values) instead of a kc_prolong pass before.
template <typename T, int R, int CI, int FUSED = 0>
__global__ __launch_bounds__(BLK) void kc_post(CycLevel<T> L, T w, T alpha = T(0),
                                               const T* __restrict__ znext = nullptr) {
    const int64_t sb = wg_item(L.slist);
    const int bs = block_state(L.sblk, sb, L.done);
    if (bs == 0) return;
    const int64_t r0 = L.srow[sb], h = L.srow[sb + 1] - r0;
    const int wd = L.ew[sb];
    const int64_t e0 = L.eoff[sb];
    for (int64_t r = threadIdx.x; r < h; r += BLK) {
        const int64_t i = r0 + r;
        if (bs == 2 && L.done[L.chart[i]] != B_RUNNING) continue;
        T s[R];
#pragma unroll
        for (int c = 0; c < R; ++c) s[c] = T(0);
        const auto e1 = [&](int64_t j) {
            if (!FUSED) return ld<T, R>(L.t, j);
            const T aq = alpha * L.q[j];
            const Vr<T, R> bj = ld<T, R>(L.b, j);
            const Vr<T, R> zc = ld<T, R>(znext, L.agg[j]);
            Vr<T, R> o;
#pragma unroll
            for (int c = 0; c < R; ++c) o.c[c] = fma(w, bj.c[c], aq * zc.c[c]);
            return o;
        };
        ell_apply<T, R, CI>(L, EllRow{e0 + r, h, wd}, i, e1, s);
        const T si = L.rs[i];
        const Vr<T, R> bi = ld<T, R>(L.b, i);
        const Vr<T, R> ei = e1(i);
        Vr<T, R> o;
#pragma unroll
        for (int c = 0; c < R; ++c) o.c[c] = si * fma(w, bi.c[c] - s[c], ei.c[c]);
        st<T, R>(L.z, i, o);
    }
}

// coarsest level: z = D^-1 b (exact for a chart coarsened to one row)
template <typename T, int R>
__global__ __launch_bounds__(BLK) void kc_coarsest(CycLevel<T> L) {
    const int64_t b = wg_item(L.blist), i = b * BLK + threadIdx.x;
    if (i >= L.n || !live_row(L, block_state(L.blk, b, L.done), i)) return;
    const T d = L.dinv[i];
    const Vr<T, R> bi = ld<T, R>(L.b, i);
    Vr<T, R> o;
#pragma unroll
    for (int c = 0; c < R; ++c) o.c[c] = d * bi.c[c];
    st<T, R>(L.z, i, o);
}

// Coarsest level, each chart's block inverted: z = inv(A_c) b over the rows of the row's chart, accumulated in FP64
// from the FP64 inverse (row-major at minv + moff[c], over the level's rows crow[c] .. crow[c + 1]); a chart with
// moff[c] < 0 (a block too large for the device factorisation) is solved by its diagonal instead.
template <typename T, int R>
__global__ __launch_bounds__(BLK) void kc_coarsest_dense(CycLevel<T> L, const int64_t* __restrict__ crow,
                                                         const int64_t* __restrict__ moff,
                                                         const double* __restrict__ minv) {
    const int64_t b = wg_item(L.blist), i = b * BLK + threadIdx.x;
    if (i >= L.n || !live_row(L, block_state(L.blk, b, L.done), i)) return;
    const int c = L.chart[i];
    const int64_t r0 = crow[c], nc = crow[c + 1] - r0, mo = moff[c];
    Vr<T, R> o;
    if (mo < 0) {
        const T d = L.dinv[i];
        const Vr<T, R> bi = ld<T, R>(L.b, i);
#pragma unroll
        for (int cc = 0; cc < R; ++cc) o.c[cc] = d * bi.c[cc];
        st<T, R>(L.z, i, o);
        return;
    }
    const double* m = minv + mo + (i - r0) * nc;
    double acc[R];
#pragma unroll
    for (int cc = 0; cc < R; ++cc) acc[cc] = 0.0;
    for (int64_t j = 0; j < nc; ++j) {
        const double v = m[j];
        const Vr<T, R> bj = ld<T, R>(L.b, r0 + j);
#pragma unroll
        for (int cc = 0; cc < R; ++cc) acc[cc] = fma(v, (double)bj.c[cc], acc[cc]);
    }
#pragma unroll
    for (int cc = 0; cc < R; ++cc) o.c[cc] = (T)acc[cc];
    st<T, R>(L.z, i, o);
}

// The coarsest operator's block of every chart factored and inverted on the device in FP64, one workgroup per chart
// (blocks of at most CHOL_MAX rows; the host marks larger ones moff[c] < 0 and they keep their diagonal).  The block
// (summed from the coarsest operator's CSR, FP32 values) goes to LDS; a right-looking Cholesky A = G G^T (pivot test
// d > 1e-13 |a_kk| on the original diagonal), the factor inverted in place from its last column (column j of
// H = G^-1 is -H22 g_j / g_jj, H22 the already inverted trailing block), and inv = H^T H written out in FP64.  A
// block that is not numerically positive definite keeps the inverse of its diagonal, and is counted in
// fallback[0]; an entry whose column leaves the chart (which the aggregation never produces) is counted in
// fallback[1] and the chart keeps its diagonal too.
constexpr int CHOL_MAX = 88;  // (88^2 + 2 * 88) doubles of LDS stay below 64 KiB
__global__ __launch_bounds__(BLK) void k_coarsest_chol(SetupCsr A, const int64_t* __restrict__ crow,
                                                       const int64_t* __restrict__ moff, double* __restrict__ minv,
                                                       unsigned long long* __restrict__ fallback) {
    extern __shared__ double sm[];
    __shared__ int bad;
    const int c = blockIdx.x;
    const int64_t r0 = crow[c], mo = moff[c];
    const int m = (int)(crow[c + 1] - r0);
    if (m == 0 || mo < 0) return;
    double* a = sm;              // m x m, row-major, lower triangle used
    double* d0 = sm + m * m;     // original diagonal
    double* col = d0 + m;        // a column of the factor
    const int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    for (int q = tid; q < m * m; q += BLK) a[q] = 0.0;
    __syncthreads();
    for (int i = tid; i < m; i += BLK)
        for (int64_t k = A.ptr[r0 + i]; k < A.ptr[r0 + i + 1]; ++k) {
            const int64_t j = (int64_t)A.cols[k] - r0;
            if (j < 0 || j >= m) bad = 2;  // (any writer; read after the barrier)
            else a[i * m + j] += (double)A.vals[k];
        }
    __syncthreads();
    for (int i = tid; i < m; i += BLK) d0[i] = a[i * m + i];
    __syncthreads();
    for (int k = 0; k < m && bad == 0; ++k) {
        if (tid == 0) {
            const double d = a[k * m + k];
            if (!(d > 1e-13 * fabs(d0[k])) || !isfinite(d)) bad = 1;
            else a[k * m + k] = sqrt(d);
        }
        __syncthreads();
        if (bad) break;
        const double g = a[k * m + k];
        for (int i = k + 1 + tid; i < m; i += BLK) a[i * m + k] /= g;
        __syncthreads();
        const int w = m - k - 1;
        for (int q = tid; q < w * w; q += BLK) {
            const int ii = k + 1 + q / w, jj = k + 1 + q % w;
            if (jj <= ii) a[ii * m + jj] -= a[ii * m + k] * a[jj * m + k];
        }
        __syncthreads();
    }
    double* out = minv + mo;
    if (bad) {
        if (tid == 0) atomicAdd(&fallback[bad == 1 ? 0 : 1], 1ull);
        for (int q = tid; q < m * m; q += BLK) {
            const int i = q / m, j = q % m;
            out[q] = (i == j && d0[i] > 0.0) ? 1.0 / d0[i] : 0.0;
        }
        return;
    }
    for (int j = m - 1; j >= 0; --j) {  // a (lower) becomes H = G^-1
        for (int p = j + 1 + tid; p < m; p += BLK) col[p] = a[p * m + j];
        __syncthreads();
        const double gjj = a[j * m + j];
        for (int i = j + 1 + tid; i < m; i += BLK) {
            double sacc = 0.0;
            for (int p = j + 1; p <= i; ++p) sacc = fma(a[i * m + p], col[p], sacc);
            a[i * m + j] = -sacc / gjj;
        }
        __syncthreads();
        if (tid == 0) a[j * m + j] = 1.0 / gjj;
        __syncthreads();
    }
    for (int q = tid; q < m * m; q += BLK) {  // inv = H^T H
        const int i = q / m, j = q % m;
        double sacc = 0.0;
        for (int p = max(i, j); p < m; ++p) sacc = fma(a[p * m + i], a[p * m + j], sacc);
        out[q] = sacc;
    }
}

// Finest level (the batch's sliced-ELL operator, or the band form of a band chart), per live tile.  b is the PCG
// residual r.
//   down: t = r - A (w D^-1 r) for the tile's rows; kc_restrict then sums t over each aggregate.
// FORM: 0 any tile, 1 tiles of band charts only, 2 the others only (each launch then carries one path's registers)
template <typename T, int R, int CI, int FORM = 0>
__global__ __launch_bounds__(BLK) void kc_resid0(BDev B, const int* __restrict__ act, int na,
                                                 const typename CoarseVal<T>::type* __restrict__ vals,
                                                 const typename CoarseVal<T>::type* __restrict__ bvals,
                                                 const T* __restrict__ dinv, T w,
                                                 const T* __restrict__ r, T* __restrict__ t_out) {
    const int slot = xcd_slot(na);
    if (slot < 0) return;
    const int t = act[slot];
    const int c = B.tile_chart[t];
    if (B.c_done[c] != B_RUNNING) return;
    const int64_t lo = B.crow[c], hi = B.crow[c + 1] - 1;
    const auto g = [&](int64_t j) {
        const T d = w * dinv[j];
        Vr<T, R> x = ld<T, R>(r, j);
#pragma unroll
        for (int cc = 0; cc < R; ++cc) x.c[cc] *= d;
        return x;
    };
    const auto fin = [&](int64_t i, const T (&acc)[R]) {
        const Vr<T, R> ri = ld<T, R>(r, i);
        Vr<T, R> o;
#pragma unroll
        for (int cc = 0; cc < R; ++cc) o.c[cc] = ri.c[cc] - acc[cc];
        st<T, R>(t_out, i, o);
    };
    if (FORM == 1 || (FORM == 0 && B.c_bw && B.c_bw[c] > 0)) {
        const int t0 = B.chart_t0[c];
        band_rows<T, R, 0>(B, c, t - t0, B.chart_t0[c + 1] - t0, bvals, g, fin);
        return;
    }
    const int s1 = B.tile_s0[t + 1];
    for (int s = B.tile_s0[t]; s < s1; ++s) {
        if (s + 1 < s1 && spmv_group<2, T, R, 0>(B, s, lo, hi, vals, g, fin)) {
            ++s;
            continue;
        }
        const int n = B.sl_n[s];
        if ((int)threadIdx.x >= n) continue;
        const int64_t i = B.sl_row0[s] + threadIdx.x;
        T acc[R];
#pragma unroll
        for (int cc = 0; cc < R; ++cc) acc[cc] = T(0);
        row_slots<T, R, CI, 0>(B, s, n, i, lo, hi, vals, g, acc);
        fin(i, acc);
    }
}

//   up, in two passes: e1 = w D^-1 r + alpha P z1 into e (kc_prolong0), then z = e1 + w D^-1 (r - A e1) with the
//   neighbours' e1 read back (kc_post0), partial r.z of the tile into part_c[par]
template <typename T, int R>
__global__ __launch_bounds__(BLK) void kc_prolong0(BDev B, const int* __restrict__ act, int na,
                                                   const T* __restrict__ dinv, T w, T alpha,
                                                   const T* __restrict__ r, const int32_t* __restrict__ agg,
                                                   const T* __restrict__ z1, T* __restrict__ e) {
    const int slot = xcd_slot_rev(na);
    if (slot < 0) return;
    const int t = act[slot];
    if (B.c_done[B.tile_chart[t]] != B_RUNNING) return;
    tile_rows2(B.tile_r0[t], B.tile_r1[t], [&](int64_t i0, int64_t i2, bool two) {
        const int64_t ix[2] = {i0, i2};
        int32_t ag[2];
        T di[2];
        Vr<T, R> ri[2], zc[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            ag[k] = agg[ix[k]];
            di[k] = w * dinv[ix[k]];
            ri[k] = ld<T, R>(r, ix[k]);
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) zc[k] = ld<T, R>(z1, ag[k]);
        Vr<T, R> o[2];
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int cc = 0; cc < R; ++cc) o[k].c[cc] = fma(di[k], ri[k].c[cc], alpha * zc[k].c[cc]);
        st<T, R>(e, i0, o[0]);
        st<T, R>(e, i2, two ? o[1] : o[0]);  // two false: row i0 once more
    });
}

template <typename T, int R, int CI>
__global__ __launch_bounds__(BLK) void kc_post0(BDev B, const int* __restrict__ act, int na, int par,
                                                const typename CoarseVal<T>::type* __restrict__ vals,
                                                const typename CoarseVal<T>::type* __restrict__ bvals,
                                                const T* __restrict__ dinv, T w,
                                                const T* __restrict__ r, const T* __restrict__ e, T* __restrict__ z) {
    __shared__ double sh[WAVES];
    const int slot = xcd_slot(na);
    if (slot < 0) return;
    const int t = act[slot];
    const int c = B.tile_chart[t];
    if (B.c_done[c] != B_RUNNING) return;
    const int64_t lo = B.crow[c], hi = B.crow[c + 1] - 1;
    double arz = 0.0;
    const auto g = [&](int64_t j) { return ld<T, R>(e, j); };
    const auto fin = [&](int64_t i, const T (&acc)[R]) {
        const T di = w * dinv[i];
        const Vr<T, R> ri = ld<T, R>(r, i);
        const Vr<T, R> ei = ld<T, R>(e, i);
        Vr<T, R> o;
#pragma unroll
        for (int cc = 0; cc < R; ++cc) {
            o.c[cc] = fma(di, ri.c[cc] - acc[cc], ei.c[cc]);
            arz += (double)ri.c[cc] * (double)o.c[cc];
        }
        st<T, R>(z, i, o);
    };
    if (B.c_bw && B.c_bw[c] > 0) {
        const int t0 = B.chart_t0[c];
        band_rows<T, R, 0>(B, c, t - t0, B.chart_t0[c + 1] - t0, bvals, g, fin);
        const double srz = block_sum_all(arz, sh);
        if (threadIdx.x == 0) B.part_c[(int64_t)par * B.ntile + t] = srz;
        return;
    }
    const int s1 = B.tile_s0[t + 1];
    for (int s = B.tile_s0[t]; s < s1; ++s) {
        if (s + 1 < s1 && spmv_group<2, T, R, 0>(B, s, lo, hi, vals, g, fin)) {
            ++s;
            continue;
        }
        const int n = B.sl_n[s];
        if ((int)threadIdx.x >= n) continue;
        const int64_t i = B.sl_row0[s] + threadIdx.x;
        T acc[R];
#pragma unroll
        for (int cc = 0; cc < R; ++cc) acc[cc] = T(0);
        row_slots<T, R, CI, 0>(B, s, n, i, lo, hi, vals, g, acc);
        fin(i, acc);
    }
    const double srz = block_sum_all(arz, sh);
    if (threadIdx.x == 0) B.part_c[(int64_t)par * B.ntile + t] = srz;
}

}  // namespace ruv
