// host_lscm.cpp -- the host side of the per-chart LSCM driver (lscm::unfold_chart with the GPU backend).
//
// Tangent frame (crates/researchuv-unwrap/src/lscm.rs:152-183, lscm_eig.rs:13-66), pins (border_targets
// :280-413, rect_point :187-204, anchor_ids :210-247, dedupe :533-542), cotan weights (:252-274) and the
// free-block assembly (:582-666).  The assembled free block is A = L (x) I2 in interleaved (u, v) order; it
// is produced here as L (nfree rows, ascending columns, the diagonal at its sorted position) plus the
// interleaved right-hand side, which is bitwise the Sparse::to_csr(2*nfree) form of the reference
// (sparse.rs:69-84) read on even rows.  Compiled without multiply-add contraction.
#include "host_lscm.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

namespace ruv {

namespace {

constexpr double EPS_DEGENERATE = 1.1754943508222875e-38;  // lscm.rs:36
constexpr double STAGNATION = 0.01;                          // lscm.rs:44
constexpr int DEFAULT_MAX_ITER = 100;                        // lscm.rs:46
constexpr int DEFAULT_MAX_ITER_GUARD = 10;                   // lscm.rs:48

struct V3 {
    double x, y, z;
};
inline V3 at(const double* p, size_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 o) { return {a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x}; }
inline double len(V3 a) { return std::sqrt(dot(a, a)); }

// f64::total_cmp key
inline int64_t total_key(double x) {
    int64_t b;
    std::memcpy(&b, &x, 8);
    return b ^ (int64_t)(((uint64_t)(b >> 63)) >> 1);
}

// rect_point (lscm.rs:187-204)
void rect_point(double f, double a, double b, double* out) {
    const double p2 = 2.0 * (a + b);
    double t = std::fmod(f, 1.0) * p2;
    const double x0 = a / 2.0, y0 = -b / 2.0;
    if (t < a) {
        out[0] = x0 - t;
        out[1] = y0;
        return;
    }
    t -= a;
    if (t < b) {
        out[0] = -a / 2.0;
        out[1] = y0 + t;
        return;
    }
    t -= b;
    if (t < a) {
        out[0] = -a / 2.0 + t;
        out[1] = b / 2.0;
        return;
    }
    t -= a;
    out[0] = a / 2.0;
    out[1] = b / 2.0 - t;
}

using PinMap = std::map<uint32_t, std::pair<double, double>>;  // BTreeMap<u32, Vec2>: insert overwrites

// border_targets (lscm.rs:280-413)
void border_targets(const double* pos, const Chart& ch, PinMap* pins) {
    const size_t n = ch.loops.size();
    const double area3d = ch.area3d;
    std::vector<std::vector<double>> fracs(n);
    std::vector<double> perims(n);
    for (size_t li = 0; li < n; ++li) {
        const auto& l = ch.loops[li];
        std::vector<double> segs;
        for (size_t k = 0; k + 1 < l.size(); ++k) segs.push_back(len(sub(at(pos, l[k + 1]), at(pos, l[k]))));
        double tot = -0.0;
        for (double s : segs) tot += s;
        const double total = std::max(tot, 1e-30);
        std::vector<double> fr{0.0};
        for (double s : segs) fr.push_back(fr.back() + s);
        for (double& f : fr) f /= total;
        fracs[li] = std::move(fr);
        double ps = -0.0;
        for (double s : segs) ps += s;
        perims[li] = ps;
    }
    if (n == 2) {
        double mx = 0.0;
        for (double p : perims) mx = std::fmax(mx, p);
        if (std::fabs(perims[0] - perims[1]) <= 0.25 * mx) {
            // cylinder (annulus) special case
            const double w = 0.5 * (perims[0] + perims[1]);
            const double h = area3d / std::max(w, 1e-30);
            auto strip = [](const std::vector<uint32_t>& l) {
                std::vector<uint32_t> o = l;
                if (o.size() > 1 && o.front() == o.back()) o.pop_back();
                return o;
            };
            const auto l0 = strip(ch.loops[0]), l1 = strip(ch.loops[1]);
            const auto& f0 = fracs[0];
            for (size_t k = 0; k < l0.size(); ++k) (*pins)[l0[k]] = {-w / 2.0 + w * f0[k], h / 2.0};
            for (uint32_t v1 : l1) {
                const V3 p1 = at(pos, v1);
                size_t best = 0;
                double best_d = INFINITY;
                for (size_t k0 = 0; k0 < l0.size(); ++k0) {
                    const double d = len(sub(at(pos, l0[k0]), p1));
                    if (d < best_d) {
                        best_d = d;
                        best = k0;
                    }
                }
                (*pins)[v1] = {-w / 2.0 + w * f0[best], -h / 2.0};
            }
            return;
        }
    }
    std::vector<std::pair<double, double>> sides(n);
    if (n == 1) {
        const double p = perims[0], a = area3d;
        if (p * p >= 16.0 * a) {
            const double s_ = p / 2.0;
            const double d = std::sqrt(std::max(s_ * s_ - 4.0 * a, 0.0));
            sides[0] = {(s_ + d) / 2.0, (s_ - d) / 2.0};
        } else {
            const double s = std::sqrt(a);
            sides[0] = {s, s};
        }
    } else {
        V3 c{0.0, 0.0, 0.0};
        for (uint32_t v : ch.vertex_ids) c = add(c, at(pos, v));
        const double nv = (double)std::max<size_t>(ch.vertex_ids.size(), 1);
        c = {c.x / nv, c.y / nv, c.z / nv};
        std::vector<double> dists(n);
        for (size_t li = 0; li < n; ++li) {
            double s = 0.0;
            for (uint32_t v : ch.loops[li]) s += len(sub(at(pos, v), c));
            dists[li] = s / (double)std::max<size_t>(ch.loops[li].size(), 1);
        }
        std::vector<size_t> order(n);
        for (size_t i = 0; i < n; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b) { return total_key(dists[a]) < total_key(dists[b]); });
        double denom = 0.0;
        for (size_t k = 0; k < n; ++k) {
            const double sign = ((n - 1 - k) % 2 == 0) ? 1.0 : -1.0;
            denom += sign * perims[order[k]] * perims[order[k]];
        }
        const double s = 4.0 * std::sqrt(area3d / std::max(std::fabs(denom), 1e-30));
        for (size_t li : order) {
            const double side = s * perims[li] / 4.0;
            sides[li] = {side, side};
        }
    }
    for (size_t li = 0; li < n; ++li) {
        const auto [a, b] = sides[li];
        const auto& l = ch.loops[li];
        const bool closed = l.size() > 1 && l.front() == l.back();
        const size_t nvl = closed ? l.size() - 1 : l.size();
        std::vector<double> lp3(3 * l.size()), lp2(2 * l.size());
        for (size_t k = 0; k < l.size(); ++k) std::memcpy(&lp3[3 * k], pos + 3 * (size_t)l[k], 3 * sizeof(double));
        tangent_frame(lp3.data(), (int64_t)l.size(), lp2.data());
        double sa3 = 0.0;
        for (size_t i = 0; i < l.size(); ++i) {
            const size_t j = (i + 1) % l.size();
            sa3 += lp2[2 * i] * lp2[2 * j + 1] - lp2[2 * i + 1] * lp2[2 * j];
        }
        sa3 *= 0.5;
        for (size_t k = 0; k < nvl; ++k) {
            const double f = sa3 > 0.0 ? std::fmod(1.0 - fracs[li][k], 1.0) : fracs[li][k];
            double uv[2];
            rect_point(f, a, b, uv);
            (*pins)[l[k]] = {uv[0], uv[1]};
        }
    }
}

// anchor_ids (lscm.rs:210-247): farthest-point anchors starting at the first min-x vertex, sorted.
std::vector<size_t> anchor_ids(const std::vector<double>& p3, int n) {
    const size_t n3 = p3.size() / 3;
    std::vector<size_t> chosen;
    if (n3 <= (size_t)n) {
        for (size_t i = 0; i < n3; ++i) chosen.push_back(i);
        return chosen;
    }
    double min_x = 1.7976931348623157e308;
    for (size_t i = 0; i < n3; ++i) min_x = std::fmin(min_x, p3[3 * i]);
    size_t st = 0;
    for (size_t i = 0; i < n3; ++i)
        if (p3[3 * i] == min_x) {
            st = i;
            break;
        }
    chosen.push_back(st);
    std::vector<double> d(n3);
    while (chosen.size() < (size_t)n) {
        for (size_t i = 0; i < n3; ++i) {
            double m = 1.7976931348623157e308;
            for (size_t c : chosen) m = std::fmin(m, len(sub(at(p3.data(), i), at(p3.data(), c))));
            d[i] = m;
        }
        for (size_t c : chosen) d[c] = -1.0;
        size_t bi = 0;
        double bd = -1.7976931348623157e308;
        for (size_t i = 0; i < n3; ++i)
            if (d[i] > bd) {
                bi = i;
                bd = d[i];
            }
        chosen.push_back(bi);
    }
    std::sort(chosen.begin(), chosen.end());
    return chosen;
}

}  // namespace

// eig3 (lscm_eig.rs:13-66).  Rotations are accumulated as V' = J V (rows p, q), exactly as the reference
// does; for matrices that need more than one rotation V is then J_n...J_1, an orthonormal frame that is not
// the eigenvector matrix.  The pins' winding follows this frame, so it is kept.
void eig3(const double m[9], double w[3], double vout[9]) {
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = m[3 * i + j];
    static const int pq[2][2] = {{0, 2}, {1, 2}};
    for (int it = 0; it < 64; ++it) {
        int p = 0, q = 1;
        double max_off = std::fabs(a[0][1]);
        for (const auto& e : pq) {
            const double x = std::fabs(a[e[0]][e[1]]);
            if (x > max_off) {
                max_off = x;
                p = e[0];
                q = e[1];
            }
        }
        if (max_off <= 1e-300) break;
        const double apq = a[p][q], app = a[p][p], aqq = a[q][q];
        const double theta = 0.5 * std::atan2(2.0 * apq, app - aqq);
        const double c = std::cos(theta), s = std::sin(theta);
        for (int r = 0; r < 3; ++r) {
            if (r == p || r == q) continue;
            const double arp = c * a[r][p] + s * a[r][q];
            const double arq = -s * a[r][p] + c * a[r][q];
            a[r][p] = arp;
            a[p][r] = arp;
            a[r][q] = arq;
            a[q][r] = arq;
        }
        a[p][p] = c * c * app + 2.0 * s * c * apq + s * s * aqq;
        a[q][q] = s * s * app - 2.0 * s * c * apq + c * c * aqq;
        a[p][q] = 0.0;
        a[q][p] = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double vp = v[p][k], vq = v[q][k];
            v[p][k] = c * vp - s * vq;
            v[q][k] = s * vp + c * vq;
        }
    }
    w[0] = a[0][0];
    w[1] = a[1][1];
    w[2] = a[2][2];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) vout[3 * i + j] = v[i][j];
}

// tangent_frame (lscm.rs:152-183)
void tangent_frame(const double* p3, int64_t n, double* p2) {
    if (n == 0) return;
    V3 c{0.0, 0.0, 0.0};
    for (int64_t i = 0; i < n; ++i) c = add(c, at(p3, i));
    c = {c.x / (double)n, c.y / (double)n, c.z / (double)n};
    double cm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        const V3 d = sub(at(p3, i), c);
        const double vi[3] = {d.x, d.y, d.z};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) cm[3 * a + b] += vi[a] * vi[b];
    }
    double w[3], v[9];
    eig3(cm, w, v);
    int idx[3] = {0, 1, 2};
    std::stable_sort(idx, idx + 3, [&](int a, int b) { return total_key(w[a]) > total_key(w[b]); });
    const V3 e1{v[idx[0]], v[3 + idx[0]], v[6 + idx[0]]};
    const V3 e2{v[idx[1]], v[3 + idx[1]], v[6 + idx[1]]};
    for (int64_t i = 0; i < n; ++i) {
        const V3 d = sub(at(p3, i), c);
        p2[2 * i] = dot(d, e1);
        p2[2 * i + 1] = dot(d, e2);
    }
}

// Tangent frame, pins and the free map of one chart (lscm.rs:513-542, 582-592).
void build_system_pins(const double* pos, const Chart& ch, ChartSystem* S) {
    const int64_t nv = (int64_t)ch.vertex_ids.size();
    S->nv = nv;
    S->p3.resize(3 * nv);
    for (int64_t i = 0; i < nv; ++i) std::memcpy(&S->p3[3 * i], pos + 3 * (size_t)ch.vertex_ids[i], 3 * sizeof(double));
    S->p2.assign(2 * nv, 0.0);
    tangent_frame(S->p3.data(), nv, S->p2.data());
    PinMap pins;
    if (ch.loops.empty()) {
        for (size_t i : anchor_ids(S->p3, 4)) pins[ch.vertex_ids[i]] = {S->p2[2 * i], S->p2[2 * i + 1]};
    } else {
        border_targets(pos, ch, &pins);
    }
    S->pin_ids.clear();
    S->pin_uv.clear();
    std::vector<char> is_free(nv, 1);
    for (const auto& [v, t] : pins) {
        const int64_t li = ch.local_of(v);
        if (li < 0 || !is_free[li]) continue;
        S->pin_ids.push_back((int32_t)li);
        S->pin_uv.push_back(t.first);
        S->pin_uv.push_back(t.second);
        is_free[li] = 0;
    }
    S->free_pos.assign(nv, -1);
    S->free_ids.clear();
    for (int64_t i = 0; i < nv; ++i)
        if (is_free[i]) {
            S->free_pos[i] = (int32_t)S->free_ids.size();
            S->free_ids.push_back((int32_t)i);
        }
    S->nfree = (int32_t)S->free_ids.size();
}

// Pins, cotan weights and the free block of one chart (lscm.rs:512-666; default path B = 1 - angle_mix).
void build_system(const double* pos, const Chart& ch, double angle_mix, ChartSystem* S) {
    build_system_pins(pos, ch, S);
    const int64_t nv = S->nv;
    const int32_t nfree = S->nfree;
    std::vector<char> is_free(nv, 0);
    std::vector<double> pin_at(2 * nv, 0.0);
    for (int32_t q = 0; q < nfree; ++q) is_free[S->free_ids[q]] = 1;
    for (size_t q = 0; q < S->pin_ids.size(); ++q) {
        pin_at[2 * S->pin_ids[q]] = S->pin_uv[2 * q];
        pin_at[2 * S->pin_ids[q] + 1] = S->pin_uv[2 * q + 1];
    }

    // edge_weights (lscm.rs:252-274): contributions per undirected key, summed in triangle order
    const size_t ntri = ch.tris.size() / 3;
    struct Contrib {
        uint64_t key;
        uint32_t ord;
        double cot;
    };
    std::vector<Contrib> con;
    con.reserve(3 * ntri);
    for (size_t f = 0; f < ntri; ++f) {
        const uint32_t t[3] = {ch.tris[3 * f], ch.tris[3 * f + 1], ch.tris[3 * f + 2]};
        for (int q = 0; q < 3; ++q) {
            const uint32_t e0 = t[q], e1 = t[(q + 1) % 3], o = t[(q + 2) % 3];
            const V3 a1 = sub(at(S->p3.data(), e1), at(S->p3.data(), o));
            const V3 a2 = sub(at(S->p3.data(), e0), at(S->p3.data(), o));
            const double la = len(a1), lb = len(a2);
            if (la <= EPS_DEGENERATE || lb <= EPS_DEGENERATE) continue;
            const double s = len(cross(a1, a2)) / (la * lb);
            if (s <= EPS_DEGENERATE) continue;
            double cot = (dot(a1, a2) / (la * lb)) / s;
            cot = cot > 0.0 ? cot : 0.0;  // f64::max(0.0), NaN -> 0
            const uint32_t i = std::min(e0, e1), j = std::max(e0, e1);
            con.push_back({((uint64_t)i << 32) | j, (uint32_t)(3 * f + q), cot});
        }
    }
    std::sort(con.begin(), con.end(), [](const Contrib& x, const Contrib& y) {
        return x.key < y.key || (x.key == y.key && x.ord < y.ord);
    });
    // per-vertex neighbour lists (ascending) with the edge's c, for edges with w > EPS
    std::vector<int64_t> deg(nv + 1, 0);
    std::vector<std::pair<uint64_t, double>> edges;  // key, c
    for (size_t q = 0; q < con.size();) {
        size_t e = q;
        double w = 0.0;
        while (e < con.size() && con[e].key == con[q].key) w += con[e++].cot;
        const uint64_t key = con[q].key;
        q = e;
        if (w <= EPS_DEGENERATE) continue;
        const uint32_t i = (uint32_t)(key >> 32), j = (uint32_t)(key & 0xffffffffu);
        const V3 d3 = sub(at(S->p3.data(), j), at(S->p3.data(), i));
        const double c = w / std::max(dot(d3, d3), EPS_DEGENERATE);
        if (c == 0.0) continue;  // an underflowed weight adds exact zeros, which Sparse::add skips
        edges.emplace_back(key, c);
        deg[i + 1]++;
        deg[j + 1]++;
    }
    for (int64_t i = 0; i < nv; ++i) deg[i + 1] += deg[i];
    std::vector<std::pair<uint32_t, double>> nb(deg[nv]);
    {
        std::vector<int64_t> cur(deg.begin(), deg.end() - 1);
        for (const auto& [key, c] : edges) {  // keys ascending -> every list ascending
            const uint32_t i = (uint32_t)(key >> 32), j = (uint32_t)(key & 0xffffffffu);
            nb[cur[i]++] = {j, c};
            nb[cur[j]++] = {i, c};
        }
    }
    S->edges = edges.size();
    // assembly: L rows over free vertices, rhs interleaved
    const double mix_b = 1.0 - angle_mix;
    Csr& L = S->l;
    L.n = nfree;
    L.ptr.assign(1, 0);
    L.vals.clear();
    L.cols.clear();
    S->rhs.assign(2 * (size_t)nfree, 0.0);
    for (int32_t p = 0; p < nfree; ++p) {
        const int32_t i = S->free_ids[p];
        double diag = 0.0;
        double ru = 0.0, rv = 0.0;
        size_t diag_at = L.vals.size();
        bool placed = false;
        for (int64_t k = deg[i]; k < deg[i + 1]; ++k) {
            const uint32_t j = nb[k].first;
            const double c = nb[k].second;
            const double du = S->p2[2 * j] - S->p2[2 * i], dv = S->p2[2 * j + 1] - S->p2[2 * i + 1];
            diag += c;
            if (is_free[j]) {
                if (!placed && (int64_t)j > i) {
                    diag_at = L.vals.size();
                    L.cols.push_back(p);
                    L.vals.push_back(0.0);
                    placed = true;
                }
                L.cols.push_back(S->free_pos[j]);
                L.vals.push_back(-c);
                ru += mix_b * c * du;
                rv += mix_b * c * dv;
            } else {
                ru += mix_b * c * du + c * pin_at[2 * j];
                rv += mix_b * c * dv + c * pin_at[2 * j + 1];
            }
        }
        if (!placed) {
            diag_at = L.vals.size();
            L.cols.push_back(p);
            L.vals.push_back(0.0);
        }
        L.vals[diag_at] = diag;
        if (diag == 0.0) {  // Sparse::add skips exact zeros: a row with no weighted edge has no diagonal
            L.vals.erase(L.vals.begin() + diag_at);
            L.cols.erase(L.cols.begin() + diag_at);
        }
        L.ptr.push_back((int32_t)L.vals.size());
        S->rhs[2 * p] = ru;
        S->rhs[2 * p + 1] = rv;
    }
}

int64_t Chart::local_of(uint32_t v) const {
    auto it = std::lower_bound(vertex_ids.begin(), vertex_ids.end(), v);
    return (it != vertex_ids.end() && *it == v) ? (int64_t)(it - vertex_ids.begin()) : -1;
}

// The driver iterations of unfold_chart (lscm.rs:602-718) given the free-vertex solution `sol` (2 per free vertex)
// of the chart's system: uv starts at the tangent frame, each iteration replaces it (damped) by the solution with
// the pins at their targets, with the reference's stop rules (lscm.rs:697-718).  The default system (B = 0, no
// area row) does not depend on uv, so every iteration re-assembles the same system, and the solution of the first
// is the solution of all.  uv_out (nv*2); stats (4): iters, err_accum, target, area3d.
void driver_iterations(const ChartSystem& S, const Chart& ch, const double* sol, int max_iter_opt, double* uv_out,
                       double* stats) {
    const int64_t nv = S.nv;
    std::vector<double> uv(S.p2);
    double err_accum = 1.0, step = 1.0, prev_err = INFINITY;
    int iters_done = 0;
    const int max_iter = max_iter_opt > 0 ? max_iter_opt : DEFAULT_MAX_ITER;
    for (int it = 0; it < max_iter; ++it) {
        iters_done = it + 1;
        std::vector<double> new_uv(uv);
        for (int32_t q = 0; q < S.nfree; ++q) {
            new_uv[2 * S.free_ids[q]] = sol[2 * q];
            new_uv[2 * S.free_ids[q] + 1] = sol[2 * q + 1];
        }
        for (size_t q = 0; q < S.pin_ids.size(); ++q) {
            new_uv[2 * S.pin_ids[q]] = S.pin_uv[2 * q];
            new_uv[2 * S.pin_ids[q] + 1] = S.pin_uv[2 * q + 1];
        }
        std::vector<double> dv(2 * nv);
        double sdv = -0.0, suv = -0.0;
        for (int64_t i = 0; i < nv; ++i) {
            dv[2 * i] = new_uv[2 * i] - uv[2 * i];
            dv[2 * i + 1] = new_uv[2 * i + 1] - uv[2 * i + 1];
            sdv += dv[2 * i] * dv[2 * i] + dv[2 * i + 1] * dv[2 * i + 1];
        }
        for (int64_t i = 0; i < nv; ++i) suv += uv[2 * i] * uv[2 * i] + uv[2 * i + 1] * uv[2 * i + 1];
        const double err = std::sqrt(sdv) / std::max(std::sqrt(suv), 1e-30);
        const double applied = err * step;
        if (prev_err > 1e-30 && err > 2.0 * prev_err) step = std::max(step * 0.5, 0.05);
        else if (it > 3) step = std::min(step * 1.2, 1.0);
        if (applied < 1e-8 || (prev_err > 1e-30 && it != 0 && applied / prev_err <= STAGNATION)) break;
        for (int64_t i = 0; i < 2 * nv; ++i) uv[i] = uv[i] + dv[i] * step;
        prev_err = applied;
        err_accum += err;
        if (max_iter * DEFAULT_MAX_ITER_GUARD < it) break;
    }
    std::memcpy(uv_out, uv.data(), sizeof(double) * 2 * nv);
    if (stats) {
        stats[0] = iters_done;
        stats[1] = err_accum;
        stats[2] = ch.area3d;
        stats[3] = ch.area3d;
    }
}

// unfold_chart (lscm.rs:512-727) with the device solve: the system of the first driver iteration solved on the
// device, the iterations run by driver_iterations.
int unfold_chart(const double* pos, const Chart& ch, const ruv_unfold_opts& o, double* uv_out, double* stats,
                 ruv_solve_info* info_out) {
    ChartSystem S;
    if (o.device_assembly) {
        int rc = device_assemble(pos, ch, o.angle_mix, &S);
        if (rc) return rc;
    } else {
        build_system(pos, ch, o.angle_mix, &S);
    }
    // gpu_solve gates (lscm.rs:478-481): size gate, slack weight
    if (2 * (int64_t)S.nfree < o.min_vars || o.mix_w <= 0.0)
        return fail(RUV_ERR_BAD_DIMS, "chart below the GPU size gate (the caller runs its CPU solve)");
    const int max_iter = o.max_iter > 0 ? o.max_iter : DEFAULT_MAX_ITER;
    std::vector<double> sol(2 * (size_t)S.nfree, 0.0);
    ruv_solve_info info{};
    if (S.nfree > 0) {
        const uint64_t mi = (uint64_t)std::max(max_iter, 1024) * 16;  // lscm.rs:501
        int rc = device_solve(S.l, 2, S.rhs.data(), mi, o.tol, o.precond, o.mode, sol.data(), &info);
        if (rc) return rc;
    }
    driver_iterations(S, ch, sol.data(), o.max_iter, uv_out, stats);
    if (info_out) *info_out = info;
    return RUV_OK;
}

}  // namespace ruv
