// host_mesh.cpp -- chart and vertex indexing for the unfold path (host C++).
//
// weld (crates/researchuv-unwrap/src/weld.rs:19-57), segment / build_charts / chart_border_loops
// (segment.rs:130-331).  The indexing contract is bit-exact: charts in order of their lowest unseen face,
// faces in DFS pop order, vertex ids ascending (local id = rank), border loops by the keep-going-straight
// walk with the reference's tie-breaks.  Geometry here is compiled without multiply-add contraction.
#include "host_mesh.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <tuple>
#include <unordered_map>

namespace ruv {

namespace {

inline uint64_t ekey(uint32_t a, uint32_t b) {
    return a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
}
inline uint32_t ka(uint64_t k) { return (uint32_t)(k >> 32); }
inline uint32_t kb(uint64_t k) { return (uint32_t)(k & 0xffffffffu); }

struct V3 {
    double x, y, z;
};
inline V3 at(const double* p, size_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 o) { return {a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x}; }
inline double len(V3 a) { return std::sqrt(dot(a, a)); }

// edge_faces (segment.rs:73-81): unique undirected edges ascending, each with its faces in insertion order.
struct EdgeFaces {
    std::vector<uint64_t> keys;
    std::vector<int64_t> off;
    std::vector<uint32_t> faces;
    int64_t find(uint64_t k) const {
        auto it = std::lower_bound(keys.begin(), keys.end(), k);
        return (it != keys.end() && *it == k) ? (int64_t)(it - keys.begin()) : -1;
    }
};

EdgeFaces edge_faces(const uint32_t* faces, int64_t nf) {
    std::vector<std::pair<uint64_t, uint32_t>> rec;
    rec.reserve(3 * nf);
    for (int64_t f = 0; f < nf; ++f)
        for (int k = 0; k < 3; ++k) rec.emplace_back(ekey(faces[3 * f + k], faces[3 * f + (k + 1) % 3]), (uint32_t)f);
    std::stable_sort(rec.begin(), rec.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    EdgeFaces ef;
    ef.faces.reserve(rec.size());
    for (size_t i = 0; i < rec.size(); ++i) {
        if (i == 0 || rec[i].first != rec[i - 1].first) {
            ef.keys.push_back(rec[i].first);
            ef.off.push_back((int64_t)i);
        }
        ef.faces.push_back(rec[i].second);
    }
    ef.off.push_back((int64_t)rec.size());
    return ef;
}

// segment.rs:104-117 (cross(a-b, c-b), vec.rs:100-102)
V3 face_normal_unit(const double* pos, const uint32_t* f) {
    V3 a = at(pos, f[0]), b = at(pos, f[1]), c = at(pos, f[2]);
    V3 n = cross(sub(a, b), sub(c, b));
    double l = len(n);
    if (l < 1e-30) return {0.0, 0.0, 1.0};
    return {n.x / l, n.y / l, n.z / l};
}

// chart_border_loops (segment.rs:227-331)
std::vector<std::vector<uint32_t>> border_loops(const double* pos, const uint32_t* faces,
                                                const std::vector<uint32_t>& comp, const EdgeFaces& ef,
                                                const std::vector<char>& bnd) {
    std::vector<uint64_t> be;
    for (uint32_t f : comp)
        for (int k = 0; k < 3; ++k) {
            uint64_t key = ekey(faces[3 * f + k], faces[3 * f + (k + 1) % 3]);
            if (bnd[ef.find(key)]) be.push_back(key);
        }
    std::sort(be.begin(), be.end());
    be.erase(std::unique(be.begin(), be.end()), be.end());
    std::vector<std::vector<uint32_t>> loops;
    if (be.empty()) return loops;
    // adjacency in edge-key order (a BTreeMap<u32, Vec<u32>> in the reference)
    std::map<uint32_t, std::vector<uint32_t>> adj;
    for (uint64_t k : be) {
        adj[ka(k)].push_back(kb(k));
        adj[kb(k)].push_back(ka(k));
    }
    std::vector<char> used(be.size(), 0);
    auto uidx = [&](uint32_t x, uint32_t y) -> size_t {
        return (size_t)(std::lower_bound(be.begin(), be.end(), ekey(x, y)) - be.begin());
    };
    for (const auto& [start, ns] : adj) {
        bool has_unused = false;
        for (uint32_t nb : ns)
            if (!used[uidx(start, nb)]) {
                has_unused = true;
                break;
            }
        if (!has_unused) continue;
        std::vector<uint32_t> lp{start};
        uint32_t cur = start;
        uint32_t first = *std::min_element(ns.begin(), ns.end());
        used[uidx(cur, first)] = 1;
        uint32_t prev = cur;
        cur = first;
        lp.push_back(cur);
        for (;;) {
            const auto& cn = adj[cur];
            std::vector<uint32_t> nx;
            for (uint32_t c : cn)
                if (c != prev && !used[uidx(cur, c)]) nx.push_back(c);
            if (nx.empty()) break;
            V3 d_in = sub(at(pos, cur), at(pos, prev));
            double ln = len(d_in);
            uint32_t next = nx[0];
            if (!(ln < 1e-30)) {
                V3 din{d_in.x / ln, d_in.y / ln, d_in.z / ln};
                bool have = false;
                double bd = 0.0;
                for (uint32_t c : nx) {
                    double d = dot(sub(at(pos, c), at(pos, cur)), din);
                    if (have && d <= bd) continue;
                    have = true;
                    bd = d;
                    next = c;
                }
            }
            used[uidx(cur, next)] = 1;
            prev = cur;
            cur = next;
            lp.push_back(cur);
            if (cur == start) break;
        }
        if (lp.size() >= 4) loops.push_back(std::move(lp));
    }
    return loops;
}

}  // namespace

// weld.rs:19-57: canonical id = order of first occurrence of the rounded key; collapsed faces dropped.
void weld(const double* pos, int64_t nv, const uint32_t* faces, int64_t nf, double tol, std::vector<double>* pos_out,
          std::vector<uint32_t>* faces_out) {
    pos_out->clear();
    faces_out->clear();
    if (tol <= 0.0) {
        pos_out->assign(pos, pos + 3 * nv);
        faces_out->assign(faces, faces + 3 * nf);
        return;
    }
    const double inv = 1.0 / tol;
    auto sat = [](double v) -> int64_t {  // Rust `as i64`: saturating, NaN -> 0
        if (v != v) return 0;
        if (v >= 9.2233720368547758e18) return INT64_MAX;
        if (v <= -9.2233720368547758e18) return INT64_MIN;
        return (int64_t)v;
    };
    struct KeyHash {
        size_t operator()(const std::tuple<int64_t, int64_t, int64_t>& k) const {
            uint64_t h = (uint64_t)std::get<0>(k) * 0x9E3779B97F4A7C15ull;
            h ^= (uint64_t)std::get<1>(k) + 0x7F4A7C159E3779B9ull + (h << 6) + (h >> 2);
            h ^= (uint64_t)std::get<2>(k) + 0x94D049BB133111EBull + (h << 6) + (h >> 2);
            return (size_t)h;
        }
    };
    std::unordered_map<std::tuple<int64_t, int64_t, int64_t>, uint32_t, KeyHash> canon;
    canon.reserve((size_t)nv * 2);
    std::vector<uint32_t> remap(nv);
    for (int64_t i = 0; i < nv; ++i) {
        auto key = std::make_tuple(sat(std::round(pos[3 * i] * inv)), sat(std::round(pos[3 * i + 1] * inv)),
                                   sat(std::round(pos[3 * i + 2] * inv)));
        auto it = canon.find(key);
        if (it != canon.end()) {
            remap[i] = it->second;
        } else {
            uint32_t c = (uint32_t)(pos_out->size() / 3);
            canon.emplace(key, c);
            pos_out->insert(pos_out->end(), pos + 3 * i, pos + 3 * i + 3);
            remap[i] = c;
        }
    }
    for (int64_t f = 0; f < nf; ++f) {
        uint32_t a = remap[faces[3 * f]], b = remap[faces[3 * f + 1]], c = remap[faces[3 * f + 2]];
        if (a != b && b != c && a != c) faces_out->insert(faces_out->end(), {a, b, c});
    }
}

// segment (segment.rs:130-140) + build_charts (:143-223)
void segment(const double* pos, int64_t nv, const uint32_t* faces, int64_t nf, double angle_min_deg, Charts* out) {
    (void)nv;
    out->charts.clear();
    out->cut.clear();
    EdgeFaces ef = edge_faces(faces, nf);
    const double th = angle_min_deg * (3.14159265358979323846 / 180.0);  // f64::to_radians
    std::vector<char> is_cut(ef.keys.size(), 0), is_bnd(ef.keys.size(), 0);
    for (size_t e = 0; e < ef.keys.size(); ++e) {
        const int64_t cnt = ef.off[e + 1] - ef.off[e];
        double ang;
        if (cnt == 1) {
            ang = 0.0;
        } else if (cnt == 2) {
            V3 n1 = face_normal_unit(pos, faces + 3 * (size_t)ef.faces[ef.off[e]]);
            V3 n2 = face_normal_unit(pos, faces + 3 * (size_t)ef.faces[ef.off[e] + 1]);
            double d = dot(n1, n2);  // f64::clamp(-1, 1): NaN passes through
            if (d < -1.0) d = -1.0;
            if (d > 1.0) d = 1.0;
            ang = std::acos(d);
        } else {
            ang = 3.14159265358979323846;
        }
        is_cut[e] = ang > th;
        is_bnd[e] = is_cut[e] || cnt == 1;
        if (is_cut[e]) out->cut.push_back(ef.keys[e]);
    }
    std::vector<char> seen(nf, 0);
    std::vector<uint32_t> stack;
    for (int64_t fi = 0; fi < nf; ++fi) {
        if (seen[fi]) continue;
        Chart ch;
        stack.assign(1, (uint32_t)fi);
        seen[fi] = 1;
        while (!stack.empty()) {
            uint32_t f = stack.back();
            stack.pop_back();
            ch.face_ids.push_back(f);
            for (int k = 0; k < 3; ++k) {
                int64_t e = ef.find(ekey(faces[3 * f + k], faces[3 * f + (k + 1) % 3]));
                if (is_cut[e]) continue;
                for (int64_t q = ef.off[e]; q < ef.off[e + 1]; ++q) {
                    uint32_t g = ef.faces[q];
                    if (!seen[g]) {
                        seen[g] = 1;
                        stack.push_back(g);
                    }
                }
            }
        }
        for (uint32_t f : ch.face_ids)
            for (int k = 0; k < 3; ++k) ch.vertex_ids.push_back(faces[3 * f + k]);
        std::sort(ch.vertex_ids.begin(), ch.vertex_ids.end());
        ch.vertex_ids.erase(std::unique(ch.vertex_ids.begin(), ch.vertex_ids.end()), ch.vertex_ids.end());
        auto local = [&](uint32_t v) {
            return (uint32_t)(std::lower_bound(ch.vertex_ids.begin(), ch.vertex_ids.end(), v) - ch.vertex_ids.begin());
        };
        ch.tris.reserve(3 * ch.face_ids.size());
        for (uint32_t f : ch.face_ids)
            for (int k = 0; k < 3; ++k) ch.tris.push_back(local(faces[3 * f + k]));
        for (uint32_t f : ch.face_ids)
            for (int k = 0; k < 3; ++k) {
                uint32_t x = faces[3 * f + k], y = faces[3 * f + (k + 1) % 3];
                if (is_bnd[ef.find(ekey(x, y))]) {
                    ch.border.push_back(x);
                    ch.border.push_back(y);
                }
            }
        std::sort(ch.border.begin(), ch.border.end());
        ch.border.erase(std::unique(ch.border.begin(), ch.border.end()), ch.border.end());
        ch.loops = border_loops(pos, faces, ch.face_ids, ef, is_bnd);
        double a = 0.0;
        for (uint32_t f : ch.face_ids) {
            V3 p0 = at(pos, faces[3 * f]), p1 = at(pos, faces[3 * f + 1]), p2 = at(pos, faces[3 * f + 2]);
            a += 0.5 * len(cross(sub(p1, p0), sub(p2, p0)));  // face_areas3d, segment.rs:61-70
        }
        ch.area3d = a;
        out->charts.push_back(std::move(ch));
    }
}

}  // namespace ruv
