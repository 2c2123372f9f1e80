// host_precond.cpp -- host-side preconditioner setup and CSR utilities of libruv_lscm.so.
//
// Restates the reference's setup algorithms (crates/researchuv-gpu/src/precond.rs, amg.rs) in C++; the
// applies run on the device.  Each function cites the reference lines it follows.
#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "ruv_internal.hpp"

namespace ruv {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }
int fail(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}

// CsrMatrix::row_get (solver.rs:36-43)
double Csr::get(int32_t i, int32_t j) const {
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k)
        if (cols[k] == j) return vals[k];
    return 0.0;
}

bool detect_kron2(const Csr& a, Csr* l) {
    if (a.n < 2 || (a.n & 1)) return false;
    const int32_t m = a.n / 2;
    Csr out;
    out.n = m;
    out.ptr.assign(1, 0);
    out.vals.reserve(a.vals.size() / 2);
    out.cols.reserve(a.cols.size() / 2);
    for (int32_t i = 0; i < m; ++i) {
        const int32_t e0 = a.ptr[2 * i], e1 = a.ptr[2 * i + 1], e2 = a.ptr[2 * i + 2];
        if (e1 - e0 != e2 - e1) return false;
        for (int32_t t = 0; t < e1 - e0; ++t) {
            const int32_t cu = a.cols[e0 + t], cv = a.cols[e1 + t];
            if ((cu & 1) || cv != cu + 1 || a.vals[e0 + t] != a.vals[e1 + t]) return false;
            out.cols.push_back(cu / 2);
            out.vals.push_back(a.vals[e0 + t]);
        }
        out.ptr.push_back((int32_t)out.vals.size());
    }
    *l = std::move(out);
    return true;
}

// tri_levels (precond.rs:227-249): a row's level is one past the deepest of its strict-lower dependencies;
// the transpose (dependencies pointing forward) is layered in descending row order.
void tri_levels(const Csr& f, bool descending, std::vector<int32_t>* rows, std::vector<int32_t>* off) {
    const int32_t n = f.n;
    std::vector<int32_t> level(n, 0);
    int32_t nl = 0;
    for (int32_t t = 0; t < n; ++t) {
        const int32_t i = descending ? n - 1 - t : t;
        int32_t lv = 0;
        for (int32_t k = f.ptr[i]; k < f.ptr[i + 1] - 1; ++k) lv = std::max(lv, level[f.cols[k]] + 1);
        level[i] = lv;
        nl = std::max(nl, lv + 1);
    }
    off->assign(nl + 1, 0);
    for (int32_t t = 0; t < n; ++t) (*off)[level[t] + 1]++;
    for (int32_t l = 0; l < nl; ++l) (*off)[l + 1] += (*off)[l];
    rows->assign(n, 0);
    std::vector<int32_t> cur(off->begin(), off->end() - 1);
    // rows within a level in the order the reference pushes them (ascending for L, descending for L^T)
    for (int32_t t = 0; t < n; ++t) {
        const int32_t i = descending ? n - 1 - t : t;
        (*rows)[cur[level[i]]++] = i;
    }
}

// transpose_lower (precond.rs:253-290): rows ascending, diagonal last.
Csr transpose_lower(const Csr& l) {
    const int32_t n = l.n;
    Csr t;
    t.n = n;
    t.ptr.assign(n + 1, 0);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = l.ptr[i]; k < l.ptr[i + 1]; ++k) t.ptr[l.cols[k] + 1]++;
    for (int32_t i = 0; i < n; ++i) t.ptr[i + 1] += t.ptr[i];
    t.cols.assign(l.cols.size(), 0);
    t.vals.assign(l.vals.size(), 0.0);
    std::vector<int32_t> cur(t.ptr.begin(), t.ptr.end() - 1);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = l.ptr[i]; k < l.ptr[i + 1]; ++k) {
            const int32_t j = l.cols[k];
            t.cols[cur[j]] = i;
            t.vals[cur[j]] = l.vals[k];
            cur[j]++;
        }
    // each transposed row starts with its diagonal (the smallest row index holding column j is j): rotate it last
    for (int32_t j = 0; j < n; ++j) {
        const int32_t a = t.ptr[j], b = t.ptr[j + 1];
        if (b - a > 1) {
            std::rotate(t.cols.begin() + a, t.cols.begin() + a + 1, t.cols.begin() + b);
            std::rotate(t.vals.begin() + a, t.vals.begin() + a + 1, t.vals.begin() + b);
        }
    }
    return t;
}

// The structure of factor_ic0's factor (precond.rs:145-219) without its values: L on A's strict-lower pattern,
// columns ascending (a stable sort, as the reference's), the diagonal last; L^T (transpose_lower) and both
// sweeps' levels (tri_levels).  For the device factorisation: src[q], the position in the values of A of each
// strict-lower entry q of L (-1 at the diagonals; positions through val_of, the place of every entry of `a` in
// the caller's value array, null: the identity); dsrc[i], that of row i's diagonal (its first entry, as
// Csr::get reads it; -1 when absent); lt_src[p], the L entry each L^T entry copies.
void ic0_pattern(const Csr& a, const std::vector<int64_t>* val_of, TriFactor* f, std::vector<int64_t>* src,
                 std::vector<int64_t>* dsrc, std::vector<int32_t>* lt_src) {
    const int32_t n = a.n;
    auto pos = [&](int32_t k) { return val_of ? (*val_of)[(size_t)k] : (int64_t)k; };
    Csr& L = f->l;
    L.n = n;
    L.ptr.assign(1, 0);
    L.cols.clear();
    L.vals.clear();
    src->clear();
    dsrc->assign((size_t)n, -1);
    std::vector<std::pair<int32_t, int32_t>> low;
    for (int32_t i = 0; i < n; ++i) {
        low.clear();
        for (int32_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k) {
            if (a.cols[k] < i) low.emplace_back(a.cols[k], k);
            else if (a.cols[k] == i && (*dsrc)[i] < 0) (*dsrc)[i] = pos(k);
        }
        std::stable_sort(low.begin(), low.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        for (const auto& [c, k] : low) {
            L.cols.push_back(c);
            src->push_back(pos(k));
        }
        L.cols.push_back(i);
        src->push_back(-1);  // the diagonal slot (from dsrc)
        L.ptr.push_back((int32_t)L.cols.size());
    }
    // L^T with each entry's L position as its value (exact in a double), then the values dropped
    L.vals.resize(L.cols.size());
    for (size_t q = 0; q < L.vals.size(); ++q) L.vals[q] = (double)q;
    f->lt = transpose_lower(L);
    lt_src->resize(f->lt.vals.size());
    for (size_t p = 0; p < f->lt.vals.size(); ++p) (*lt_src)[p] = (int32_t)f->lt.vals[p];
    L.vals.assign(L.cols.size(), 0.0);
    f->lt.vals.assign(f->lt.cols.size(), 0.0);
    tri_levels(f->l, false, &f->fwd_rows, &f->fwd_off);
    tri_levels(f->lt, true, &f->bwd_rows, &f->bwd_off);
}

// The position of every entry of permute_csr(a, ...) in a's values (the same permutation applied to positions).
std::vector<int64_t> permuted_positions(const Csr& a, const std::vector<int32_t>& to_perm,
                                        const std::vector<int32_t>& to_orig) {
    Csr ai = a;
    for (size_t k = 0; k < ai.vals.size(); ++k) ai.vals[k] = (double)k;
    const Csr pa = permute_csr(ai, to_perm, to_orig);
    std::vector<int64_t> out(pa.vals.size());
    for (size_t k = 0; k < out.size(); ++k) out[k] = (int64_t)pa.vals[k];
    return out;
}

// graph_coloring (precond.rs:52-77): greedy, index order, smallest colour not used by a neighbour.
std::vector<uint32_t> graph_coloring(const Csr& a) {
    const int32_t n = a.n;
    std::vector<uint32_t> color(n, UINT32_MAX);
    std::vector<uint32_t> used;
    for (int32_t v = 0; v < n; ++v) {
        used.clear();
        for (int32_t k = a.ptr[v]; k < a.ptr[v + 1]; ++k) {
            const int32_t c = a.cols[k];
            if (c != v && color[c] != UINT32_MAX) used.push_back(color[c]);
        }
        std::sort(used.begin(), used.end());
        used.erase(std::unique(used.begin(), used.end()), used.end());
        uint32_t k = 0;
        for (uint32_t c : used) {
            if (c == k) ++k;
            else break;
        }
        color[v] = k;
    }
    return color;
}

// color_permutation (precond.rs:81-101)
void color_permutation(const std::vector<uint32_t>& color, std::vector<int32_t>* to_perm, std::vector<int32_t>* to_orig) {
    const size_t n = color.size();
    uint32_t nc = 0;
    for (uint32_t c : color) nc = std::max(nc, c + 1);
    std::vector<int64_t> cur(nc + 1, 0);
    for (uint32_t c : color) cur[c + 1]++;
    for (uint32_t k = 0; k < nc; ++k) cur[k + 1] += cur[k];
    to_perm->assign(n, 0);
    to_orig->assign(n, 0);
    for (size_t v = 0; v < n; ++v) {
        const int64_t pos = cur[color[v]]++;
        (*to_perm)[v] = (int32_t)pos;
        (*to_orig)[pos] = (int32_t)v;
    }
}

// permute_csr (precond.rs:104-123): P A P^T with rows following to_orig, columns sorted.
Csr permute_csr(const Csr& a, const std::vector<int32_t>& to_perm, const std::vector<int32_t>& to_orig) {
    Csr out;
    out.n = a.n;
    out.ptr.assign(1, 0);
    out.vals.reserve(a.vals.size());
    out.cols.reserve(a.cols.size());
    std::vector<std::pair<int32_t, double>> ent;
    for (int32_t pos = 0; pos < a.n; ++pos) {
        const int32_t o = to_orig[pos];
        ent.clear();
        for (int32_t k = a.ptr[o]; k < a.ptr[o + 1]; ++k) ent.emplace_back(to_perm[a.cols[k]], a.vals[k]);
        std::sort(ent.begin(), ent.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        for (const auto& e : ent) {
            out.cols.push_back(e.first);
            out.vals.push_back(e.second);
        }
        out.ptr.push_back((int32_t)out.vals.size());
    }
    return out;
}

}  // namespace ruv
