// host_amg.cpp -- aggregation AMG setup on the host (crates/researchuv-gpu/src/amg.rs:127-455).
//
// The hierarchy is the reference's: matched aggregation (pairs for Amg, SA_AGG = 4 members for AmgSa), the
// tentative or smoothed prolongator P = (I - omega D^-1 A) B truncated to P_MAX_ROW entries per row (own
// aggregate kept), Galerkin coarse operators P^T A P, recursion until a level has at most COARSEST unknowns
// or aggregation stalls.  The coarsest level is factored densely (Cholesky, FP64) and inverted, so its
// device solve is one dense product.  The cycle applied on the device is a symmetric V(1,1) with damped
// Jacobi (see pcg.hip), a fixed SPD operator, so plain PCG stays valid.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "ruv_internal.hpp"

namespace ruv {

namespace {
constexpr int COARSEST_N = 512;      // amg.rs:50
constexpr double OMEGA = 2.0 / 3.0;  // amg.rs:53
constexpr int SA_AGG = 4;            // amg.rs:245
constexpr int P_MAX_ROW = 32;        // amg.rs:299

// merge_sorted (amg.rs:366-376): sort by column, sum duplicates in order
std::vector<std::pair<int32_t, double>> merge_sorted(std::vector<std::pair<int32_t, double>> e) {
    std::stable_sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<std::pair<int32_t, double>> out;
    out.reserve(e.size());
    for (const auto& x : e) {
        if (!out.empty() && out.back().first == x.first) out.back().second += x.second;
        else out.push_back(x);
    }
    return out;
}
}  // namespace

// aggregate (amg.rs:254-291): vertex order scan; each unmatched vertex takes its strongest unmatched
// neighbours (|A(v,u)| descending, ties by smallest u) up to max_members - 1.
void amg_aggregate(const Csr& a, int max_members, std::vector<int32_t>* agg_out, int32_t* n_coarse) {
    const int32_t n = a.n;
    std::vector<int32_t>& agg = *agg_out;
    agg.assign(n, -1);
    int32_t next = 0;
    std::vector<std::pair<double, int32_t>> un;
    for (int32_t v = 0; v < n; ++v) {
        if (agg[v] != -1) continue;
        un.clear();
        std::vector<std::pair<int32_t, double>> cands;
        for (int32_t k = a.ptr[v]; k < a.ptr[v + 1]; ++k)
            if (a.cols[k] != v) cands.emplace_back(a.cols[k], std::fabs(a.vals[k]));
        std::sort(cands.begin(), cands.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        for (const auto& [u, s] : cands)
            if (agg[u] == -1) un.emplace_back(s, u);
        std::stable_sort(un.begin(), un.end(), [](const auto& x, const auto& y) {
            if (x.first != y.first) return x.first > y.first;
            return x.second < y.second;
        });
        const size_t take = std::min<size_t>((size_t)(max_members - 1), un.size());
        std::vector<int32_t> picks;
        for (size_t i = 0; i < take; ++i) picks.push_back(un[i].second);
        std::sort(picks.begin(), picks.end());
        agg[v] = next;
        for (int32_t u : picks) agg[u] = next;
        ++next;
    }
    *n_coarse = next;
}

// prolongator (amg.rs:308-362)
Csr amg_prolongator(const Csr& a, const std::vector<int32_t>& agg, bool smooth) {
    const int32_t n = a.n;
    Csr p;
    p.n = n;
    p.ptr.assign(n + 1, 0);
    for (int32_t v = 0; v < n; ++v) {
        if (!smooth) {
            p.vals.push_back(1.0);
            p.cols.push_back(agg[v]);
        } else {
            const double d = a.get(v, v);
            std::vector<std::pair<int32_t, double>> row;
            for (int32_t k = a.ptr[v]; k < a.ptr[v + 1]; ++k) row.emplace_back(agg[a.cols[k]], -OMEGA * a.vals[k] / d);
            row.emplace_back(agg[v], 1.0);
            auto merged = merge_sorted(std::move(row));
            if ((int)merged.size() <= P_MAX_ROW) {
                for (const auto& [c, w] : merged) {
                    p.cols.push_back(c);
                    p.vals.push_back(w);
                }
            } else {
                std::vector<size_t> order(merged.size());
                for (size_t i = 0; i < order.size(); ++i) order[i] = i;
                std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
                    const double vx = std::fabs(merged[x].second), vy = std::fabs(merged[y].second);
                    if (vx != vy) return vx > vy;
                    return merged[x].first < merged[y].first;
                });
                std::vector<char> keep(merged.size(), 0);
                int budget = P_MAX_ROW - 1;
                for (size_t k : order) {
                    if (merged[k].first == agg[v]) keep[k] = 1;
                    else if (budget > 0) {
                        keep[k] = 1;
                        --budget;
                    }
                }
                for (size_t k = 0; k < merged.size(); ++k)
                    if (keep[k]) {
                        p.cols.push_back(merged[k].first);
                        p.vals.push_back(merged[k].second);
                    }
            }
        }
        p.ptr[v + 1] = (int32_t)p.vals.size();
    }
    return p;
}

// transpose_rect (amg.rs:433-455): coarse rows, fine indices ascending
Csr transpose_rect(const Csr& p, int32_t n_coarse) {
    Csr r;
    r.n = n_coarse;
    r.ptr.assign(n_coarse + 1, 0);
    for (int32_t c : p.cols) r.ptr[c + 1]++;
    for (int32_t k = 0; k < n_coarse; ++k) r.ptr[k + 1] += r.ptr[k];
    r.vals.assign(p.vals.size(), 0.0);
    r.cols.assign(p.cols.size(), 0);
    std::vector<int32_t> cur(r.ptr.begin(), r.ptr.end() - 1);
    for (int32_t v = 0; v < p.n; ++v)
        for (int32_t k = p.ptr[v]; k < p.ptr[v + 1]; ++k) {
            const int32_t g = p.cols[k];
            r.cols[cur[g]] = v;
            r.vals[cur[g]] = p.vals[k];
            cur[g]++;
        }
    return r;
}

// galerkin_p (amg.rs:384-429): W = A P (merged rows), A_c = R W over the rows of R = P^T
Csr amg_galerkin(const Csr& a, const Csr& p, int32_t n_coarse) {
    const int32_t n = a.n;
    std::vector<std::vector<std::pair<int32_t, double>>> w(n);
    for (int32_t i = 0; i < n; ++i) {
        std::vector<std::pair<int32_t, double>> ents;
        for (int32_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k) {
            const int32_t j = a.cols[k];
            const double aij = a.vals[k];
            for (int32_t q = p.ptr[j]; q < p.ptr[j + 1]; ++q) ents.emplace_back(p.cols[q], aij * p.vals[q]);
        }
        w[i] = merge_sorted(std::move(ents));
    }
    const Csr r = transpose_rect(p, n_coarse);
    Csr c;
    c.n = n_coarse;
    c.ptr.assign(n_coarse + 1, 0);
    for (int32_t g = 0; g < n_coarse; ++g) {
        std::vector<std::pair<int32_t, double>> ents;
        for (int32_t k = r.ptr[g]; k < r.ptr[g + 1]; ++k) {
            const int32_t j = r.cols[k];
            const double rg = r.vals[k];
            for (const auto& [col, v] : w[j]) ents.emplace_back(col, rg * v);
        }
        for (const auto& [col, v] : merge_sorted(std::move(ents))) {
            c.cols.push_back(col);
            c.vals.push_back(v);
        }
        c.ptr[g + 1] = (int32_t)c.vals.size();
    }
    return c;
}

namespace {
// Dense Cholesky of the coarsest operator (the loop of factor_exact, precond.rs:447-470) and its inverse
// A^-1 = L^-T L^-1, row-major n x n.  false when a pivot is not positive.
bool dense_inverse(const Csr& a, std::vector<double>* inv) {
    const int32_t n = a.n;
    std::vector<double> m((size_t)n * n, 0.0);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k) m[(size_t)i * n + a.cols[k]] = a.vals[k];
    for (int32_t j = 0; j < n; ++j) {
        double d = m[(size_t)j * n + j];
        for (int32_t k = 0; k < j; ++k) d -= m[(size_t)j * n + k] * m[(size_t)j * n + k];
        if (!(d > 1e-300)) return false;
        const double r = std::sqrt(d);
        m[(size_t)j * n + j] = r;
        for (int32_t i = j + 1; i < n; ++i) {
            double s = m[(size_t)i * n + j];
            for (int32_t k = 0; k < j; ++k) s -= m[(size_t)i * n + k] * m[(size_t)j * n + k];
            m[(size_t)i * n + j] = s / r;
        }
    }
    // Linv = L^-1 (lower), column by column
    std::vector<double> li((size_t)n * n, 0.0);
    for (int32_t c = 0; c < n; ++c) {
        li[(size_t)c * n + c] = 1.0 / m[(size_t)c * n + c];
        for (int32_t i = c + 1; i < n; ++i) {
            double s = 0.0;
            for (int32_t k = c; k < i; ++k) s += m[(size_t)i * n + k] * li[(size_t)k * n + c];
            li[(size_t)i * n + c] = -s / m[(size_t)i * n + i];
        }
    }
    inv->assign((size_t)n * n, 0.0);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int32_t k = i; k < n; ++k) s += li[(size_t)k * n + i] * li[(size_t)k * n + j];
            (*inv)[(size_t)i * n + j] = s;
            (*inv)[(size_t)j * n + i] = s;
        }
    return true;
}
}  // namespace

// hierarchy (amg.rs:127-164) with a dense-inverse coarsest level.
bool amg_hierarchy(const Csr& a, bool smooth, std::vector<AmgHostLevel>* levels, std::string* why) {
    levels->clear();
    if (a.n == 0) {
        *why = "empty operator";
        return false;
    }
    Csr cur = a;
    for (;;) {
        const int32_t n = cur.n;
        for (int32_t i = 0; i < n; ++i)
            if (std::fabs(cur.get(i, i)) < 1e-300) {
                *why = "zero diagonal on level " + std::to_string(levels->size());
                return false;
            }
        std::vector<int32_t> agg;
        int32_t nc = 0;
        amg_aggregate(cur, smooth ? SA_AGG : 2, &agg, &nc);
        AmgHostLevel L;
        L.diag_only = nc >= n;
        const bool coarsest = n <= COARSEST_N || L.diag_only;
        L.op = cur;
        if (coarsest) {
            if (!L.diag_only) {
                if (!dense_inverse(cur, &L.dense_inv)) {
                    *why = "coarsest level is not positive definite";
                    return false;
                }
                L.has_factor = true;
            }
            levels->push_back(std::move(L));
            break;
        }
        L.p = amg_prolongator(cur, agg, smooth);
        L.r = transpose_rect(L.p, nc);
        L.n_coarse = nc;
        L.agg = agg;
        Csr next = amg_galerkin(cur, L.p, nc);
        levels->push_back(std::move(L));
        cur = std::move(next);
        if (levels->size() > 40) {  // amg.rs:159-161
            *why = "hierarchy deeper than 40 levels";
            return false;
        }
    }
    return true;
}

}  // namespace ruv
