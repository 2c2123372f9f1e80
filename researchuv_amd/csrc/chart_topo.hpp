// chart_topo.hpp -- integer topology of one chart for the device assembly and its adjoint (host side).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace ruv {

struct ChartTopo {
    int64_t nv = 0, ntri = 0, ne = 0;
    std::vector<int32_t> ei, ej;    // undirected edges (i < j), ascending
    std::vector<int64_t> ooff;      // ne + 1: opposite vertices of each edge, triangle order
    std::vector<int32_t> opp;
    std::vector<int64_t> aoff;      // nv + 1: ascending neighbour lists
    std::vector<int32_t> anb;
    std::vector<int64_t> aedge;     // edge id of each neighbour entry
    std::vector<int64_t> toff;      // nv + 1: incident triangles, ascending
    std::vector<int32_t> tlist;
    std::vector<int32_t> free_ids, free_pos;
    std::vector<int64_t> tri_edge;  // 3 per triangle: edge id of corner q's edge (tris[q], tris[q+1])
};

// tris: 3 local ids per triangle; is_pin: nv flags.
inline void build_topo(int64_t nv, const uint32_t* tris, int64_t ntri, const std::vector<char>& is_pin, ChartTopo* T) {
    T->nv = nv;
    T->ntri = ntri;
    struct Rec {
        uint64_t key;
        uint32_t ord;
        int32_t o;
    };
    std::vector<Rec> con;
    con.reserve(3 * ntri);
    for (int64_t f = 0; f < ntri; ++f)
        for (int q = 0; q < 3; ++q) {
            const uint32_t e0 = tris[3 * f + q], e1 = tris[3 * f + (q + 1) % 3], o = tris[3 * f + (q + 2) % 3];
            const uint32_t i = std::min(e0, e1), j = std::max(e0, e1);
            con.push_back({((uint64_t)i << 32) | j, (uint32_t)(3 * f + q), (int32_t)o});
        }
    std::sort(con.begin(), con.end(),
              [](const Rec& a, const Rec& b) { return a.key < b.key || (a.key == b.key && a.ord < b.ord); });
    T->ei.clear();
    T->ej.clear();
    T->opp.clear();
    T->ooff.assign(1, 0);
    T->tri_edge.assign(3 * ntri, 0);
    for (size_t q = 0; q < con.size(); ++q) {
        if (q == 0 || con[q].key != con[q - 1].key) {
            if (q) T->ooff.push_back((int64_t)T->opp.size());
            T->ei.push_back((int32_t)(con[q].key >> 32));
            T->ej.push_back((int32_t)(con[q].key & 0xffffffffu));
        }
        T->opp.push_back(con[q].o);
        T->tri_edge[con[q].ord] = (int64_t)T->ei.size() - 1;
    }
    T->ooff.push_back((int64_t)T->opp.size());
    T->ne = (int64_t)T->ei.size();
    T->aoff.assign(nv + 1, 0);
    for (int64_t e = 0; e < T->ne; ++e) {
        T->aoff[T->ei[e] + 1]++;
        T->aoff[T->ej[e] + 1]++;
    }
    for (int64_t i = 0; i < nv; ++i) T->aoff[i + 1] += T->aoff[i];
    T->anb.assign(T->aoff[nv], 0);
    T->aedge.assign(T->aoff[nv], 0);
    {
        std::vector<int64_t> cur(T->aoff.begin(), T->aoff.end() - 1);
        for (int64_t e = 0; e < T->ne; ++e) {  // edges ascending -> every list ascending
            T->anb[cur[T->ei[e]]] = T->ej[e];
            T->aedge[cur[T->ei[e]]++] = e;
            T->anb[cur[T->ej[e]]] = T->ei[e];
            T->aedge[cur[T->ej[e]]++] = e;
        }
    }
    T->toff.assign(nv + 1, 0);
    for (int64_t f = 0; f < ntri; ++f)
        for (int q = 0; q < 3; ++q) T->toff[tris[3 * f + q] + 1]++;
    for (int64_t i = 0; i < nv; ++i) T->toff[i + 1] += T->toff[i];
    T->tlist.assign(T->toff[nv], 0);
    {
        std::vector<int64_t> cur(T->toff.begin(), T->toff.end() - 1);
        for (int64_t f = 0; f < ntri; ++f)
            for (int q = 0; q < 3; ++q) {
                const uint32_t v = tris[3 * f + q];
                // a degenerate triangle listing v twice is counted once per occurrence; the gather skips repeats
                T->tlist[cur[v]++] = (int32_t)f;
            }
    }
    T->free_pos.assign(nv, -1);
    T->free_ids.clear();
    for (int64_t i = 0; i < nv; ++i)
        if (!is_pin[i]) {
            T->free_pos[i] = (int32_t)T->free_ids.size();
            T->free_ids.push_back((int32_t)i);
        }
}

}  // namespace ruv
