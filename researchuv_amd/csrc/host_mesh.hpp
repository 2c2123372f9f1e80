// host_mesh.hpp -- mesh welding and charting (host side of the unfold path).
#pragma once

#include <cstdint>
#include <vector>

namespace ruv {

// segment::Chart (segment.rs:32-46)
struct Chart {
    std::vector<uint32_t> face_ids;
    std::vector<uint32_t> vertex_ids;         // ascending source ids; local id = rank
    std::vector<uint32_t> tris;               // 3 local ids per face, in face_ids order
    std::vector<uint32_t> border;             // ascending source ids on cut / boundary edges
    std::vector<std::vector<uint32_t>> loops; // source ids, first repeated last when closed
    double area3d = 0.0;
    int64_t local_of(uint32_t v) const;       // Chart::local_of (segment.rs:55-57); -1 when absent
};

struct Charts {
    std::vector<Chart> charts;
    std::vector<uint64_t> cut;  // cut edge keys (min << 32 | max), ascending
};

void weld(const double* pos, int64_t nv, const uint32_t* faces, int64_t nf, double tol, std::vector<double>* pos_out,
          std::vector<uint32_t>* faces_out);
void segment(const double* pos, int64_t nv, const uint32_t* faces, int64_t nf, double angle_min_deg, Charts* out);

}  // namespace ruv
