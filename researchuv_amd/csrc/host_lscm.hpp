// host_lscm.hpp -- per-chart LSCM system on the host (frame, pins, weights, assembly) and the driver.
#pragma once

#include <cstdint>
#include <vector>

#include "host_mesh.hpp"
#include "ruv_internal.hpp"

namespace ruv {

struct ChartSystem {
    int64_t nv = 0;
    int32_t nfree = 0;
    size_t edges = 0;               // weighted undirected edges
    std::vector<double> p3, p2;     // chart-local positions and tangent frame
    std::vector<int32_t> pin_ids;   // local ids, ascending source id
    std::vector<double> pin_uv;     // 2 per pin
    std::vector<int32_t> free_pos;  // local id -> free index or -1
    std::vector<int32_t> free_ids;  // free index -> local id
    Csr l;                          // L on the free vertices (A = L (x) I2)
    std::vector<double> rhs;        // 2 * nfree, interleaved (u, v)
};

// Many charts' systems laid out as one block-diagonal batch (see host_batch.cpp).
struct LscmBatch {
    int64_t nchart = 0, nrows = 0, nv_total = 0;
    std::vector<int64_t> row_ptr{0};
    std::vector<int32_t> cols;
    std::vector<double> vals;
    std::vector<double> rhs;
    std::vector<int64_t> chart_row_off{0};
    std::vector<int32_t> chart_mesh;   // source mesh of each chart
    std::vector<int64_t> chart_nv;     // chart vertex count
    std::vector<int64_t> chart_npins;
};

void eig3(const double m[9], double w[3], double v[9]);
void tangent_frame(const double* p3, int64_t n, double* p2);
void build_system_pins(const double* pos, const Chart& ch, ChartSystem* S);
void build_system(const double* pos, const Chart& ch, double angle_mix, ChartSystem* S);
// the same system with L and the rhs assembled on the device (assemble.hip)
int device_assemble(const double* pos, const Chart& ch, double angle_mix, ChartSystem* S);
int unfold_chart(const double* pos, const Chart& ch, const ruv_unfold_opts& o, double* uv_out, double* stats,
                 ruv_solve_info* info);
void driver_iterations(const ChartSystem& S, const Chart& ch, const double* sol, int max_iter_opt, double* uv_out,
                       double* stats);

}  // namespace ruv
