// dev_assembly.hpp -- a chart's topology, geometry and assembled system resident on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "chart_topo.hpp"
#include "ruv_internal.hpp"

namespace ruv {

struct DevAssembly {
    hipStream_t stream = nullptr;
    std::vector<void*> mem;
    int64_t nv = 0, ne = 0, ntri = 0, nnz = 0;
    int32_t nfree = 0;
    int32_t *ei = nullptr, *ej = nullptr, *opp = nullptr, *anb = nullptr, *tlist = nullptr, *free_ids = nullptr,
            *free_pos = nullptr, *cnt = nullptr, *d_rp = nullptr, *cols = nullptr;
    int64_t *ooff = nullptr, *aoff = nullptr, *aedge = nullptr, *toff = nullptr, *tri_edge = nullptr;
    uint32_t* tris = nullptr;
    double *p3 = nullptr, *p2 = nullptr, *pin_at = nullptr, *c = nullptr, *w = nullptr, *dd = nullptr,
           *vals = nullptr, *rhs = nullptr;
    std::vector<int32_t> rp;  // host copy of the row pointers
    ~DevAssembly();
};

int dev_assemble(const ChartTopo& T, const std::vector<double>& p3, const std::vector<double>& p2,
                 const std::vector<double>& pin_at, const std::vector<uint32_t>& tris, double mix_b, DevAssembly* A);
int dev_download(DevAssembly& A, Csr* L, std::vector<double>* rhs, size_t* edges);

}  // namespace ruv
