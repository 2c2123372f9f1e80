// batch_impl.hpp -- host state of one batch (shared by batch.hip and batch_amg.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "batch_kernels.hpp"
#include "ruv_internal.hpp"

namespace ruv {

struct Bamg;  // device-built aggregation hierarchy (batch_amg.hip)
void bamg_destroy(Bamg* a);

// Device blocks kept for reuse: a released block serves a later request of at least its size and at most twice
// it.  Rebuilds of a hierarchy ask for the same sizes again, so after the first build they allocate nothing.
// Every user of a pool works on one stream, so a block reused after its release is ordered after its last use.
struct DevPool {
    std::multimap<size_t, void*> free_blocks;
    std::map<void*, size_t> owned;
    hipError_t get(void** p, size_t bytes) {
        bytes = std::max<size_t>(256, (bytes + 255) & ~(size_t)255);
        auto it = free_blocks.lower_bound(bytes);
        if (it != free_blocks.end() && it->first <= 2 * bytes) {
            *p = it->second;
            free_blocks.erase(it);
            return hipSuccess;
        }
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) return e;
        owned[q] = bytes;
        *p = q;
        return hipSuccess;
    }
    void put(void* p) {
        auto it = owned.find(p);
        if (it != owned.end()) free_blocks.insert({it->second, p});
    }
    ~DevPool() {
        for (auto& kv : owned) (void)hipFree(kv.first);
    }
};

#define HIPCHK(call)                                                                                         \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess) return fail(RUV_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// launch kinds of the profiling marks, numbered as ruv_batch_step_profile reports them
enum : int {
    BK_DIR = 0, BK_SPMV = 1, BK_UPDATE = 2, BK_RESID0 = 3, BK_RESTRICT = 4, BK_DOWN = 5, BK_COARSEST = 6,
    BK_PROLONG = 7, BK_POST = 8, BK_PROLONG0 = 9, BK_POST0 = 10
};

struct BatchImpl {
    int R = 2, mode = 0, precond = 0, nchart = 0, ntile = 0, nslice = 0;
    int ci = 0;          // general slices: 0 = 16-bit column offsets, 1 = 32-bit columns
    int spt = 2;         // slices per tile (chosen by the autotuner)
    int nt = 0;          // step kernel loads the matrix values non-temporally (chosen by the autotuner)
    int nstencil = 0;    // slices stored in stencil form
    int64_t nrows = 0, nnz = 0, nell = 0, ncol = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<void*> mem;
    BDev B{};
    // host copies of the slice table and chart row offsets (tile building)
    std::vector<int64_t> h_sl_row0;
    std::vector<int32_t> h_sl_n;
    std::vector<int32_t> chart_s0;  // first slice of each chart, nchart + 1
    std::vector<int64_t> crow;      // first row of each chart, nchart + 1
    std::vector<int> chart_t0;
    std::vector<int64_t> h_tile_r0, h_tile_r1;  // rows of each tile (host copy)
    std::vector<int64_t> h_sl_off;               // host copies of the storage form (value updates)
    std::vector<int32_t> h_sl_w, h_dtab;
    std::vector<char> h_stencil;
    // tiles (rebuilt by the autotuner)
    int32_t *d_tile_s0 = nullptr, *d_tile_chart = nullptr, *d_chart_t0 = nullptr;
    int64_t *d_tile_r0 = nullptr, *d_tile_r1 = nullptr;
    double *d_part_a = nullptr, *d_part_b = nullptr, *d_part_c = nullptr, *d_part_o = nullptr, *d_part_d = nullptr;
    int* act = nullptr;
    int* h_act = nullptr;  // pinned staging of the active-tile lists: two halves of ntile
    // matrix
    float* ev32 = nullptr;
    double* ev64 = nullptr;
    // band form of the band charts (batch_kernels.hpp): per chart grid width (0: not in band form), offset mask and
    // first value (nchart + 1); nbv values in FP32 (bv32, mixed) or FP64 (bv64)
    std::vector<int32_t> h_bw, h_bmask;
    std::vector<int64_t> h_bbase;
    int nband = 0;
    int64_t nbv = 0;
    float* bv32 = nullptr;
    double* bv64 = nullptr;
    double fin_slots = 0.0;  // value slots a finest-operator pass streams: ELL slots of the other charts, band slots
    int64_t* ell_ranges = nullptr;  // (first, count) ranges of the sliced-ELL entries of the charts not in band form
    int64_t n_ell_ranges = 0;
    float* dinv32 = nullptr;
    double* dinv64 = nullptr;
    // vectors
    double *x64 = nullptr, *r64 = nullptr;
    void *xv = nullptr, *rv = nullptr, *zv = nullptr, *pv0 = nullptr, *pv1 = nullptr, *qv = nullptr;
    int* h_done = nullptr;
    int* h_done_b = nullptr;           // second pinned flag buffer (the look one chunk late)
    hipEvent_t fev[2] = {nullptr, nullptr};
    int act_half = 0;                  // half of h_act the last list was staged in
    int na_band = 0;                   // leading tiles of the last list that belong to band charts
    // 1: look at which charts run one chunk late while most tiles run, so the device never waits for a look
    // (measured on the 4096-patch batch: 154.5 ms against 152.9 ms for looking right after each chunk)
    int look_lag = 0;
    int32_t* h_lists = nullptr;  // pinned staging of the AMG's running-chart launch lists: two halves (grown, kept)
    size_t h_lists_cap = 0;
    int* h_final = nullptr;
    unsigned* h_iters = nullptr;
    double* h_true = nullptr;
    std::vector<uint64_t> iters_total;
    double inner_tol = 1e-4;  // inner FP32 solve: reduce ||r|| by this much (two rounds reach 1e-8)
    // mixed refinement: a chart also stops only once its estimated forward error is at most uv_tol max|x| (a tenth
    // of the 1e-6 UV bound: on the 4096-patch batch the estimate fell short of the true error by up to 2.3x, and
    // 2.5e-7 left a chart at 7.5e-7; 0 turns the test off)
    double uv_tol = 1e-7;
    // mixed refinement: a round that can end a chart aims its inner solve at final_aim times the chart's outer
    // target (0: at the target itself, like the other rounds' floor); 0.35 measured best of 0.15 .. 0.6
    double final_aim = 0.35;
    // mixed: the first inner solve stops for the charts still running once fewer than this fraction of its tiles
    // run; they go on in the next round (from their last direction) beside the others, instead of a tail of few
    // charts (0.02, 0.05, 0.1 measured alike; 0: off)
    double first_cut = 0.05;
    int max_rounds = 40;
    // rounds after the first go on from the previous round's last direction (1) or start afresh (0)
    int continue_rounds = 1;
    // ... when |refined - recursive residual| <= cont_gap * |recursive residual| (measured: the 4096-chart batch's
    // charts are nearly all below 0.03, the 2048^2 stretched grid's single chart between 0.1 and 0.3, where going
    // on took 1017 steps against 292 for a fresh start)
    double cont_gap = 0.05;
    // the AMG's coarse launches list the running charts' slices once fewer than this fraction of tiles run
    double list_below = 0.5;
    int amg_chunk = 4;  // AMG-preconditioned steps between convergence looks
    // step-kernel accounting (optional): an event pair around every launch, summed after each solve
    bool ktime = false;
    std::vector<hipEvent_t> kev;
    size_t kev_used = 0;
    int64_t k_launches = 0;
    double k_seconds = 0.0, k_bytes = 0.0, k_stored = 0.0;
    std::vector<double> chart_step_bytes;   // algorithmic bytes of one SpMV of chart c (nonzeros only)
    std::vector<double> chart_step_stored;  // bytes of one SpMV of chart c as stored (padding slots included)
    // the finest operator's bytes besides its values: general slices' stored columns (per nonzero / per stored
    // slot) and the slice tables
    double fin_colb = 0.0, fin_colb_st = 0.0, fin_meta = 0.0;
    // launch profiling (optional): with marks set, the step's launches each record an event after them and their
    // kind, level and bytes (kind < 0: not a kernel; mev[0] is recorded before the first)
    struct LaunchMark {
        int kind, level;
        double bytes, stored;
    };
    std::vector<LaunchMark>* marks = nullptr;
    std::vector<hipEvent_t> mev;
    hipError_t mark(int kind, int level, double bytes, double stored) {
        if (!marks) return hipSuccess;
        const size_t k = marks->size() + 1;
        while (mev.size() <= k) {
            hipEvent_t e = nullptr;
            const hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
            if (r != hipSuccess) return r;
            mev.push_back(e);
        }
        marks->push_back({kind, level, bytes, stored});
        return hipEventRecord(mev[k], stream);
    }
    DevPool pool;                          // setup and hierarchy blocks (precond AMG)
    Bamg* amg = nullptr;                   // precond AMG: the hierarchy
    int64_t* amg_ptr = nullptr;            // precond AMG: the finest operator in FP64 CSR for (re)builds,
    int32_t* amg_cols = nullptr;           // and each row's finest-level matching group (its tile)
    float* amg_vals = nullptr;
    int32_t* amg_group = nullptr;

    template <typename T>
    hipError_t alloc(T** p, size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) mem.push_back(q);
        *p = (T*)q;
        return e;
    }
    template <typename T>
    hipError_t upload(T** p, const T* h, size_t n) {
        hipError_t e = alloc(p, n);
        if (e == hipSuccess && n) e = hipMemcpyAsync(*p, h, n * sizeof(T), hipMemcpyHostToDevice, stream);
        return e;
    }
    void free_tiles() {
        for (void* p : {(void*)d_tile_s0, (void*)d_tile_chart, (void*)d_chart_t0, (void*)d_tile_r0, (void*)d_tile_r1,
                        (void*)d_part_a, (void*)d_part_b, (void*)d_part_c, (void*)d_part_o, (void*)d_part_d,
                        (void*)act})
            if (p) (void)hipFree(p);
        d_tile_s0 = d_tile_chart = d_chart_t0 = nullptr;
        d_tile_r0 = d_tile_r1 = nullptr;
        d_part_a = d_part_b = d_part_c = d_part_o = d_part_d = nullptr;
        act = nullptr;
        if (h_act) (void)hipHostFree(h_act);
        h_act = nullptr;
    }
    ~BatchImpl() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (amg) bamg_destroy(amg);
        free_tiles();
        for (void* p : mem) (void)hipFree(p);
        for (void* p : {(void*)h_done, (void*)h_done_b, (void*)h_final, (void*)h_iters, (void*)h_true,
                        (void*)h_lists})
            if (p) (void)hipHostFree(p);
        for (hipEvent_t e : fev)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : kev) (void)hipEventDestroy(e);
        for (hipEvent_t e : mev) (void)hipEventDestroy(e);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

}  // namespace ruv
