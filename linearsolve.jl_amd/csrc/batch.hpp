// batch.hpp — handle of a batch of small dense systems resident on the device.
#pragma once
#include <vector>

#include "common.hpp"

struct lsx_batch {
    lsx_ctx* ctx = nullptr;
    int dtype = LSX_F64;           // LSX_F64, or LSX_F32 (FP32 factors + FP64 refinement)
    int64_t count = 0, total_a = 0, total_v = 0;
    double flops = 0.0;            // sum of luflop(n_i) + 2 n_i^2
    hipStream_t stream = nullptr;
    int32_t* order = nullptr;      // systems by decreasing n
    int32_t* n = nullptr;
    uint8_t* cfg = nullptr;        // per-system panel-width index (tile table)
    std::vector<int32_t> hn;       // sizes (host)
    int64_t* aoff = nullptr;
    int64_t* voff = nullptr;
    double* A = nullptr;           // FP64: factors (physical row order) of the last factorization;
                                   // FP32 path: the FP64 matrices (refinement residuals)
    float* A32 = nullptr;          // FP32 path: factors
    double* b = nullptr;
    double* x = nullptr;
    int64_t* ipiv = nullptr;
    int32_t* perm = nullptr;
    int64_t* info = nullptr;
    int32_t* iters = nullptr;      // FP32 path: refinement steps per system
    double* berr = nullptr;        // FP32 path: final backward error per system
    double tol = 1e-14;            // FP32 path: refinement target (normwise backward error)
    int maxref = 10;
    bool use_graph = true;
    hipGraphExec_t graph_exec = nullptr;
    uint32_t* counter = nullptr;
    int64_t* hinfo = nullptr;      // pinned
    lsx::Profiler prof;
};

namespace lsx {
// Factor and solve every system: A copied from dA_src (device or host) when copy_a.
int batch_run(lsx_batch* B, const double* dA_src, const double* db, double* dx, bool copy_a);
}  // namespace lsx
