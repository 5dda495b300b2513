// batch.hpp — handle of a batch of small dense systems resident on the device.
#pragma once
#include "common.hpp"

struct lsx_batch {
    lsx_ctx* ctx = nullptr;
    int64_t count = 0, total_a = 0, total_v = 0;
    double flops = 0.0;            // sum of luflop(n_i) + 2 n_i^2
    hipStream_t stream = nullptr;
    int32_t* order = nullptr;      // systems by decreasing n
    int32_t* n = nullptr;
    int64_t* aoff = nullptr;
    int64_t* voff = nullptr;
    double* A = nullptr;           // factors (physical row order) of the last factorization
    double* b = nullptr;
    double* x = nullptr;
    int64_t* ipiv = nullptr;
    int32_t* perm = nullptr;
    int64_t* info = nullptr;
    uint32_t* counter = nullptr;
    int64_t* hinfo = nullptr;      // pinned
    lsx::Profiler prof;
};

namespace lsx {
// Factor and solve every system: A copied from dA_src (device or host) when copy_a.
int batch_run(lsx_batch* B, const double* dA_src, const double* db, double* dx, bool copy_a);
}  // namespace lsx
