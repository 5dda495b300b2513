// gmres.hpp — device-resident CSR matrix handle and the SimpleGMRES launcher.
#pragma once
#include "common.hpp"

namespace lsx {
constexpr int PMAX_G = 256;         // max workgroups of one GMRES solve (8192 rows each)
constexpr int GMRES_MAX_CAP = 128;  // max Krylov dimension held in LDS

struct GmresOut {
    int32_t retcode;
    int32_t pad;
    int64_t iters;
    double resid;
    uint64_t err;
};

struct GmresParams {
    int memory;
    int restart;
    int64_t maxiters;
    double abstol, reltol;
};
}  // namespace lsx

struct lsx_csr {
    lsx_ctx* ctx = nullptr;
    int64_t N = 0, nnz = 0;
    int32_t* rowptr = nullptr;
    int32_t* col = nullptr;
    double* val = nullptr;
    hipStream_t stream = nullptr;
    double* ws = nullptr;        // Krylov basis + correction
    size_t ws_bytes = 0;
    uint64_t* red = nullptr;     // all-reduce slots + barrier counter
    lsx::GmresOut* out = nullptr;
    lsx::GmresOut* hout = nullptr;  // pinned
    double* db = nullptr;        // b, x, pl, pr staging (device)
    double* dx = nullptr;
    double* dpl = nullptr;
    double* dpr = nullptr;
};

namespace lsx {
int gmres_launch(lsx_csr* A, const GmresParams& p, const double* db, double* dx, const double* dpl,
                 const double* dpr, hipStream_t st);
int csr_spmv_dev(lsx_csr* A, const double* dx, double* dy, hipStream_t st);
}  // namespace lsx
