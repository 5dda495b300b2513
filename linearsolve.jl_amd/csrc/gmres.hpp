// gmres.hpp — device-resident CSR matrix handle and the SimpleGMRES launcher.
#pragma once
#include "common.hpp"

namespace lsx {
constexpr int PMAX_G = 256;         // max workgroups of one GMRES solve (8192 rows each)
constexpr int GMRES_MAX_CAP = 128;     // max Krylov dimension whose bookkeeping is held in LDS
constexpr int GMRES_MAX_CAP_G = 2048;  // max Krylov dimension of one solve (bookkeeping in HBM)

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

// Everything one solve's workgroups read: the matrix, vectors, workspaces and exchange words.
struct GmresArgs {
    int64_t N;
    const int32_t* rowptr;
    const int32_t* col;
    const double* val;
    const double* b;
    double* x;
    const double* pl;   // nullable: Jacobi / diagonal left preconditioner
    const double* pr;   // nullable: diagonal right preconditioner
    double* V;          // (cap + 1) x N basis
    double* xr;         // N: correction of the current pass (restart)
    int memory;
    int restart;
    int cap;            // size of the bookkeeping (memory, or maxiters when not restarted)
    double* Rg = nullptr;  // non-null: each workgroup's bookkeeping (R, Givens pairs, z) in HBM at
                        // Rg + w gmres_book_words(cap) instead of LDS
    int64_t maxiters;
    double abstol, reltol;
    int P;
    uint64_t* red;      // LL all-reduce slots: 2 parities x P x 2 words; word 4 P holds the barrier
                        // counter, words 4 P + 8 ... 4 P + 11 the totals workgroup 0 publishes
    uint32_t* bar;      // barrier counter (red + 4 P)
    GmresOut* out;
    int64_t timeout;
};

// Launch shape of one solve of N rows.
struct GmresPlan {
    int P;          // workgroups
    int rpt;        // rows per thread (1, 4 or 16)
    int cap;        // Krylov dimension
    bool rglob;     // the bookkeeping is in HBM (it does not fit the LDS beside q and the basis slice)
    size_t dynlds;  // dynamic LDS bytes per workgroup
};

// Doubles of one workgroup's bookkeeping: the packed triangle R, the Givens pairs and z.
__host__ __device__ inline int64_t gmres_book_words(int64_t cap) { return cap * (cap + 1) / 2 + 3 * cap + 1; }

// 64-bit words of a solve's exchange area (all-reduce slots, barrier counter, published totals);
// zero before use.
__host__ __device__ inline int64_t gmres_red_words(int P) { return 4 * (int64_t)P + 32; }
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
    double* rg = nullptr;        // per-workgroup bookkeeping in HBM (Krylov dimensions beyond the LDS)
    size_t rg_bytes = 0;
    uint64_t* red = nullptr;     // all-reduce slots + barrier counter
    lsx::GmresOut* out = nullptr;
    lsx::GmresOut* hout = nullptr;  // pinned
    double* db = nullptr;        // b, x, pl, pr staging (device)
    double* dx = nullptr;
    double* dpl = nullptr;
    double* dpr = nullptr;
    // batched (block-diagonal) solves: column indices local to each row's block, per-block
    // exchange words, launch tables
    int32_t* col_local = nullptr;
    int64_t local_bs = 0;        // block size col_local was made for (0: none)
    uint64_t* bred = nullptr;    // nblocks x gmres_red_words(1)
    lsx::GmresArgs* bitems = nullptr;
    int* bfirst = nullptr;
    lsx::GmresOut* bout = nullptr;
    uint32_t* bticket = nullptr;
    int* bflag = nullptr;        // off-block entry found
    int64_t bcap = 0;            // blocks the batched buffers hold
    double* bpart = nullptr;     // partial sums of the stopping norm
    int64_t bkey[9] = {};        // parameters the device launch table was built for
};

namespace lsx {
int gmres_launch(lsx_csr* A, const GmresParams& p, const double* db, double* dx, const double* dpl,
                 const double* dpr, hipStream_t st);
// Workgroups, rows per thread and LDS of a solve; LSX_EINVAL (message in ctx) when it does not fit.
// hbm: the Krylov dimension may exceed what the LDS holds (single solves, gmres_launch); the
// multi-solve launches keep the bookkeeping in LDS.
int gmres_plan(lsx_ctx* ctx, int64_t N, const GmresParams& p, GmresPlan* out, bool hbm = false);
// Solves of one rows-per-thread class in one launch on `pool` workgroups (at least the widest
// solve's count): rank r of solve i is work number first[i] + r (first has nitems + 1 entries,
// first[nitems] = total); workgroups take numbers from *ticket as they start and after each solve,
// so every solve but the last one started has all its ranks running.  Zeroes each solve's
// exchange area, *out and *ticket first.  items/first are device copies; dynlds the largest of
// the solves' plans.
int gmres_launch_multi(lsx_ctx* ctx, hipStream_t st, const GmresArgs* items, const int* first, int nitems,
                       int pool, int rpt, size_t dynlds, uint32_t* ticket);
int csr_spmv_dev(lsx_csr* A, const double* dx, double* dy, hipStream_t st);
// Checks a device CSR structure on stream st: *dflag (device) becomes nonzero when rowptr or the
// column indices are out of range.  The caller reads the flag after the stream.
int csr_validate(hipStream_t st, int64_t N, int64_t nnz, const int32_t* rowptr, const int32_t* col, int* dflag);
}  // namespace lsx
