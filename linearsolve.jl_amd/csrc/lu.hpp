// lu.hpp — resident dense-LU handle and the FP64 getrf/getrs driver entry points.
#pragma once
#include "common.hpp"

namespace lsx {
constexpr int PW = 32;        // panel width held in registers by the panel kernel
constexpr int NBO = 128;      // outer block: K of the trailing MFMA update
constexpr int PMAX = 128;     // max workgroups cooperating on one panel (512 rows each)
constexpr int REC = 4 + 2 * PW;   // words of one FP64 exchange record: key lo/hi, row, pad, data
constexpr int RECS = 72;          // record stride in the exchange buffer (words)
constexpr int DB = 64;            // diagonal block of the explicit inverses and triangular solves
constexpr int TRHS = 16;          // right-hand sides per triangular-solve launch
constexpr int64_t LGMAX_BYTES = 144 * 1024;  // LDS of the deferred L-interchange pass
}  // namespace lsx

struct lsx_lu {
    lsx_ctx* ctx = nullptr;
    int64_t n = 0;
    int64_t cap = 0;                // order the buffers hold (n <= cap; FP64 handles can be reshaped)
    bool own_F = true;              // F allocated here (false: the caller points F at its memory)
    bool lookahead = true;          // panel of the next block on stream2 (false: one stream)
    int dtype = LSX_F64;
    int64_t ld = 0;
    hipStream_t stream = nullptr;   // trailing updates and everything outside getrf
    hipStream_t stream2 = nullptr;  // getrf lookahead: panel factorization of the next block
    hipStream_t stream_upd = nullptr;  // getrf trailing updates on the CUs stream2 does not use
                                       // (null: no CU partition, updates run on stream)
    hipEvent_t ev_start = nullptr;  // getrf input ready on stream
    hipStream_t cur = nullptr;      // stream the launch helpers use at the moment
    hipEvent_t ev_next = nullptr;   // next block's columns updated (stream -> stream2)
    hipEvent_t ev_panel = nullptr;  // next block's panel factored, inverses and row map made (stream2 -> stream)
    hipEvent_t ev_perm = nullptr;   // next block's interchange list made (stream2 -> stream)
    // FP64 factors (also the FP64 copy of A for the mixed path's residuals)
    double* F = nullptr;
    int64_t* ipiv = nullptr;     // device, 1-based
    int* perm = nullptr;         // device: row i of P*A is row perm[i] of A
    int64_t* dinfo = nullptr;    // device: [0] info, [1] error flags
    int64_t* hinfo = nullptr;    // pinned host mirror of dinfo
    uint64_t* slots = nullptr;   // panel exchange records
    int64_t* plist = nullptr;    // permutation lists: inner [0, 4*NBO), outer by block parity
                                 // [4*NBO, 8*NBO) and [8*NBO, 12*NBO)
    int* pcount = nullptr;       // [0] inner T, [1], [2] outer T by block parity
    int* pinv = nullptr;         // per outer block: inverse of the row map after that block
                                 // (the deferred row interchanges of the L columns)
    uint64_t* tmsg = nullptr;    // triangular-solve block messages (nblk * TMW words)
    double* DL = nullptr;        // inverses of the 64 x 64 diagonal blocks of L (nblk x 64 x 64)
    double* DU = nullptr;        // ... and of U
    uint32_t seq = 1;
    uint32_t tseq = 1;
    // solve workspaces
    double* wb = nullptr;
    double* wy = nullptr;
    int64_t wcap = 0;  // columns of wb/wy
    // residual check: staging copy of a host A, statistics of the residual kernel
    double* astage = nullptr;
    double* rpart = nullptr;
    double* hrpart = nullptr;     // pinned
    // mixed precision state
    float* F32 = nullptr;         // low-precision working factors (FP32 storage)
    // tournament-pivoting block factorization (calu.hpp) of the mixed path
    uint64_t* calu_cand = nullptr;  // candidate records (flag-in-word)
    uint64_t* calu_bc = nullptr;    // broadcast L11\U11, U12, pivot rows (flag-in-word)
    uint32_t* calu_ctr = nullptr;   // arrival counter
    uint32_t calu_cbase = 0;      // counter value the next launch starts from
    uint32_t calu_seq = 1;        // sequence number of the next launch's first panel
    int64_t ld32 = 0;
    double* rwork = nullptr;      // residual / correction workspace (n x cap)
    void* aux = nullptr;          // mixed-path extra state
    bool factored = false;
    int64_t info = 0;
    lsx::Profiler prof;
};

namespace lsx {
// Clears the error word of a handle's device-side waits (once, at the start of a public entry
// point; the waits of every later launch OR into it).
inline int lu_clear_err(lsx_lu* lu) {
    return hipMemsetAsync(lu->dinfo + 1, 0, sizeof(int64_t), lu->stream) == hipSuccess ? LSX_OK : LSX_EHIP;
}
// Buffers for order lu->cap (0: lu->n).  With own_F false F is not allocated; with lookahead
// false getrf runs on lu->stream alone and no CU-partitioned streams are made.
int lu64_alloc(lsx_lu* lu);
// An FP64 handle takes a new order n <= cap (factors of the previous order are dropped).
int lu64_reshape(lsx_lu* lu, int64_t n);
void lu64_free(lsx_lu* lu);
// Factors lu->F in place (already holding A).  Returns info >= 0 or an lsx_status < 0.
int lu64_getrf(lsx_lu* lu);
// X = A \ B with device buffers (B n x nrhs, ldb; X n x nrhs, ldx; may alias).
int lu64_getrs_dev(lsx_lu* lu, int64_t nrhs, const double* dB, int64_t ldb, double* dX, int64_t ldx);
// X = A^{-T} B (device buffers, may alias).
int lu64_getrs_t_dev(lsx_lu* lu, int64_t nrhs, const double* dB, int64_t ldb, double* dX, int64_t ldx);
int lu_ensure_work(lsx_lu* lu, int64_t nrhs);
// Mixed-precision (FP32 / BF16) factorization state, lu_mixed.hip.
int mixed_alloc(lsx_lu* lu);
void mixed_free(lsx_lu* lu);
int mixed_getrf(lsx_lu* lu);
int mixed_get_factors(lsx_lu* lu, double* F, int64_t ldf);
int mixed_getrs_dev(lsx_lu* lu, int64_t nrhs, const double* dB, int64_t ldb, double* dX, int64_t ldx);
}  // namespace lsx
