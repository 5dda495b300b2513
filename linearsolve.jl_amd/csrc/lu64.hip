// lu64.hip — FP64 partial-pivot LU (getrf) and triangular solves (getrs) for gfx950.
//
// Algorithm: right-looking blocked LU with two block levels, the structure of LAPACK dgetrf
// and of the reference's own restatement src/blocked_lufact.jl:658-679:
//   outer block NBO = 128 columns (the K of the trailing MFMA update, which holds ~2/3 n^3 flops);
//   inside it, inner panels of PW = 32 columns are factored by a register-resident kernel whose
//   workgroups each hold 512 rows x 32 columns (k_panel<T, 512>; one workgroup of 256 rows when
//   the panel has at most 256), exchanging one pivot candidate record per column through
//   flag-in-word messages (no fences, one memory round trip per column).
// Pivot search: maximum |a| with NaN ignored, first (smallest) row on ties
// (src/blocked_lufact.jl:38-54); zero pivot -> info = first such column, factorization
// continues (src/generic_lufact.jl:71-141).  Scaling multiplies by the reciprocal of the pivot
// (blocked_lufact.jl:110-113) unless |pivot| < DBL_MIN (then divides, as LAPACK dgetf2).
#include <cfloat>
#include <utility>
#include <type_traits>

#include "gemm.hpp"
#include "lu_driver.hpp"

namespace lsx {

// ------------------------------------------------------------------------------------------
// Host driver (the blocked algorithm itself is lu_driver.hpp, shared with the mixed path).
static void gemm_f64(lsx_lu* lu, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                     int64_t ldc, int64_t M, int64_t N, int64_t K, int mode) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    const bool big = mode == 0 && cdiv(M, 128) * cdiv(N, 128) >= 2 * (int64_t)lu->ctx->num_cus;
    const int pi = lu->prof.begin(lu->cur, big ? PK_GEMM : PK_GEMM_SMALL, 2.0 * M * N * K,
                                  sizeof(double) * (2.0 * M * N + (double)M * K + (double)K * N));
    gemm_dispatch<double>(lu->cur, lu->ctx->num_cus, A, lda, B, ldb, C, ldc, M, N, K, mode);
    lu->prof.end(lu->cur, pi);
}

int lu64_getrf(lsx_lu* lu) {
    auto g = [lu](const double* A, int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc, int64_t M,
                  int64_t N, int64_t K, int mode) { gemm_f64(lu, A, lda, B, ldb, C, ldc, M, N, K, mode); };
    return getrf_blocked<double>(lu, lu->F, lu->ld, lu->DL, lu->DU, g, g);
}

int lu_ensure_work(lsx_lu* lu, int64_t nrhs) {
    if (nrhs <= lu->wcap) return LSX_OK;
    lsx_ctx* ctx = lu->ctx;
    if (lu->wb) (void)hipFree(lu->wb);
    if (lu->wy) (void)hipFree(lu->wy);
    lu->wb = lu->wy = nullptr;
    lu->wcap = 0;
    const int64_t cap = std::max<int64_t>(nrhs, TRHS);
    LSX_HIP(ctx, hipMalloc(&lu->wb, sizeof(double) * std::max(lu->n, lu->cap) * cap));
    LSX_HIP(ctx, hipMalloc(&lu->wy, sizeof(double) * std::max(lu->n, lu->cap) * cap));
    lu->wcap = cap;
    return LSX_OK;
}

int lu64_getrs_dev(lsx_lu* lu, int64_t nrhs, const double* dB, int64_t ldb, double* dX, int64_t ldx) {
    return getrs_blocked<double>(lu, lu->F, lu->ld, lu->DL, lu->DU, nrhs, dB, ldb, dX, ldx);
}

int lu64_getrs_t_dev(lsx_lu* lu, int64_t nrhs, const double* dB, int64_t ldb, double* dX, int64_t ldx) {
    return getrs_t_blocked<double>(lu, lu->F, lu->ld, lu->DL, lu->DU, nrhs, dB, ldb, dX, ldx);
}

}  // namespace lsx
