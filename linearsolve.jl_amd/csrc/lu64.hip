// lu64.hip — FP64 partial-pivot LU (getrf) and triangular solves (getrs) for gfx950.
//
// Algorithm: right-looking blocked LU with two block levels, the structure of LAPACK dgetrf
// and of the reference's own restatement src/blocked_lufact.jl:658-679:
//   outer block NBO = 128 columns (the K of the trailing MFMA update, which holds ~2/3 n^3 flops);
//   inside it, inner panels of PW = 32 columns are factored by a register-resident kernel whose
//   workgroups each hold 1024 rows x 32 columns, exchanging one pivot candidate record per
//   column through flag-in-word messages (no fences, one memory round trip per column).
// Pivot search: maximum |a| with NaN ignored, first (smallest) row on ties
// (src/blocked_lufact.jl:38-54); zero pivot -> info = first such column, factorization
// continues (src/generic_lufact.jl:71-141).  Scaling multiplies by the reciprocal of the pivot
// (blocked_lufact.jl:110-113) unless |pivot| < DBL_MIN (then divides, as LAPACK dgetf2).
#include <cfloat>
#include <utility>
#include <type_traits>

#include "lu_driver.hpp"

namespace lsx {

typedef double v4d __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------------------------------
// Trailing update on FP64 MFMA (v_mfma_f64_16x16x4f64), LDS-staged:
//   mode 0: C -= A * B      mode 1: C = A * B (in place on B when M <= 128: a workgroup reads all
//                           K rows of its B columns before it writes C)
// A: M x K (lda), B: K x N (ldb), C: M x N (ldc), column-major.  128 x 128 tile per workgroup,
// four waves of 64 x 64, K staged 16 at a time with double buffering.  The MFMA computes the
// transposed tile (B^T A^T) so each lane's results run down a column of C (128-byte stores).
constexpr int GBM = 128, GBN = 128, GBK = 16;
__global__ __launch_bounds__(256) void k_gemm_f64(const double* __restrict Ag, int64_t lda,
                                                  const double* Bg, int64_t ldb, double* Cg,
                                                  int64_t ldc, int M, int N, int K, int mode) {
    const int TM = (M + GBM - 1) / GBM, TN = (N + GBN - 1) / GBN, T = TM * TN;
    // XCD-aware: consecutive tiles of the column-major tile order land on one XCD
    const int per = (T + 7) / 8;
    const int b = blockIdx.x;
    const int tile = (b & 7) * per + (b >> 3);
    if (tile >= T) return;
    const int tm = tile % TM, tn = tile / TM;
    const int row0 = tm * GBM, col0 = tn * GBN;

    __shared__ double As[2][GBK][GBM + 1];
    __shared__ double Bs[2][GBK][GBN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    v4d acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    double ra[8], rb[8];
    auto gload = [&](int k0) {
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int idx = it * 256 + tid;
            const int r = idx & (GBM - 1), k = idx >> 7;
            const int gr = row0 + r, gk = k0 + k;
            ra[it] = (gr < M && gk < K) ? Ag[(int64_t)gk * lda + gr] : 0.0;
            const int kb = idx & (GBK - 1), c = idx >> 4;
            const int gc = col0 + c, gk2 = k0 + kb;
            rb[it] = (gc < N && gk2 < K) ? Bg[(int64_t)gc * ldb + gk2] : 0.0;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int idx = it * 256 + tid;
            As[buf][idx >> 7][idx & (GBM - 1)] = ra[it];
            Bs[buf][idx & (GBK - 1)][idx >> 4] = rb[it];
        }
    };
    const int nk = (K + GBK - 1) / GBK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int s = 0; s < nk; ++s) {
        const int buf = s & 1;
        if (s + 1 < nk) gload((s + 1) * GBK);
#pragma unroll
        for (int k4 = 0; k4 < GBK / 4; ++k4) {
            const int kr = k4 * 4 + (lane >> 4);
            double x[4], y[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) x[i] = Bs[buf][kr][wn * 64 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 4; ++j) y[j] = As[buf][kr][wm * 64 + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[i], y[j], acc[i][j], 0, 0, 0);
        }
        if (s + 1 < nk) {
            sstore(buf ^ 1);
            __syncthreads();
        }
    }
    if (mode == 1) __syncthreads();  // in place on B: every wave has finished reading B
    // lane holds C(row = rbase + (lane & 15), col = cbase + 4q + (lane >> 4)).  Sixteen C values
    // at a time are loaded before any of them is stored, so the loads overlap.
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double cv[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = row0 + wm * 64 + j * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int col = col0 + wn * 64 + i * 16 + 4 * q + (lane >> 4);
                cv[j][q] = (mode == 0 && row < M && col < N) ? Cg[(int64_t)col * ldc + row] : 0.0;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = row0 + wm * 64 + j * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int col = col0 + wn * 64 + i * 16 + 4 * q + (lane >> 4);
                if (row < M && col < N) Cg[(int64_t)col * ldc + row] = cv[j][q] - acc[i][j][q] * (mode == 0 ? 1.0 : -1.0);
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// Host driver (the blocked algorithm itself is lu_driver.hpp, shared with the mixed path).
static void gemm_f64(lsx_lu* lu, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                     int64_t ldc, int64_t M, int64_t N, int64_t K, int mode) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    const int64_t T = cdiv(M, GBM) * cdiv(N, GBN);
    const int64_t grid = 8 * cdiv(T, 8);
    const int pi = lu->prof.begin(lu->cur, PK_GEMM, 2.0 * M * N * K,
                                  sizeof(double) * (2.0 * M * N + (double)M * K + (double)K * N));
    hipLaunchKernelGGL(k_gemm_f64, dim3((unsigned)grid), dim3(256), 0, lu->cur, A, lda, B, ldb, C, ldc,
                       (int)M, (int)N, (int)K, mode);
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
    LSX_HIP(ctx, hipMalloc(&lu->wb, sizeof(double) * lu->n * cap));
    LSX_HIP(ctx, hipMalloc(&lu->wy, sizeof(double) * lu->n * cap));
    lu->wcap = cap;
    return LSX_OK;
}

int lu64_getrs_dev(lsx_lu* lu, int64_t nrhs, const double* dB, int64_t ldb, double* dX, int64_t ldx) {
    return getrs_blocked<double>(lu, lu->F, lu->ld, lu->DL, lu->DU, nrhs, dB, ldb, dX, ldx);
}

}  // namespace lsx
