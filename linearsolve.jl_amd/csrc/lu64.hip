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
//   MODE 0: C -= A * B      MODE 1: C = A * B (in place on B when M <= BM: a workgroup reads all
//                           K rows of its B columns before it writes C)
// A: M x K (lda), B: K x N (ldb), C: M x N (ldc), column-major.  BM x BN tile per workgroup
// (128 x 128, or 64 x 64 when the big tiles cannot fill the GPU), four waves of BM/2 x BN/2,
// K staged 16 at a time with double buffering.  The MFMA computes the transposed tile
// (B^T A^T) so each lane's results run down a column of C (128-byte stores).
constexpr int GBK = 16;
constexpr int GGM = 8;  // tile-rows per group of the tile order (L2 reuse of A and B blocks)
template <int MODE, int BM, int BN>
__global__ __launch_bounds__(256, 2) void k_gemm_f64(const double* __restrict Ag, int64_t lda,
                                                     const double* Bg, int64_t ldb, double* Cg,
                                                     int64_t ldc, int M, int N, int K) {
    constexpr int WI = BN / 32, WJ = BM / 32;       // 16 x 16 blocks per wave along C columns / rows
    constexpr int LA = BM * GBK / 256, LB = BN * GBK / 256;  // staged elements per thread
    const int TM = (M + BM - 1) / BM, TN = (N + BN - 1) / BN, T = TM * TN;
    // XCD-aware: workgroup b runs on XCD b mod 8, which takes a contiguous range of the tile
    // order; in that order groups of GGM tile-rows are swept column by column, so the tiles
    // running at once on an XCD share few A and B blocks in its L2
    const int per = (T + 7) / 8;
    const int b = blockIdx.x;
    const int tile = (b & 7) * per + (b >> 3);
    if (tile >= T) return;
    const int grp = tile / (GGM * TN), first = grp * GGM, gsz = min(TM - first, GGM);
    const int rem = tile - grp * GGM * TN;
    const int tm = first + rem % gsz, tn = rem / gsz;
    const int row0 = tm * BM, col0 = tn * BN;

    __shared__ double As[2][GBK][BM + 1];
    __shared__ double Bs[2][GBK][BN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    v4d acc[WI][WJ];
#pragma unroll
    for (int i = 0; i < WI; ++i)
#pragma unroll
        for (int j = 0; j < WJ; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    // lane holds C(row = rbase + (lane & 15), col = cbase + 4q + (lane >> 4)) of each 16 x 16 block
    auto crow = [&](int j) { return row0 + wm * (BM / 2) + j * 16 + (lane & 15); };
    auto ccol = [&](int i, int q) { return col0 + wn * (BN / 2) + i * 16 + 4 * q + (lane >> 4); };

    double ra[LA], rb[LB];
    auto gload = [&](int k0) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int idx = it * 256 + tid;
            const int r = idx % BM, k = idx / BM;
            const int gr = row0 + r, gk = k0 + k;
            ra[it] = (gr < M && gk < K) ? Ag[(int64_t)gk * lda + gr] : 0.0;
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int idx = it * 256 + tid;
            const int kb = idx & (GBK - 1), c = idx >> 4;
            const int gc = col0 + c, gk2 = k0 + kb;
            rb[it] = (gc < N && gk2 < K) ? Bg[(int64_t)gc * ldb + gk2] : 0.0;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int idx = it * 256 + tid;
            As[buf][idx / BM][idx % BM] = ra[it];
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int idx = it * 256 + tid;
            Bs[buf][idx & (GBK - 1)][idx >> 4] = rb[it];
        }
    };
    // the first WJ*4 C values of the epilogue, prefetched into the staging registers during the
    // last K stage (MODE 0, when they fit)
    constexpr bool PREF = (MODE == 0) && (WJ * 4 <= LA + LB);
    auto cpre = [&](int j, int q) -> double& { const int e = j * 4 + q; return e < LA ? ra[e] : rb[e - LA]; };
    const int nk = (K + GBK - 1) / GBK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int s = 0; s < nk; ++s) {
        const int buf = s & 1;
        if (s + 1 < nk) {
            gload((s + 1) * GBK);
        } else if (PREF) {
#pragma unroll
            for (int j = 0; j < WJ; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = crow(j), col = ccol(0, q);
                    cpre(j, q) = (row < M && col < N) ? Cg[(int64_t)col * ldc + row] : 0.0;
                }
        }
#pragma unroll
        for (int k4 = 0; k4 < GBK / 4; ++k4) {
            const int kr = k4 * 4 + (lane >> 4);
            double x[WI], y[WJ];
#pragma unroll
            for (int i = 0; i < WI; ++i) x[i] = Bs[buf][kr][wn * (BN / 2) + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < WJ; ++j) y[j] = As[buf][kr][wm * (BM / 2) + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < WI; ++i)
#pragma unroll
                for (int j = 0; j < WJ; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[i], y[j], acc[i][j], 0, 0, 0);
        }
        if (s + 1 < nk) {
            sstore(buf ^ 1);
            __syncthreads();
        }
    }
    if (MODE == 1) __syncthreads();  // in place on B: every wave has finished reading B
    // epilogue: the C values of one 16-column strip are loaded before any of them is stored
#pragma unroll
    for (int i = 0; i < WI; ++i) {
        double cv[WJ][4];
#pragma unroll
        for (int j = 0; j < WJ; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = crow(j), col = ccol(i, q);
                if (MODE == 1) cv[j][q] = 0.0;
                else if (PREF && i == 0) cv[j][q] = cpre(j, q);
                else cv[j][q] = (row < M && col < N) ? Cg[(int64_t)col * ldc + row] : 0.0;
            }
#pragma unroll
        for (int j = 0; j < WJ; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = crow(j), col = ccol(i, q);
                if (row < M && col < N) Cg[(int64_t)col * ldc + row] = (MODE == 0) ? cv[j][q] - acc[i][j][q] : acc[i][j][q];
            }
    }
}

// ------------------------------------------------------------------------------------------
// Host driver (the blocked algorithm itself is lu_driver.hpp, shared with the mixed path).
template <int MODE, int BM, int BN>
static void launch_gemm_f64(hipStream_t st, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                            int64_t ldc, int64_t M, int64_t N, int64_t K) {
    const int64_t T = cdiv(M, BM) * cdiv(N, BN);
    hipLaunchKernelGGL((k_gemm_f64<MODE, BM, BN>), dim3((unsigned)(8 * cdiv(T, 8))), dim3(256), 0, st, A, lda, B,
                       ldb, C, ldc, (int)M, (int)N, (int)K);
}

static void gemm_f64(lsx_lu* lu, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                     int64_t ldc, int64_t M, int64_t N, int64_t K, int mode) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    // 128 x 128 tiles when they give every CU two workgroups, 64 x 64 otherwise
    const bool big = mode == 0 && cdiv(M, 128) * cdiv(N, 128) >= 2 * (int64_t)lu->ctx->num_cus;
    const int pi = lu->prof.begin(lu->cur, big ? PK_GEMM : PK_GEMM_SMALL, 2.0 * M * N * K,
                                  sizeof(double) * (2.0 * M * N + (double)M * K + (double)K * N));
    if (mode == 0) {
        if (big) launch_gemm_f64<0, 128, 128>(lu->cur, A, lda, B, ldb, C, ldc, M, N, K);
        else launch_gemm_f64<0, 64, 64>(lu->cur, A, lda, B, ldb, C, ldc, M, N, K);
    } else {
        // in place on B needs one row of tiles: M <= 64 (the TRSM products, h1/h2 <= 64)
        if (M <= 64) launch_gemm_f64<1, 64, 64>(lu->cur, A, lda, B, ldb, C, ldc, M, N, K);
        else launch_gemm_f64<1, 128, 128>(lu->cur, A, lda, B, ldb, C, ldc, M, N, K);
    }
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
