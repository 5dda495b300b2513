// gemm.hpp — the trailing-update GEMM of the LU factorizations on MFMA, templated on the
// element type (FP64 for the plain path, FP32 for the mixed path).
#pragma once
#include "common.hpp"

namespace lsx {

typedef double gv4d __attribute__((ext_vector_type(4)));
typedef float gv4f __attribute__((ext_vector_type(4)));

// MFMA 16 x 16 x 4 of the element type and where its results sit: lane l / slot q holds the
// transposed tile's D(i, j) with j = l & 15 (a row of C) and i = ccol(l, q) (a column of C).
// The FP64 and FP32 instructions place the four slots differently (measured on the device).
template <typename T> struct GemmMma;
template <> struct GemmMma<double> {
    typedef gv4d acc_t;
    __device__ static acc_t mma(double x, double y, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0); }
    __device__ static int ccol(int lane, int q) { return 4 * q + (lane >> 4); }
};
template <> struct GemmMma<float> {
    typedef gv4f acc_t;
    __device__ static acc_t mma(float x, float y, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0); }
    __device__ static int ccol(int lane, int q) { return 4 * (lane >> 4) + q; }
};

// Trailing update on MFMA (FP64 v_mfma_f64_16x16x4f64 or FP32 v_mfma_f32_16x16x4f32), LDS-staged:
//   MODE 0: C -= A * B      MODE 1: C = A * B (in place on B when M <= BM: a workgroup reads all
//                           K rows of its B columns before it writes C)
// A: M x K (lda), B: K x N (ldb), C: M x N (ldc), column-major.  BM x BN tile per workgroup
// (128 x 128, or 64 x 64 when the big tiles cannot fill the GPU), NT / 64 waves of
// BM/2 x BN/(NT/128) (two along the rows), K staged 16 at a time with double buffering.  The MFMA computes the transposed tile
// (B^T A^T) so each lane's results run down a column of C (128-byte stores).
constexpr int GBK = 16;
constexpr int GGM = 8;  // tile-rows per group of the tile order (L2 reuse of A and B blocks)
template <typename T, int MODE, int BM, int BN, int NT = 256>
__global__ __launch_bounds__(NT, 2 * NT / 256) void k_gemm(const T* __restrict Ag, int64_t lda, const T* Bg, int64_t ldb,
                                                           T* Cg, int64_t ldc, int M, int N, int K) {
    typedef GemmMma<T> G;
    constexpr int WNW = NT / 128;                    // waves along C columns (two along rows)
    constexpr int WI = BN / (16 * WNW), WJ = BM / 32;  // 16 x 16 blocks per wave along C columns / rows
    constexpr int LA = BM * GBK / NT, LB = BN * GBK / NT;  // staged elements per thread
    const int TM = (M + BM - 1) / BM, TN = (N + BN - 1) / BN, NTL = TM * TN;
    // XCD-aware: workgroup b runs on XCD b mod 8, which takes a contiguous range of the tile
    // order; in that order groups of GGM tile-rows are swept column by column, so the tiles
    // running at once on an XCD share few A and B blocks in its L2
    const int per = (NTL + 7) / 8;
    const int b = blockIdx.x;
    const int tile = (b & 7) * per + (b >> 3);
    if (tile >= NTL) return;
    const int grp = tile / (GGM * TN), first = grp * GGM, gsz = min(TM - first, GGM);
    const int rem = tile - grp * GGM * TN;
    const int tm = first + rem % gsz, tn = rem / gsz;
    const int row0 = tm * BM, col0 = tn * BN;

    __shared__ T As[2][GBK][BM + 1];
    __shared__ T Bs[2][GBK][BN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;  // wave (wm, wn): rows wm BM/2 .., columns wn BN/WNW ..
    typename G::acc_t acc[WI][WJ];
#pragma unroll
    for (int i = 0; i < WI; ++i)
#pragma unroll
        for (int j = 0; j < WJ; ++j) acc[i][j] = typename G::acc_t{T(0), T(0), T(0), T(0)};

    // lane holds C(row = rbase + (lane & 15), col = cbase + G::ccol(lane, q)) of each 16 x 16 block
    auto crow = [&](int j) { return row0 + wm * (BM / 2) + j * 16 + (lane & 15); };
    auto ccol = [&](int i, int q) { return col0 + wn * (BN / WNW) + i * 16 + G::ccol(lane, q); };

    T ra[LA], rb[LB];
    auto gload = [&](int k0) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int idx = it * NT + tid;
            const int r = idx % BM, k = idx / BM;
            const int gr = row0 + r, gk = k0 + k;
            ra[it] = (gr < M && gk < K) ? Ag[(int64_t)gk * lda + gr] : T(0);
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int idx = it * NT + tid;
            const int kb = idx & (GBK - 1), c = idx >> 4;
            const int gc = col0 + c, gk2 = k0 + kb;
            rb[it] = (gc < N && gk2 < K) ? Bg[(int64_t)gc * ldb + gk2] : T(0);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int idx = it * NT + tid;
            As[buf][idx / BM][idx % BM] = ra[it];
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int idx = it * NT + tid;
            Bs[buf][idx & (GBK - 1)][idx >> 4] = rb[it];
        }
    };
    // the first WJ*4 C values of the epilogue, prefetched into the staging registers during the
    // last K stage (MODE 0, when they fit)
    constexpr bool PREF = (MODE == 0) && (WJ * 4 <= LA + LB);
    auto cpre = [&](int j, int q) -> T& { const int e = j * 4 + q; return e < LA ? ra[e] : rb[e - LA]; };
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
                    cpre(j, q) = (row < M && col < N) ? Cg[(int64_t)col * ldc + row] : T(0);
                }
        }
#pragma unroll
        for (int k4 = 0; k4 < GBK / 4; ++k4) {
            const int kr = k4 * 4 + (lane >> 4);
            T x[WI], y[WJ];
#pragma unroll
            for (int i = 0; i < WI; ++i) x[i] = Bs[buf][kr][wn * (BN / WNW) + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < WJ; ++j) y[j] = As[buf][kr][wm * (BM / 2) + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < WI; ++i)
#pragma unroll
                for (int j = 0; j < WJ; ++j)
                    acc[i][j] = G::mma(x[i], y[j], acc[i][j]);
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
        T cv[WJ][4];
#pragma unroll
        for (int j = 0; j < WJ; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = crow(j), col = ccol(i, q);
                if (MODE == 1) cv[j][q] = T(0);
                else if (PREF && i == 0) cv[j][q] = cpre(j, q);
                else cv[j][q] = (row < M && col < N) ? Cg[(int64_t)col * ldc + row] : T(0);
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

template <typename T, int MODE, int BM, int BN, int NT = 256>
static void launch_gemm(hipStream_t st, const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc, int64_t M,
                        int64_t N, int64_t K) {
    const int64_t T_ = cdiv(M, BM) * cdiv(N, BN);
    hipLaunchKernelGGL((k_gemm<T, MODE, BM, BN, NT>), dim3((unsigned)(8 * cdiv(T_, 8))), dim3(NT), 0, st, A, lda, B,
                       ldb, C, ldc, (int)M, (int)N, (int)K);
}

// Threads per tile: eight waves (64 x 32 each on a 128 x 128 tile, 32 x 16 on a 64 x 64 tile),
// two workgroups per CU.  Four waves of 64 x 64 per 128 x 128 tile need 255 registers, so a CU
// held eight waves; these take 128, so it holds sixteen (FP64 trailing update 812 -> 722 us).
constexpr int GEMM_NT = 512;

// C -= A B (mode 0) or C = A B in place on B (mode 1, M <= 64) on stream st: 128 x 128 tiles
// when they give every CU two workgroups, 64 x 64 otherwise.  Returns whether big tiles ran.
template <typename T>
static bool gemm_dispatch(hipStream_t st, int num_cus, const T* A, int64_t lda, const T* B, int64_t ldb, T* C,
                          int64_t ldc, int64_t M, int64_t N, int64_t K, int mode) {
    const bool big = mode == 0 && cdiv(M, 128) * cdiv(N, 128) >= 2 * (int64_t)num_cus;
    if (mode == 0) {
        if (big) launch_gemm<T, 0, 128, 128, GEMM_NT>(st, A, lda, B, ldb, C, ldc, M, N, K);
        else launch_gemm<T, 0, 64, 64, GEMM_NT>(st, A, lda, B, ldb, C, ldc, M, N, K);
    } else {
        if (M <= 64) launch_gemm<T, 1, 64, 64, GEMM_NT>(st, A, lda, B, ldb, C, ldc, M, N, K);
        else launch_gemm<T, 1, 128, 128>(st, A, lda, B, ldb, C, ldc, M, N, K);
    }
    return big;
}

}  // namespace lsx
