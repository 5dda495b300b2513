// gemm_split.hpp — the trailing update of the mixed-precision LU on BF16 MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation), with each FP32 operand split into S BF16 pieces
// (S = 3 for FP32 factors, S = 2 for BF16 factors):
//   x = x0 + x1 + x2 (+ r),  x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1),
// every piece within 2^-9 of what is left, so |r| <= 2^-9S |x|.  A product x y is formed from the
// pieces' products whose order (i + j) is below S:
//   S = 1: x0 y0                                  (one BF16 product, relative error ~2^-8)
//   S = 2: x0 y0 + x0 y1 + x1 y0                  (three products, ~3 2^-18)
//   S = 3: the six products of order <= 2         (~4 2^-27 = 2^-25, below FP32's 2^-24)
// (the BF16x3 / BF16x6 schemes of Henry, Tang and Heinecke, "Leveraging the bfloat16 artificial
// intelligence datatype for higher-precision computations", ARITH 2019).  Products of two BF16
// pieces are exact in FP32, so S = 3 gives the FP32 trailing update at BF16 MFMA rates; it is
// bound by the FP32 read and write of C (8 bytes per element per update), not by the MFMA.
//
// C -= A B, A: M x K (lda), B: K x N (ldb), C: M x N (ldc), column-major FP32.  128 x 128 tiles,
// eight waves of 64 x 32 (two per SIMD, so one wave's loads overlap another's MFMAs); K staged 32
// at a time, split into BF16 pieces as it is stored in LDS
// (A as [row][k], B as [col][k], so a lane's MFMA operand is one 16-byte read).  The MFMA puts
// four consecutive rows of one C column in a lane, so C is read and written as float4.
#pragma once
#include "common.hpp"

namespace lsx {

typedef __bf16 sv8bf __attribute__((ext_vector_type(8)));
typedef float sv4f __attribute__((ext_vector_type(4)));

constexpr int SBM = 128, SBN = 128, SBK = 32, SPAD = 8;

// The S BF16 pieces of eight FP32 values.
template <int S>
__device__ __forceinline__ void split8(const float (&x)[8], sv8bf (&p)[S]) {
    float r[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) r[e] = x[e];
#pragma unroll
    for (int s = 0; s < S; ++s) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const __bf16 h = (__bf16)r[e];
            p[s][e] = h;
            if (s + 1 < S) r[e] = r[e] - (float)h;  // exact: h is r rounded to 8 bits
        }
    }
}

template <int S>
__global__ __launch_bounds__(512, 1) void k_gemm_split(const float* __restrict Ag, int64_t lda,
                                                      const float* __restrict Bg, int64_t ldb, float* Cg,
                                                      int64_t ldc, int M, int N, int K) {
    constexpr int GGM_ = 8;
    const int TM = (M + SBM - 1) / SBM, TN = (N + SBN - 1) / SBN, NTL = TM * TN;
    // XCD-aware tile order: workgroup b runs on XCD b mod 8, which takes a contiguous range of the
    // order; in it, groups of 8 tile-rows are swept column by column (A and B blocks reused in L2)
    const int per = (NTL + 7) / 8;
    const int b = blockIdx.x;
    const int tile = (b & 7) * per + (b >> 3);
    if (tile >= NTL) return;
    const int grp = tile / (GGM_ * TN), first = grp * GGM_, gsz = min(TM - first, GGM_);
    const int rem = tile - grp * GGM_ * TN;
    const int tm = first + rem % gsz, tn = rem / gsz;
    const int row0 = tm * SBM, col0 = tn * SBN;

    __shared__ __align__(16) __bf16 As[2][S][SBM][SBK + SPAD];
    __shared__ __align__(16) __bf16 Bs[2][S][SBN][SBK + SPAD];
    // eight waves, 2 x 4 over the tile: wave (wm, wn) computes rows wm*64 .. +64, columns wn*32 .. +32
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    sv4f acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = sv4f{0.f, 0.f, 0.f, 0.f};

    // staging: one (row, group of 8 k) pair of A and one (column, group of 8 k) pair of B per thread
    float ra[8], rb[8];
    const bool bvec = ((ldb & 3) == 0) && ((reinterpret_cast<uintptr_t>(Bg) & 15) == 0);
    const int ar = tid & (SBM - 1), akg = tid >> 7;  // A: rows across lanes (coalesced per k)
    const int bkb = tid & 3, bc = tid >> 2;          // B: 8 consecutive k of one column per thread
    auto gload = [&](int k0) {
        const int gr = row0 + ar;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int gk = k0 + akg * 8 + e;
            ra[e] = (gr < M && gk < K) ? Ag[(int64_t)gk * lda + gr] : 0.f;
        }
        const int gc = col0 + bc;
        const float* bp = Bg + (int64_t)gc * ldb + k0 + bkb * 8;
        if (bvec && gc < N && k0 + bkb * 8 + 7 < K) {
            const sv4f u0 = *reinterpret_cast<const sv4f*>(bp);
            const sv4f u1 = *reinterpret_cast<const sv4f*>(bp + 4);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                rb[e] = u0[e];
                rb[e + 4] = u1[e];
            }
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int gk = k0 + bkb * 8 + e;
                rb[e] = (gc < N && gk < K) ? bp[e] : 0.f;
            }
        }
    };
    auto sstore = [&](int buf) {
        sv8bf pa[S], pb[S];
        split8<S>(ra, pa);
        split8<S>(rb, pb);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            *reinterpret_cast<sv8bf*>(&As[buf][s][ar][akg * 8]) = pa[s];
            *reinterpret_cast<sv8bf*>(&Bs[buf][s][bc][bkb * 8]) = pb[s];
        }
    };
    // epilogue layout: lane holds rows rb + 4 (lane >> 4) + q (q < 4) of column cb + (lane & 15) of
    // each 16 x 16 block; the C values are loaded during the last K stage
    const bool vec = ((ldc & 3) == 0) && ((reinterpret_cast<uintptr_t>(Cg) & 15) == 0);
    sv4f cv[4][2];
    auto cload = [&]() {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = col0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = row0 + wm * 64 + i * 16 + 4 * (lane >> 4);
                const float* cp = Cg + (int64_t)col * ldc + row;
                if (col < N && vec && row + 3 < M) {
                    cv[i][j] = *reinterpret_cast<const sv4f*>(cp);
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q) cv[i][j][q] = (col < N && row + q < M) ? cp[q] : 0.f;
                }
            }
        }
    };
    const int nk = (K + SBK - 1) / SBK;
    gload(0);
    sstore(0);
    __syncthreads();
    const int kq = 8 * (lane >> 4);
    for (int st = 0; st < nk; ++st) {
        const int buf = st & 1;
        if (st + 1 < nk) gload((st + 1) * SBK);
        else cload();
        sv8bf x[4][S], y[2][S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                x[i][s] = *reinterpret_cast<const sv8bf*>(&As[buf][s][wm * 64 + i * 16 + (lane & 15)][kq]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                y[j][s] = *reinterpret_cast<const sv8bf*>(&Bs[buf][s][wn * 32 + j * 16 + (lane & 15)][kq]);
        }
        // the small terms first, the leading x0 y0 last; each term over all eight accumulators, so
        // consecutive MFMAs are independent
        constexpr int NTERM = S == 1 ? 1 : (S == 2 ? 3 : 6);
        constexpr int TA[6] = {2, 1, 0, 1, 0, 0}, TB[6] = {0, 1, 2, 0, 1, 0};  // piece of x / of y
#pragma unroll
        for (int t = 6 - NTERM; t < 6; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x[i][TA[t]], y[j][TB[t]], acc[i][j], 0, 0, 0);
        if (st + 1 < nk) {
            sstore(buf ^ 1);
            __syncthreads();
        }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int col = col0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = row0 + wm * 64 + i * 16 + 4 * (lane >> 4);
            float* cp = Cg + (int64_t)col * ldc + row;
            sv4f nv;
#pragma unroll
            for (int q = 0; q < 4; ++q) nv[q] = cv[i][j][q] - acc[i][j][q];
            if (col < N && vec && row + 3 < M) {
                *reinterpret_cast<sv4f*>(cp) = nv;
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (col < N && row + q < M) cp[q] = nv[q];
            }
        }
    }
}

// C -= A B with S BF16 pieces per operand on stream st.
template <int S>
static void launch_gemm_split(hipStream_t st, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                              int64_t ldc, int64_t M, int64_t N, int64_t K) {
    const int64_t T_ = cdiv(M, SBM) * cdiv(N, SBN);
    const dim3 grid((unsigned)(8 * cdiv(T_, 8)));
    hipLaunchKernelGGL((k_gemm_split<S>), grid, dim3(512), 0, st, A, lda, B, ldb, C, ldc, (int)M, (int)N, (int)K);
}

}  // namespace lsx
