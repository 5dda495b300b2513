// lu_driver.hpp — the blocked right-looking getrf and the getrs of the resident factors,
// templated on the working type; the trailing-update GEMM is supplied by the caller (FP64 MFMA
// for the plain path, FP32 or BF16 MFMA for the mixed path).
//
// Structure of LAPACK dgetrf and of src/blocked_lufact.jl:658-679, with two block levels:
// outer blocks of NBO columns (K of the trailing update) split into inner panels of PW columns.
#pragma once
#include <algorithm>

#include "lu_kernels.hpp"

namespace lsx {

template <typename T>
static void launch_panel_t(lsx_lu* lu, T* F, int64_t ld, int64_t j0, int W) {
    const int64_t m_rows = lu->n - j0;
    PanelArgs<T> a;
    a.A = F;
    a.ld = ld;
    a.m = lu->n;
    a.j0 = j0;
    a.W = W;
    a.ipiv = lu->ipiv;
    a.dinfo = lu->dinfo;
    a.slots = lu->slots;
    a.seq = lu->seq;
    a.timeout = lu->ctx->wall_khz * 5000;  // 5 s
    lu->seq += (uint32_t)W + 1u;
    const int pi = lu->prof.begin(lu->stream, PK_PANEL, (double)m_rows * W * W,
                                  2.0 * sizeof(T) * (double)m_rows * W);
    if (m_rows <= 256) {
        a.P = 1;
        hipLaunchKernelGGL((k_panel<T, 256>), dim3(1), dim3(256), 0, lu->stream, a);
    } else {
        a.P = (int)cdiv(m_rows, 512);
        hipLaunchKernelGGL((k_panel<T, 512>), dim3(a.P), dim3(512), 0, lu->stream, a);
    }
    lu->prof.end(lu->stream, pi);
}

template <typename T>
static void launch_laswp_t(lsx_lu* lu, T* F, int64_t ld, int64_t k0, int kb, const int64_t* list,
                           const int* count, int64_t c0, int64_t c1, int trsm) {
    if (c1 <= c0) return;
    const int64_t nb = cdiv(c1 - c0, LCW);
    const int pi = lu->prof.begin(lu->stream, PK_LASWP, trsm ? (double)kb * kb * (c1 - c0) : 0.0,
                                  sizeof(T) * 4.0 * kb * (c1 - c0));
    hipLaunchKernelGGL(k_laswp_trsm<T>, dim3((unsigned)nb), dim3(256), 0, lu->stream, F, ld, k0, kb, list, count,
                       c0, c1, trsm);
    lu->prof.end(lu->stream, pi);
}

// Factors F (n x n, ld) in place; pivots into lu->ipiv, row map into lu->perm, diagonal-block
// inverses into DL / DU.  gemm_upd does the inner and trailing updates (mode 0: C -= A B),
// gemm_trsm the products of the blocked TRSM (mode 1: C = A B in place, mode 0).
template <typename T, typename GU, typename GT>
int getrf_blocked(lsx_lu* lu, T* F, int64_t ld, T* DL, T* DU, GU gemm_upd, GT gemm_trsm) {
    lsx_ctx* ctx = lu->ctx;
    const int64_t n = lu->n;
    if (lu->seq > 0xF0000000u) {
        LSX_HIP(ctx, hipMemsetAsync(lu->slots, 0, sizeof(uint64_t) * 2 * (PMAX + 1) * RECS, lu->stream));
        lu->seq = 1;
    }
    LSX_HIP(ctx, hipMemsetAsync(lu->dinfo, 0, 2 * sizeof(int64_t), lu->stream));
    hipLaunchKernelGGL(k_iota, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, lu->stream, lu->perm, n);
    int64_t* il = lu->plist;
    int64_t* ol = lu->plist + 4 * NBO;
    int* ic = lu->pcount;
    int* oc = lu->pcount + 1;
    for (int64_t jo = 0; jo < n; jo += NBO) {
        const int jbo = (int)std::min<int64_t>(NBO, n - jo);
        // inner panels of PW columns inside the outer block
        for (int64_t ji = jo; ji < jo + jbo; ji += PW) {
            const int W = (int)std::min<int64_t>(PW, jo + jbo - ji);
            launch_panel_t<T>(lu, F, ld, ji, W);
            const bool left = ji > jo, right = ji + W < jo + jbo;
            if (left || right) {
                hipLaunchKernelGGL(k_perm_build, dim3(1), dim3(256), 0, lu->stream, lu->ipiv, ji, W, il, ic,
                                   (int*)nullptr);
                if (left) launch_laswp_t<T>(lu, F, ld, ji, W, il, ic, jo, ji, 0);
                if (right) {
                    launch_laswp_t<T>(lu, F, ld, ji, W, il, ic, ji + W, jo + jbo, 1);
                    gemm_upd(F + ji * ld + (ji + W), ld, F + (ji + W) * ld + ji, ld, F + (ji + W) * ld + (ji + W), ld,
                             n - ji - W, jo + jbo - ji - W, W, 0);
                }
            }
        }
        // inverses of the diagonal 64-blocks of this outer block (final from here on)
        const int64_t b0 = jo / DB, nbk = cdiv(jbo, DB);
        {
            const int pi = lu->prof.begin(lu->stream, PK_OTHER, 0.0, 0.0);
            hipLaunchKernelGGL(k_diag_inv<T>, dim3((unsigned)nbk, 2), dim3(DB), 0, lu->stream, F, ld, n, b0, DL, DU);
            lu->prof.end(lu->stream, pi);
        }
        hipLaunchKernelGGL(k_perm_build, dim3(1), dim3(256), 0, lu->stream, lu->ipiv, jo, jbo, ol, oc, lu->perm);
        if (jo > 0) launch_laswp_t<T>(lu, F, ld, jo, jbo, ol, oc, 0, jo, 0);
        if (jo + jbo < n) {
            const int64_t c1 = jo + jbo, nr = n - c1;
            launch_laswp_t<T>(lu, F, ld, jo, jbo, ol, oc, c1, n, 0);
            // U12 = L11^{-1} A12 with L11 = [L1 0; L2 L3] (64-blocks):
            //   X1 = L1^{-1} A12[0:64];  A12[64:] -= L2 X1;  X2 = L3^{-1} A12[64:]
            T* A12 = F + c1 * ld + jo;
            const int64_t h1 = std::min<int64_t>(DB, jbo), h2 = jbo - h1;
            gemm_trsm(DL + b0 * DB * DB, DB, A12, ld, A12, ld, h1, nr, h1, 1);
            if (h2 > 0) {
                gemm_trsm(F + jo * ld + jo + DB, ld, A12, ld, A12 + DB, ld, h2, nr, DB, 0);
                gemm_trsm(DL + (b0 + 1) * DB * DB, DB, A12 + DB, ld, A12 + DB, ld, h2, nr, h2, 1);
            }
            // trailing update A22 -= L21 U12
            gemm_upd(F + jo * ld + c1, ld, A12, ld, F + c1 * ld + c1, ld, nr, nr, jbo, 0);
        }
    }
    LSX_HIP(ctx, hipGetLastError());
    LSX_HIP(ctx, hipMemcpyAsync(lu->hinfo, lu->dinfo, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, lu->stream));
    LSX_HIP(ctx, hipStreamSynchronize(lu->stream));
    lu->prof.collect();
    if (lu->hinfo[1] != 0)
        return fail(ctx, LSX_ETIMEOUT, "getrf: a panel exchange wait exceeded its bound (flags %lld)",
                    (long long)lu->hinfo[1]);
    return (int)std::min<int64_t>(lu->hinfo[0], INT32_MAX);
}

// X = A \ B with the resident factors (device buffers; FP64 right-hand sides, arithmetic in
// FP64 whatever the factor type).  The lower solve writes lu->wy; X may alias B.
template <typename T>
int getrs_blocked(lsx_lu* lu, const T* F, int64_t ld, const T* DL, const T* DU, int64_t nrhs, const double* dB,
                  int64_t ldb, double* dX, int64_t ldx) {
    lsx_ctx* ctx = lu->ctx;
    const int64_t n = lu->n;
    const int nblk = (int)cdiv(n, DB);
    if (lu->tseq > 0xF0000000u) {
        LSX_HIP(ctx, hipMemsetAsync(lu->tflags, 0, sizeof(uint32_t) * 2 * nblk, lu->stream));
        lu->tseq = 1;
    }
    LSX_HIP(ctx, hipMemsetAsync(lu->dinfo + 1, 0, sizeof(int64_t), lu->stream));
    const int64_t timeout = ctx->wall_khz * 5000;
    for (int64_t c0 = 0; c0 < nrhs; c0 += TRHS) {
        const int nr = (int)std::min<int64_t>(TRHS, nrhs - c0);
        const uint32_t s1 = lu->tseq++, s2 = lu->tseq++;
        const int pi = lu->prof.begin(lu->stream, PK_TRSV, 2.0 * n * n * nr, sizeof(T) * (double)n * n);
        hipLaunchKernelGGL((k_trsv<T, true>), dim3(nblk), dim3(256), 0, lu->stream, F, ld, n, lu->perm, DL,
                           dB + c0 * ldb, ldb, lu->wy, n, nr, lu->tflags, s1, timeout, lu->dinfo + 1);
        hipLaunchKernelGGL((k_trsv<T, false>), dim3(nblk), dim3(256), 0, lu->stream, F, ld, n, lu->perm, DU,
                           (const double*)lu->wy, n, dX + c0 * ldx, ldx, nr, lu->tflags + nblk, s2, timeout,
                           lu->dinfo + 1);
        lu->prof.end(lu->stream, pi);
    }
    LSX_HIP(ctx, hipGetLastError());
    return LSX_OK;
}

}  // namespace lsx
