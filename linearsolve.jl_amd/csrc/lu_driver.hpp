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
    a.timeout = wait_ticks(lu->ctx, 5.0);
    lu->seq += (uint32_t)W + 1u;
    const int pi = lu->prof.begin(lu->cur, PK_PANEL, (double)m_rows * W * W,
                                  2.0 * sizeof(T) * (double)m_rows * W);
    if (m_rows <= 256) {
        a.P = 1;
        hipLaunchKernelGGL((k_panel<T, 256>), dim3(1), dim3(256), 0, lu->cur, a);
    } else {
        a.P = (int)cdiv(m_rows, 512);
        hipLaunchKernelGGL((k_panel<T, 512>), dim3(a.P), dim3(512), 0, lu->cur, a);
    }
    lu->prof.end(lu->cur, pi);
}

template <typename T>
static void launch_laswp_t(lsx_lu* lu, T* F, int64_t ld, int64_t k0, int kb, const int64_t* list,
                           const int* count, int64_t c0, int64_t c1, int trsm) {
    if (c1 <= c0) return;
    const int64_t nb = cdiv(c1 - c0, LCW);
    const int pi = lu->prof.begin(lu->cur, PK_LASWP, trsm ? (double)kb * kb * (c1 - c0) : 0.0,
                                  sizeof(T) * 4.0 * kb * (c1 - c0));
    hipLaunchKernelGGL(k_laswp_trsm<T>, dim3((unsigned)nb), dim3(256), 0, lu->cur, F, ld, k0, kb, list, count,
                       c0, c1, trsm);
    lu->prof.end(lu->cur, pi);
}

// Factors F (n x n, ld) in place; pivots into lu->ipiv, row map into lu->perm, diagonal-block
// inverses into DL / DU.  gemm_upd does the inner and trailing updates (mode 0: C -= A B),
// gemm_trsm the products of the blocked TRSM (mode 1: C = A B in place, mode 0); both launch on
// lu->cur.
//
// Lookahead: the trailing update of outer block b first updates the columns of block b + 1 on
// lu->stream, then block b + 1 is factored on lu->stream2 while lu->stream updates the remaining
// columns.  The row interchanges of a block are applied right of it at once and, when a column
// segment fits in LDS, left of it (the L columns) in one pass at the end.
// The exact panels of a block (below) unless block_fact(jo, jbo) factors the whole block column
// itself, rows in place, and writes its pivots into lu->ipiv (the mixed path's tournament
// pivoting); then the block's interchanges are applied to its own columns here.
struct NoBlockFact {
    bool operator()(int64_t, int) const { return false; }
};

template <typename T, typename GU, typename GT, typename BF = NoBlockFact>
int getrf_blocked(lsx_lu* lu, T* F, int64_t ld, T* DL, T* DU, GU gemm_upd, GT gemm_trsm, BF block_fact = BF()) {
    lsx_ctx* ctx = lu->ctx;
    const int64_t n = lu->n;
    const hipStream_t s0 = lu->stream_upd ? lu->stream_upd : lu->stream, s1 = lu->stream2;
    const bool defer = lu->pinv != nullptr && (n - NBO) * (int64_t)sizeof(T) <= LGMAX_BYTES;
    if (s0 != lu->stream) {  // the input (copied on lu->stream) before anything on s0
        LSX_HIP(ctx, hipEventRecord(lu->ev_start, lu->stream));
        LSX_HIP(ctx, hipStreamWaitEvent(s0, lu->ev_start, 0));
    }
    lu->cur = s0;
    if (lu->seq > 0xF0000000u) {
        LSX_HIP(ctx, hipMemsetAsync(lu->slots, 0, sizeof(uint64_t) * 2 * (PMAX + 1) * RECS, s0));
        lu->seq = 1;
    }
    LSX_HIP(ctx, hipMemsetAsync(lu->dinfo, 0, 2 * sizeof(int64_t), s0));
    hipLaunchKernelGGL(k_iota, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s0, lu->perm, n);
    int64_t* il = lu->plist;
    int* ic = lu->pcount;

    // factorization of outer block [jo, jo + jbo) (its columns already updated): inner panels,
    // their interchanges and updates inside the block, diagonal-block inverses, the block's
    // interchange list (parity par) and the row map
    auto panel_block = [&](int64_t jo, int jbo, int par) -> int {
        if (block_fact(jo, jbo)) {
            // rows were factored in place: the block's interchange list (from here on the update
            // stream may apply it to the columns right of the block), its interchanges on its own
            // columns, then the inverses of its diagonal blocks (in the final row order)
            hipLaunchKernelGGL(k_perm_build, dim3(1), dim3(256), 0, lu->cur, lu->ipiv, jo, jbo,
                               lu->plist + (4 + 4 * par) * NBO, lu->pcount + 1 + par, lu->perm);
            LSX_HIP(ctx, hipEventRecord(lu->ev_perm, lu->cur));
            launch_laswp_t<T>(lu, F, ld, jo, jbo, lu->plist + (4 + 4 * par) * NBO, lu->pcount + 1 + par, jo,
                              jo + jbo, 0);
            const int64_t b0 = jo / DB, nbk = cdiv(jbo, DB);
            const int pi = lu->prof.begin(lu->cur, PK_OTHER, 0.0, 0.0);
            hipLaunchKernelGGL(k_diag_inv<T>, dim3((unsigned)nbk, 2), dim3(256), 0, lu->cur, F, ld, n, b0, DL, DU);
            lu->prof.end(lu->cur, pi);
            return LSX_OK;
        }
        for (int64_t ji = jo; ji < jo + jbo; ji += PW) {
            const int W = (int)std::min<int64_t>(PW, jo + jbo - ji);
            launch_panel_t<T>(lu, F, ld, ji, W);
            const bool left = ji > jo, right = ji + W < jo + jbo;
            if (left || right) {
                hipLaunchKernelGGL(k_perm_build, dim3(1), dim3(256), 0, lu->cur, lu->ipiv, ji, W, il, ic,
                                   (int*)nullptr);
                if (left) launch_laswp_t<T>(lu, F, ld, ji, W, il, ic, jo, ji, 0);
                if (right) {
                    launch_laswp_t<T>(lu, F, ld, ji, W, il, ic, ji + W, jo + jbo, 1);
                    gemm_upd(F + ji * ld + (ji + W), ld, F + (ji + W) * ld + ji, ld, F + (ji + W) * ld + (ji + W), ld,
                             n - ji - W, jo + jbo - ji - W, W, 0);
                }
            }
        }
        // the block's interchange list first: the update stream applies it to the columns right
        // of the block while the diagonal-block inverses are formed
        hipLaunchKernelGGL(k_perm_build, dim3(1), dim3(256), 0, lu->cur, lu->ipiv, jo, jbo,
                           lu->plist + (4 + 4 * par) * NBO, lu->pcount + 1 + par, lu->perm);
        LSX_HIP(ctx, hipEventRecord(lu->ev_perm, lu->cur));
        const int64_t b0 = jo / DB, nbk = cdiv(jbo, DB);
        const int pi = lu->prof.begin(lu->cur, PK_OTHER, 0.0, 0.0);
        hipLaunchKernelGGL(k_diag_inv<T>, dim3((unsigned)nbk, 2), dim3(256), 0, lu->cur, F, ld, n, b0, DL, DU);
        lu->prof.end(lu->cur, pi);
        return LSX_OK;
    };
    // the inverse of the row map after block jo, for the deferred L-column gather; launched after
    // the block's ev_panel so that the update stream does not wait for it
    auto inv_perm = [&](int64_t jo) {
        if (defer)
            hipLaunchKernelGGL(k_inv_perm, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, lu->cur, lu->perm,
                               lu->pinv + (jo / NBO) * n, n);
    };
    // interchanges, U12 = L11^{-1} A12 and A22 -= L21 U12 of block [jo, jo + jbo) on columns
    // [c0, c1); wait_inv: the block's diagonal-block inverses may still be in the making on the
    // panel stream (ev_panel), so the products wait for them after the interchanges
    auto update = [&](int64_t jo, int jbo, int par, int64_t c0, int64_t c1, bool wait_inv) -> int {
        if (c1 <= c0) return LSX_OK;
        const int64_t r1 = jo + jbo, nc = c1 - c0;
        launch_laswp_t<T>(lu, F, ld, jo, jbo, lu->plist + (4 + 4 * par) * NBO, lu->pcount + 1 + par, c0, c1, 0);
        if (wait_inv && lu->cur != s1) LSX_HIP(ctx, hipStreamWaitEvent(lu->cur, lu->ev_panel, 0));
        // L11 = [L1 0; L2 L3] (64-blocks): X1 = L1^{-1} A12[0:64];  A12[64:] -= L2 X1;  X2 = L3^{-1} A12[64:]
        T* A12 = F + c0 * ld + jo;
        const int64_t b0 = jo / DB;
        const int64_t h1 = std::min<int64_t>(DB, jbo), h2 = jbo - h1;
        gemm_trsm(DL + b0 * DB * DB, DB, A12, ld, A12, ld, h1, nc, h1, 1);
        if (h2 > 0) {
            gemm_trsm(F + jo * ld + jo + DB, ld, A12, ld, A12 + DB, ld, h2, nc, DB, 0);
            gemm_trsm(DL + (b0 + 1) * DB * DB, DB, A12 + DB, ld, A12 + DB, ld, h2, nc, h2, 1);
        }
        gemm_upd(F + jo * ld + r1, ld, A12, ld, F + c0 * ld + r1, ld, n - r1, nc, jbo, 0);
        return LSX_OK;
    };

    int rc_ = panel_block(0, (int)std::min<int64_t>(NBO, n), 0);
    if (rc_) return rc_;
    LSX_HIP(ctx, hipEventRecord(lu->ev_panel, s0));
    inv_perm(0);
    int par = 0;
    for (int64_t jo = 0; jo < n; jo += NBO, par ^= 1) {
        const int jbo = (int)std::min<int64_t>(NBO, n - jo);
        const int64_t c1 = jo + jbo;
        lu->cur = s0;
        if (!defer && jo > 0)
            launch_laswp_t<T>(lu, F, ld, jo, jbo, lu->plist + (4 + 4 * par) * NBO, lu->pcount + 1 + par, 0, jo, 0);
        if (c1 >= n) break;
        const int jbn = (int)std::min<int64_t>(NBO, n - c1);
        if ((rc_ = update(jo, jbo, par, c1, c1 + jbn, true))) return rc_;
        LSX_HIP(ctx, hipEventRecord(lu->ev_next, s0));
        LSX_HIP(ctx, hipStreamWaitEvent(s1, lu->ev_next, 0));
        lu->cur = s1;
        if ((rc_ = panel_block(c1, jbn, par ^ 1))) return rc_;
        LSX_HIP(ctx, hipEventRecord(lu->ev_panel, s1));
        inv_perm(c1);
        lu->cur = s0;
        if ((rc_ = update(jo, jbo, par, c1 + jbn, n, false))) return rc_;
        // the next block's interchange list (ev_perm) is all the next iteration's first step
        // needs; its products wait for the inverses (ev_panel) themselves
        LSX_HIP(ctx, hipStreamWaitEvent(s0, lu->ev_perm, 0));
    }
    lu->cur = s0;
    LSX_HIP(ctx, hipStreamWaitEvent(s0, lu->ev_panel, 0));  // the last block's inverses and row map
    if (defer && s1 != s0) {  // and every block's inverse row map (launched on s1 after its ev_panel)
        LSX_HIP(ctx, hipEventRecord(lu->ev_next, s1));
        LSX_HIP(ctx, hipStreamWaitEvent(s0, lu->ev_next, 0));
    }
    if (defer && n > NBO) {
        const int64_t ncols = (cdiv(n, NBO) - 1) * NBO;
        const size_t lds = sizeof(T) * (size_t)(n - NBO);
        const int pi = lu->prof.begin(s0, PK_LASWP, 0.0, 2.0 * sizeof(T) * (double)n * n / 2);
        hipLaunchKernelGGL(k_left_gather<T>, dim3((unsigned)ncols), dim3(256), lds, s0, F, ld, n,
                           (const int*)lu->perm, (const int*)lu->pinv);
        lu->prof.end(s0, pi);
    }
    LSX_HIP(ctx, hipGetLastError());
    LSX_HIP(ctx, hipMemcpyAsync(lu->hinfo, lu->dinfo, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, s0));
    LSX_HIP(ctx, hipStreamSynchronize(s0));
    lu->cur = lu->stream;
    lu->prof.collect();
    if (lu->hinfo[1] != 0)
        return fail(ctx, LSX_ETIMEOUT, "getrf: a panel exchange wait exceeded its bound (flags %lld)",
                    (long long)lu->hinfo[1]);
    return (int)std::min<int64_t>(lu->hinfo[0], INT32_MAX);
}

// X = A \ B with the resident factors (device buffers; FP64 right-hand sides, arithmetic in
// FP64 whatever the factor type).  The lower solve writes lu->wy; X may alias B.  A wait that
// gives up sets lu->dinfo[1], which the public entry point clears once (lu_clear_err) and reads
// after its last launch, so a timeout in any of its solves is reported.
template <typename T>
int getrs_blocked(lsx_lu* lu, const T* F, int64_t ld, const T* DL, const T* DU, int64_t nrhs, const double* dB,
                  int64_t ldb, double* dX, int64_t ldx) {
    lsx_ctx* ctx = lu->ctx;
    const int64_t n = lu->n;
    const int nblk = (int)cdiv(n, DB);
    if (lu->tseq > 0xF0000000u) {
        LSX_HIP(ctx, hipMemsetAsync(lu->tmsg, 0, sizeof(uint64_t) * nblk * TMW, lu->stream));
        lu->tseq = 1;
    }
    const int64_t timeout = wait_ticks(ctx, 5.0);
    for (int64_t c0 = 0; c0 < nrhs; c0 += TRHS) {
        const int nr = (int)std::min<int64_t>(TRHS, nrhs - c0);
        const uint32_t s1 = lu->tseq++, s2 = lu->tseq++;
        const int pi = lu->prof.begin(lu->stream, PK_TRSV, 2.0 * n * n * nr, sizeof(T) * (double)n * n);
        if (nr == 1) {
            hipLaunchKernelGGL((k_trsv<T, true, 1>), dim3(nblk), dim3(256), 0, lu->stream, F, ld, n, lu->perm, DL,
                               dB + c0 * ldb, ldb, lu->wy, n, nr, lu->tmsg, s1, timeout, lu->dinfo + 1);
            hipLaunchKernelGGL((k_trsv<T, false, 1>), dim3(nblk), dim3(256), 0, lu->stream, F, ld, n, lu->perm, DU,
                               (const double*)lu->wy, n, dX + c0 * ldx, ldx, nr, lu->tmsg, s2, timeout,
                               lu->dinfo + 1);
        } else {
            hipLaunchKernelGGL((k_trsv<T, true, TRHS>), dim3(nblk), dim3(256), 0, lu->stream, F, ld, n, lu->perm,
                               DL, dB + c0 * ldb, ldb, lu->wy, n, nr, lu->tmsg, s1, timeout, lu->dinfo + 1);
            hipLaunchKernelGGL((k_trsv<T, false, TRHS>), dim3(nblk), dim3(256), 0, lu->stream, F, ld, n, lu->perm,
                               DU, (const double*)lu->wy, n, dX + c0 * ldx, ldx, nr, lu->tmsg, s2, timeout,
                               lu->dinfo + 1);
        }
        lu->prof.end(lu->stream, pi);
    }
    LSX_HIP(ctx, hipGetLastError());
    return LSX_OK;
}

// X = A^{-T} B with the resident factors (one right-hand side per launch pair; FP64 values).
// The first solve writes lu->wy; X may alias B.
template <typename T>
int getrs_t_blocked(lsx_lu* lu, const T* F, int64_t ld, const T* DL, const T* DU, int64_t nrhs, const double* dB,
                    int64_t ldb, double* dX, int64_t ldx) {
    lsx_ctx* ctx = lu->ctx;
    const int64_t n = lu->n;
    const int nblk = (int)cdiv(n, DB);
    if (lu->tseq > 0xF0000000u) {
        LSX_HIP(ctx, hipMemsetAsync(lu->tmsg, 0, sizeof(uint64_t) * nblk * TMW, lu->stream));
        lu->tseq = 1;
    }
    const int64_t timeout = wait_ticks(ctx, 5.0);
    for (int64_t s = 0; s < nrhs; ++s) {
        const uint32_t s1 = lu->tseq++, s2 = lu->tseq++;
        const int pi = lu->prof.begin(lu->stream, PK_TRSV, 2.0 * n * n, sizeof(T) * (double)n * n);
        hipLaunchKernelGGL((k_trsv_t<T, true>), dim3(nblk), dim3(256), 0, lu->stream, F, ld, n, lu->perm, DU,
                           dB + s * ldb, lu->wy, lu->tmsg, s1, timeout, lu->dinfo + 1);
        hipLaunchKernelGGL((k_trsv_t<T, false>), dim3(nblk), dim3(256), 0, lu->stream, F, ld, n, lu->perm, DL,
                           (const double*)lu->wy, dX + s * ldx, lu->tmsg, s2, timeout, lu->dinfo + 1);
        lu->prof.end(lu->stream, pi);
    }
    LSX_HIP(ctx, hipGetLastError());
    return LSX_OK;
}

}  // namespace lsx
