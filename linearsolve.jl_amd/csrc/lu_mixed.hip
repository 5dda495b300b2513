// lu_mixed.hip — mixed-precision LU: factor with FP32 arithmetic (trailing products of three BF16
// pieces per operand, FP32 accumulation: FP32 accuracy at BF16 MFMA rates) or with BF16 trailing
// products (two pieces per operand, relative error ~2^-17), with tournament pivoting, then recover
// FP64 accuracy by iterative refinement (IR) or GMRES-IR with FP64 residuals.
//
// The API shape follows OpenBLAS32MixedLUFactorization (src/openblas.jl:461-543: cast A to
// FP32, sgetrf, sgetrs, cast back), which stops at FP32 accuracy; here the low-precision factors
// precondition an FP64 refinement until the normwise backward error
//   max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)
// reaches the requested tolerance.  FP64 is used only for the residual r = b - A x, the update
// x += d and (GMRES-IR) the Krylov orthogonalization, as in Carson & Higham (SISC 2018).
#include <algorithm>
#include <cmath>
#include <vector>

#include "gemm.hpp"
#include "gemm_split.hpp"
#include "calu.hpp"
#include "lu_driver.hpp"
#include "residual.hpp"

namespace lsx {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

struct MixedState {
    float* DL = nullptr;     // FP32 inverses of diagonal blocks
    float* DU = nullptr;
    double* anorm = nullptr; // device: ||A||_inf
    double* r = nullptr;     // n x cap residuals / corrections
    double* xw = nullptr;    // n x cap iterate
    double* bw = nullptr;    // n x cap right-hand sides
    double* part = nullptr;  // reduction partials
    double* hpart = nullptr; // pinned host copy of partials
    int64_t cap = 0;
    // GMRES-IR workspace
    double* V = nullptr;     // n x (m + 1)
    double* w = nullptr;
    double* hvec = nullptr;  // device m + 1
    int gm = 0;
    double anorm_h = 0.0;
};

static MixedState* ms(lsx_lu* lu) { return reinterpret_cast<MixedState*>(lu->aux); }

// ---------------------------------------------------------------------------------------- kernels
__global__ void k_cast_f32(const double* __restrict src, int64_t lds, float* __restrict dst, int64_t ldd, int64_t n) {
    const int64_t col = blockIdx.y;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[col * ldd + i] = (float)src[col * lds + i];
}

// Y += alpha X, column-major n x nrhs.
__global__ void k_axpy_cols(double* __restrict Y, int64_t ldy, const double* __restrict X, int64_t ldx, int64_t n,
                            int nrhs, double alpha) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int s = 0; s < nrhs; ++s) Y[s * ldy + i] += alpha * X[s * ldx + i];
}

// Partial dot products of columns 0..k-1 of V (n x k, ld) with w, per 256-row block:
// part[blk * k + j] (k_part_sum adds them up in a fixed order).
__global__ __launch_bounds__(256) void k_vtw(const double* __restrict V, int64_t ld, int k, const double* __restrict w,
                                             int64_t n, double* __restrict part) {
    __shared__ double red[4][64];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const double wi = i < n ? w[i] : 0.0;
    for (int j = 0; j < k; ++j) {
        double v = i < n ? V[j * ld + i] * wi : 0.0;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][j & 63] = v;
        __syncthreads();
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * k + j] = red[0][j & 63] + red[1][j & 63] + red[2][j & 63] + red[3][j & 63];
        __syncthreads();
    }
}

// w -= V h (n x k)
__global__ void k_wmvh(double* __restrict w, const double* __restrict V, int64_t ld, int k, const double* __restrict h,
                       int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = w[i];
    for (int j = 0; j < k; ++j) s -= V[j * ld + i] * h[j];
    w[i] = s;
}

__global__ void k_scale_copy(double* __restrict dst, const double* __restrict src, double a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = a * src[i];
}

// ---------------------------------------------------------------------------------------- host
int mixed_alloc(lsx_lu* lu) {
    lsx_ctx* ctx = lu->ctx;
    const int64_t n = lu->n;
    MixedState* m = new (std::nothrow) MixedState();
    if (!m) return LSX_ENOMEM;
    lu->aux = m;
    lu->ld32 = pad_ld(n);
    const int64_t nblk = cdiv(n, DB);
    if (hipMalloc(&lu->F32, sizeof(float) * lu->ld32 * n) != hipSuccess)
        return fail(ctx, LSX_ENOMEM, "lsx_lu_create: cannot allocate the FP32 factors");
    LSX_HIP(ctx, hipMalloc(&m->DL, sizeof(float) * DB * DB * nblk));
    LSX_HIP(ctx, hipMalloc(&m->DU, sizeof(float) * DB * DB * nblk));
    LSX_HIP(ctx, hipMalloc(&m->anorm, sizeof(double)));
    const int64_t pw = std::max<int64_t>(resid_part_words(n), cdiv(n, 256) * 64);
    LSX_HIP(ctx, hipMalloc(&m->part, sizeof(double) * pw));
    LSX_HIP(ctx, hipHostMalloc(&m->hpart, sizeof(double) * pw, hipHostMallocDefault));
    // tournament pivoting: candidate records, broadcast, arrival counter
    LSX_HIP(ctx, hipMalloc(&lu->calu_cand, sizeof(uint64_t) * CCAND));
    LSX_HIP(ctx, hipMalloc(&lu->calu_bc, sizeof(uint64_t) * CBC));
    LSX_HIP(ctx, hipMalloc(&lu->calu_ctr, 256));
    LSX_HIP(ctx, hipMemsetAsync(lu->calu_cand, 0, sizeof(uint64_t) * CCAND, lu->stream));
    LSX_HIP(ctx, hipMemsetAsync(lu->calu_bc, 0, sizeof(uint64_t) * CBC, lu->stream));
    LSX_HIP(ctx, hipMemsetAsync(lu->calu_ctr, 0, 256, lu->stream));
    LSX_HIP(ctx, hipStreamSynchronize(lu->stream));
    lu->calu_cbase = 0;
    lu->calu_seq = 1;
    return LSX_OK;
}

void mixed_free(lsx_lu* lu) {
    MixedState* m = ms(lu);
    if (lu->stream) (void)hipStreamSynchronize(lu->stream);
    (void)hipFree(lu->F32);
    lu->F32 = nullptr;
    if (!m) return;
    (void)hipFree(m->DL);
    (void)hipFree(m->DU);
    (void)hipFree(m->anorm);
    (void)hipFree(m->r);
    (void)hipFree(m->xw);
    (void)hipFree(m->bw);
    (void)hipFree(m->part);
    if (m->hpart) (void)hipHostFree(m->hpart);
    (void)hipFree(m->V);
    (void)hipFree(m->w);
    (void)hipFree(m->hvec);
    (void)hipFree(lu->calu_cand);
    (void)hipFree(lu->calu_bc);
    (void)hipFree(lu->calu_ctr);
    lu->calu_cand = lu->calu_bc = nullptr;
    lu->calu_ctr = nullptr;
    delete m;
    lu->aux = nullptr;
}

static void gemm_f32(lsx_lu* lu, const float* A, int64_t lda, const float* B, int64_t ldb, float* C, int64_t ldc,
                     int64_t M, int64_t N, int64_t K, int mode) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    const bool big = mode == 0 && cdiv(M, 128) * cdiv(N, 128) >= 2 * (int64_t)lu->ctx->num_cus;
    const int pi = lu->prof.begin(lu->cur, big ? PK_GEMM : PK_GEMM_SMALL, 2.0 * M * N * K, sizeof(float) * (2.0 * M * N));
    gemm_dispatch<float>(lu->cur, lu->ctx->num_cus, A, lda, B, ldb, C, ldc, M, N, K, mode);
    lu->prof.end(lu->cur, pi);
}

// Trailing update of the mixed factorization on BF16 MFMA: FP32 factors with the products of
// three BF16 pieces per operand (bf16x6, FP32 accuracy), BF16 factors with two pieces per
// operand (bf16x3, relative error about 3 2^-18 per product).
static void gemm_lowp(lsx_lu* lu, const float* A, int64_t lda, const float* B, int64_t ldb, float* C, int64_t ldc,
                      int64_t M, int64_t N, int64_t K) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    const bool big = cdiv(M, 128) * cdiv(N, 128) >= 2 * (int64_t)lu->ctx->num_cus;
    const int pi = lu->prof.begin(lu->cur, big ? PK_GEMM : PK_GEMM_SMALL, 2.0 * M * N * K, sizeof(float) * (2.0 * M * N));
    if (lu->dtype == LSX_BF16) launch_gemm_split<2>(lu->cur, A, lda, B, ldb, C, ldc, M, N, K);
    else launch_gemm_split<3>(lu->cur, A, lda, B, ldb, C, ldc, M, N, K);
    lu->prof.end(lu->cur, pi);
}

// Block column [jo, jo + jbo) by tournament pivoting (k_block_calu) on lu->cur; false when the
// block has more rows than one launch holds (the exact panels take it then).
static bool calu_block(lsx_lu* lu, int64_t jo, int jbo) {
    const int64_t m = lu->n - jo;
    const int P = (int)cdiv(m, CBS);
    if (P > CALU_PMAX) return false;
    const int np = (int)cdiv(jbo, CW);
    CaluArgs a;
    a.A = lu->F32;
    a.ld = lu->ld32;
    a.n = lu->n;
    a.jo = jo;
    a.JB = jbo;
    a.P = P;
    a.ipiv = lu->ipiv;
    a.dinfo = lu->dinfo;
    a.cand = lu->calu_cand;
    a.bc = lu->calu_bc;
    a.ctr = lu->calu_ctr;
    a.cbase = lu->calu_cbase;
    a.seq = lu->calu_seq;
    a.timeout = wait_ticks(lu->ctx, 5.0);
    lu->calu_cbase += (uint32_t)(np * P);
    lu->calu_seq += (uint32_t)np;
    const int pi = lu->prof.begin(lu->cur, PK_PANEL, (double)m * jbo * jbo, 2.0 * sizeof(float) * (double)m * jbo);
    hipLaunchKernelGGL(k_block_calu, dim3(P), dim3(CBS), 0, lu->cur, a);
    lu->prof.end(lu->cur, pi);
    return true;
}

int mixed_getrf(lsx_lu* lu) {
    lsx_ctx* ctx = lu->ctx;
    MixedState* m = ms(lu);
    const int64_t n = lu->n;
    // ||A||_inf in FP64 (for the backward error), then the working copy in FP32
    const int64_t nb = resid_blocks(n);
    resid_launch(lu->stream, lu->F, lu->ld, n, nullptr, 0, 1, nullptr, 0, nullptr, 0, m->part, RES_ROWSUM);
    LSX_HIP(ctx, hipMemcpyAsync(m->hpart, m->part, sizeof(double) * nb, hipMemcpyDeviceToHost, lu->stream));
    const int pc = lu->prof.begin(lu->stream, PK_CONVERT, 0.0, 12.0 * (double)n * n);
    hipLaunchKernelGGL(k_cast_f32, dim3((unsigned)std::min<int64_t>(cdiv(n, 256), 64), (unsigned)n), dim3(256), 0,
                       lu->stream, lu->F, lu->ld, lu->F32, lu->ld32, n);
    lu->prof.end(lu->stream, pc);
    auto gt = [lu](const float* A, int64_t lda, const float* B, int64_t ldb, float* C, int64_t ldc, int64_t M,
                   int64_t N, int64_t K, int mode) { gemm_f32(lu, A, lda, B, ldb, C, ldc, M, N, K, mode); };
    auto gu = [lu](const float* A, int64_t lda, const float* B, int64_t ldb, float* C, int64_t ldc, int64_t M,
                   int64_t N, int64_t K, int) { gemm_lowp(lu, A, lda, B, ldb, C, ldc, M, N, K); };
    if (lu->calu_seq > 0xF0000000u || lu->calu_cbase > 0xF0000000u) {
        LSX_HIP(ctx, hipMemsetAsync(lu->calu_ctr, 0, 256, lu->stream));
        LSX_HIP(ctx, hipMemsetAsync(lu->calu_cand, 0, sizeof(uint64_t) * CCAND, lu->stream));
        LSX_HIP(ctx, hipMemsetAsync(lu->calu_bc, 0, sizeof(uint64_t) * CBC, lu->stream));
        lu->calu_cbase = 0;
        lu->calu_seq = 1;
    }
    auto bf = [lu](int64_t jo, int jbo) { return calu_block(lu, jo, jbo); };
    const int rc = getrf_blocked<float>(lu, lu->F32, lu->ld32, m->DL, m->DU, gu, gt, bf);
    double an = 0.0;
    for (int64_t i = 0; i < nb; ++i) an = nanmax(an, m->hpart[i]);
    m->anorm_h = an;
    return rc;
}

// dst (n x ncols, ldd, FP64) = src (FP32, lds)
__global__ void k_widen(const float* __restrict src, int64_t lds, double* __restrict dst, int64_t ldd, int64_t n) {
    const int64_t col = blockIdx.y;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[col * ldd + i] = (double)src[col * lds + i];
}

int mixed_get_factors(lsx_lu* lu, double* F, int64_t ldf) {
    // widen the FP32 factors on the device: straight into F when it is memory of this GPU, else
    // through lu->wb in blocks of columns, one copy per block
    lsx_ctx* ctx = lu->ctx;
    const int64_t n = lu->n;
    const hipStream_t st = lu->stream;
    const unsigned gx = (unsigned)std::min<int64_t>(cdiv(n, 256), 64);
    if (is_device_ptr(ctx->device, F)) {
        hipLaunchKernelGGL(k_widen, dim3(gx, (unsigned)n), dim3(256), 0, st, (const float*)lu->F32, lu->ld32, F, ldf, n);
        LSX_HIP(ctx, hipGetLastError());
    } else {
        const int64_t cols = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)(64 << 20) / (8 * n)));
        int rc = lu_ensure_work(lu, cols);
        if (rc) return rc;
        for (int64_t c0 = 0; c0 < n; c0 += cols) {
            const int64_t nc = std::min<int64_t>(cols, n - c0);
            hipLaunchKernelGGL(k_widen, dim3(gx, (unsigned)nc), dim3(256), 0, st, (const float*)(lu->F32 + c0 * lu->ld32),
                               lu->ld32, lu->wb, n, n);
            LSX_HIP(ctx, hipGetLastError());
            LSX_HIP(ctx, hipMemcpy2DAsync(F + c0 * ldf, sizeof(double) * ldf, lu->wb, sizeof(double) * n,
                                          sizeof(double) * n, nc, hipMemcpyDeviceToHost, st));
        }
    }
    LSX_HIP(ctx, hipStreamSynchronize(st));
    return LSX_OK;
}

int mixed_getrs_dev(lsx_lu* lu, int64_t nrhs, const double* dB, int64_t ldb, double* dX, int64_t ldx) {
    MixedState* m = ms(lu);
    return getrs_blocked<float>(lu, lu->F32, lu->ld32, m->DL, m->DU, nrhs, dB, ldb, dX, ldx);
}

static int ensure_mixed_work(lsx_lu* lu, int64_t nrhs) {
    MixedState* m = ms(lu);
    lsx_ctx* ctx = lu->ctx;
    if (nrhs <= m->cap) return LSX_OK;
    (void)hipFree(m->r);
    (void)hipFree(m->xw);
    (void)hipFree(m->bw);
    const int64_t cap = std::max<int64_t>(nrhs, TRHS);
    LSX_HIP(ctx, hipMalloc(&m->r, sizeof(double) * lu->n * cap));
    LSX_HIP(ctx, hipMalloc(&m->xw, sizeof(double) * lu->n * cap));
    LSX_HIP(ctx, hipMalloc(&m->bw, sizeof(double) * lu->n * cap));
    m->cap = cap;
    return lu_ensure_work(lu, cap);
}

// r = b - A x for columns [c0, c0 + nr) and the per-column backward errors (NaN stays NaN, so
// it never passes a tolerance test); with bnorm null, bmax receives max |b| per column.
static int residual(lsx_lu* lu, int64_t c0, int nr, const double* bnorm, double* be, double* bmax = nullptr) {
    lsx_ctx* ctx = lu->ctx;
    MixedState* m = ms(lu);
    const int64_t n = lu->n;
    const int pi = lu->prof.begin(lu->stream, PK_REFINE, 2.0 * n * n * nr, 8.0 * (double)n * n);
    resid_launch(lu->stream, lu->F, lu->ld, n, m->xw + c0 * n, n, nr, m->bw + c0 * n, n, m->r + c0 * n, n, m->part,
                 RES_RESIDUAL);
    lu->prof.end(lu->stream, pi);
    LSX_HIP(ctx, hipGetLastError());
    LSX_HIP(ctx, hipMemcpyAsync(m->hpart, m->part, sizeof(double) * resid_part_words(n), hipMemcpyDeviceToHost,
                                lu->stream));
    LSX_HIP(ctx, hipStreamSynchronize(lu->stream));
    double stv[RNR][RSTAT];
    resid_combine(m->hpart, n, nr, stv);
    for (int s = 0; s < nr; ++s) {
        if (bmax) bmax[s] = stv[s][3];
        if (!bnorm) continue;
        const double rn = stv[s][0], xn = stv[s][2];
        const double den = m->anorm_h * xn + bnorm[c0 + s];
        be[s] = den > 0.0 ? rn / den : (rn > 0.0 || rn != rn ? rn / den : 0.0);
    }
    return LSX_OK;
}

// out[j] = sum of the block partials part[b k + j] over blocks b (k workgroups of one wave: each
// lane sums its blocks in order, then a fixed shuffle tree; the same order on every call).
__global__ __launch_bounds__(64) void k_part_sum(const double* __restrict part, int64_t nb, int k,
                                                 double* __restrict out) {
    const int j = blockIdx.x, l = threadIdx.x;
    double s = 0.0;
    for (int64_t b = l; b < nb; b += 64) s += part[b * k + j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (l == 0) out[j] = s;
}

// Deterministic dot products of V's first k columns with w into the device array h (no host
// round trip).
static int vtw_dev(lsx_lu* lu, const double* V, int k, const double* w, double* h) {
    MixedState* m = ms(lu);
    const int64_t n = lu->n, nb = cdiv(n, 256);
    hipLaunchKernelGGL(k_vtw, dim3((unsigned)nb), dim3(256), 0, lu->stream, V, n, k, w, n, m->part);
    hipLaunchKernelGGL(k_part_sum, dim3((unsigned)k), dim3(64), 0, lu->stream, (const double*)m->part, nb, k, h);
    return hipGetLastError() == hipSuccess ? LSX_OK : fail(lu->ctx, LSX_EHIP, "GMRES-IR: dot-product launch failed");
}

// Solves A d = r (column s of m->r, result into m->r column s) by GMRES in FP64 with the
// low-precision LU as left preconditioner and classical Gram-Schmidt applied twice.  The dot
// products and the updates stay on the device; each Arnoldi step copies its 2k + 3 reduced
// values back once, for the Givens rotations and the stopping test on the host.
// Returns the number of iterations or < 0.
static int gmres_ir_correction(lsx_lu* lu, int64_t s, int maxit, double rtol) {
    lsx_ctx* ctx = lu->ctx;
    MixedState* m = ms(lu);
    const int64_t n = lu->n;
    const int M = std::max(2, std::min(maxit, 50));
    if (m->gm < M) {
        (void)hipFree(m->V);
        (void)hipFree(m->w);
        (void)hipFree(m->hvec);
        LSX_HIP(ctx, hipMalloc(&m->V, sizeof(double) * n * (M + 1)));
        LSX_HIP(ctx, hipMalloc(&m->w, sizeof(double) * n));
        LSX_HIP(ctx, hipMalloc(&m->hvec, sizeof(double) * 3 * (M + 2)));
        m->gm = M;
    }
    // device scalars: hA (first Gram-Schmidt pass), hB (second), hN (||w||^2), one copy per step
    double* hA = m->hvec;
    double* hB = hA + (M + 2);
    double* hN = hB + (M + 2);
    double* r = m->r + s * n;
    const unsigned g = (unsigned)cdiv(n, 256);
    // z0 = M^-1 r
    int rc = mixed_getrs_dev(lu, 1, r, n, m->V, n);
    if (rc) return rc;
    double beta;
    rc = vtw_dev(lu, m->V, 1, m->V, hN);
    if (rc) return rc;
    LSX_HIP(ctx, hipMemcpyAsync(&beta, hN, sizeof(double), hipMemcpyDeviceToHost, lu->stream));
    LSX_HIP(ctx, hipStreamSynchronize(lu->stream));
    beta = std::sqrt(beta);
    if (beta == 0.0) {
        LSX_HIP(ctx, hipMemsetAsync(r, 0, sizeof(double) * n, lu->stream));
        return 0;
    }
    hipLaunchKernelGGL(k_scale_copy, dim3(g), dim3(256), 0, lu->stream, m->V, (const double*)m->V, 1.0 / beta, n);
    std::vector<double> H((size_t)(M + 1) * M, 0.0), cs(M), sn(M), z(M + 1, 0.0), h(M + 1), hh(3 * (M + 2));
    z[0] = beta;
    int k = 0;
    for (; k < M; ++k) {
        // w = M^-1 A v_k
        resid_launch(lu->stream, lu->F, lu->ld, n, m->V + k * n, n, 1, nullptr, 0, m->w, n, m->part, RES_PRODUCT);
        rc = mixed_getrs_dev(lu, 1, m->w, n, m->w, n);
        if (rc) return rc;
        // CGS2 on the device
        rc = vtw_dev(lu, m->V, k + 1, m->w, hA);
        if (rc) return rc;
        hipLaunchKernelGGL(k_wmvh, dim3(g), dim3(256), 0, lu->stream, m->w, (const double*)m->V, n, k + 1,
                           (const double*)hA, n);
        rc = vtw_dev(lu, m->V, k + 1, m->w, hB);
        if (rc) return rc;
        hipLaunchKernelGGL(k_wmvh, dim3(g), dim3(256), 0, lu->stream, m->w, (const double*)m->V, n, k + 1,
                           (const double*)hB, n);
        rc = vtw_dev(lu, m->w, 1, m->w, hN);
        if (rc) return rc;
        LSX_HIP(ctx, hipMemcpyAsync(hh.data(), hA, sizeof(double) * 3 * (M + 2), hipMemcpyDeviceToHost, lu->stream));
        LSX_HIP(ctx, hipStreamSynchronize(lu->stream));
        for (int j = 0; j <= k; ++j) h[j] = hh[j] + hh[(M + 2) + j];
        const double hn = std::sqrt(hh[2 * (M + 2)]);
        // Givens on column k of H
        for (int j = 0; j < k; ++j) {
            const double t = cs[j] * h[j] + sn[j] * h[j + 1];
            h[j + 1] = -sn[j] * h[j] + cs[j] * h[j + 1];
            h[j] = t;
        }
        const double rho = std::hypot(h[k], hn);
        cs[k] = rho > 0 ? h[k] / rho : 1.0;
        sn[k] = rho > 0 ? hn / rho : 0.0;
        h[k] = rho;
        for (int j = 0; j <= k; ++j) H[(size_t)j * M + k] = h[j];
        z[k + 1] = -sn[k] * z[k];
        z[k] = cs[k] * z[k];
        if (hn > 0.0 && k + 1 < M)
            hipLaunchKernelGGL(k_scale_copy, dim3(g), dim3(256), 0, lu->stream, m->V + (k + 1) * n, (const double*)m->w,
                               1.0 / hn, n);
        if (std::fabs(z[k + 1]) <= rtol * beta || hn == 0.0) {
            ++k;
            break;
        }
    }
    // y = H^-1 z, d = V y
    std::vector<double> y(k);
    for (int i = k - 1; i >= 0; --i) {
        double t = z[i];
        for (int j = i + 1; j < k; ++j) t -= H[(size_t)i * M + j] * y[j];
        y[i] = H[(size_t)i * M + i] != 0.0 ? t / H[(size_t)i * M + i] : 0.0;
    }
    for (int i = 0; i < k; ++i) y[i] = -y[i];
    LSX_HIP(ctx, hipMemcpyAsync(m->hvec, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, lu->stream));
    LSX_HIP(ctx, hipMemsetAsync(r, 0, sizeof(double) * n, lu->stream));
    hipLaunchKernelGGL(k_wmvh, dim3(g), dim3(256), 0, lu->stream, r, (const double*)m->V, n, k, (const double*)m->hvec, n);
    LSX_HIP(ctx, hipGetLastError());
    return k;
}

}  // namespace lsx

using namespace lsx;

extern "C" int lsx_lu_refine(lsx_lu* lu, int64_t nrhs, const double* B, int64_t ldb, double* X, int64_t ldx,
                             double tol, int32_t max_iters, int32_t mode, lsx_refine_stats* stats) {
    if (!lu || !B || !X || nrhs < 1) return LSX_EINVAL;
    lsx_ctx* ctx = lu->ctx;
    if (!lu->factored) return fail(ctx, LSX_ESTATE, "lsx_lu_refine: no factorization");
    if (lu->dtype == LSX_F64) return fail(ctx, LSX_EINVAL, "lsx_lu_refine: the handle holds FP64 factors; use lsx_lu_solve");
    if (mode < 0 || mode > 2) return fail(ctx, LSX_EINVAL, "lsx_lu_refine: mode must be 0, 1 or 2");
    DeviceGuard g(ctx);
    MixedState* m = ms(lu);
    const int64_t n = lu->n;
    int rc = ensure_mixed_work(lu, nrhs);
    if (rc) return rc;
    rc = lu_clear_err(lu);
    if (rc) return fail(ctx, rc, "lsx_lu_refine: cannot clear the wait error word");
    LSX_HIP(ctx, hipMemcpy2DAsync(m->bw, sizeof(double) * n, B, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                                  hipMemcpyDefault, lu->stream));
    // ||b_j||_inf (host), from the residual kernel with x = 0
    std::vector<double> bnorm(nrhs, 0.0), be(TRHS, 0.0);
    LSX_HIP(ctx, hipMemsetAsync(m->xw, 0, sizeof(double) * n * nrhs, lu->stream));
    int total_it = 0, outer = 0, mode_used = (mode == 1) ? 1 : 0, escalated = 0;
    double worst = 0.0;
    bool all_conv = true;
    for (int64_t c0 = 0; c0 < nrhs; c0 += TRHS) {
        const int nr = (int)std::min<int64_t>(TRHS, nrhs - c0);
        rc = residual(lu, c0, nr, nullptr, nullptr, bnorm.data() + c0);  // x = 0: max |b| per column
        if (rc) return rc;
        // x0 = LU \ b
        rc = mixed_getrs_dev(lu, nr, m->bw + c0 * n, n, m->xw + c0 * n, n);
        if (rc) return rc;
        int use_gmres = (mode == 1);
        double prev = INFINITY;
        int stall = 0;
        bool conv = false;
        for (int it = 0; it <= max_iters; ++it) {
            rc = residual(lu, c0, nr, bnorm.data(), be.data());
            if (rc) return rc;
            double worst_c = 0.0;
            for (int s = 0; s < nr; ++s) worst_c = nanmax(worst_c, be[s]);
            ++outer;
            if (worst_c <= tol) {
                conv = true;
                worst = nanmax(worst, worst_c);
                break;
            }
            if (it == max_iters) {
                worst = nanmax(worst, worst_c);
                break;
            }
            if (mode == 2) {
                stall = (worst_c > 0.5 * prev) ? stall + 1 : 0;
                if (stall >= 2 && !use_gmres) {
                    // IR stalls: the factors are too coarse for plain refinement
                    use_gmres = 1;
                    mode_used = 1;
                    stall = 0;
                } else if (stall >= 1 && use_gmres && lu->dtype == LSX_BF16 && !escalated) {
                    // GMRES-IR stalls too (u_bf16 * cond(A) >= 1): refactor in FP32, keep x.  One
                    // GMRES-IR step that does not halve the backward error decides it: an FP32
                    // factorization costs about as much as 60 inner GMRES iterations.
                    // The refactorization clears the wait error word, so a wait that gave up in
                    // the solves so far is reported first.
                    LSX_HIP(ctx, hipMemcpyAsync(lu->hinfo + 1, lu->dinfo + 1, sizeof(int64_t),
                                                hipMemcpyDeviceToHost, lu->stream));
                    LSX_HIP(ctx, hipStreamSynchronize(lu->stream));
                    if (lu->hinfo[1]) return fail(ctx, LSX_ETIMEOUT, "lsx_lu_refine: a block wait exceeded its bound");
                    const int saved = lu->dtype;
                    lu->dtype = LSX_F32;
                    rc = mixed_getrf(lu);
                    lu->dtype = saved;
                    if (rc < 0) return rc;
                    if (rc > 0) {
                        // the FP32 factors are singular: no refinement is possible (LAPACK info)
                        lu->info = rc;
                        return rc;
                    }
                    escalated = 1;
                    use_gmres = 0;
                    stall = 0;
                }
            }
            prev = worst_c;
            if (use_gmres) {
                for (int s = 0; s < nr; ++s) {
                    const int k = gmres_ir_correction(lu, c0 + s, 30, 1e-6);
                    if (k < 0) return k;
                    total_it += k;
                }
            } else {
                rc = mixed_getrs_dev(lu, nr, m->r + c0 * n, n, m->r + c0 * n, n);
                if (rc) return rc;
                ++total_it;
            }
            const int pi = lu->prof.begin(lu->stream, PK_REFINE, 2.0 * n * nr, 24.0 * n * nr);
            hipLaunchKernelGGL(k_axpy_cols, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, lu->stream, m->xw + c0 * n, n,
                               (const double*)(m->r + c0 * n), n, n, nr, 1.0);
            lu->prof.end(lu->stream, pi);
        }
        all_conv = all_conv && conv;
    }
    LSX_HIP(ctx, hipMemcpy2DAsync(X, sizeof(double) * ldx, m->xw, sizeof(double) * n, sizeof(double) * n, nrhs,
                                  hipMemcpyDefault, lu->stream));
    LSX_HIP(ctx, hipMemcpyAsync(lu->hinfo + 1, lu->dinfo + 1, sizeof(int64_t), hipMemcpyDeviceToHost, lu->stream));
    LSX_HIP(ctx, hipStreamSynchronize(lu->stream));
    lu->prof.collect();
    if (lu->hinfo[1]) return fail(ctx, LSX_ETIMEOUT, "lsx_lu_refine: a block wait exceeded its bound");
    if (stats) {
        stats->mode_used = mode_used;
        stats->iterations = total_it;
        stats->outer = outer;
        stats->converged = all_conv ? 1 : 0;
        stats->backward_error = worst;
        stats->escalated = escalated;
        stats->pad = 0;
    }
    return LSX_OK;
}
