// api.hip — the C-ABI entry points of liblsx.so (declared in include/lsx.h).
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "lu.hpp"
#include "residual.hpp"

namespace lsx {
// Scratch memory of lsx_residual_norm_noalloc: the per-block statistics of one residual launch
// for n <= LSX_NOALLOC_MAX_N, then the two running sums of squares.  A file-scope device array,
// so that entry point never allocates; every GPU the library runs on holds its own copy
// (656 KB) until the process ends.  Calls on one GPU share it: overlapping calls race.
constexpr int64_t NOALLOC_PART_WORDS = resid_part_words(LSX_NOALLOC_MAX_N);
constexpr int64_t NOALLOC_WORDS = NOALLOC_PART_WORDS + 32;  // + the sums (own 256-byte line)
static __device__ __align__(256) double scratch[NOALLOC_WORDS];

int fail(lsx_ctx* ctx, int code, const char* fmt, ...) {
    if (!ctx) return code;
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->err = buf;
    return code;
}

int lu64_alloc(lsx_lu* lu) {
    lsx_ctx* ctx = lu->ctx;
    if (lu->cap < lu->n) lu->cap = lu->n;
    const int64_t n = lu->cap;  // every buffer holds the largest order
    lu->ld = pad_ld(lu->n);
    const int nblk = (int)cdiv(n, DB);
    LSX_HIP(ctx, hipStreamCreateWithFlags(&lu->stream, hipStreamNonBlocking));
    // getrf lookahead streams.  The panel workgroups wait on each other, so they get CUs of their
    // own: 4 per XCD (32 in all, enough for the 512-row workgroups of n <= 16384) and the
    // trailing updates the rest.  The CU set {per*x + (x + 8j) mod per} puts 4 CUs on every XCD
    // whether CU mask bits are numbered XCD by XCD or round-robin over the XCDs.
    // (only for n >= 2048: below that the lookahead gains nothing, and a CU-masked stream takes a
    // hardware queue of its own, of which a process has few)
    const int per = ctx->num_cus / 8;
    bool masked = false;
    if (lu->lookahead && n >= 2048 && ctx->num_cus % 8 == 0 && per % 8 == 0 && per >= 32 && cdiv(n, 512) <= 32) {
        const int words = (ctx->num_cus + 31) / 32;
        std::vector<uint32_t> ms(words, 0u), mc(words, 0u);
        for (int x = 0; x < 8; ++x)
            for (int j = 0; j < 4; ++j) {
                const int i = per * x + (x + 8 * j) % per;
                ms[i / 32] |= 1u << (i % 32);
            }
        for (int i = 0; i < ctx->num_cus; ++i)
            if (!((ms[i / 32] >> (i % 32)) & 1u)) mc[i / 32] |= 1u << (i % 32);
        masked = hipExtStreamCreateWithCUMask(&lu->stream2, (uint32_t)words, ms.data()) == hipSuccess &&
                 hipExtStreamCreateWithCUMask(&lu->stream_upd, (uint32_t)words, mc.data()) == hipSuccess;
        if (!masked) {  // no CU partition: plain streams
            (void)hipGetLastError();
            if (lu->stream2 && lu->stream2 != lu->stream) (void)hipStreamDestroy(lu->stream2);
            if (lu->stream_upd) (void)hipStreamDestroy(lu->stream_upd);
            lu->stream2 = lu->stream_upd = nullptr;
        }
    }
    if (!lu->lookahead) lu->stream2 = lu->stream;
    else if (!masked) LSX_HIP(ctx, hipStreamCreateWithFlags(&lu->stream2, hipStreamNonBlocking));
    LSX_HIP(ctx, hipEventCreateWithFlags(&lu->ev_start, hipEventDisableTiming));
    LSX_HIP(ctx, hipEventCreateWithFlags(&lu->ev_next, hipEventDisableTiming));
    LSX_HIP(ctx, hipEventCreateWithFlags(&lu->ev_panel, hipEventDisableTiming));
    LSX_HIP(ctx, hipEventCreateWithFlags(&lu->ev_perm, hipEventDisableTiming));
    lu->cur = lu->stream;
    if (lu->own_F && hipMalloc(&lu->F, sizeof(double) * pad_ld(n) * n) != hipSuccess)
        return fail(ctx, LSX_ENOMEM, "lsx_lu_create: cannot allocate %lld x %lld FP64 factors", (long long)n,
                    (long long)n);
    LSX_HIP(ctx, hipMalloc(&lu->ipiv, sizeof(int64_t) * n));
    LSX_HIP(ctx, hipMalloc(&lu->perm, sizeof(int) * n));
    LSX_HIP(ctx, hipMalloc(&lu->dinfo, sizeof(int64_t) * 2));
    LSX_HIP(ctx, hipHostMalloc(&lu->hinfo, sizeof(int64_t) * 2, hipHostMallocDefault));
    LSX_HIP(ctx, hipMalloc(&lu->slots, sizeof(uint64_t) * 2 * (PMAX + 1) * RECS));
    LSX_HIP(ctx, hipMalloc(&lu->plist, sizeof(int64_t) * 12 * NBO));
    LSX_HIP(ctx, hipMalloc(&lu->pcount, sizeof(int) * 3));
    // inverse row maps of the deferred L-column interchanges (used when a column segment of the
    // working type fits in LDS; FP32 is the smaller type)
    if ((n - NBO) * (int64_t)sizeof(float) <= LGMAX_BYTES && n > NBO)
        LSX_HIP(ctx, hipMalloc(&lu->pinv, sizeof(int) * cdiv(n, NBO) * n));
    LSX_HIP(ctx, hipMalloc(&lu->tmsg, sizeof(uint64_t) * nblk * 2 * DB * TRHS));
    LSX_HIP(ctx, hipMalloc(&lu->DL, sizeof(double) * DB * DB * nblk));
    LSX_HIP(ctx, hipMalloc(&lu->DU, sizeof(double) * DB * DB * nblk));
    LSX_HIP(ctx, hipMemsetAsync(lu->slots, 0, sizeof(uint64_t) * 2 * (PMAX + 1) * RECS, lu->stream));
    LSX_HIP(ctx, hipMemsetAsync(lu->tmsg, 0, sizeof(uint64_t) * nblk * 2 * DB * TRHS, lu->stream));
    LSX_HIP(ctx, hipMemsetAsync(lu->dinfo, 0, sizeof(int64_t) * 2, lu->stream));
    LSX_HIP(ctx, hipStreamSynchronize(lu->stream));
    return LSX_OK;
}

int lu64_reshape(lsx_lu* lu, int64_t n) {
    if (lu->dtype != LSX_F64 || n < 1 || n > lu->cap)
        return fail(lu->ctx, LSX_EINVAL, "lu reshape: order %lld outside [1, %lld]", (long long)n, (long long)lu->cap);
    lu->n = n;
    lu->ld = pad_ld(n);
    lu->factored = false;
    return LSX_OK;
}

void lu64_free(lsx_lu* lu) {
    if (lu->stream) (void)hipStreamSynchronize(lu->stream);
    if (lu->own_F) (void)hipFree(lu->F);
    (void)hipFree(lu->ipiv);
    (void)hipFree(lu->perm);
    (void)hipFree(lu->dinfo);
    if (lu->hinfo) (void)hipHostFree(lu->hinfo);
    (void)hipFree(lu->slots);
    (void)hipFree(lu->plist);
    (void)hipFree(lu->pcount);
    (void)hipFree(lu->pinv);
    (void)hipFree(lu->tmsg);
    (void)hipFree(lu->DL);
    (void)hipFree(lu->DU);
    (void)hipFree(lu->wb);
    (void)hipFree(lu->wy);
    (void)hipFree(lu->astage);
    (void)hipFree(lu->rpart);
    if (lu->hrpart) (void)hipHostFree(lu->hrpart);
    if (lu->stream2) (void)hipStreamSynchronize(lu->stream2);
    if (lu->stream_upd) (void)hipStreamSynchronize(lu->stream_upd);
    if (lu->ev_start) (void)hipEventDestroy(lu->ev_start);
    if (lu->ev_next) (void)hipEventDestroy(lu->ev_next);
    if (lu->ev_panel) (void)hipEventDestroy(lu->ev_panel);
    if (lu->ev_perm) (void)hipEventDestroy(lu->ev_perm);
    if (lu->stream2 && lu->stream2 != lu->stream) (void)hipStreamDestroy(lu->stream2);
    if (lu->stream_upd) (void)hipStreamDestroy(lu->stream_upd);
    if (lu->stream) (void)hipStreamDestroy(lu->stream);
}
}  // namespace lsx

using namespace lsx;

extern "C" {

int lsx_version(void) { return 1; }

int lsx_ctx_create(int device, lsx_ctx** out) {
    if (!out) return LSX_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return LSX_EHIP;
    if (device < 0 || device >= count) return LSX_EINVAL;
    lsx_ctx* c = new (std::nothrow) lsx_ctx();
    if (!c) return LSX_ENOMEM;
    c->device = device;
    DeviceGuard g(c);
    (void)hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
        c->wall_khz = khz;
    // the device address of the no-allocation scratch array on this GPU (each GPU holds its own
    // copy); host code uses this address, never the array's name
    const hipError_t e = hipGetSymbolAddress((void**)&c->scratch, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !c->scratch) {
        (void)hipGetLastError();
        delete c;
        return LSX_EHIP;
    }
    *out = c;
    return LSX_OK;
}

int lsx_ctx_destroy(lsx_ctx* ctx) {
    if (!ctx) return LSX_EINVAL;
    delete ctx;
    return LSX_OK;
}

const char* lsx_last_error(const lsx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int lsx_ctx_set_wait_bound(lsx_ctx* ctx, double seconds) {
    if (!ctx || !(seconds >= 0.0) || seconds > 3600.0) return LSX_EINVAL;
    ctx->wait_bound_s = seconds;
    return LSX_OK;
}

int lsx_malloc(lsx_ctx* ctx, size_t bytes, void** dptr) {
    if (!ctx || !dptr) return LSX_EINVAL;
    DeviceGuard g(ctx);
    if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess)
        return fail(ctx, LSX_ENOMEM, "lsx_malloc: cannot allocate %zu bytes", bytes);
    return LSX_OK;
}

int lsx_free(lsx_ctx* ctx, void* dptr) {
    if (!ctx) return LSX_EINVAL;
    DeviceGuard g(ctx);
    LSX_HIP(ctx, hipFree(dptr));
    return LSX_OK;
}

int lsx_memcpy(lsx_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return LSX_EINVAL;
    DeviceGuard g(ctx);
    LSX_HIP(ctx, hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    return LSX_OK;
}

int lsx_synchronize(lsx_ctx* ctx) {
    if (!ctx) return LSX_EINVAL;
    DeviceGuard g(ctx);
    LSX_HIP(ctx, hipDeviceSynchronize());
    return LSX_OK;
}

int lsx_lu_create(lsx_ctx* ctx, int64_t n, int dtype, lsx_lu** out) {
    if (!ctx || !out) return LSX_EINVAL;
    *out = nullptr;
    if (n < 1 || n > (int64_t)PMAX * 512)
        return fail(ctx, LSX_EINVAL, "lsx_lu_create: n = %lld outside [1, %d]", (long long)n, PMAX * 512);
    if (dtype != LSX_F64 && dtype != LSX_F32 && dtype != LSX_BF16)
        return fail(ctx, LSX_EINVAL, "lsx_lu_create: unknown dtype %d", dtype);
    DeviceGuard g(ctx);
    lsx_lu* lu = new (std::nothrow) lsx_lu();
    if (!lu) return LSX_ENOMEM;
    lu->ctx = ctx;
    lu->n = n;
    lu->dtype = dtype;
    int rc = lu64_alloc(lu);
    if (rc == LSX_OK && dtype != LSX_F64) rc = mixed_alloc(lu);
    if (rc != LSX_OK) {
        if (dtype != LSX_F64) mixed_free(lu);
        lu64_free(lu);
        delete lu;
        return rc;
    }
    *out = lu;
    return LSX_OK;
}

int lsx_lu_destroy(lsx_lu* lu) {
    if (!lu) return LSX_EINVAL;
    DeviceGuard g(lu->ctx);
    if (lu->dtype != LSX_F64) mixed_free(lu);
    lu64_free(lu);
    delete lu;
    return LSX_OK;
}

int lsx_lu_factor(lsx_lu* lu, const double* A, int64_t lda) {
    if (!lu || !A) return LSX_EINVAL;
    lsx_ctx* ctx = lu->ctx;
    if (lda < lu->n) return fail(ctx, LSX_EINVAL, "lsx_lu_factor: lda %lld < n %lld", (long long)lda, (long long)lu->n);
    DeviceGuard g(ctx);
    lu->factored = false;
    LSX_HIP(ctx, hipMemcpy2DAsync(lu->F, sizeof(double) * lu->ld, A, sizeof(double) * lda, sizeof(double) * lu->n,
                                  lu->n, hipMemcpyDefault, lu->stream));
    int rc = (lu->dtype == LSX_F64) ? lu64_getrf(lu) : mixed_getrf(lu);
    if (rc < 0) return rc;
    lu->info = rc;
    lu->factored = true;
    return rc;
}

int lsx_lu_get_factors(lsx_lu* lu, double* F, int64_t ldf, int64_t* ipiv) {
    if (!lu) return LSX_EINVAL;
    lsx_ctx* ctx = lu->ctx;
    if (!lu->factored) return fail(ctx, LSX_ESTATE, "lsx_lu_get_factors: no factorization");
    DeviceGuard g(ctx);
    if (F) {
        if (ldf < lu->n) return fail(ctx, LSX_EINVAL, "lsx_lu_get_factors: ldf < n");
        if (lu->dtype == LSX_F64)
            LSX_HIP(ctx, hipMemcpy2DAsync(F, sizeof(double) * ldf, lu->F, sizeof(double) * lu->ld,
                                          sizeof(double) * lu->n, lu->n, hipMemcpyDefault, lu->stream));
        else {
            int rc = mixed_get_factors(lu, F, ldf);
            if (rc) return rc;
        }
    }
    if (ipiv) LSX_HIP(ctx, hipMemcpyAsync(ipiv, lu->ipiv, sizeof(int64_t) * lu->n, hipMemcpyDefault, lu->stream));
    LSX_HIP(ctx, hipStreamSynchronize(lu->stream));
    return LSX_OK;
}

int lsx_lu_solve(lsx_lu* lu, int64_t nrhs, const double* B, int64_t ldb, double* X, int64_t ldx) {
    if (!lu || !B || !X || nrhs < 0) return LSX_EINVAL;
    lsx_ctx* ctx = lu->ctx;
    if (!lu->factored) return fail(ctx, LSX_ESTATE, "lsx_lu_solve: no factorization");
    if (ldb < lu->n || ldx < lu->n) return fail(ctx, LSX_EINVAL, "lsx_lu_solve: ldb/ldx < n");
    if (nrhs == 0) return LSX_OK;
    DeviceGuard g(ctx);
    int rc = lu_ensure_work(lu, nrhs);
    if (rc) return rc;
    const int64_t n = lu->n;
    LSX_HIP(ctx, hipMemsetAsync(lu->dinfo + 1, 0, sizeof(int64_t), lu->stream));
    LSX_HIP(ctx, hipMemcpy2DAsync(lu->wb, sizeof(double) * n, B, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                                  hipMemcpyDefault, lu->stream));
    rc = (lu->dtype == LSX_F64) ? lu64_getrs_dev(lu, nrhs, lu->wb, n, lu->wb, n) : mixed_getrs_dev(lu, nrhs, lu->wb, n, lu->wb, n);
    if (rc) return rc;
    LSX_HIP(ctx, hipMemcpy2DAsync(X, sizeof(double) * ldx, lu->wb, sizeof(double) * n, sizeof(double) * n, nrhs,
                                  hipMemcpyDefault, lu->stream));
    LSX_HIP(ctx, hipMemcpyAsync(lu->hinfo + 1, lu->dinfo + 1, sizeof(int64_t), hipMemcpyDeviceToHost, lu->stream));
    LSX_HIP(ctx, hipStreamSynchronize(lu->stream));
    lu->prof.collect();
    if (lu->hinfo[1]) return fail(ctx, LSX_ETIMEOUT, "lsx_lu_solve: a block wait exceeded its bound");
    return LSX_OK;
}

int lsx_lu_solve_adjoint(lsx_lu* lu, int64_t nrhs, const double* B, int64_t ldb, double* X, int64_t ldx) {
    if (!lu || !B || !X || nrhs < 0) return LSX_EINVAL;
    lsx_ctx* ctx = lu->ctx;
    if (!lu->factored) return fail(ctx, LSX_ESTATE, "lsx_lu_solve_adjoint: no factorization");
    if (ldb < lu->n || ldx < lu->n) return fail(ctx, LSX_EINVAL, "lsx_lu_solve_adjoint: ldb/ldx < n");
    if (lu->dtype != LSX_F64)
        return fail(ctx, LSX_EINVAL, "lsx_lu_solve_adjoint: adjoint solves use FP64 handles");
    if (nrhs == 0) return LSX_OK;
    DeviceGuard g(ctx);
    int rc = lu_ensure_work(lu, nrhs);
    if (rc) return rc;
    const int64_t n = lu->n;
    LSX_HIP(ctx, hipMemsetAsync(lu->dinfo + 1, 0, sizeof(int64_t), lu->stream));
    LSX_HIP(ctx, hipMemcpy2DAsync(lu->wb, sizeof(double) * n, B, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                                  hipMemcpyDefault, lu->stream));
    rc = lu64_getrs_t_dev(lu, nrhs, lu->wb, n, lu->wb, n);
    if (rc) return rc;
    LSX_HIP(ctx, hipMemcpy2DAsync(X, sizeof(double) * ldx, lu->wb, sizeof(double) * n, sizeof(double) * n, nrhs,
                                  hipMemcpyDefault, lu->stream));
    LSX_HIP(ctx, hipMemcpyAsync(lu->hinfo + 1, lu->dinfo + 1, sizeof(int64_t), hipMemcpyDeviceToHost, lu->stream));
    LSX_HIP(ctx, hipStreamSynchronize(lu->stream));
    lu->prof.collect();
    if (lu->hinfo[1]) return fail(ctx, LSX_ETIMEOUT, "lsx_lu_solve_adjoint: a block wait exceeded its bound");
    return LSX_OK;
}

int lsx_lu_residual_norm(lsx_lu* lu, const double* A, int64_t lda, int64_t nrhs, const double* B, int64_t ldb,
                         const double* X, int64_t ldx, double* norms) {
    if (!lu || !A || !B || !X || !norms || nrhs < 0) return LSX_EINVAL;
    lsx_ctx* ctx = lu->ctx;
    const int64_t n = lu->n;
    if (lda < n || ldb < n || ldx < n) return fail(ctx, LSX_EINVAL, "lsx_lu_residual_norm: lda/ldb/ldx < n");
    norms[0] = norms[1] = 0.0;
    if (nrhs == 0) return LSX_OK;
    DeviceGuard g(ctx);
    int rc = lu_ensure_work(lu, nrhs);
    if (rc) return rc;
    const hipStream_t st = lu->stream;
    // A is read where it lies when it is memory of this device, else staged once
    const double* dA = A;
    int64_t ldA = lda;
    if (!is_device_ptr(ctx->device, A)) {
        if (!lu->astage && hipMalloc(&lu->astage, sizeof(double) * std::max(lu->n, lu->cap) * std::max(lu->n, lu->cap)) != hipSuccess) {
            lu->astage = nullptr;
            return fail(ctx, LSX_ENOMEM, "lsx_lu_residual_norm: cannot stage A");
        }
        LSX_HIP(ctx, hipMemcpy2DAsync(lu->astage, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                                      hipMemcpyDefault, st));
        dA = lu->astage;
        ldA = n;
    }
    if (!lu->rpart) {
        const size_t words = (size_t)resid_part_words(std::max(lu->n, lu->cap));
        if (hipMalloc(&lu->rpart, sizeof(double) * words) != hipSuccess ||
            hipHostMalloc(&lu->hrpart, sizeof(double) * words, hipHostMallocDefault) != hipSuccess)
            return fail(ctx, LSX_ENOMEM, "lsx_lu_residual_norm: cannot allocate the statistics buffers");
    }
    // B into wb (the residual overwrites it), X into wy
    LSX_HIP(ctx, hipMemcpy2DAsync(lu->wb, sizeof(double) * n, B, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                                  hipMemcpyDefault, st));
    LSX_HIP(ctx, hipMemcpy2DAsync(lu->wy, sizeof(double) * n, X, sizeof(double) * ldx, sizeof(double) * n, nrhs,
                                  hipMemcpyDefault, st));
    double ssr = 0.0, ssb = 0.0;
    for (int64_t c0 = 0; c0 < nrhs; c0 += RNR) {
        const int nr = (int)std::min<int64_t>(RNR, nrhs - c0);
        resid_launch(st, dA, ldA, n, lu->wy + c0 * n, n, nr, lu->wb + c0 * n, n, lu->wb + c0 * n, n, lu->rpart,
                     RES_RESIDUAL);
        LSX_HIP(ctx, hipGetLastError());
        LSX_HIP(ctx, hipMemcpyAsync(lu->hrpart, lu->rpart, sizeof(double) * resid_part_words(n), hipMemcpyDeviceToHost, st));
        LSX_HIP(ctx, hipStreamSynchronize(st));
        double stv[RNR][RSTAT];
        resid_combine(lu->hrpart, n, nr, stv);
        for (int s = 0; s < nr; ++s) {
            ssr += stv[s][1];
            ssb += stv[s][4];
        }
    }
    norms[0] = std::sqrt(ssr);
    norms[1] = std::sqrt(ssb);
    return LSX_OK;
}

int lsx_residual_norm_noalloc(lsx_ctx* ctx, int64_t n, const double* A, int64_t lda, int64_t nrhs,
                              const double* B, int64_t ldb, const double* X, int64_t ldx, double* norms) {
    if (!ctx || !A || !B || !X || !norms || n < 1 || nrhs < 0) return LSX_EINVAL;
    if (n > LSX_NOALLOC_MAX_N)
        return fail(ctx, LSX_EINVAL, "lsx_residual_norm_noalloc: n = %lld exceeds the scratch bound %d",
                    (long long)n, LSX_NOALLOC_MAX_N);
    if (lda < n || ldb < n || ldx < n)
        return fail(ctx, LSX_EINVAL, "lsx_residual_norm_noalloc: lda/ldb/ldx < n");
    if (!ctx->scratch) return fail(ctx, LSX_ESTATE, "lsx_residual_norm_noalloc: no scratch address");
    DeviceGuard g(ctx);
    // nothing is staged (that would allocate): A, B and X must be memory of this GPU
    if (!is_device_ptr(ctx->device, A) || !is_device_ptr(ctx->device, B) || !is_device_ptr(ctx->device, X))
        return fail(ctx, LSX_EINVAL, "lsx_residual_norm_noalloc: A, B and X must be device memory of GPU %d",
                    ctx->device);
    if (nrhs == 0) {
        norms[0] = norms[1] = 0.0;
        return LSX_OK;
    }
    double* part = ctx->scratch;
    double* acc = ctx->scratch + NOALLOC_PART_WORDS;
    const hipStream_t st = nullptr;  // the null stream: ordered after the caller's blocking work
    for (int64_t c0 = 0; c0 < nrhs; c0 += RNR) {
        const int nr = (int)std::min<int64_t>(RNR, nrhs - c0);
        // statistics only (no R): every block writes all of its statistics words, so nothing in
        // the scratch needs clearing, and the first sum launch starts the totals at zero
        resid_launch(st, A, lda, n, X + c0 * ldx, ldx, nr, B + c0 * ldb, ldb, nullptr, 0, part, RES_RESIDUAL);
        hipLaunchKernelGGL(k_resid_sumsq, dim3(1), dim3(64), 0, st, (const double*)part, resid_blocks(n), nr, acc,
                           c0 == 0 ? 1 : 0);
        LSX_HIP(ctx, hipGetLastError());
    }
    double h[2];
    LSX_HIP(ctx, hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, st));
    LSX_HIP(ctx, hipStreamSynchronize(st));
    norms[0] = std::sqrt(h[0]);
    norms[1] = std::sqrt(h[1]);
    return LSX_OK;
}

int lsx_lu_profile(lsx_lu* lu, int32_t enable, lsx_profile* out) {
    if (!lu) return LSX_EINVAL;
    DeviceGuard g(lu->ctx);
    LSX_HIP(lu->ctx, hipStreamSynchronize(lu->stream));
    lu->prof.collect();
    if (out) *out = lu->prof.acc;
    if (enable && !lu->prof.on) lu->prof.reset();
    lu->prof.on = enable != 0;
    return LSX_OK;
}

}  // extern "C"
