/*
 * lsx.h — C-ABI of the MI355X (gfx950) dense-LU / SimpleGMRES library (liblsx.so).
 *
 * Conventions (those of the reference's Julia arrays and LAPACK):
 *   - matrices are column-major with an explicit leading dimension;
 *   - pivots are 1-based int64 LAPACK swap sequences (ipiv[k] = row interchanged with row k);
 *   - every pointer argument may point to host memory or to device memory of the context's GPU
 *     (the library detects which), except where a function says "device";
 *   - device arrays are read and written on the library's own non-blocking streams, which do not
 *     wait for the caller's streams: an input must be complete, and an output no longer in use,
 *     when the call is made (synchronize the stream that produced it first);
 *   - every function returns int32: 0 = success, > 0 = LAPACK-style info (1-based index of the
 *     first exactly-zero pivot; the factorization still ran to completion), < 0 = error
 *     (one of lsx_status below); lsx_last_error() then returns a message.  No C++ exception
 *     crosses this boundary.
 *   - one handle = one HIP stream; handles are independent and may be used from different host
 *     threads; a single handle must not be used by two threads at once.
 *
 * Reference citations are paths in LinearSolve.jl v5.12.0.
 */
#ifndef LSX_H
#define LSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    LSX_OK = 0,
    LSX_EINVAL = -1,   /* bad argument */
    LSX_EHIP = -2,     /* HIP runtime error */
    LSX_ENOMEM = -3,   /* device allocation failed or over the HBM budget */
    LSX_ESTATE = -4,   /* call out of order (e.g. solve before a successful factor) */
    LSX_ETIMEOUT = -5, /* a device-side wait exceeded its bound (results invalid) */
    LSX_EAGAIN = -6    /* result not ready yet (serving poll) */
} lsx_status;

typedef enum { LSX_F64 = 0, LSX_F32 = 1, LSX_BF16 = 2 } lsx_dtype;

/* SciMLBase.ReturnCode values used by the hot paths (simplegmres.jl:603-606,
 * factorization.jl:716-722). */
typedef enum {
    LSX_RET_DEFAULT = 0,
    LSX_RET_SUCCESS = 1,
    LSX_RET_MAXITERS = 2,
    LSX_RET_INFEASIBLE = 3,
    LSX_RET_FAILURE = 4
} lsx_retcode;

typedef struct lsx_ctx lsx_ctx;
typedef struct lsx_lu lsx_lu;
typedef struct lsx_batch lsx_batch;
typedef struct lsx_csr lsx_csr;
typedef struct lsx_serve lsx_serve;

/* ---------------------------------------------------------------- context and memory ----- */
/* A context binds one GPU.  Stands in for the device side of LinearCache.cacheval
 * (src/common.jl:281-306). */
int lsx_ctx_create(int device, lsx_ctx** out);
int lsx_ctx_destroy(lsx_ctx* ctx);
const char* lsx_last_error(const lsx_ctx* ctx);
int lsx_version(void);
/* Bound, in seconds, of every device-side wait of later calls on this context: the panel
 * exchange of getrf, the block chain of the triangular solves and the all-reduces of GMRES.
 * 0 restores the defaults (5 s for the LU waits, 20 s for GMRES).  When a wait reaches its bound
 * the call returns LSX_ETIMEOUT; every other wait of the same call then gives up at once, so the
 * call ends promptly.  No reference counterpart (LAPACK and SparseArrays never wait on other
 * processors); for callers that must bound a call's time. */
int lsx_ctx_set_wait_bound(lsx_ctx* ctx, double seconds);
/* Device memory helpers for hosts without a GPU array package (device pointers). */
int lsx_malloc(lsx_ctx* ctx, size_t bytes, void** dptr);
int lsx_free(lsx_ctx* ctx, void* dptr);
int lsx_memcpy(lsx_ctx* ctx, void* dst, const void* src, size_t bytes);
int lsx_synchronize(lsx_ctx* ctx);

/* ---------------------------------------------------------------- dense LU --------------- */
/* Handle of one n x n factorization kept resident on the device: replaces the LU object that
 * LUFactorization stores in cache.cacheval (src/factorization.jl:714, init_cacheval :840-849).
 * dtype LSX_F64 = plain partial-pivot LU (LAPACK dgetrf semantics); LSX_F32 / LSX_BF16 = the
 * mixed-precision factorization used by lsx_lu_refine: FP32 factors with tournament pivoting
 * on block columns of at most 16384 rows (its pivots need not be LAPACK's; taller ones take exact
 * partial pivoting), their trailing products on BF16 MFMA from three (LSX_F32: FP32 accuracy)
 * or two (LSX_BF16: relative error ~2^-17) BF16 pieces per operand. */
int lsx_lu_create(lsx_ctx* ctx, int64_t n, int dtype, lsx_lu** out);
int lsx_lu_destroy(lsx_lu* lu);

/* getrf: copies A (host or device, lda >= n) into the handle and factors it.
 * Replaces lu(A, check=false) / LAPACK.getrf!(factors, ipiv) at src/factorization.jl:689-697
 * (ccall path src/openblas.jl:144-151).  Returns info (> 0: first zero pivot, as LAPACK). */
int lsx_lu_factor(lsx_lu* lu, const double* A, int64_t lda);

/* Copies the resident factors (L unit-lower below the diagonal, U on and above) and the
 * 1-based pivots out; either pointer may be NULL.  ipiv has n entries. */
int lsx_lu_get_factors(lsx_lu* lu, double* F, int64_t ldf, int64_t* ipiv);

/* getrs: X = A \ B for nrhs right-hand sides with the resident factors.  B and X may alias.
 * Replaces _smart_lu_ldiv!/ldiv!(F, b) -> LAPACK dgetrs at src/factorization.jl:601-611,727-728
 * (ccall path src/openblas.jl:267-275). */
int lsx_lu_solve(lsx_lu* lu, int64_t nrhs, const double* B, int64_t ldb, double* X, int64_t ldx);

/* Adjoint (= transpose, real data) solve with the same resident factors: X = A^{-T} B, i.e.
 * LAPACK dgetrs with trans = 'T'.  Replaces the factorization reuse of solve!(cache; adjoint =
 * true) -> _adjoint_factorization_solve -> F' \ b (src/common.jl:1012-1027,
 * src/adjoint_factorization.jl:162-176).  FP64 handles only.  B and X may alias. */
int lsx_lu_solve_adjoint(lsx_lu* lu, int64_t nrhs, const double* B, int64_t ldb, double* X, int64_t ldx);

/* norms[0] = ||B - A X||_2 and norms[1] = ||B||_2 (2-norms over all nrhs columns), computed in
 * FP64 on the device.  A is the matrix before factorization (host or device; a host A is staged
 * into a device copy the handle keeps).  Replaces the GEMV of the post-solve residual check
 * _check_residual_safety (src/factorization.jl:127-156: mul!(buf, A, y); axpy!(-1, b, buf);
 * norm(buf) > abstol + reltol * norm(b) -> ReturnCode.APosterioriSafetyFailure), run when
 * residualsafety = true and the solve refactorized (:646). */
int lsx_lu_residual_norm(lsx_lu* lu, const double* A, int64_t lda, int64_t nrhs, const double* B, int64_t ldb,
                         const double* X, int64_t ldx, double* norms);

/* The same two norms as lsx_lu_residual_norm (bit for bit: the same kernel and summation order)
 * with no handle and no device allocation, for callers that cannot allocate: code running
 * between the kernels of a captured or tight serving loop, harnesses that forbid hipMalloc.  No
 * counterpart in the reference.  A, B and X must be device memory of ctx's GPU (a host array
 * would need a staging copy) and n <= LSX_NOALLOC_MAX_N; otherwise LSX_EINVAL.  norms is host
 * memory.  The work runs on the GPU's null stream (ordered after the caller's work on blocking
 * streams) and the call returns when it is done.  Its partial sums live in a scratch array of
 * 656 KB that each GPU holds from lsx_ctx_create until the process ends; calls on one GPU share
 * it, so calls that overlap on one GPU (from several host threads) race. */
#define LSX_NOALLOC_MAX_N 65536
int lsx_residual_norm_noalloc(lsx_ctx* ctx, int64_t n, const double* A, int64_t lda, int64_t nrhs,
                              const double* B, int64_t ldb, const double* X, int64_t ldx, double* norms);

/* Per-kernel-class device timing of a handle (HIP events around each launch while enabled).
 * Classes: 0 panel, 1 row interchanges + TRSM, 2 trailing MFMA update (128 x 128 tiles),
 * 3 permutation lists, 4 triangular solves, 5 refinement residual/update, 6 low-precision
 * conversion, 7 other, 8 small MFMA products (64 x 64 tiles: narrow updates and the TRSM
 * products).  flops/bytes are the algorithmic work of the launches timed. */
#define LSX_PROFILE_CLASSES 9
typedef struct {
    int64_t launches[LSX_PROFILE_CLASSES];
    double ms[LSX_PROFILE_CLASSES];
    double flops[LSX_PROFILE_CLASSES];
    double bytes[LSX_PROFILE_CLASSES];
} lsx_profile;

/* enable = 1 starts timing (and clears the counters), 0 stops; out (nullable) receives the
 * counters accumulated so far. */
int lsx_lu_profile(lsx_lu* lu, int32_t enable, lsx_profile* out);

/* Statistics of the last lsx_lu_refine call. */
typedef struct {
    int32_t mode_used;      /* 0 = iterative refinement, 1 = GMRES-IR */
    int32_t iterations;     /* refinement steps (IR) or total GMRES iterations (GMRES-IR) */
    int32_t outer;          /* refinement sweeps */
    int32_t converged;      /* 1 if the backward error reached tol */
    double backward_error;  /* max over columns of ||b-Ax||_inf/(||A||_inf ||x||_inf+||b||_inf) */
    int32_t escalated;      /* 1 if mode 2 refactored in FP32 after BF16 refinement stalled */
    int32_t pad;
} lsx_refine_stats;

/* Mixed-precision solve of the factored lu (LSX_F32/LSX_BF16) with FP64 refinement to a normwise
 * backward error <= tol.  Returns 0, or LAPACK-style info > 0 when a refactorization in FP32
 * (mode 2, below) finds an exactly-zero pivot (no solution then), or < 0 on error.  mode 0 = IR, 1 = GMRES-IR, 2 = IR, then GMRES-IR if IR stalls, then
 * (BF16 factors only) a refactorization in FP32 if GMRES-IR stalls too.  A (host or device) is the original FP64 matrix when it is not already held by the
 * handle (pass NULL after lsx_lu_factor: the handle keeps its FP64 copy).  No reference
 * counterpart does refinement: the API shape follows OpenBLAS32MixedLUFactorization
 * (src/openblas.jl:461-543), whose FP32 result is refined here to FP64 accuracy. */
int lsx_lu_refine(lsx_lu* lu, int64_t nrhs, const double* B, int64_t ldb, double* X, int64_t ldx,
                  double tol, int32_t max_iters, int32_t mode, lsx_refine_stats* stats);

/* ---------------------------------------------------------------- batched dense LU ------- */
/* count independent systems of sizes n[i] (1 <= n[i] <= 512), packed back to back:
 * A_i column-major at offset sum_{j<i} n_j^2, b_i / x_i at offset sum_{j<i} n_j.
 * Replaces the per-block lu!/ldiv! loop of BlockDiagonalFactorization
 * (ext/LinearSolveBlockDiagonalsExt.jl:101-114,123-185).  dtype LSX_F64 = LAPACK dgetrf/dgetrs
 * per block; LSX_F32 = FP32 factorization with FP64 iterative refinement of each solution
 * (no reference counterpart: the FP32 analogue of the blockwise path). */
int lsx_batch_create(lsx_ctx* ctx, int64_t count, const int64_t* n, int dtype, lsx_batch** out);
int lsx_batch_destroy(lsx_batch* batch);
/* Factor and solve every system in one persistent launch (replayed from a HIP graph).
 * info[i] (nullable, host or device) receives each system's LAPACK info; ipiv (nullable)
 * receives the pivots packed like b.  Returns the number of systems with info > 0. */
int lsx_batch_factor_solve(lsx_batch* batch, const double* A, const double* b, double* x,
                           int64_t* info, int64_t* ipiv);
/* Size-binned tile table: a system of size n uses panel width bw[k] (8, 16 or 32) of the first bin
 * with n <= upper[k]; sizes beyond the last bin keep the built-in choice.  Chosen by the autotuner (linearsolve.jl_amd/autotune.py); no reference
 * counterpart. */
int lsx_batch_set_tiles(lsx_batch* batch, int32_t nbins, const int32_t* upper, const int32_t* bw);
/* LSX_F32 batches: refinement target (normwise backward error, default 1e-14) and step limit
 * (default 10).  No reference counterpart. */
int lsx_batch_set_refinement(lsx_batch* batch, double tol, int32_t max_iters);
/* Per-system refinement steps and final backward error of the last factor_solve (LSX_F32;
 * LSX_F64 batches report 0 steps and -1).  Either pointer may be NULL.  No reference counterpart. */
int lsx_batch_stats(lsx_batch* batch, int32_t* iters, double* backward_error);

/* ---------------------------------------------------------------- sparse / SimpleGMRES --- */
/* CSR matrix resident on the device (int32 row pointers and column indices, FP64 values). */
int lsx_csr_create(lsx_ctx* ctx, int64_t N, int64_t nnz, const int32_t* rowptr,
                   const int32_t* colind, const double* vals, lsx_csr** out);
int lsx_csr_destroy(lsx_csr* A);
/* y = A x (device or host vectors). */
int lsx_csr_spmv(lsx_csr* A, const double* x, double* y);

typedef struct {
    int32_t retcode;   /* lsx_retcode */
    int32_t pad;
    int64_t iters;     /* LinearSolution.iters */
    double resid;      /* LinearSolution.resid: the preconditioned residual estimate rNorm */
} lsx_gmres_stats;

/* SimpleGMRES (src/simplegmres.jl:375-616, scalar layout): restarted GMRES with modified
 * Gram-Schmidt Arnoldi and symmetric Givens rotations.  pl_diag / pr_diag: diagonal left/right
 * preconditioners (ldiv! divides elementwise), NULL = identity.  x receives the solution
 * (warm_start = false: the iteration starts from zero).  The Krylov workspace stays resident on
 * the handle between calls. */
int lsx_gmres_solve(lsx_csr* A, const double* b, double* x, int32_t memory, int32_t restart,
                    double abstol, double reltol, int64_t maxiters, const double* pl_diag,
                    const double* pr_diag, lsx_gmres_stats* stats);

/* The largest Krylov dimension a solve on A can hold: memory with restart, or maxiters without.
 * 2048 for every N: the basis is in HBM, and each solving workgroup keeps the packed triangle R,
 * the Givens pairs and z in its LDS beside its rows of q and the basis when they fit (up to 128
 * vectors for N <= 512, about 70 for N > 2048), else in an HBM area of its own.  (The batched
 * layout, lsx_gmres_solve_batched, keeps them in LDS and refuses dimensions beyond that.)
 * Replaces the growth of the workspace that SimpleGMRES and KrylovJL_GMRES do without restart
 * (src/simplegmres.jl:222-235 _gmres_grow_pass!/_gmres_grow_krylov!, Krylov.jl's gmres with
 * restart = false under src/iterative_wrappers.jl:693): a caller asks first and restarts at
 * this dimension when maxiters exceeds it. */
int lsx_gmres_max_memory(lsx_csr* A, int32_t* memory);

/* SimpleGMRES with blocksize > 0 (the batched layout, src/simplegmres.jl:190-373): A is block
 * diagonal with N / blocksize uniform blocks, and every block runs its own GMRES.  Each block
 * stops when its residual estimate reaches eps = abstol + reltol ||Pl \ b||_2 (the norm of the
 * whole vector), the reference's test on its slowest block; stats carry the slowest block's
 * iterations and rNorm, Infeasible if any block is, else Success if all are solved, else
 * MaxIters.  The blocks run as one multi-solve launch.  Returns LSX_EINVAL when A has entries
 * outside the diagonal blocks or blocksize exceeds 8192. */
int lsx_gmres_solve_batched(lsx_csr* A, int64_t blocksize, const double* b, double* x, int32_t memory,
                            int32_t restart, double abstol, double reltol, int64_t maxiters,
                            const double* pl_diag, const double* pr_diag, lsx_gmres_stats* stats);

/* ---------------------------------------------------------------- serving ---------------- */
typedef struct {
    int32_t kind;          /* 0 = dense LU solve, 1 = SimpleGMRES */
    int32_t dtype;         /* dense: LSX_F64 */
    int64_t n;             /* dense order, or sparse N */
    const double* A;       /* dense: n x n column-major (lda = n), host or device */
    const double* b;       /* length n, host or device */
    int64_t nnz;           /* sparse CSR (int32 indices), host or device */
    const int32_t* rowptr;
    const int32_t* colind;
    const double* vals;
    const double* pl_diag; /* sparse: diagonal left preconditioner or NULL */
    int32_t memory, restart;
    double abstol, reltol;
    int64_t maxiters;
    /* synth != 0: the inputs are generated on the device instead of read from A/b/CSR:
     *   dense:  A = I - g J, J iid N(0,1)/sqrt(n), g ~ U[1e-3, 1e-1], b iid N(0,1) from seed;
     *   sparse: 2D 5-point convection-diffusion on a grid_m x grid_m grid with Peclet pe
     *           (diagonal 4, west/south -1 - pe h/2, east/north -1 + pe h/2, h = 1/(grid_m+1)),
     *           b iid N(0,1) from seed, Jacobi preconditioner (pl_diag ignored). */
    int32_t synth;
    int32_t grid_m;
    uint64_t seed;
    double pe;
} lsx_request;

typedef struct {
    int32_t status;     /* 0 done, LSX_EAGAIN pending, < 0 error */
    int32_t retcode;    /* lsx_retcode */
    int64_t info;       /* dense: LAPACK info */
    int64_t iters;      /* sparse: GMRES iterations */
    double resid;       /* sparse: rNorm */
    double t_submit, t_start, t_done; /* seconds on the serving clock */
} lsx_result;

/* A continuous-batching request loop over one GPU.  hbm_budget bytes are allocated once: a quarter
 * (at most 2 GiB) is the arena of the batched launches, the rest a heap from which each admitted
 * request's device memory is carved (admission = the carve succeeds, and the request's
 * cooperating workgroups fit beside the running ones).  Dense systems with n <= 512 are grouped
 * into batched launches, larger ones run on reusable LU handles, SimpleGMRES requests are grouped
 * into multi-solve launches.  nstreams (2..8) lanes serve dense and sparse requests, split evenly;
 * the batched lane comes on top.  The request's host buffers must stay valid until the result is
 * polled.  No counterpart in the reference (LinearSolve.jl solves one problem per call). */
int lsx_serve_create(lsx_ctx* ctx, size_t hbm_budget, int32_t nstreams, lsx_serve** out);
int lsx_serve_destroy(lsx_serve* s);
int lsx_serve_submit(lsx_serve* s, const lsx_request* req, double* x_out, uint64_t* ticket);
/* Returns LSX_OK with *res filled when the ticket is done, LSX_EAGAIN while pending. */
int lsx_serve_poll(lsx_serve* s, uint64_t ticket, lsx_result* res);
int lsx_serve_drain(lsx_serve* s);
double lsx_serve_clock(lsx_serve* s);
/* Serving counters: [0] submitted, [1] completed, [2] batched launches, [3] requests served by
 * batched launches, [4] peak bytes admitted, [5] requests waiting for admission now. */
int lsx_serve_stats(lsx_serve* s, int64_t* out6);

#ifdef __cplusplus
}
#endif
#endif
