/*
 * lsx_oracle.c — CPU restatement of the reference's dense-LU and SimpleGMRES hot paths.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke entry point and
 * the CPU-baseline leg of the benchmark may load it.  The product (the HIP library) never calls
 * it, links it or falls back to it.
 *
 * What it restates (citations are paths inside LinearSolve.jl v5.12.0):
 *   ora_dgetrf   blocked right-looking getrf: src/blocked_lufact.jl:658-679 (driver), :38-54
 *                (pivot search, first index of the maximum, NaN ignored), :93-122 (panel getf2),
 *                :126-141 (row interchanges outside the panel), :146-178 (unit-lower TRSM),
 *                :186-620 (Schur update C -= L21*U12); info semantics of src/generic_lufact.jl:71-141
 *                (first zero pivot, factorization runs to completion).  The reference's
 *                LUFactorization calls LAPACK dgetrf (OpenBLAS_jll, src/factorization.jl:633,697);
 *                blocked_lufact.jl is the reference's own statement of the same algorithm.
 *   ora_dgetrs   back-solve of src/factorization.jl:433-491 (_naive_lu_ldiv!): apply the swap
 *                sequence ipiv[1..n] in order, unit-lower forward sweep, upper backward sweep
 *                (vector: divide by U[j,j]; matrix: multiply by inv(U[j,j])).
 *   ora_gmres    SimpleGMRES, scalar layout: src/simplegmres.jl:110-129 (reinit), :140-162
 *                (_sym_givens), :253-324 (orth step, norm, Givens apply/new, z update, next v),
 *                :326-368 (back-substitution, correction), :396-463 (_init_cacheval), :465-616
 *                (restart loop, stopping tests, retcodes).  Pl and Pr are diagonal or identity
 *                (ldiv! with a Diagonal divides elementwise).  SpMV is CSR with columns summed in
 *                ascending order, the same order SparseArrays' CSC mul! accumulates a row.
 *
 * Pinning (which tests check what) is listed in the tests that load this file.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define A_(i, j) A[(size_t)(j) * (size_t)lda + (size_t)(i)]

/* blocked_lufact.jl:38-54: amax by '>' select (NaN compares false), then first index of amax. */
static int64_t find_pivot(const double* A, int64_t lda, int64_t k, int64_t m, double* amax_out) {
    double amax = 0.0;
    for (int64_t i = k; i < m; ++i) {
        double a = fabs(A_(i, k));
        amax = (a > amax) ? a : amax;
    }
    int64_t kp = k;
    if (amax != 0.0) {
        for (int64_t i = k; i < m; ++i)
            if (fabs(A_(i, k)) == amax) { kp = i; break; }
    }
    *amax_out = amax;
    return kp;
}

/* blocked_lufact.jl:93-122 (and :58-89 when the panel is the whole matrix). ipiv is 1-based. */
static int64_t panel(double* A, int64_t lda, int64_t* ipiv, int64_t m, int64_t j0, int64_t j1,
                     int64_t info) {
    for (int64_t k = j0; k <= j1; ++k) {
        double amax;
        int64_t kp = find_pivot(A, lda, k, m, &amax);
        ipiv[k] = kp + 1;
        if (A_(kp, k) != 0.0) {
            if (k != kp)
                for (int64_t j = j0; j <= j1; ++j) {
                    double t = A_(k, j); A_(k, j) = A_(kp, j); A_(kp, j) = t;
                }
            double inv = 1.0 / A_(k, k);
            for (int64_t i = k + 1; i < m; ++i) A_(i, k) *= inv;
        } else if (info == 0) {
            info = k + 1;
        }
        for (int64_t j = k + 1; j <= j1; ++j) {
            double akj = A_(k, j);
            for (int64_t i = k + 1; i < m; ++i) A_(i, j) = A_(i, j) - A_(i, k) * akj;
        }
    }
    return info;
}

/* blocked_lufact.jl:126-141: interchanges of ipiv[k0..k1] on columns c0..c1. */
static void swap_rows(double* A, int64_t lda, const int64_t* ipiv, int64_t k0, int64_t k1,
                      int64_t c0, int64_t c1) {
    for (int64_t j = c0; j <= c1; ++j)
        for (int64_t k = k0; k <= k1; ++k) {
            int64_t kp = ipiv[k] - 1;
            if (kp != k) { double t = A_(k, j); A_(k, j) = A_(kp, j); A_(kp, j) = t; }
        }
}

/* blocked_lufact.jl:146-178: U12 := L11 \ A12, L11 unit lower. */
static void trsm_unit_lower(double* A, int64_t lda, int64_t j0, int64_t j1, int64_t c0, int64_t c1) {
    for (int64_t c = c0; c <= c1; ++c)
        for (int64_t k = j0; k <= j1; ++k) {
            double bk = A_(k, c);
            for (int64_t i = k + 1; i <= j1; ++i) A_(i, c) = A_(i, c) - A_(i, k) * bk;
        }
}

/* blocked_lufact.jl:186-620: A[j1+1:m, c0:c1] -= A[j1+1:m, j0:j1] * A[j0:j1, c0:c1]. */
static void schur(double* A, int64_t lda, int64_t m, int64_t j0, int64_t j1, int64_t c0, int64_t c1) {
    for (int64_t c = c0; c <= c1; ++c)
        for (int64_t k = j0; k <= j1; ++k) {
            double b = A_(k, c);
            double* __restrict cc = &A_(0, c);
            const double* __restrict lk = &A_(0, k);
            for (int64_t i = j1 + 1; i < m; ++i) cc[i] -= lk[i] * b;
        }
}

/* blocked_lufact.jl:658-679 (driver).  At minmn <= 8 the reference runs the unblocked sweep
 * (:725-726); here that is one panel covering every pivot column, the same operations.
 * nb <= 0 selects the reference's default panel width (:16).  Returns info. */
int64_t ora_dgetrf(int64_t m, int64_t n, double* A, int64_t lda, int64_t* ipiv, int64_t nb) {
    int64_t minmn = m < n ? m : n;
    if (minmn <= 0) return 0;
    if (nb <= 0) nb = minmn <= 160 ? 8 : 16;
    if (minmn <= 8) nb = minmn;
    int64_t info = 0;
    for (int64_t j0 = 0; j0 < minmn; j0 += nb) {
        int64_t jb = (nb < minmn - j0) ? nb : (minmn - j0);
        int64_t j1 = j0 + jb - 1;
        info = panel(A, lda, ipiv, m, j0, j1, info);
        swap_rows(A, lda, ipiv, j0, j1, 0, j0 - 1);
        if (j1 < n - 1) {
            swap_rows(A, lda, ipiv, j0, j1, j1 + 1, n - 1);
            trsm_unit_lower(A, lda, j0, j1, j1 + 1, n - 1);
            if (j1 < m - 1) schur(A, lda, m, j0, j1, j1 + 1, n - 1);
        }
    }
    return info;
}

/* factorization.jl:433-491: x = U \ (L \ (P b)), in place on B (n x nrhs, column-major). */
void ora_dgetrs(int64_t n, int64_t nrhs, const double* LU, int64_t ldlu, const int64_t* ipiv,
                double* B, int64_t ldb) {
#define F_(i, j) LU[(size_t)(j) * (size_t)ldlu + (size_t)(i)]
#define B_(i, j) B[(size_t)(j) * (size_t)ldb + (size_t)(i)]
    for (int64_t i = 0; i < n; ++i) {
        int64_t p = ipiv[i] - 1;
        if (p != i)
            for (int64_t c = 0; c < nrhs; ++c) { double t = B_(i, c); B_(i, c) = B_(p, c); B_(p, c) = t; }
    }
    for (int64_t c = 0; c < nrhs; ++c) {
        for (int64_t j = 0; j < n; ++j) {
            double bj = B_(j, c);
            for (int64_t i = j + 1; i < n; ++i) B_(i, c) = B_(i, c) - F_(i, j) * bj;
        }
        for (int64_t j = n - 1; j >= 0; --j) {
            if (nrhs == 1) B_(j, c) /= F_(j, j);
            else B_(j, c) *= 1.0 / F_(j, j);
            double bj = B_(j, c);
            for (int64_t i = 0; i < j; ++i) B_(i, c) = B_(i, c) - F_(i, j) * bj;
        }
    }
#undef F_
#undef B_
}

/* ------------------------------------------------------------------------------------------ */
/* SimpleGMRES (scalar layout).                                                               */

static void spmv(int64_t N, const int32_t* rp, const int32_t* ci, const double* v, const double* x,
                 double* y) {
    for (int64_t i = 0; i < N; ++i) {
        double s = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += v[k] * x[ci[k]];
        y[i] = s;
    }
}
static double dot(int64_t N, const double* a, const double* b) {
    double s = 0.0;
    for (int64_t i = 0; i < N; ++i) s += a[i] * b[i];
    return s;
}
static double nrm2(int64_t N, const double* a) { return sqrt(dot(N, a, a)); }
static double sgn(double a) { return a > 0 ? 1.0 : (a < 0 ? -1.0 : 0.0); }

/* simplegmres.jl:140-162 */
void ora_sym_givens(double a, double b, double* c, double* s, double* rho) {
    if (b == 0.0) {
        *c = (a == 0.0) ? 1.0 : sgn(a); *s = 0.0; *rho = fabs(a);
    } else if (a == 0.0) {
        *c = 0.0; *s = sgn(b); *rho = fabs(b);
    } else if (fabs(b) > fabs(a)) {
        double t = a / b;
        *s = sgn(b) / sqrt(1.0 + t * t); *c = *s * t; *rho = b / *s;
    } else {
        double t = b / a;
        *c = sgn(a) / sqrt(1.0 + t * t); *s = *c * t; *rho = a / *c;
    }
}

/* Return codes mirror SciMLBase.ReturnCode as used at simplegmres.jl:603-606. */
enum { ORA_DEFAULT = 0, ORA_SUCCESS = 1, ORA_MAXITERS = 2, ORA_INFEASIBLE = 3 };

/*
 * ora_gmres: x (length N, output), returns retcode; *iters_out, *rnorm_out as the reference's
 * LinearSolution iters/resid.  pl/pr: diagonal preconditioners (NULL = identity).
 * trace (nullable, length >= maxiters): rNorm after each inner iteration.
 * restart=0 grows the Krylov basis up to maxiters (memory is then only the initial size).
 */
int ora_gmres(int64_t N, const int32_t* rp, const int32_t* ci, const double* val, const double* b,
              double* x, int memory, int restart, double abstol, double reltol, int64_t maxiters,
              const double* pl, const double* pr, int64_t* iters_out, double* rnorm_out,
              double* trace) {
    /* _init_cacheval :420-462 */
    if (memory > maxiters) memory = (int)maxiters;
    int64_t cap = restart ? memory : (maxiters > memory ? maxiters : memory);
    if (cap < 1) cap = 1;
    double* w = calloc((size_t)N, sizeof(double));
    double* q = calloc((size_t)N, sizeof(double));
    double* p = calloc((size_t)N, sizeof(double));
    double* dx = calloc((size_t)N, sizeof(double));
    double* V = calloc((size_t)N * (size_t)(cap + 1), sizeof(double));
    double* s = calloc((size_t)cap + 1, sizeof(double));
    double* c = calloc((size_t)cap + 1, sizeof(double));
    double* z = calloc((size_t)cap + 2, sizeof(double));
    double* R = calloc((size_t)(cap + 1) * (size_t)(cap + 2) / 2 + 1, sizeof(double));
    for (int64_t i = 0; i < N; ++i) x[i] = 0.0;
    /* r0 = Pl \ b  (warm_start = false) */
    for (int64_t i = 0; i < N; ++i) q[i] = pl ? b[i] / pl[i] : b[i];
    double beta = nrm2(N, q);
    double eps_ = abstol + reltol * beta;
    double* xr = restart ? dx : x;
    int status = ORA_DEFAULT;
    double rNorm = beta;
    int64_t iter = 0;
    if (beta == 0.0) {
        status = ORA_SUCCESS;
        goto done;
    }
    {
        int npass = 0;
        const double btol = pow(DBL_EPSILON, 0.75);
        int breakdown = 0, inconsistent = 0;
        int solved = rNorm <= eps_;
        int64_t inner_maxiters = maxiters;
        int tired = iter >= maxiters;
        int64_t inner_iter = 0;
        while (!(solved || tired || breakdown)) {
            int64_t nr = 0;
            if (restart) {
                for (int64_t i = 0; i < N; ++i) xr[i] = 0.0;
                if (npass >= 1) {
                    spmv(N, rp, ci, val, x, w);
                    for (int64_t i = 0; i < N; ++i) w[i] = b[i] - w[i];
                    for (int64_t i = 0; i < N; ++i) q[i] = pl ? w[i] / pl[i] : w[i];
                }
            }
            /* _gmres_start! :242-247 */
            double b0 = nrm2(N, q);
            z[0] = b0;
            for (int64_t i = 0; i < N; ++i) V[i] = q[i] / b0;
            npass++;
            inner_iter = 0;
            int inner_tired = 0;
            while (!(solved || inner_tired || breakdown)) {
                inner_iter++;
                const double* vk = V + (size_t)(inner_iter - 1) * (size_t)N;
                const double* pp = vk;
                if (pr) { for (int64_t i = 0; i < N; ++i) p[i] = vk[i] / pr[i]; pp = p; }
                spmv(N, rp, ci, val, pp, w);
                for (int64_t i = 0; i < N; ++i) q[i] = pl ? w[i] / pl[i] : w[i];
                for (int64_t i = 1; i <= inner_iter; ++i) { /* MGS :253-257 */
                    const double* vi = V + (size_t)(i - 1) * (size_t)N;
                    double h = dot(N, vi, q);
                    R[nr + i - 1] = h;
                    for (int64_t t = 0; t < N; ++t) q[t] = q[t] - h * vi[t];
                }
                double Hbis = nrm2(N, q);
                for (int64_t i = 1; i <= inner_iter - 1; ++i) { /* :272-277 */
                    double Rtmp = c[i - 1] * R[nr + i - 1] + s[i - 1] * R[nr + i];
                    R[nr + i] = s[i - 1] * R[nr + i - 1] - c[i - 1] * R[nr + i];
                    R[nr + i - 1] = Rtmp;
                }
                ora_sym_givens(R[nr + inner_iter - 1], Hbis, &c[inner_iter - 1], &s[inner_iter - 1],
                               &R[nr + inner_iter - 1]);
                double zk1 = s[inner_iter - 1] * z[inner_iter - 1]; /* :300-304 */
                z[inner_iter - 1] = c[inner_iter - 1] * z[inner_iter - 1];
                rNorm = fabs(zk1);
                if (trace) trace[iter + inner_iter - 1] = rNorm;
                nr += inner_iter;
                int resid_decrease_mach = (rNorm + 1.0 <= 1.0);
                int resid_decrease_lim = rNorm <= eps_;
                breakdown = Hbis <= btol;
                solved = resid_decrease_lim || resid_decrease_mach;
                inner_tired = restart ? (inner_iter >= (memory < inner_maxiters ? memory : inner_maxiters))
                                      : (inner_iter >= inner_maxiters);
                if (!(solved || inner_tired || breakdown)) {
                    double* vn = V + (size_t)inner_iter * (size_t)N;
                    for (int64_t i = 0; i < N; ++i) vn[i] = q[i] / Hbis;
                    z[inner_iter] = zk1;
                }
            }
            /* back-substitution :576-585 */
            double* y = z;
            for (int64_t i = inner_iter; i >= 1; --i) {
                int64_t pos = nr + i - inner_iter;
                for (int64_t j = inner_iter; j >= i + 1; --j) {
                    y[i - 1] = y[i - 1] - R[pos - 1] * y[j - 1];
                    pos = pos - j + 1;
                }
                if (fabs(R[pos - 1]) <= btol) { y[i - 1] = 0.0; inconsistent = 1; }
                else y[i - 1] = y[i - 1] / R[pos - 1];
            }
            for (int64_t i = 1; i <= inner_iter; ++i) { /* :587-589 */
                const double* vi = V + (size_t)(i - 1) * (size_t)N;
                for (int64_t t = 0; t < N; ++t) xr[t] = xr[t] + y[i - 1] * vi[t];
            }
            if (pr) {
                for (int64_t t = 0; t < N; ++t) p[t] = xr[t];
                for (int64_t t = 0; t < N; ++t) xr[t] = p[t] / pr[t];
            }
            if (restart) for (int64_t t = 0; t < N; ++t) x[t] = x[t] + xr[t];
            inner_maxiters -= inner_iter;
            iter += inner_iter;
            tired = iter >= maxiters;
        }
        if (tired) status = ORA_MAXITERS;
        if (solved) status = ORA_SUCCESS;
        if (inconsistent) status = ORA_INFEASIBLE;
    }
done:
    *iters_out = iter;
    *rnorm_out = rNorm;
    free(w); free(q); free(p); free(dx); free(V); free(s); free(c); free(z); free(R);
    return status;
}
