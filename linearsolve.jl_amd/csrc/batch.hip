// batch.hip — thousands of independent small dense systems (n <= 512), each factored with
// partial pivoting and solved, in one persistent launch.
//
// Reference: the blockwise factorization of BlockDiagonalFactorization
// (ext/LinearSolveBlockDiagonalsExt.jl:101-114: lu!(B_i, pivot; check = false) per block,
// success = all(issuccess); :123-185 per-block ldiv!).  Per system the arithmetic is that of
// LAPACK dgetrf/dgetrs (pivot = first index of max |a|, NaN ignored; info = first zero pivot).
//
// Scheduling: systems are sorted by decreasing size on the host (longest first); a persistent
// grid of workgroups takes them from an atomic counter, one workgroup per system.
// Per system (BT threads): right-looking blocked LU with panels of BW columns held in LDS; rows
// are never moved in memory: a per-system permutation perm[] (logical row -> physical row)
// absorbs the interchanges, so a swap outside the panel is a swap of two perm entries.  The
// trailing update A22 -= L21 U12 runs on FP64 MFMA with L21 from LDS and U12 from L2.
// The factors stay in the workspace in physical row order together with perm (for re-solves);
// pivots are reported as the LAPACK swap sequence.
#include <algorithm>
#include <cfloat>
#include <new>
#include <numeric>
#include <vector>

#include "batch.hpp"

namespace lsx {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int BT = 256;       // threads per system
constexpr int BW = 16;        // panel width
constexpr int BNMAX = 512;    // largest system

struct BatchArgs {
    int64_t count;
    const int32_t* order;     // systems by decreasing n
    const int32_t* n;
    const int64_t* aoff;      // offset of A_i (n_i^2 doubles)
    const int64_t* voff;      // offset of b_i / x_i / ipiv_i
    double* A;                // workspace: factored in place
    const double* b;
    double* x;
    int64_t* ipiv;            // LAPACK swap sequence (1-based), packed like b
    int32_t* perm;            // logical -> physical rows, packed like b
    int64_t* info;
    uint32_t* counter;
};

struct BatchShared {
    double P[BNMAX][BW + 1];   // panel rows (logical order j0..n)
    double wkey[BT / 64];
    int wrow[BT / 64];
    int prow;
    double xs[BNMAX];          // solve vector
    int perm[BNMAX];
};

__device__ __forceinline__ bool bbetter(double ka, int ra, double kb, int rb) {
    return ka > kb || (ka == kb && ra < rb);
}

__global__ __launch_bounds__(BT) void k_batch_lu(BatchArgs a) {
    __shared__ BatchShared sh;
    __shared__ int s_item;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    while (true) {
        if (tid == 0) s_item = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        const int64_t item = s_item;
        if (item >= a.count) break;
        const int sys = a.order[item];
        const int n = a.n[sys];
        double* A = a.A + a.aoff[sys];
        const int64_t vo = a.voff[sys];
        int64_t info = 0;
        for (int i = tid; i < n; i += BT) sh.perm[i] = i;
        __syncthreads();
        for (int j0 = 0; j0 < n; j0 += BW) {
            const int jb = min(BW, n - j0);
            const int m = n - j0;
            // load the panel (logical rows j0..n-1 through perm)
            for (int idx = tid; idx < m * BW; idx += BT) {
                const int r = idx % m, c = idx / m;
                sh.P[r][c] = (c < jb) ? A[(int64_t)(j0 + c) * n + sh.perm[j0 + r]] : 0.0;
            }
            __syncthreads();
            for (int c = 0; c < jb; ++c) {
                // pivot: first row of max |a| among logical rows j0+c..n-1 (NaN never wins)
                double key = -1.0;
                int krow = 0x7fffffff;
                for (int r = c + tid; r < m; r += BT) {
                    const double t = fabs(sh.P[r][c]);
                    const double kk = (t == t) ? t : 0.0;
                    if (bbetter(kk, r, key, krow)) { key = kk; krow = r; }
                }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    const double ok = __shfl_xor(key, off);
                    const int orow = __shfl_xor(krow, off);
                    if (bbetter(ok, orow, key, krow)) { key = ok; krow = orow; }
                }
                if (lane == 0) { sh.wkey[wave] = key; sh.wrow[wave] = krow; }
                __syncthreads();
                if (tid == 0) {
                    double bk = sh.wkey[0];
                    int br = sh.wrow[0];
                    for (int i = 1; i < BT / 64; ++i)
                        if (bbetter(sh.wkey[i], sh.wrow[i], bk, br)) { bk = sh.wkey[i]; br = sh.wrow[i]; }
                    sh.prow = br;
                    const int k = j0 + c;
                    if (a.ipiv) a.ipiv[vo + k] = (int64_t)(j0 + br) + 1;
                    if (br != c) {
                        const int t = sh.perm[k];
                        sh.perm[k] = sh.perm[j0 + br];
                        sh.perm[j0 + br] = t;
                    }
                }
                __syncthreads();
                const int pr = sh.prow;
                if (pr != c && tid < BW) {
                    const double t = sh.P[c][tid];
                    sh.P[c][tid] = sh.P[pr][tid];
                    sh.P[pr][tid] = t;
                }
                __syncthreads();
                const double pv = sh.P[c][c];
                if (pv == 0.0 && info == 0) info = j0 + c + 1;
                for (int r = c + 1 + tid; r < m; r += BT) {
                    double l = sh.P[r][c];
                    if (pv != 0.0) {
                        if (fabs(pv) >= DBL_MIN) l *= (1.0 / pv);
                        else l /= pv;
                    }
                    sh.P[r][c] = l;
                    for (int q = c + 1; q < jb; ++q) sh.P[r][q] = fma(-l, sh.P[c][q], sh.P[r][q]);
                }
                __syncthreads();
            }
            // write the panel back (physical rows through the updated perm)
            for (int idx = tid; idx < m * jb; idx += BT) {
                const int r = idx % m, c = idx / m;
                A[(int64_t)(j0 + c) * n + sh.perm[j0 + r]] = sh.P[r][c];
            }
            const int c1 = j0 + jb;
            if (c1 < n) {
                // U12 = L11^{-1} A12, one thread per column
                for (int col = c1 + tid; col < n; col += BT) {
                    double u[BW];
#pragma unroll
                    for (int i = 0; i < BW; ++i) u[i] = (i < jb) ? A[(int64_t)col * n + sh.perm[j0 + i]] : 0.0;
#pragma unroll
                    for (int kk = 0; kk < BW; ++kk)
#pragma unroll
                        for (int i = kk + 1; i < BW; ++i) u[i] = fma(-sh.P[i][kk], u[kk], u[i]);
#pragma unroll
                    for (int i = 0; i < BW; ++i)
                        if (i < jb) A[(int64_t)col * n + sh.perm[j0 + i]] = u[i];
                }
                __syncthreads();
                // A22 -= L21 U12 on MFMA: 16 x 16 tiles, waves round-robin.  The MFMA computes
                // the transposed tile, so lane l / slot q holds C(rb + (l & 15), cb + 4q + (l >> 4)).
                const int mr = n - c1, nc = n - c1;
                const int tr = (mr + 15) / 16, tc = (nc + 15) / 16;
                for (int t = wave; t < tr * tc; t += BT / 64) {
                    const int rb = (t % tr) * 16, cb = (t / tr) * 16;
                    const int lr = rb + (lane & 15);                 // local row of C
                    const bool rok = lr < mr;
                    const int prow = rok ? sh.perm[c1 + lr] : 0;
                    v4d acc;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int lc = cb + 4 * q + (lane >> 4);
                        acc[q] = (rok && lc < nc) ? A[(int64_t)(c1 + lc) * n + prow] : 0.0;
                    }
#pragma unroll
                    for (int k4 = 0; k4 < BW / 4; ++k4) {
                        const int kk = k4 * 4 + (lane >> 4);
                        const int lc = cb + (lane & 15);
                        // X(i, k) = -U12(k, col i), Y(k, j) = L21(row j, k)
                        const double xv = (lc < nc && kk < jb) ? -A[(int64_t)(c1 + lc) * n + sh.perm[j0 + kk]] : 0.0;
                        const double yv = (rok && kk < jb) ? sh.P[jb + lr][kk] : 0.0;
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xv, yv, acc, 0, 0, 0);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int lc = cb + 4 * q + (lane >> 4);
                        if (rok && lc < nc) A[(int64_t)(c1 + lc) * n + prow] = acc[q];
                    }
                }
            }
            __syncthreads();
        }
        // solve: x = U^{-1} L^{-1} P b, blocked by BW rows
        for (int i = tid; i < n; i += BT) sh.xs[i] = a.b[vo + sh.perm[i]];
        __syncthreads();
        for (int k0 = 0; k0 < n; k0 += BW) {
            const int kb = min(BW, n - k0);
            if (tid == 0) {
                for (int k = k0; k < k0 + kb; ++k) {
                    const double xk = sh.xs[k];
                    for (int i = k + 1; i < k0 + kb; ++i) sh.xs[i] = fma(-A[(int64_t)k * n + sh.perm[i]], xk, sh.xs[i]);
                }
            }
            __syncthreads();
            for (int i = k0 + kb + tid; i < n; i += BT) {
                double s = sh.xs[i];
                const int pi = sh.perm[i];
                for (int k = k0; k < k0 + kb; ++k) s = fma(-A[(int64_t)k * n + pi], sh.xs[k], s);
                sh.xs[i] = s;
            }
            __syncthreads();
        }
        for (int k1 = n; k1 > 0; k1 -= BW) {
            const int k0 = max(0, k1 - BW);
            if (tid == 0) {
                for (int k = k1 - 1; k >= k0; --k) {
                    const double xk = sh.xs[k] / A[(int64_t)k * n + sh.perm[k]];
                    sh.xs[k] = xk;
                    for (int i = k0; i < k; ++i) sh.xs[i] = fma(-A[(int64_t)k * n + sh.perm[i]], xk, sh.xs[i]);
                }
            }
            __syncthreads();
            for (int i = tid; i < k0; i += BT) {
                double s = sh.xs[i];
                const int pi = sh.perm[i];
                for (int k = k0; k < k1; ++k) s = fma(-A[(int64_t)k * n + pi], sh.xs[k], s);
                sh.xs[i] = s;
            }
            __syncthreads();
        }
        for (int i = tid; i < n; i += BT) {
            a.x[vo + i] = sh.xs[i];
            a.perm[vo + i] = sh.perm[i];
        }
        if (tid == 0) a.info[sys] = info;
        __syncthreads();
    }
}

__global__ void k_reset_counter(uint32_t* c) { *c = 0u; }

int batch_launch_raw(lsx_ctx* ctx, hipStream_t st, int64_t count, const int32_t* order, const int32_t* n,
                     const int64_t* aoff, const int64_t* voff, double* A, const double* b, double* x, int64_t* ipiv,
                     int32_t* perm, int64_t* info, uint32_t* counter) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, counter);
    BatchArgs a;
    a.count = count;
    a.order = order;
    a.n = n;
    a.aoff = aoff;
    a.voff = voff;
    a.A = A;
    a.b = b;
    a.x = x;
    a.ipiv = ipiv;
    a.perm = perm;
    a.info = info;
    a.counter = counter;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(count, (int64_t)ctx->num_cus * 2));
    hipLaunchKernelGGL(k_batch_lu, dim3(grid), dim3(BT), 0, st, a);
    LSX_HIP(ctx, hipGetLastError());
    return LSX_OK;
}

int batch_run(lsx_batch* B, const double* dA_src, const double* db, double* dx, bool copy_a) {
    lsx_ctx* ctx = B->ctx;
    if (copy_a)
        LSX_HIP(ctx, hipMemcpyAsync(B->A, dA_src, sizeof(double) * B->total_a, hipMemcpyDefault, B->stream));
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, B->stream, B->counter);
    BatchArgs a;
    a.count = B->count;
    a.order = B->order;
    a.n = B->n;
    a.aoff = B->aoff;
    a.voff = B->voff;
    a.A = B->A;
    a.b = db;
    a.x = dx;
    a.ipiv = B->ipiv;
    a.perm = B->perm;
    a.info = B->info;
    a.counter = B->counter;
    const int grid = std::max<int64_t>(1, std::min<int64_t>(B->count, (int64_t)ctx->num_cus * 2));
    const int pi = B->prof.begin(B->stream, PK_OTHER, B->flops, sizeof(double) * 2.0 * (double)B->total_a);
    hipLaunchKernelGGL(k_batch_lu, dim3(grid), dim3(BT), 0, B->stream, a);
    B->prof.end(B->stream, pi);
    LSX_HIP(ctx, hipGetLastError());
    return LSX_OK;
}

}  // namespace lsx

using namespace lsx;

static void batch_free(lsx_batch* B) {
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    (void)hipFree(B->order);
    (void)hipFree(B->n);
    (void)hipFree(B->aoff);
    (void)hipFree(B->voff);
    (void)hipFree(B->A);
    (void)hipFree(B->b);
    (void)hipFree(B->x);
    (void)hipFree(B->ipiv);
    (void)hipFree(B->perm);
    (void)hipFree(B->info);
    (void)hipFree(B->counter);
    if (B->hinfo) (void)hipHostFree(B->hinfo);
    if (B->stream) (void)hipStreamDestroy(B->stream);
}

extern "C" {

int lsx_batch_create(lsx_ctx* ctx, int64_t count, const int64_t* n, int dtype, lsx_batch** out) {
    if (!ctx || !out || !n || count < 1) return LSX_EINVAL;
    *out = nullptr;
    if (dtype != LSX_F64) return fail(ctx, LSX_EINVAL, "lsx_batch_create: only LSX_F64 systems are batched");
    std::vector<int32_t> hn(count);
    std::vector<int64_t> aoff(count), voff(count);
    int64_t ta = 0, tv = 0;
    double flops = 0.0;
    for (int64_t i = 0; i < count; ++i) {
        if (n[i] < 1 || n[i] > BNMAX)
            return fail(ctx, LSX_EINVAL, "lsx_batch_create: n[%lld] = %lld outside [1, %d]", (long long)i,
                        (long long)n[i], BNMAX);
        hn[i] = (int32_t)n[i];
        aoff[i] = ta;
        voff[i] = tv;
        ta += n[i] * n[i];
        tv += n[i];
        const double nn = (double)n[i];
        flops += nn + nn * (nn - 1) / 2 + (nn - 1) * nn * (2 * nn - 1) / 3 + 2 * nn * nn;
    }
    std::vector<int32_t> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return hn[x] > hn[y]; });
    DeviceGuard g(ctx);
    lsx_batch* B = new (std::nothrow) lsx_batch();
    if (!B) return LSX_ENOMEM;
    B->ctx = ctx;
    B->count = count;
    B->total_a = ta;
    B->total_v = tv;
    B->flops = flops;
    auto bad = [&](int rc) {
        batch_free(B);
        delete B;
        return rc;
    };
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&B->order, sizeof(int32_t) * count) != hipSuccess ||
        hipMalloc(&B->n, sizeof(int32_t) * count) != hipSuccess ||
        hipMalloc(&B->aoff, sizeof(int64_t) * count) != hipSuccess ||
        hipMalloc(&B->voff, sizeof(int64_t) * count) != hipSuccess ||
        hipMalloc(&B->A, sizeof(double) * ta) != hipSuccess || hipMalloc(&B->b, sizeof(double) * tv) != hipSuccess ||
        hipMalloc(&B->x, sizeof(double) * tv) != hipSuccess || hipMalloc(&B->ipiv, sizeof(int64_t) * tv) != hipSuccess ||
        hipMalloc(&B->perm, sizeof(int32_t) * tv) != hipSuccess ||
        hipMalloc(&B->info, sizeof(int64_t) * count) != hipSuccess ||
        hipMalloc(&B->counter, 256) != hipSuccess ||
        hipHostMalloc(&B->hinfo, sizeof(int64_t) * count, hipHostMallocDefault) != hipSuccess)
        return bad(fail(ctx, LSX_ENOMEM, "lsx_batch_create: device allocation failed"));
    if (hipMemcpy(B->order, order.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->n, hn.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->aoff, aoff.data(), sizeof(int64_t) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->voff, voff.data(), sizeof(int64_t) * count, hipMemcpyHostToDevice) != hipSuccess)
        return bad(fail(ctx, LSX_EHIP, "lsx_batch_create: upload failed"));
    *out = B;
    return LSX_OK;
}

int lsx_batch_destroy(lsx_batch* B) {
    if (!B) return LSX_EINVAL;
    DeviceGuard g(B->ctx);
    batch_free(B);
    delete B;
    return LSX_OK;
}

int lsx_batch_factor_solve(lsx_batch* B, const double* A, const double* b, double* x, int64_t* info,
                           int64_t* ipiv) {
    if (!B || !A || !b || !x) return LSX_EINVAL;
    lsx_ctx* ctx = B->ctx;
    DeviceGuard g(ctx);
    LSX_HIP(ctx, hipMemcpyAsync(B->b, b, sizeof(double) * B->total_v, hipMemcpyDefault, B->stream));
    int rc = batch_run(B, A, B->b, B->x, true);
    if (rc) return rc;
    LSX_HIP(ctx, hipMemcpyAsync(x, B->x, sizeof(double) * B->total_v, hipMemcpyDefault, B->stream));
    LSX_HIP(ctx, hipMemcpyAsync(B->hinfo, B->info, sizeof(int64_t) * B->count, hipMemcpyDeviceToHost, B->stream));
    if (info) LSX_HIP(ctx, hipMemcpyAsync(info, B->info, sizeof(int64_t) * B->count, hipMemcpyDefault, B->stream));
    if (ipiv) LSX_HIP(ctx, hipMemcpyAsync(ipiv, B->ipiv, sizeof(int64_t) * B->total_v, hipMemcpyDefault, B->stream));
    LSX_HIP(ctx, hipStreamSynchronize(B->stream));
    B->prof.collect();
    int nfail = 0;
    for (int64_t i = 0; i < B->count; ++i) nfail += B->hinfo[i] > 0;
    return nfail;
}

}  // extern "C"
