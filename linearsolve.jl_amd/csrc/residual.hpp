// residual.hpp — FP64 residuals R = B - A X (and products A X, row sums of |A|) of a dense
// column-major A, with per-workgroup column statistics, for the residual-safety check and the
// iterative refinement of the mixed-precision LU.
//
// HBM-bound: A is read once per call (8 n^2 bytes for up to RNR right-hand sides).  One
// workgroup of 16 waves per 64 rows; wave w sums the columns of its sixteenth of A for the
// workgroup's rows (one row per lane, 8 column loads in flight), and the sixteen partial sums of a
// row are added in wave order, so every result is reproducible bit for bit.
#pragma once
#include "common.hpp"

namespace lsx {

constexpr int RNR = 16;   // right-hand sides per launch
constexpr int RWAVES = 16;
constexpr int RROWS = 64;
// statistics of one workgroup's rows, per right-hand side s:
//   [0] max |r|, [1] sum r^2, [2] max |x|, [3] max |b|, [4] sum b^2
constexpr int RSTAT = 5;

enum { RES_RESIDUAL = 0, RES_ROWSUM = 1, RES_PRODUCT = 2 };

// max that propagates NaN (a NaN residual must not read as a small one)
__host__ __device__ inline double nanmax(double a, double b) { return (a > b || a != a) ? a : b; }

// mode RES_RESIDUAL: R = B - A X (R may be null: statistics only), statistics into
//                    part[blk][RSTAT][RNR];
// mode RES_PRODUCT:  R = A X;
// mode RES_ROWSUM:   part[blk] = max over the block's rows of sum_k |A[i, k]| (X, B, R unused).
template <int NR>
__global__ __launch_bounds__(1024) void k_resid(const double* __restrict A, int64_t lda, int64_t n,
                                                const double* __restrict X, int64_t ldx, int nrhs,
                                                const double* B, int64_t ldb, double* R,
                                                int64_t ldr, double* __restrict part, int mode) {
    // B and R may alias (each element is read, then written, by the same thread)
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t row = (int64_t)blockIdx.x * RROWS + lane;
    const bool rv = row < n;
    const int64_t chunk = (n + RWAVES - 1) / RWAVES;
    const int64_t c0 = (int64_t)w * chunk, c1 = c0 + chunk < n ? c0 + chunk : n;
    double acc[NR];
#pragma unroll
    for (int s = 0; s < NR; ++s) acc[s] = 0.0;
    const double* ar = A + (rv ? row : 0);
    if (mode == RES_ROWSUM) {
        for (int64_t j = c0; j < c1; j += 8) {
            double a[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = (rv && j + u < c1) ? ar[(j + u) * lda] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[0] += fabs(a[u]);
        }
    } else {
        for (int64_t j = c0; j < c1; j += 8) {
            double a[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = (rv && j + u < c1) ? ar[(j + u) * lda] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (j + u < c1) {
#pragma unroll
                    for (int s = 0; s < NR; ++s)
                        if (s < nrhs) acc[s] = fma(a[u], X[(int64_t)s * ldx + j + u], acc[s]);
                }
            }
        }
    }
    __shared__ double Ps[RWAVES][RROWS];
    __shared__ double Ax[RROWS][NR + 1];
    const int ns = mode == RES_ROWSUM ? 1 : NR;
    for (int s = 0; s < ns; ++s) {
        Ps[w][lane] = acc[s];
        __syncthreads();
        if (w == 0) {
            double t = 0.0;
#pragma unroll
            for (int q = 0; q < RWAVES; ++q) t += Ps[q][lane];
            Ax[lane][s] = t;
        }
        __syncthreads();
    }
    if (mode == RES_ROWSUM) {
        if (tid == 0) {
            double m = 0.0;
            for (int i = 0; i < RROWS; ++i)
                if ((int64_t)blockIdx.x * RROWS + i < n) m = nanmax(m, Ax[i][0]);
            part[blockIdx.x] = m;
        }
        return;
    }
    // one (row, right-hand side) per thread: 64 rows x 16 slots = 1024 threads
    const int s = w;
    double r = 0.0, x = 0.0, b = 0.0;
    const bool act = rv && s < nrhs && s < NR;
    if (act) {
        const double ax = Ax[lane][s];
        if (mode == RES_PRODUCT) {
            r = ax;
        } else {
            b = B[(int64_t)s * ldb + row];
            r = b - ax;
            x = X[(int64_t)s * ldx + row];
        }
        if (R) R[(int64_t)s * ldr + row] = r;
    }
    if (mode == RES_PRODUCT) return;
    // statistics of the block per right-hand side: wave s holds column s, one row per lane
    double mr = fabs(r), sr = r * r, mx = fabs(x), mb = fabs(b), sb = b * b;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mr = nanmax(mr, __shfl_xor(mr, off));
        mx = nanmax(mx, __shfl_xor(mx, off));
        mb = nanmax(mb, __shfl_xor(mb, off));
        sr += __shfl_xor(sr, off);
        sb += __shfl_xor(sb, off);
    }
    if (lane == 0 && s < RNR) {
        double* p = part + (int64_t)blockIdx.x * RSTAT * RNR;
        p[0 * RNR + s] = mr;
        p[1 * RNR + s] = sr;
        p[2 * RNR + s] = mx;
        p[3 * RNR + s] = mb;
        p[4 * RNR + s] = sb;
    }
}

constexpr int64_t resid_blocks(int64_t n) { return (n + RROWS - 1) / RROWS; }
// doubles of the statistics area of one launch
constexpr int64_t resid_part_words(int64_t n) { return resid_blocks(n) * RSTAT * RNR; }

// One launch over nrhs <= RNR right-hand sides (mode as k_resid).
inline void resid_launch(hipStream_t st, const double* A, int64_t lda, int64_t n, const double* X, int64_t ldx,
                         int nrhs, const double* B, int64_t ldb, double* R, int64_t ldr, double* part, int mode) {
    const dim3 grid((unsigned)resid_blocks(n));
    if (mode == RES_ROWSUM || nrhs == 1)
        hipLaunchKernelGGL(k_resid<1>, grid, dim3(1024), 0, st, A, lda, n, X, ldx, nrhs, B, ldb, R, ldr, part, mode);
    else
        hipLaunchKernelGGL(k_resid<RNR>, grid, dim3(1024), 0, st, A, lda, n, X, ldx, nrhs, B, ldb, R, ldr, part, mode);
}

// Column statistics of a launch, combined on the host in block order: for s < nrhs,
// out[s] = {max |r|, sum r^2, max |x|, max |b|, sum b^2}.
inline void resid_combine(const double* hpart, int64_t n, int nrhs, double (*out)[RSTAT]) {
    const int64_t nb = resid_blocks(n);
    for (int s = 0; s < nrhs; ++s) {
        double v[RSTAT] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int64_t b = 0; b < nb; ++b) {
            const double* p = hpart + b * RSTAT * RNR;
            v[0] = nanmax(v[0], p[0 * RNR + s]);
            v[1] += p[1 * RNR + s];
            v[2] = nanmax(v[2], p[2 * RNR + s]);
            v[3] = nanmax(v[3], p[3 * RNR + s]);
            v[4] += p[4 * RNR + s];
        }
        for (int k = 0; k < RSTAT; ++k) out[s][k] = v[k];
    }
}

// The two sums of squares of a launch's statistics (||R||^2 and ||B||^2 over its nrhs columns),
// added to acc[0], acc[1] (first != 0: acc starts at zero), in the order resid_combine and the
// host loops over columns use, so the device and host totals agree bit for bit.  One workgroup.
static __global__ __launch_bounds__(64) void k_resid_sumsq(const double* __restrict part, int64_t nb, int nrhs,
                                                    double* __restrict acc, int first) {
    __shared__ double t[2][RNR];
    const int s = threadIdx.x;
    if (s < nrhs && s < RNR) {
        double sr = 0.0, sb = 0.0;
        for (int64_t b = 0; b < nb; ++b) {
            const double* p = part + b * RSTAT * RNR;
            sr += p[1 * RNR + s];
            sb += p[4 * RNR + s];
        }
        t[0][s] = sr;
        t[1][s] = sb;
    }
    __syncthreads();
    if (s == 0) {
        double a0 = first ? 0.0 : acc[0], a1 = first ? 0.0 : acc[1];
        for (int k = 0; k < nrhs && k < RNR; ++k) {
            a0 += t[0][k];
            a1 += t[1][k];
        }
        acc[0] = a0;
        acc[1] = a1;
    }
}

}  // namespace lsx
