// lu_kernels.hpp — LU building blocks templated on the working type (double for the plain
// FP64 path, float for the mixed-precision factorization): panel, row-interchange lists,
// interchange + TRSM, inverses of diagonal blocks, triangular solves.
//
// Pivot search: maximum |a| with NaN ignored, first (smallest) row on ties
// (src/blocked_lufact.jl:38-54); zero pivot -> info = first such column, factorization
// continues (src/generic_lufact.jl:71-141).  Scaling multiplies by the reciprocal of the pivot
// (blocked_lufact.jl:110-113) unless |pivot| < the smallest normal number (then divides, as
// LAPACK dgetf2).
#pragma once
#include <cfloat>
#include <limits>

#include "lu.hpp"

namespace lsx {

// ------------------------------------------------------------------------------------------
// Panel factorization: rows [j0, m) x columns [j0, j0+W), W <= PW.
template <typename T>
struct PanelArgs {
    T* A;
    int64_t ld;
    int64_t m;
    int64_t j0;
    int W;
    int P;
    int64_t* ipiv;
    int64_t* dinfo;
    uint64_t* slots;
    uint32_t seq;
    int64_t timeout;
};

__device__ __forceinline__ bool better(double ka, int64_t ra, double kb, int64_t rb) {
    return ka > kb || (ka == kb && ra < rb);
}

// Words (32-bit) per element and flag-tagged message writes of one element.
template <typename T> struct Words;
template <> struct Words<double> {
    static constexpr int W = 2;
    __device__ static void put(uint64_t* p, double x, uint32_t seq) {
        st_relaxed_u64(p, ll_pack((uint32_t)__double2loint(x), seq));
        st_relaxed_u64(p + 1, ll_pack((uint32_t)__double2hiint(x), seq));
    }
    __device__ static double key(const uint32_t* w) { return __hiloint2double((int)w[1], (int)w[0]); }
};
template <> struct Words<float> {
    static constexpr int W = 1;
    __device__ static void put(uint64_t* p, float x, uint32_t seq) {
        st_relaxed_u64(p, ll_pack((uint32_t)__float_as_uint(x), seq));
    }
    __device__ static double key(const uint32_t* w) { return (double)__uint_as_float(w[0]); }
};
// Exchange record (32-bit words): key at [0, WPE), row at [WPE], row data from word 4.
template <typename T> constexpr int rec_words() { return 4 + PW * Words<T>::W; }

// Each thread owns one row of the panel in registers, in a rotating layout: at column c the
// physical slot p holds logical column c + p (p < PW - c, the not yet eliminated columns) and
// the finished multipliers L[r, 0..c) sit in the last c slots.  Column c is therefore always
// v[0] and every register access is static, with a compact (not unrolled) column loop.
// All rows rotate in lockstep, so row records exchanged between threads share the layout.
template <typename T, int BS>
__global__ __launch_bounds__(BS) void k_panel(PanelArgs<T> a) {
    constexpr int NW = BS / 64;
    constexpr int WPE = Words<T>::W;
    constexpr int RECT = rec_words<T>();
    const int tid = threadIdx.x, w = blockIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r = a.j0 + (int64_t)w * BS + tid;
    const bool valid = r < a.m;
    T* col0 = a.A + a.j0 * a.ld;
    T v[PW];
#pragma unroll
    for (int q = 0; q < PW; ++q) v[q] = (valid && q < a.W) ? col0[q * a.ld + r] : T(0);

    __shared__ double s_wkey[2][NW];
    __shared__ int64_t s_wrow[2][NW];
    __shared__ __align__(16) uint32_t s_rec[2][(PMAX + 1) * RECT];

    const int P = a.P;
#pragma unroll 1
    for (int c = 0; c < a.W; ++c) {
        const int par = c & 1;
        const int64_t k = a.j0 + c;
        double key = -1.0;
        if (valid && r >= k) {
            const double t = fabs((double)v[0]);
            key = (t == t) ? t : 0.0;  // NaN never wins (treated as 0)
        }
        int64_t row = r;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ok = __shfl_xor(key, off);
            const int64_t orow = __shfl_xor((long long)row, off);
            if (better(ok, orow, key, row)) { key = ok; row = orow; }
        }
        if (lane == 0) { s_wkey[par][wave] = key; s_wrow[par][wave] = row; }
        __syncthreads();
        double bkey = s_wkey[par][0];
        int64_t brow = s_wrow[par][0];
#pragma unroll
        for (int i = 1; i < NW; ++i) {
            const double kk = s_wkey[par][i];
            const int64_t rr = s_wrow[par][i];
            if (better(kk, rr, bkey, brow)) { bkey = kk; brow = rr; }
        }
        uint32_t* rec = s_rec[par];
        int pwin = 0;
        int64_t grow = brow;
        if (P == 1) {
            if (valid && r == brow) {
                T* d = reinterpret_cast<T*>(rec + 4);
#pragma unroll
                for (int q = 0; q < PW; ++q) d[q] = v[q];
            }
            if (tid == c) {
                T* d = reinterpret_cast<T*>(rec + RECT + 4);
#pragma unroll
                for (int q = 0; q < PW; ++q) d[q] = v[q];
            }
            __syncthreads();
        } else {
            const uint32_t seq = a.seq + (uint32_t)c + 1u;
            uint64_t* g = a.slots + (size_t)par * (PMAX + 1) * RECS;
            if (valid && r == brow) {
                uint64_t* my = g + (size_t)w * RECS;
                Words<T>::put(my, (T)bkey, seq);
                st_relaxed_u64(my + WPE, ll_pack((uint32_t)brow, seq));
#pragma unroll
                for (int q = 0; q < PW; ++q) Words<T>::put(my + 4 + WPE * q, v[q], seq);
            }
            if (w == 0 && tid == c) {
                uint64_t* kr = g + (size_t)P * RECS;
#pragma unroll
                for (int q = 0; q < PW; ++q) Words<T>::put(kr + 4 + WPE * q, v[q], seq);
            }
            // Every thread reads a share of the words of all P records and of the row-k record:
            // all of its loads are issued at once, then the ones not yet current are re-polled.
            constexpr int RW = WPE + 1 + WPE * PW;  // words read per candidate record
            constexpr int MAXW = (PMAX * RW + 2 * PW + BS - 1) / BS;
            const int total = P * RW + WPE * PW;
            uint64_t x[MAXW];
            auto word = [&](int idx, int& dst) -> const uint64_t* {
                int rr = P, ww = 4 + (idx - P * RW);
                if (idx < P * RW) {
                    rr = idx / RW;
                    const int wi = idx - rr * RW;
                    ww = wi <= WPE ? wi : wi - (WPE + 1) + 4;
                }
                dst = rr * RECT + ww;
                return g + (size_t)rr * RECS + ww;
            };
#pragma unroll
            for (int i = 0; i < MAXW; ++i) {
                const int idx = tid + i * BS;
                int d;
                x[i] = (idx < total) ? ld_relaxed_u64(word(idx, d)) : 0;
            }
#pragma unroll
            for (int i = 0; i < MAXW; ++i) {
                const int idx = tid + i * BS;
                if (idx >= total) continue;
                int d;
                const uint64_t* src = word(idx, d);
                if ((uint32_t)(x[i] >> 32) != seq) {
                    const int64_t t0 = wall_clock64();
                    while (true) {
                        x[i] = ld_relaxed_u64(src);
                        if ((uint32_t)(x[i] >> 32) == seq) break;
                        if (wall_clock64() - t0 > a.timeout) {
                            atomicOr((unsigned long long*)&a.dinfo[1], 1ull);
                            break;
                        }
                    }
                }
                rec[d] = (uint32_t)x[i];
            }
            __syncthreads();
            double gk = Words<T>::key(rec);
            int64_t gr = (int64_t)rec[WPE];
            for (int p = 1; p < P; ++p) {
                const double kk = Words<T>::key(rec + p * RECT);
                const int64_t rr = (int64_t)rec[p * RECT + WPE];
                if (better(kk, rr, gk, gr)) { gk = kk; gr = rr; pwin = p; }
            }
            grow = gr;
        }
        const T* prow = reinterpret_cast<const T*>(rec + pwin * RECT + 4);
        const T* krow = reinterpret_cast<const T*>(rec + P * RECT + 4);
        const T pv = prow[0];
        if (grow != k) {
            // row k takes the pivot row, the pivot row takes row k
            const T* src = (valid && r == k) ? prow : ((valid && r == grow) ? krow : nullptr);
            if (src) {
#pragma unroll
                for (int q = 0; q < PW; ++q) v[q] = src[q];
            }
        }
        if (w == 0 && tid == 0) {
            a.ipiv[k] = grow + 1;
            if (pv == T(0) && a.dinfo[0] == 0) a.dinfo[0] = k + 1;
        }
        // scale, rank-1 update of the remaining columns, rotate by one slot
        T l = v[0];
        const bool upd = valid && r > k;
        if (upd && pv != T(0)) {
            if (fabs(pv) >= std::numeric_limits<T>::min()) l *= (T(1) / pv);
            else l /= pv;
        }
        const int lim = upd ? PW - 1 - c : 0;  // slots 1..lim hold columns still to eliminate
        // in groups of 8 slots so that the pivot-row LDS reads are not all hoisted at once
        // (keeps the 1024-thread variant within 128 VGPRs)
#pragma unroll
        for (int p0 = 0; p0 < PW; p0 += 8) {
#pragma unroll
            for (int p = p0; p < p0 + 8 && p < PW - 1; ++p) {
                const T nv = v[p + 1];
                v[p] = (p < lim) ? fma(-l, prow[p + 1], nv) : nv;
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        v[PW - 1] = l;
    }
    if (valid) {
        // after W rotations slot p holds logical column (p + W) mod PW
#pragma unroll
        for (int p = 0; p < PW; ++p) {
            const int q = (p + a.W) & (PW - 1);
            if (q < a.W) col0[q * a.ld + r] = v[p];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Row permutation of the LAPACK swap sequence ipiv[k0 .. k0+kb) (ipiv[k] >= k, 1-based) as a
// list of (destination row, source row) pairs: after the swaps, row dest[i] holds the former
// row src[i].  Optionally applies the same permutation to perm[] (row i of P*A = row perm[i]).
// One workgroup of 256 threads; kb <= 128.
static __global__ __launch_bounds__(256) void k_perm_build(const int64_t* __restrict ipiv, int64_t k0, int kb,
                                                    int64_t* __restrict list, int* __restrict count,
                                                    int* __restrict perm) {
    __shared__ int64_t s_p[NBO];
    __shared__ int s_first[NBO];
    __shared__ int s_pos[NBO];
    __shared__ int s_slot[NBO];
    __shared__ int64_t s_rows[2 * NBO];
    __shared__ int s_lab[2 * NBO];
    __shared__ int s_T;
    __shared__ int s_pv[2 * NBO];
    const int t = threadIdx.x;
    const int64_t top = k0 + kb;
    if (t < kb) s_p[t] = ipiv[k0 + t] - 1;
    __syncthreads();
    int fidx = -1;
    if (t < kb) {
        const int64_t p = s_p[t];
        if (p >= top) {
            fidx = t;
            for (int u = 0; u < t; ++u)
                if (s_p[u] == p) { fidx = u; break; }
        }
        s_first[t] = (fidx == t) ? 1 : 0;
    }
    __syncthreads();
    if (t == 0) {
        int cnt = 0;
        for (int u = 0; u < kb; ++u) {
            s_slot[u] = s_first[u] ? (kb + cnt) : -1;
            if (s_first[u]) s_rows[kb + cnt] = s_p[u];
            cnt += s_first[u];
        }
        s_T = kb + cnt;
    }
    if (t < kb) s_rows[t] = k0 + t;
    __syncthreads();
    if (t < kb) s_pos[t] = (fidx < 0) ? (int)(s_p[t] - k0) : s_slot[fidx];
    const int T = s_T;
    if (t < T) s_lab[t] = t;
    __syncthreads();
    if (t == 0) {
        for (int u = 0; u < kb; ++u) {
            const int q = s_pos[u];
            const int x = s_lab[u];
            s_lab[u] = s_lab[q];
            s_lab[q] = x;
        }
    }
    __syncthreads();
    if (t < T) {
        list[t] = s_rows[t];
        list[2 * NBO + t] = s_rows[s_lab[t]];
    }
    if (t == 0) *count = T;
    if (perm) {
        if (t < T) s_pv[t] = perm[s_rows[s_lab[t]]];
        __syncthreads();
        if (t < T) perm[s_rows[t]] = s_pv[t];
    }
}

static __global__ void k_iota(int* __restrict p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int)i;
}

// ------------------------------------------------------------------------------------------
// Applies a permutation list to columns [c0, c1) and, if trsm (kb <= PW), replaces the top kb
// rows (k0 .. k0+kb) of those columns by L11^{-1} times them, L11 = unit lower of A[k0:, k0:]
// staged in LDS (blocked_lufact.jl:126-178).  Each workgroup handles LCW columns.
constexpr int LCW = 32;
template <typename T>
__global__ __launch_bounds__(256) void k_laswp_trsm(T* __restrict A, int64_t ld, int64_t k0, int kb,
                                                    const int64_t* __restrict list,
                                                    const int* __restrict count, int64_t c0,
                                                    int64_t c1, int trsm) {
    __shared__ int64_t s_dst[2 * NBO];
    __shared__ int64_t s_src[2 * NBO];
    __shared__ T s_buf[2 * NBO][LCW + 1];
    __shared__ T s_L[PW][PW + 1];
    const int t = threadIdx.x;
    const int NT = *count;
    for (int i = t; i < NT; i += 256) {
        s_dst[i] = list[i];
        s_src[i] = list[2 * NBO + i];
    }
    if (trsm)
        for (int idx = t; idx < kb * kb; idx += 256) {
            const int kk = idx / kb, i = idx - kk * kb;
            s_L[i][kk] = A[(k0 + kk) * ld + k0 + i];
        }
    __syncthreads();
    const int64_t cb = c0 + (int64_t)blockIdx.x * LCW;
    const int ncol = (int)min((int64_t)LCW, c1 - cb);
    // gather: consecutive threads walk consecutive list entries of one column
    for (int idx = t; idx < NT * LCW; idx += 256) {
        const int cc = idx / NT, i = idx - cc * NT;
        if (cc < ncol) s_buf[i][cc] = A[(cb + cc) * ld + s_src[i]];
    }
    __syncthreads();
    const int first = trsm ? kb : 0;
    for (int idx = t; idx < (NT - first) * LCW; idx += 256) {
        const int cc = idx / (NT - first), i = first + idx - cc * (NT - first);
        if (cc < ncol) A[(cb + cc) * ld + s_dst[i]] = s_buf[i][cc];
    }
    if (!trsm) return;
    // forward substitution with the unit-lower kb x kb block, one thread per column
    if (t < LCW) {
        for (int kk = 0; kk < kb - 1; ++kk) {
            const T xk = s_buf[kk][t];
            for (int i = kk + 1; i < kb; ++i) s_buf[i][t] = fma(-s_L[i][kk], xk, s_buf[i][t]);
        }
    }
    __syncthreads();
    for (int idx = t; idx < kb * LCW; idx += 256) {
        const int c2 = idx / kb, i = idx - c2 * kb;
        if (c2 < ncol) A[(cb + c2) * ld + k0 + i] = s_buf[i][c2];
    }
}

// ------------------------------------------------------------------------------------------
// Inverses of the 64 x 64 diagonal blocks of L (unit lower) and U (upper) of the factored
// matrix, used by the blocked TRSM of the outer update and by the triangular solves.
// grid = (number of blocks, 2): y = 0 inverts L blocks, y = 1 U blocks.  One thread per
// column of the inverse (columns are independent), the block staged in LDS.
template <typename T>
__global__ __launch_bounds__(64) void k_diag_inv(const T* __restrict F, int64_t ld, int64_t n,
                                                 int64_t b0, T* __restrict DL, T* __restrict DU) {
    const int64_t blk = b0 + blockIdx.x;
    const int64_t r0 = blk * DB;
    const int nb = (int)min((int64_t)DB, n - r0);
    const int c = threadIdx.x;
    __shared__ T D[DB][DB + 1];   // D[i][k]
    __shared__ T X[DB][DB + 1];   // X[i][c]
    for (int k = 0; k < DB; ++k) D[c][k] = (c < nb && k < nb) ? F[(r0 + k) * ld + r0 + c] : (c == k ? T(1) : T(0));
    __syncthreads();
    if (blockIdx.y == 0) {
        // column c of L^{-1}: x_c = 1, x_i = -sum_{k=c}^{i-1} L[i,k] x_k (i > c)
        for (int i = 0; i < DB; ++i) X[i][c] = (i == c) ? T(1) : T(0);
        for (int i = c + 1; i < DB; ++i) {
            T acc = T(0);
            for (int k = c; k < i; ++k) acc = fma(D[i][k], X[k][c], acc);
            X[i][c] = -acc;
        }
    } else {
        // column c of U^{-1}: x_c = 1/U[c,c], x_i = -(sum_{k=i+1}^{c} U[i,k] x_k) / U[i,i] (i < c)
        for (int i = 0; i < DB; ++i) X[i][c] = T(0);
        X[c][c] = T(1) / D[c][c];
        for (int i = c - 1; i >= 0; --i) {
            T acc = T(0);
            for (int k = i + 1; k <= c; ++k) acc = fma(D[i][k], X[k][c], acc);
            X[i][c] = -acc / D[i][i];
        }
    }
    __syncthreads();
    T* out = (blockIdx.y == 0 ? DL : DU) + blk * DB * DB;
    for (int i = 0; i < DB; ++i) out[c * DB + i] = X[i][c];  // column-major 64 x 64
}

// ------------------------------------------------------------------------------------------
// Triangular solves (getrs), one workgroup per DB-row block, blocks chained by flags.
//   LOWER: Out = L^{-1} (P In)  (unit lower; the permutation gathers In by perm)
//   UPPER: Out = U^{-1} In
// Each block accumulates the contributions of the already solved blocks as they are published
// (four k-quarters of partial sums per row), then applies the inverse of its diagonal block.
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void k_trsv(const T* __restrict F, int64_t ld, int64_t n,
                                              const int* __restrict perm, const T* __restrict Dinv,
                                              const double* In, int64_t ldin, double* Out, int64_t ldout,
                                              int nrhs, uint32_t* flags, uint32_t seq, int64_t timeout,
                                              int64_t* err) {
    const int nblk = (int)((n + DB - 1) / DB);
    const int bw = LOWER ? (int)blockIdx.x : nblk - 1 - (int)blockIdx.x;
    const int64_t r0 = (int64_t)bw * DB;
    const int tid = threadIdx.x, rr = tid & 63, g = tid >> 6;
    const int64_t row = r0 + rr;
    __shared__ double Ds[DB][DB + 1];
    __shared__ double Ys[DB][TRHS + 1];
    __shared__ double Ps[4][DB][TRHS + 1];
    for (int idx = tid; idx < DB * DB; idx += 256) {
        const int i = idx & (DB - 1), k = idx >> 6;
        Ds[i][k] = (double)Dinv[(int64_t)bw * DB * DB + k * DB + i];
    }
    double acc[TRHS];
#pragma unroll
    for (int s = 0; s < TRHS; ++s) acc[s] = 0.0;
    const int jbeg = LOWER ? 0 : nblk - 1, jend = LOWER ? bw : bw, jstep = LOWER ? 1 : -1;
    for (int j = jbeg; j != jend; j += jstep) {
        double lv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int64_t col = (int64_t)j * DB + 16 * g + q;
            lv[q] = (row < n && col < n) ? (double)F[col * ld + row] : 0.0;
        }
        if (tid == 0 && ld_relaxed_u32(&flags[j]) != seq) {
            const int64_t t0 = wall_clock64();
            while (ld_relaxed_u32(&flags[j]) != seq) {
                if (wall_clock64() - t0 > timeout) { atomicOr((unsigned long long*)err, 2ull); break; }
            }
        }
        __syncthreads();
        for (int idx = tid; idx < DB * TRHS; idx += 256) {
            const int kk = idx & (DB - 1), s = idx >> 6;
            const int64_t gr = (int64_t)j * DB + kk;
            Ys[kk][s] = (s < nrhs && gr < n) ? ld_coherent_f64(Out + (int64_t)s * ldout + gr) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q)
#pragma unroll
            for (int s = 0; s < TRHS; ++s) acc[s] = fma(lv[q], Ys[16 * g + q][s], acc[s]);
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < TRHS; ++s) Ps[g][rr][s] = acc[s];
    __syncthreads();
    // right-hand side minus contributions, one (row, rhs) per thread-slot
    for (int idx = tid; idx < DB * TRHS; idx += 256) {
        const int i = idx & (DB - 1), s = idx >> 6;
        const int64_t gr = r0 + i;
        double b = 0.0;
        if (gr < n && s < nrhs) b = LOWER ? In[(int64_t)s * ldin + perm[gr]] : In[(int64_t)s * ldin + gr];
        Ys[i][s] = b - (Ps[0][i][s] + Ps[1][i][s] + Ps[2][i][s] + Ps[3][i][s]);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int s = g + 4 * q;
        double y = 0.0;
#pragma unroll 8
        for (int k = 0; k < DB; ++k) y = fma(Ds[rr][k], Ys[k][s], y);
        if (row < n && s < nrhs) Out[(int64_t)s * ldout + row] = y;
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&flags[bw], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace lsx
