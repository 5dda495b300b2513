// lu64.hip — FP64 partial-pivot LU (getrf) and triangular solves (getrs) for gfx950.
//
// Algorithm: right-looking blocked LU with two block levels, the structure of LAPACK dgetrf
// and of the reference's own restatement src/blocked_lufact.jl:658-679:
//   outer block NBO = 128 columns (the K of the trailing MFMA update, which holds ~2/3 n^3 flops);
//   inside it, inner panels of PW = 32 columns are factored by a register-resident kernel whose
//   workgroups each hold 1024 rows x 32 columns, exchanging one pivot candidate record per
//   column through flag-in-word messages (no fences, one memory round trip per column).
// Pivot search: maximum |a| with NaN ignored, first (smallest) row on ties
// (src/blocked_lufact.jl:38-54); zero pivot -> info = first such column, factorization
// continues (src/generic_lufact.jl:71-141).  Scaling multiplies by the reciprocal of the pivot
// (blocked_lufact.jl:110-113) unless |pivot| < DBL_MIN (then divides, as LAPACK dgetf2).
#include <cfloat>
#include <utility>
#include <type_traits>

#include "lu.hpp"

namespace lsx {

typedef double v4d __attribute__((ext_vector_type(4)));

template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// ------------------------------------------------------------------------------------------
// Panel factorization: rows [j0, m) x columns [j0, j0+W), W <= PW.
struct PanelArgs {
    double* A;
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

template <int BS>
__global__ __launch_bounds__(BS) void k_panel(PanelArgs a) {
    constexpr int NW = BS / 64;
    const int tid = threadIdx.x, w = blockIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r = a.j0 + (int64_t)w * BS + tid;
    const bool valid = r < a.m;
    double* col0 = a.A + a.j0 * a.ld;
    double v[PW];
#pragma unroll
    for (int q = 0; q < PW; ++q) v[q] = (valid && q < a.W) ? col0[q * a.ld + r] : 0.0;

    __shared__ double s_wkey[2][NW];
    __shared__ int64_t s_wrow[2][NW];
    __shared__ __align__(16) uint32_t s_rec[2][(PMAX + 1) * REC];

    const int P = a.P;
    // Columns are unrolled at compile time so v[] stays in registers.
    static_for<PW>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        if (c >= a.W) return;
        const int par = c & 1;
        const int64_t k = a.j0 + c;
        double key = -1.0;
        if (valid && r >= k) {
            const double t = fabs(v[c]);
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
        if (P == 1) {
            if (valid && r == brow) {
                rec[0] = __double2loint(bkey);
                rec[1] = __double2hiint(bkey);
                rec[2] = (uint32_t)brow;
                double* d = reinterpret_cast<double*>(rec + 4);
#pragma unroll
                for (int q = 0; q < PW; ++q) d[q] = v[q];
            }
            if (tid == c) {
                double* d = reinterpret_cast<double*>(rec + REC + 4);
#pragma unroll
                for (int q = 0; q < PW; ++q) d[q] = v[q];
            }
            __syncthreads();
        } else {
            const uint32_t seq = a.seq + (uint32_t)c + 1u;
            uint64_t* g = a.slots + (size_t)par * (PMAX + 1) * RECS;
            if (valid && r == brow) {
                uint64_t* my = g + (size_t)w * RECS;
                st_relaxed_u64(my + 0, ll_pack((uint32_t)__double2loint(bkey), seq));
                st_relaxed_u64(my + 1, ll_pack((uint32_t)__double2hiint(bkey), seq));
                st_relaxed_u64(my + 2, ll_pack((uint32_t)brow, seq));
#pragma unroll
                for (int q = 0; q < PW; ++q) {
                    st_relaxed_u64(my + 4 + 2 * q, ll_pack((uint32_t)__double2loint(v[q]), seq));
                    st_relaxed_u64(my + 5 + 2 * q, ll_pack((uint32_t)__double2hiint(v[q]), seq));
                }
            }
            if (w == 0 && tid == c) {
                uint64_t* kr = g + (size_t)P * RECS;
#pragma unroll
                for (int q = 0; q < PW; ++q) {
                    st_relaxed_u64(kr + 4 + 2 * q, ll_pack((uint32_t)__double2loint(v[q]), seq));
                    st_relaxed_u64(kr + 5 + 2 * q, ll_pack((uint32_t)__double2hiint(v[q]), seq));
                }
            }
            // Every thread polls a share of the words of all P records and of the row-k record.
            constexpr int RW = 3 + 2 * PW;  // words read per candidate record
            const int total = P * RW + 2 * PW;
            for (int idx = tid; idx < total; idx += BS) {
                int rr, ww;
                if (idx < P * RW) {
                    rr = idx / RW;
                    const int wi = idx - rr * RW;
                    ww = wi < 3 ? wi : wi + 1;
                } else {
                    rr = P;
                    ww = 4 + (idx - P * RW);
                }
                const uint64_t* src = g + (size_t)rr * RECS + ww;
                uint64_t x = ld_relaxed_u64(src);
                if ((uint32_t)(x >> 32) != seq) {
                    const int64_t t0 = wall_clock64();
                    while (true) {
                        x = ld_relaxed_u64(src);
                        if ((uint32_t)(x >> 32) == seq) break;
                        if (wall_clock64() - t0 > a.timeout) {
                            atomicOr((unsigned long long*)&a.dinfo[1], 1ull);
                            break;
                        }
                    }
                }
                rec[rr * REC + ww] = (uint32_t)x;
            }
            __syncthreads();
            double gk = __hiloint2double((int)rec[1], (int)rec[0]);
            int64_t gr = (int64_t)rec[2];
            for (int p = 1; p < P; ++p) {
                const double kk = __hiloint2double((int)rec[p * REC + 1], (int)rec[p * REC]);
                const int64_t rr = (int64_t)rec[p * REC + 2];
                if (better(kk, rr, gk, gr)) { gk = kk; gr = rr; pwin = p; }
            }
            brow = gr;
        }
        const int64_t grow = (P == 1) ? (int64_t)rec[2] : brow;
        const double* prow = reinterpret_cast<const double*>(rec + pwin * REC + 4);
        const double* krow = reinterpret_cast<const double*>(rec + P * REC + 4);
        const double pv = prow[c];
        if (grow != k) {
            if (valid && r == k) {
#pragma unroll
                for (int q = 0; q < PW; ++q) v[q] = prow[q];
            } else if (valid && r == grow) {
#pragma unroll
                for (int q = 0; q < PW; ++q) v[q] = krow[q];
            }
        }
        if (w == 0 && tid == 0) {
            a.ipiv[k] = grow + 1;
            if (pv == 0.0 && a.dinfo[0] == 0) a.dinfo[0] = k + 1;
        }
        if (valid && r > k) {
            if (pv != 0.0) {
                if (fabs(pv) >= DBL_MIN) v[c] *= (1.0 / pv);
                else v[c] /= pv;
            }
            const double l = v[c];
#pragma unroll
            for (int q = c + 1; q < PW; ++q) v[q] = fma(-l, prow[q], v[q]);
        }
    });
    if (valid) {
#pragma unroll
        for (int q = 0; q < PW; ++q)
            if (q < a.W) col0[q * a.ld + r] = v[q];
    }
}

// ------------------------------------------------------------------------------------------
// Row permutation of the LAPACK swap sequence ipiv[k0 .. k0+kb) (ipiv[k] >= k, 1-based) as a
// list of (destination row, source row) pairs: after the swaps, row dest[i] holds the former
// row src[i].  Optionally applies the same permutation to perm[] (row i of P*A = row perm[i]).
// One workgroup of 256 threads; kb <= 128.
__global__ __launch_bounds__(256) void k_perm_build(const int64_t* __restrict ipiv, int64_t k0, int kb,
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

__global__ void k_iota(int* __restrict p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int)i;
}

// ------------------------------------------------------------------------------------------
// Applies a permutation list to columns [c0, c1) and, if trsm, replaces the top kb rows
// (k0 .. k0+kb) of those columns by L11^{-1} times them, L11 = unit lower of A[k0:, k0:]
// (blocked_lufact.jl:126-178).  Each workgroup handles LCW columns.
constexpr int LCW = 32;
__global__ __launch_bounds__(256) void k_laswp_trsm(double* __restrict A, int64_t ld, int64_t k0, int kb,
                                                    const int64_t* __restrict list,
                                                    const int* __restrict count, int64_t c0,
                                                    int64_t c1, int trsm) {
    __shared__ int64_t s_dst[2 * NBO];
    __shared__ int64_t s_src[2 * NBO];
    __shared__ double s_buf[2 * NBO][LCW + 1];
    const int t = threadIdx.x;
    const int T = *count;
    for (int i = t; i < T; i += 256) {
        s_dst[i] = list[i];
        s_src[i] = list[2 * NBO + i];
    }
    __syncthreads();
    const int64_t cb = c0 + (int64_t)blockIdx.x * LCW;
    const int ncol = (int)min((int64_t)LCW, c1 - cb);
    // gather: consecutive threads walk consecutive list entries of one column
    for (int idx = t; idx < T * LCW; idx += 256) {
        const int cc = idx / T, i = idx - cc * T;
        if (cc < ncol) s_buf[i][cc] = A[(cb + cc) * ld + s_src[i]];
    }
    __syncthreads();
    const int first = trsm ? kb : 0;
    for (int idx = t; idx < (T - first) * LCW; idx += 256) {
        const int cc = idx / (T - first), i = first + idx - cc * (T - first);
        if (cc < ncol) A[(cb + cc) * ld + s_dst[i]] = s_buf[i][cc];
    }
    if (!trsm) return;
    // forward substitution with the unit-lower kb x kb block, columns in parallel
    const int cc = t % LCW, g = t / LCW;  // 8 row groups
    for (int kk = 0; kk < kb - 1; ++kk) {
        __syncthreads();
        const double xk = s_buf[kk][cc];
        for (int i = kk + 1 + g; i < kb; i += 256 / LCW)
            s_buf[i][cc] = fma(-A[(k0 + kk) * ld + k0 + i], xk, s_buf[i][cc]);
    }
    __syncthreads();
    for (int idx = t; idx < kb * LCW; idx += 256) {
        const int c2 = idx / kb, i = idx - c2 * kb;
        if (c2 < ncol) A[(cb + c2) * ld + k0 + i] = s_buf[i][c2];
    }
}

// ------------------------------------------------------------------------------------------
// Trailing update C -= A * B on FP64 MFMA (v_mfma_f64_16x16x4f64), LDS-staged.
// A: M x K, B: K x N, C: M x N, all column-major sharing ld.  128 x 128 tile per workgroup,
// four waves of 64 x 64, K staged 16 at a time with double buffering.  The MFMA computes the
// transposed tile (B^T A^T) so each lane's results run down a column of C (128-byte stores).
constexpr int GBM = 128, GBN = 128, GBK = 16;
__global__ __launch_bounds__(256) void k_gemm_f64(const double* __restrict Ag,
                                                  const double* __restrict Bg, double* __restrict Cg,
                                                  int64_t ld, int M, int N, int K) {
    const int TM = (M + GBM - 1) / GBM, TN = (N + GBN - 1) / GBN, T = TM * TN;
    // XCD-aware: consecutive tiles of the column-major tile order land on one XCD
    const int per = (T + 7) / 8;
    const int b = blockIdx.x;
    const int tile = (b & 7) * per + (b >> 3);
    if (tile >= T) return;
    const int tm = tile % TM, tn = tile / TM;
    const int row0 = tm * GBM, col0 = tn * GBN;

    __shared__ double As[2][GBK][GBM + 1];
    __shared__ double Bs[2][GBK][GBN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    v4d acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    double ra[8], rb[8];
    auto gload = [&](int k0) {
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int idx = it * 256 + tid;
            const int r = idx & (GBM - 1), k = idx >> 7;
            const int gr = row0 + r, gk = k0 + k;
            ra[it] = (gr < M && gk < K) ? Ag[(int64_t)gk * ld + gr] : 0.0;
            const int kb = idx & (GBK - 1), c = idx >> 4;
            const int gc = col0 + c, gk2 = k0 + kb;
            rb[it] = (gc < N && gk2 < K) ? Bg[(int64_t)gc * ld + gk2] : 0.0;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int idx = it * 256 + tid;
            As[buf][idx >> 7][idx & (GBM - 1)] = ra[it];
            Bs[buf][idx & (GBK - 1)][idx >> 4] = rb[it];
        }
    };
    const int nk = (K + GBK - 1) / GBK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int s = 0; s < nk; ++s) {
        const int buf = s & 1;
        if (s + 1 < nk) gload((s + 1) * GBK);
#pragma unroll
        for (int k4 = 0; k4 < GBK / 4; ++k4) {
            const int kr = k4 * 4 + (lane >> 4);
            double x[4], y[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) x[i] = Bs[buf][kr][wn * 64 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 4; ++j) y[j] = As[buf][kr][wm * 64 + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[i], y[j], acc[i][j], 0, 0, 0);
        }
        if (s + 1 < nk) {
            sstore(buf ^ 1);
            __syncthreads();
        }
    }
    // lane holds C(row = rbase + (lane & 15), col = cbase + 4q + (lane >> 4))
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = row0 + wm * 64 + j * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int col = col0 + wn * 64 + i * 16 + 4 * q + (lane >> 4);
                if (row < M && col < N) {
                    double* cp = Cg + (int64_t)col * ld + row;
                    *cp = *cp - acc[i][j][q];
                }
            }
        }
}

// ------------------------------------------------------------------------------------------
// Triangular solves (getrs), one workgroup per TRB-row block, blocks chained by flags.
// Lower: Y = L^{-1} (P B) with unit-lower L (factorization.jl:433-457 order of operations).
__global__ __launch_bounds__(256) void k_trsv_lower(const double* __restrict F, int64_t ld, int64_t n,
                                                    const int* __restrict perm,
                                                    const double* __restrict B, int64_t ldb,
                                                    double* __restrict Y, int64_t ldy, int nrhs,
                                                    uint32_t* flags, uint32_t seq, int64_t timeout,
                                                    int64_t* err) {
    const int w = blockIdx.x;
    const int64_t r0 = (int64_t)w * TRB;
    const int tid = threadIdx.x, rr = tid & 63, g = tid >> 6;
    const int64_t row = r0 + rr;
    __shared__ double Ys[TRB][TRHS + 1];
    __shared__ double Xs[TRB][TRHS + 1];
    double acc[4];
    const int pr = (row < n) ? perm[row] : 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int s = g + 4 * q;
        acc[q] = (row < n && s < nrhs) ? B[(int64_t)s * ldb + pr] : 0.0;
    }
    for (int j = 0; j < w; ++j) {
        if (tid == 0) {
            if (ld_relaxed_u32(&flags[j]) != seq) {
                const int64_t t0 = wall_clock64();
                while (ld_relaxed_u32(&flags[j]) != seq) {
                    if (wall_clock64() - t0 > timeout) { atomicOr((unsigned long long*)err, 2ull); break; }
                }
            }
        }
        __syncthreads();
        for (int idx = tid; idx < TRB * TRHS; idx += 256) {
            const int kk = idx & (TRB - 1), s = idx >> 6;
            const int64_t gr = (int64_t)j * TRB + kk;
            Ys[kk][s] = (s < nrhs && gr < n) ? ld_coherent_f64(Y + (int64_t)s * ldy + gr) : 0.0;
        }
        __syncthreads();
        if (row < n) {
            const double* lp = F + ((int64_t)j * TRB) * ld + row;
#pragma unroll 8
            for (int kk = 0; kk < TRB; ++kk) {
                const double l = lp[(int64_t)kk * ld];
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] = fma(-l, Ys[kk][g + 4 * q], acc[q]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) Xs[rr][g + 4 * q] = acc[q];
    const int nb = (int)min((int64_t)TRB, n - r0);
    for (int kk = 0; kk < nb - 1; ++kk) {
        __syncthreads();
        if (rr > kk && rr < nb) {
            const double l = F[(r0 + kk) * ld + row];
#pragma unroll
            for (int q = 0; q < 4; ++q) Xs[rr][g + 4 * q] = fma(-l, Xs[kk][g + 4 * q], Xs[rr][g + 4 * q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int s = g + 4 * q;
        if (row < n && s < nrhs) Y[(int64_t)s * ldy + row] = Xs[rr][s];
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&flags[w], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Upper: X = U^{-1} Y, blocks processed from the bottom (workgroup w owns block nblk-1-w).
// X may alias Y: a block's input rows are read only by the workgroup that overwrites them.
__global__ __launch_bounds__(256) void k_trsv_upper(const double* __restrict F, int64_t ld, int64_t n,
                                                    const double* Y, int64_t ldy, double* X, int64_t ldx,
                                                    int nrhs, uint32_t* flags, uint32_t seq,
                                                    int64_t timeout, int64_t* err) {
    const int nblk = (int)((n + TRB - 1) / TRB);
    const int bw = nblk - 1 - (int)blockIdx.x;
    const int64_t r0 = (int64_t)bw * TRB;
    const int tid = threadIdx.x, rr = tid & 63, g = tid >> 6;
    const int64_t row = r0 + rr;
    __shared__ double Ys[TRB][TRHS + 1];
    __shared__ double Xs[TRB][TRHS + 1];
    double acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int s = g + 4 * q;
        acc[q] = (row < n && s < nrhs) ? Y[(int64_t)s * ldy + row] : 0.0;
    }
    for (int j = nblk - 1; j > bw; --j) {
        if (tid == 0) {
            if (ld_relaxed_u32(&flags[j]) != seq) {
                const int64_t t0 = wall_clock64();
                while (ld_relaxed_u32(&flags[j]) != seq) {
                    if (wall_clock64() - t0 > timeout) { atomicOr((unsigned long long*)err, 2ull); break; }
                }
            }
        }
        __syncthreads();
        for (int idx = tid; idx < TRB * TRHS; idx += 256) {
            const int kk = idx & (TRB - 1), s = idx >> 6;
            const int64_t gr = (int64_t)j * TRB + kk;
            Ys[kk][s] = (s < nrhs && gr < n) ? ld_coherent_f64(X + (int64_t)s * ldx + gr) : 0.0;
        }
        __syncthreads();
        if (row < n) {
            const int64_t cend = min((int64_t)TRB, n - (int64_t)j * TRB);
            const double* up = F + ((int64_t)j * TRB) * ld + row;
            for (int kk = 0; kk < cend; ++kk) {
                const double u = up[(int64_t)kk * ld];
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] = fma(-u, Ys[kk][g + 4 * q], acc[q]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) Xs[rr][g + 4 * q] = acc[q];
    const int nb = (int)min((int64_t)TRB, n - r0);
    for (int kk = nb - 1; kk >= 0; --kk) {
        __syncthreads();
        if (rr == kk) {
            const double d = F[(r0 + kk) * ld + r0 + kk];
#pragma unroll
            for (int q = 0; q < 4; ++q) Xs[kk][g + 4 * q] = Xs[kk][g + 4 * q] / d;
        }
        __syncthreads();
        if (rr < kk) {
            const double u = F[(r0 + kk) * ld + row];
#pragma unroll
            for (int q = 0; q < 4; ++q) Xs[rr][g + 4 * q] = fma(-u, Xs[kk][g + 4 * q], Xs[rr][g + 4 * q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int s = g + 4 * q;
        if (row < n && s < nrhs) X[(int64_t)s * ldx + row] = Xs[rr][s];
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&flags[bw], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------------------------------------
// Host driver.
static int launch_panel(lsx_lu* lu, int64_t j0, int W) {
    const int64_t m_rows = lu->n - j0;
    PanelArgs a;
    a.A = lu->F;
    a.ld = lu->ld;
    a.m = lu->n;
    a.j0 = j0;
    a.W = W;
    a.ipiv = lu->ipiv;
    a.dinfo = lu->dinfo;
    a.slots = lu->slots;
    a.seq = lu->seq;
    a.timeout = lu->ctx->wall_khz * 5000;  // 5 s
    lu->seq += (uint32_t)W + 1u;
    if (m_rows <= 256) {
        a.P = 1;
        hipLaunchKernelGGL(k_panel<256>, dim3(1), dim3(256), 0, lu->stream, a);
    } else if (m_rows <= 512) {
        a.P = 1;
        hipLaunchKernelGGL(k_panel<512>, dim3(1), dim3(512), 0, lu->stream, a);
    } else {
        a.P = (int)cdiv(m_rows, 1024);
        hipLaunchKernelGGL(k_panel<1024>, dim3(a.P), dim3(1024), 0, lu->stream, a);
    }
    return LSX_OK;
}

static void launch_gemm(lsx_lu* lu, const double* A, const double* B, double* C, int64_t M, int64_t N,
                        int64_t K) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    const int64_t T = cdiv(M, GBM) * cdiv(N, GBN);
    const int64_t grid = 8 * cdiv(T, 8);
    hipLaunchKernelGGL(k_gemm_f64, dim3((unsigned)grid), dim3(256), 0, lu->stream, A, B, C, lu->ld,
                       (int)M, (int)N, (int)K);
}

static void launch_laswp(lsx_lu* lu, int64_t k0, int kb, const int64_t* list, const int* count,
                         int64_t c0, int64_t c1, int trsm) {
    if (c1 <= c0) return;
    const int64_t nb = cdiv(c1 - c0, LCW);
    hipLaunchKernelGGL(k_laswp_trsm, dim3((unsigned)nb), dim3(256), 0, lu->stream, lu->F, lu->ld, k0, kb,
                       list, count, c0, c1, trsm);
}

int lu64_getrf(lsx_lu* lu) {
    lsx_ctx* ctx = lu->ctx;
    const int64_t n = lu->n, ld = lu->ld;
    double* F = lu->F;
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
        for (int64_t ji = jo; ji < jo + jbo; ji += PW) {
            const int W = (int)std::min<int64_t>(PW, jo + jbo - ji);
            launch_panel(lu, ji, W);
            const bool left = ji > jo, right = ji + W < jo + jbo;
            if (left || right) {
                hipLaunchKernelGGL(k_perm_build, dim3(1), dim3(256), 0, lu->stream, lu->ipiv, ji, W, il, ic,
                                   (int*)nullptr);
                if (left) launch_laswp(lu, ji, W, il, ic, jo, ji, 0);
                if (right) {
                    launch_laswp(lu, ji, W, il, ic, ji + W, jo + jbo, 1);
                    launch_gemm(lu, F + ji * ld + (ji + W), F + (ji + W) * ld + ji, F + (ji + W) * ld + (ji + W),
                                n - ji - W, jo + jbo - ji - W, W);
                }
            }
        }
        hipLaunchKernelGGL(k_perm_build, dim3(1), dim3(256), 0, lu->stream, lu->ipiv, jo, jbo, ol, oc,
                           lu->perm);
        if (jo > 0) launch_laswp(lu, jo, jbo, ol, oc, 0, jo, 0);
        if (jo + jbo < n) {
            launch_laswp(lu, jo, jbo, ol, oc, jo + jbo, n, 1);
            launch_gemm(lu, F + jo * ld + (jo + jbo), F + (jo + jbo) * ld + jo, F + (jo + jbo) * ld + (jo + jbo),
                        n - jo - jbo, n - jo - jbo, jbo);
        }
    }
    LSX_HIP(ctx, hipGetLastError());
    LSX_HIP(ctx, hipMemcpyAsync(lu->hinfo, lu->dinfo, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, lu->stream));
    LSX_HIP(ctx, hipStreamSynchronize(lu->stream));
    if (lu->hinfo[1] != 0)
        return fail(ctx, LSX_ETIMEOUT, "getrf: a panel exchange wait exceeded its bound (flags %lld)",
                    (long long)lu->hinfo[1]);
    return (int)std::min<int64_t>(lu->hinfo[0], INT32_MAX);
}

int lu_ensure_work(lsx_lu* lu, int64_t nrhs) {
    if (nrhs <= lu->wcap) return LSX_OK;
    lsx_ctx* ctx = lu->ctx;
    if (lu->wb) (void)hipFree(lu->wb);
    if (lu->wy) (void)hipFree(lu->wy);
    lu->wb = lu->wy = nullptr;
    lu->wcap = 0;
    const int64_t cap = std::max<int64_t>(nrhs, TRHS);
    LSX_HIP(ctx, hipMalloc(&lu->wb, sizeof(double) * lu->n * cap));
    LSX_HIP(ctx, hipMalloc(&lu->wy, sizeof(double) * lu->n * cap));
    lu->wcap = cap;
    return LSX_OK;
}

int lu64_getrs_dev(lsx_lu* lu, int64_t nrhs, const double* dB, int64_t ldb, double* dX, int64_t ldx) {
    lsx_ctx* ctx = lu->ctx;
    const int64_t n = lu->n;
    const int nblk = (int)cdiv(n, TRB);
    if (lu->tseq > 0xF0000000u) {
        LSX_HIP(ctx, hipMemsetAsync(lu->tflags, 0, sizeof(uint32_t) * 2 * nblk, lu->stream));
        lu->tseq = 1;
    }
    LSX_HIP(ctx, hipMemsetAsync(lu->dinfo + 1, 0, sizeof(int64_t), lu->stream));
    const int64_t timeout = ctx->wall_khz * 5000;
    for (int64_t c0 = 0; c0 < nrhs; c0 += TRHS) {
        const int nr = (int)std::min<int64_t>(TRHS, nrhs - c0);
        const uint32_t s1 = lu->tseq++, s2 = lu->tseq++;
        hipLaunchKernelGGL(k_trsv_lower, dim3(nblk), dim3(256), 0, lu->stream, lu->F, lu->ld, n, lu->perm,
                           dB + c0 * ldb, ldb, lu->wy, n, nr, lu->tflags, s1, timeout, lu->dinfo + 1);
        hipLaunchKernelGGL(k_trsv_upper, dim3(nblk), dim3(256), 0, lu->stream, lu->F, lu->ld, n,
                           (const double*)lu->wy, n, dX + c0 * ldx, ldx, nr, lu->tflags + nblk, s2, timeout,
                           lu->dinfo + 1);
    }
    LSX_HIP(ctx, hipGetLastError());
    return LSX_OK;
}

}  // namespace lsx
