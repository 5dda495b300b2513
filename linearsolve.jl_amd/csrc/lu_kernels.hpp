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
#include <climits>
#include <limits>

#include "lu.hpp"
#include "wave.hpp"

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

// Words (32-bit) per element of an exchange record, and word i of an element.
template <typename T> struct Words;
template <> struct Words<double> {
    static constexpr int W = 2;
    __device__ static uint32_t word(double x, int i) { return (uint32_t)(i ? __double2hiint(x) : __double2loint(x)); }
};
template <> struct Words<float> {
    static constexpr int W = 1;
    __device__ static uint32_t word(float x, int) { return __float_as_uint(x); }
};
// Exchange record (32-bit words): [0,1] key (lo, hi), [2] logical row, [3] pad, [4, ...) the
// candidate row in the panel's register layout.  16-byte multiple, so the row data is aligned.
template <typename T> constexpr int rec_words() { return 4 + PW * Words<T>::W; }

// Each thread owns one row of the panel in registers, in a rotating layout: at column c the
// physical slot p holds logical column c + p (p < PW - c, the not yet eliminated columns) and
// the finished values (multipliers L[r, 0..c), or the U row of a pivot row) sit in the last c
// slots.  Column c is therefore always v[0] and every register access is static, with a
// compact (not unrolled) column loop.
// Rows do not move while the panel is factored: each thread carries the logical row index lr
// of the row it holds.  The interchange of column k relabels two threads (the pivot row's
// becomes k, the one labelled k takes the pivot row's old label) and marks the pivot row
// finished; at the end every row is written to its logical position.  Ties of the pivot search
// are broken by the logical index, as an explicit row interchange would.
// P workgroups of BS threads cover rows [j0, m); with P > 1 each column exchanges one record
// per workgroup (its best candidate and that row's data) as flag-in-word messages.  There the
// rank-1 update of column c is deferred into column c + 1: it is applied at once to the new
// pivot column (needed for the pivot search) and to the published candidate row, and to the
// rest of the row while the exchange is in flight.
template <typename T, int BS>
__global__ __launch_bounds__(BS) void k_panel(PanelArgs<T> a) {
    constexpr int NW = BS / 64;
    constexpr int RW = rec_words<T>();
    constexpr int WPE = Words<T>::W;
    const int tid = threadIdx.x, w = blockIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r = a.j0 + (int64_t)w * BS + tid;
    const bool valid = r < a.m;
    T* col0 = a.A + a.j0 * a.ld;
    T v[PW];
#pragma unroll
    for (int q = 0; q < PW; ++q) v[q] = (valid && q < a.W) ? col0[q * a.ld + r] : T(0);
    int lr = (int)r;
    bool live = valid;
    int64_t info0 = 0;
    if (w == 0 && tid == 0) info0 = a.dinfo[0];

    __shared__ __align__(16) uint32_t s_cand[2][NW][RW];
    __shared__ double s_candl[2][NW];
    __shared__ __align__(16) uint32_t s_rec[2][PMAX * RW];

    // interchange of column k with pivot row grow (relabelling), pivots and info.  A live row
    // always wins the pivot search, so grow lies in [k, m); only a record left stale by a wait
    // that gave up (the factorization then reports LSX_ETIMEOUT) could hold another label, and it
    // is replaced by k so that every later row index stays inside the matrix.
    auto pivot_bookkeeping = [&](int k, int grow, T pv) {
        if (grow < k || (int64_t)grow >= a.m) grow = k;
        if (lr == grow) {
            live = false;
            lr = k;
        } else if (lr == k) {
            lr = grow;
        }
        if (w == 0 && tid == 0) {
            a.ipiv[k] = (int64_t)grow + 1;
            if (pv == T(0) && info0 == 0) {
                info0 = k + 1;
                a.dinfo[0] = info0;
            }
        }
    };
    auto scale = [&](T x, T pv) -> T {
        if (live && pv != T(0)) {
            if (fabs(pv) >= std::numeric_limits<T>::min()) x *= (T(1) / pv);
            else x /= pv;
        }
        return x;
    };

    const int P = a.P;
    if (P == 1) {
        // one workgroup: everything through LDS, update right after each pivot
#pragma unroll 1
        for (int c = 0; c < a.W; ++c) {
            const int par = c & 1;
            const int k = (int)(a.j0 + c);
            const uint64_t u = live ? pivot_key(v[0]) : 0;
            const int wl = wave_argmax(u, lr);
            if (lane == wl) {
                uint32_t* d = s_cand[par][wave];
                *reinterpret_cast<uint64_t*>(d) = u;
                d[2] = (uint32_t)lr;
                T* dv = reinterpret_cast<T*>(d + 4);
#pragma unroll
                for (int q = 0; q < PW; ++q) dv[q] = v[q];
            }
            __syncthreads();
            int wi;
            {
                const uint64_t cu = lane < NW ? *reinterpret_cast<const uint64_t*>(s_cand[par][lane]) : 0;
                const int cl = lane < NW ? (int)s_cand[par][lane][2] : INT_MAX;
                wi = wave_argmax(cu, cl);
            }
            const uint32_t* win = s_cand[par][wi];
            const T* prow = reinterpret_cast<const T*>(win + 4);
            const int grow = (int)win[2];
            const T pv = prow[0];
            pivot_bookkeeping(k, grow, pv);
            const T l = scale(v[0], pv);
            const int lim = live ? PW - 1 - c : 0;  // slots 1..lim hold columns still to eliminate
            // in groups of 8 slots: the group's pivot-row values are all read (unconditionally, so
            // the reads overlap) before the updates, which select instead of branch
#pragma unroll
            for (int p0 = 0; p0 < PW - 1; p0 += 8) {
                T pr[8];
#pragma unroll
                for (int p = p0; p < p0 + 8 && p < PW - 1; ++p) pr[p - p0] = prow[p + 1];
#pragma unroll
                for (int p = p0; p < p0 + 8 && p < PW - 1; ++p) {
                    const T nv = v[p + 1];
                    const T u = fma(-l, pr[p - p0], nv);
                    v[p] = (p < lim) ? u : nv;
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            v[PW - 1] = l;
        }
    } else {
        // pending update (column c - 1): pend = this row takes it, multiplier lp, pivot row pprev
        bool pend = false;
        T lp = T(0);
        const T* pprev = reinterpret_cast<const T*>(s_rec[1] + 4);  // read, unused, before column 1
#pragma unroll 1
        for (int c = 0; c < a.W; ++c) {
            const int par = c & 1;
            const int k = (int)(a.j0 + c);
            const int plim = PW - c;  // the pending update covers slots [0, plim)
            {
                const T u0 = fma(-lp, pprev[1], v[0]);
                v[0] = pend ? u0 : v[0];
            }
            const uint64_t u = live ? pivot_key(v[0]) : 0;
            const int wl = wave_argmax(u, lr);
            if (lane == wl) {  // raw row (slot 0 already current) and its pending multiplier
                uint32_t* d = s_cand[par][wave];
                *reinterpret_cast<uint64_t*>(d) = u;
                d[2] = (uint32_t)lr;
                T* dv = reinterpret_cast<T*>(d + 4);
#pragma unroll
                for (int q = 0; q < PW; ++q) dv[q] = v[q];
                s_candl[par][wave] = (double)lp;
            }
            __syncthreads();
            int wi;
            {
                const uint64_t cu = lane < NW ? *reinterpret_cast<const uint64_t*>(s_cand[par][lane]) : 0;
                const int cl = lane < NW ? (int)s_cand[par][lane][2] : INT_MAX;
                wi = wave_argmax(cu, cl);
            }
            const uint32_t seq = a.seq + (uint32_t)c + 1u;
            uint64_t* g = a.slots + (size_t)par * (PMAX + 1) * RECS;
            if (tid < RW) {
                // word tid of this workgroup's record, the candidate row brought up to date
                const uint32_t* win = s_cand[par][wi];
                uint32_t word = win[tid];
                if (tid >= 4 && c > 0) {
                    const int q = (tid - 4) / WPE;
                    if (q > 0 && q < plim) {
                        const T x = fma(-(T)s_candl[par][wi], pprev[q + 1], reinterpret_cast<const T*>(win + 4)[q]);
                        word = Words<T>::word(x, (tid - 4) % WPE);
                    }
                }
                st_relaxed_u64(g + (size_t)w * RECS + tid, ll_pack(word, seq));
            }
            // the rest of the pending update, and the rotation by one slot, while the records
            // travel (slot 0, column c, is kept in c0 until the pivot is known)
            const T c0 = v[0];
#pragma unroll
            for (int p0 = 1; p0 < PW; p0 += 8) {
#pragma unroll
                for (int p = p0; p < p0 + 8 && p < PW; ++p) {
                    const T nv = v[p];
                    v[p - 1] = (pend && p < plim) ? fma(-lp, pprev[p + 1], nv) : nv;
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            // every thread reads a share of the words of all P records, CW words per chunk (one
            // chunk for P <= 32): all words of the chunk that are not yet current are re-read
            // together, round by round (the update above leaves the records time to arrive).
            constexpr int CW = (32 * RW + BS - 1) / BS;
            const int total = P * RW;
#pragma unroll 1
            for (int base = 0; base < total; base += CW * BS) {
                uint64_t x[CW];
                uint32_t stale = 0;  // bit i: word i of this chunk still to read
#pragma unroll
                for (int i = 0; i < CW; ++i)
                    if (base + tid + i * BS < total) stale |= 1u << i;
                const int64_t t0 = wall_clock64();
                while (stale) {
#pragma unroll
                    for (int i = 0; i < CW; ++i) {
                        const int idx = base + tid + i * BS;
                        const int rr = idx / RW;
                        if (stale & (1u << i)) x[i] = ld_relaxed_u64(g + (size_t)rr * RECS + (idx - rr * RW));
                    }
#pragma unroll
                    for (int i = 0; i < CW; ++i) {
                        if ((stale & (1u << i)) && (uint32_t)(x[i] >> 32) == seq) {
                            s_rec[par][base + tid + i * BS] = (uint32_t)x[i];
                            stale &= ~(1u << i);
                        }
                    }
                    if (stale) {
                        // give up at the bound, or at once when another wait of this
                        // factorization already has (its results are invalid anyway)
                        if (wall_clock64() - t0 > a.timeout) {
                            atomicOr((unsigned long long*)&a.dinfo[1], 1ull);
                            break;
                        }
                        if (ld_relaxed_u64((const uint64_t*)&a.dinfo[1]) != 0) break;
                        __builtin_amdgcn_s_sleep(2);
                    }
                }
            }
            __syncthreads();
            // winner among the P records: lane l looks at records l and l + 64
            uint64_t bu = 0;
            int bl = INT_MAX, bp = 0;
#pragma unroll
            for (int h = 0; h < PMAX / 64; ++h) {
                const int pidx = lane + 64 * h;
                if (pidx < P) {
                    const uint32_t* rp = s_rec[par] + pidx * RW;
                    const uint64_t pu = *reinterpret_cast<const uint64_t*>(rp);
                    const int pl = (int)rp[2];
                    if (pu > bu || (pu == bu && pl < bl)) { bu = pu; bl = pl; bp = pidx; }
                }
            }
            const int wlane = wave_argmax(bu, bl);
            const int pw = __shfl(bp, wlane);
            const uint32_t* win = s_rec[par] + pw * RW;
            const T* prow = reinterpret_cast<const T*>(win + 4);
            const int grow = (int)win[2];
            const T pv = prow[0];
            pivot_bookkeeping(k, grow, pv);
            const T l = scale(c0, pv);
            v[PW - 1] = l;
            pend = live;
            lp = l;
            pprev = prow;
        }
        // the update of the last column touches only columns >= W (padding): nothing pending
    }
    if (valid) {
        // after W rotations slot p holds logical column (p + W) mod PW; the row goes to its
        // logical position
        T* dst = col0 + lr;
#pragma unroll
        for (int p = 0; p < PW; ++p) {
            const int q = (p + a.W) & (PW - 1);
            if (q < a.W) dst[q * a.ld] = v[p];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Row permutation of the LAPACK swap sequence ipiv[k0 .. k0+kb) (ipiv[k] >= k, 1-based) as a
// list of (destination row, source row) pairs: after the swaps, row dest[i] holds the former
// row src[i].  Optionally applies the same permutation to perm[] (row i of P*A = row perm[i]).
// One workgroup of 256 threads; kb <= 128.
// Positions 0..kb-1 stand for rows k0..k0+kb-1, positions kb..T-1 for the distinct pivot rows
// below the block (first occurrences, in order); swap u exchanges positions u and pos[u] >= u.
// Position q is last written by the last swap v with pos[v] = q, which moved there what
// position v held just before swap v, so every final label follows such links backwards in time
// independently of the others: no sequential replay of the swaps.
static __global__ __launch_bounds__(256) void k_perm_build(const int64_t* __restrict ipiv, int64_t k0, int kb,
                                                    int64_t* __restrict list, int* __restrict count,
                                                    int* __restrict perm) {
    __shared__ int64_t s_p[NBO];
    __shared__ int s_first[NBO];
    __shared__ int s_pos[NBO];
    __shared__ int s_prev[NBO];         // last earlier swap with the same target position
    __shared__ int s_last[2 * NBO];     // last swap targeting each position
    __shared__ int64_t s_rows[2 * NBO];
    __shared__ int s_pv[2 * NBO];
    const int t = threadIdx.x;
    const int64_t top = k0 + kb;
    if (t < kb) s_p[t] = ipiv[k0 + t] - 1;
    if (t < 2 * NBO) s_last[t] = -1;
    __syncthreads();
    // every loop below runs to its end (no early exit), so its LDS reads are independent
    int fidx = t;
    if (t < kb) {  // first swap with the same pivot row (below the block)
        const int64_t p = s_p[t];
        for (int u = t - 1; u >= 0; --u)
            if (s_p[u] == p) fidx = u;
        s_first[t] = (p >= top && fidx == t) ? 1 : 0;
        s_rows[t] = k0 + t;
    }
    __syncthreads();
    int extra = 0, before = 0;
    for (int u = 0; u < kb; ++u) {
        const int f = s_first[u];
        extra += f;
        before += (u < fidx) ? f : 0;
    }
    const int T = kb + extra;
    if (t < kb) {  // target position of swap t
        const int64_t p = s_p[t];
        const int pos = p >= top ? kb + before : (int)(p - k0);
        if (p >= top && fidx == t) s_rows[kb + before] = p;
        s_pos[t] = pos;
        atomicMax(&s_last[pos], t);
    }
    __syncthreads();
    if (t < kb) {
        const int pos = s_pos[t];
        int prev = -1;
        for (int u = 0; u < t; ++u)
            if (s_pos[u] == pos) prev = u;
        s_prev[t] = prev;
    }
    __syncthreads();
    int lab = t;
    if (t < T) {
        // label of position t: position q at time tm took, from the last swap v < tm targeting q,
        // what position v held just before swap v
        int q = t < kb ? s_pos[t] : t, tm = t < kb ? t : kb;
        while (true) {
            int v = s_last[q];
            while (v >= tm) v = s_prev[v];
            if (v < 0) break;
            q = v;
            tm = v;
        }
        lab = q;
        list[t] = s_rows[t];
        list[2 * NBO + t] = s_rows[lab];
    }
    if (t == 0) *count = T;
    if (perm) {
        if (t < T) s_pv[t] = perm[s_rows[lab]];
        __syncthreads();
        if (t < T) perm[s_rows[t]] = s_pv[t];
    }
}

// Inverse of a row map: inv[perm[r]] = r.
static __global__ void k_inv_perm(const int* __restrict perm, int* __restrict inv, int64_t n) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) inv[perm[r]] = (int)r;
}

// Deferred row interchanges of the L columns.  While the matrix is factored, the interchanges of
// an outer block are applied to the columns right of it only; column c of outer block b keeps
// its rows [r0, n) (r0 = end of block b) in the row order the matrix had after block b.  At the
// end, row i of that segment takes the stored row inv_b[perm[i]]: perm is the final row map and
// inv_b the inverse of the row map after block b.  One workgroup per column, the segment staged
// in LDS (dynamic, (n - NBO) elements at most).
template <typename T>
__global__ __launch_bounds__(256) void k_left_gather(T* __restrict F, int64_t ld, int64_t n,
                                                     const int* __restrict perm, const int* __restrict pinv) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    T* s = reinterpret_cast<T*>(s_raw);
    const int64_t c = blockIdx.x;
    const int64_t b = c / NBO;
    const int64_t r0 = (b + 1) * NBO;
    if (r0 >= n) return;
    T* col = F + c * ld;
    // eight loads of a thread in flight before their LDS stores, and eight row indices (two
    // dependent loads each) before the gathers
    for (int64_t i0 = r0 + threadIdx.x; i0 < n; i0 += 8 * 256) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = i0 + u * 256 < n ? col[i0 + u * 256] : T(0);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + u * 256 < n) s[i0 + u * 256 - r0] = v[u];
    }
    __syncthreads();
    const int* inv = pinv + b * n;
    for (int64_t i0 = r0 + threadIdx.x; i0 < n; i0 += 8 * 256) {
        int pr[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) pr[u] = i0 + u * 256 < n ? perm[i0 + u * 256] : 0;
#pragma unroll
        for (int u = 0; u < 8; ++u) pr[u] = i0 + u * 256 < n ? inv[pr[u]] : 0;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + u * 256 < n) col[i0 + u * 256] = s[pr[u] - r0];
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
        for (int idx = t; idx < PW * PW; idx += 256) {
            const int kk = idx / PW, i = idx - kk * PW;
            s_L[i][kk] = (i < kb && kk < kb) ? A[(k0 + kk) * ld + k0 + i] : T(0);
        }
    __syncthreads();
    const int64_t cb = c0 + (int64_t)blockIdx.x * LCW;
    const int ncol = (int)min((int64_t)LCW, c1 - cb);
    // gather: consecutive threads walk consecutive list entries of one column; eight loads of a
    // thread in flight before their LDS stores (a store right after each load would wait for
    // the scattered rows one at a time)
    for (int base = t; base < NT * LCW; base += 8 * 256) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = base + u * 256, cc = idx / NT, i = idx - cc * NT;
            v[u] = (idx < NT * LCW && cc < ncol) ? A[(cb + cc) * ld + s_src[i]] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = base + u * 256, cc = idx / NT, i = idx - cc * NT;
            if (idx < NT * LCW && cc < ncol) s_buf[i][cc] = v[u];
        }
    }
    __syncthreads();
    const int first = trsm ? kb : 0;
    for (int idx = t; idx < (NT - first) * LCW; idx += 256) {
        const int cc = idx / (NT - first), i = first + idx - cc * (NT - first);
        if (cc < ncol) A[(cb + cc) * ld + s_dst[i]] = s_buf[i][cc];
    }
    if (!trsm) return;
    // forward substitution with the unit-lower kb x kb block (kb <= PW): thread (col, group)
    // updates rows group + 8j of its column, one barrier per pivot row
    {
        const int col = t & (LCW - 1), grp = t >> 5;
        for (int kk = 0; kk < kb - 1; ++kk) {
            const T xk = s_buf[kk][col];
#pragma unroll
            for (int j = 0; j < PW / 8; ++j) {
                const int i = grp + 8 * j;
                if (i > kk && i < kb) s_buf[i][col] = fma(-s_L[i][kk], xk, s_buf[i][col]);
            }
            __syncthreads();
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
// grid = (number of blocks, 2): y = 0 inverts L blocks, y = 1 U blocks.  Four threads per
// column c of the inverse (columns are independent): thread p of the quad keeps the entries
// x[4 m + p] in registers and forms the substitution sum over them; the quad's four partial sums
// are added by two lane exchanges, so every step of the substitution is a chain of at most 16
// products instead of 64.  The block is staged in LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_diag_inv(const T* __restrict F, int64_t ld, int64_t n,
                                                  int64_t b0, T* __restrict DL, T* __restrict DU) {
    const int64_t blk = b0 + blockIdx.x;
    const int64_t r0 = blk * DB;
    const int nb = (int)min((int64_t)DB, n - r0);
    const int t = threadIdx.x, c = t >> 2, p = t & 3;
    __shared__ T D[DB][DB + 1];   // D[i][k]
    {
        // all sixteen loads of a thread issued before any LDS store
        T v[DB * DB / 256];
#pragma unroll
        for (int u = 0; u < DB * DB / 256; ++u) {
            const int idx = t + 256 * u, i = idx & (DB - 1), k = idx / DB;
            v[u] = (i < nb && k < nb) ? F[(r0 + k) * ld + r0 + i] : (i == k ? T(1) : T(0));
        }
#pragma unroll
        for (int u = 0; u < DB * DB / 256; ++u) {
            const int idx = t + 256 * u;
            D[idx & (DB - 1)][idx / DB] = v[u];
        }
    }
    __syncthreads();
    T xk[DB / 4];  // x[4 m + p]
#pragma unroll
    for (int m = 0; m < DB / 4; ++m) xk[m] = T(0);
    auto quad_sum = [&](T v) {
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        return v;
    };
    if (blockIdx.y == 0) {
        // column c of L^{-1}: x_c = 1, x_i = -sum_{k=c}^{i-1} L[i,k] x_k (i > c); x_k = 0 for k < c
#pragma unroll
        for (int i = 0; i < DB; ++i) {
            T a0 = T(0), a1 = T(0);
#pragma unroll
            for (int m = 0; 4 * m < i; ++m) {
                const T u = fma(D[i][4 * m + p], xk[m], (m & 1) ? a1 : a0);
                if (m & 1) a1 = (4 * m + p < i) ? u : a1;
                else a0 = (4 * m + p < i) ? u : a0;
            }
            const T sum = quad_sum(a0 + a1);
            const T xi = (i < c) ? T(0) : (i == c ? T(1) : -sum);
            if (p == (i & 3)) xk[i >> 2] = xi;
        }
    } else {
        // column c of U^{-1}: x_c = 1/U[c,c], x_i = -(sum_{k=i+1}^{c} U[i,k] x_k) / U[i,i] (i < c);
        // x_k = 0 for k > c
#pragma unroll
        for (int i = DB - 1; i >= 0; --i) {
            T a0 = T(0), a1 = T(0);
#pragma unroll
            for (int m = (i + 1) / 4; m < DB / 4; ++m) {
                const T u = fma(D[i][4 * m + p], xk[m], (m & 1) ? a1 : a0);
                if (m & 1) a1 = (4 * m + p > i) ? u : a1;
                else a0 = (4 * m + p > i) ? u : a0;
            }
            const T sum = quad_sum(a0 + a1);
            const T xi = (i > c) ? T(0) : (i == c ? T(1) / D[c][c] : -sum / D[i][i]);
            if (p == (i & 3)) xk[i >> 2] = xi;
        }
    }
    T* out = (blockIdx.y == 0 ? DL : DU) + blk * DB * DB;
#pragma unroll
    for (int m = 0; m < DB / 4; ++m) out[c * DB + 4 * m + p] = xk[m];  // column-major 64 x 64
}

// ------------------------------------------------------------------------------------------
// Triangular solves (getrs), one workgroup per DB-row block.
//   LOWER: Out = L^{-1} (P In)  (unit lower; the permutation gathers In by perm)
//   UPPER: Out = U^{-1} In
// Each block publishes its solution (DB x nrhs FP64 values) as flag-in-word messages: two
// 64-bit words per value, its 32-bit halves each with the launch's sequence number.  The blocks
// after it read those words directly, so a dependency costs one memory round trip (no flags, no
// fences).  A block accumulates the contributions of the blocks before it in solve order (four
// k-quarters of partial sums per row), the message reads of the next block issued before the
// current one is used, then applies the inverse of its diagonal block.
constexpr int TMW = 2 * DB * TRHS;  // message words of one block
// NR: right-hand-side slots handled (1, or TRHS for nrhs > 1); the arithmetic of each value is
// the same for both.  The block's right-hand side is read before the chain reaches the block.
template <typename T, bool LOWER, int NR>
__global__ __launch_bounds__(256) void k_trsv(const T* __restrict F, int64_t ld, int64_t n,
                                              const int* __restrict perm, const T* __restrict Dinv,
                                              const double* In, int64_t ldin, double* Out, int64_t ldout,
                                              int nrhs, uint64_t* msg, uint32_t seq, int64_t timeout,
                                              int64_t* err) {
    constexpr int WPT = (2 * DB * NR + 255) / 256;  // message words per thread
    constexpr int BPT = (DB * NR + 255) / 256;      // right-hand-side values per thread
    const int nblk = (int)((n + DB - 1) / DB);
    const int bw = LOWER ? (int)blockIdx.x : nblk - 1 - (int)blockIdx.x;
    const int64_t r0 = (int64_t)bw * DB;
    const int tid = threadIdx.x, rr = tid & 63, g = tid >> 6;
    const int64_t row = r0 + rr;
    __shared__ double Ds[DB][DB + 1];
    __shared__ double Ys[2][DB][NR + 1];
    __shared__ double Ps[4][DB][NR + 1];
    double bin[BPT];
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int idx = tid + 256 * i, k = idx & (DB - 1), s = idx >> 6;
        const int64_t gr = r0 + k;
        bin[i] = 0.0;
        if (idx < DB * NR && gr < n && s < nrhs)
            bin[i] = LOWER ? In[(int64_t)s * ldin + perm[gr]] : In[(int64_t)s * ldin + gr];
    }
    {
        T dv[DB * DB / 256];  // all loads in flight before the LDS stores
#pragma unroll
        for (int u = 0; u < DB * DB / 256; ++u) dv[u] = Dinv[(int64_t)bw * DB * DB + tid + 256 * u];
#pragma unroll
        for (int u = 0; u < DB * DB / 256; ++u) {
            const int idx = tid + 256 * u, i = idx & (DB - 1), k = idx >> 6;
            Ds[i][k] = (double)dv[u];
        }
    }
    double acc[NR];
#pragma unroll
    for (int s = 0; s < NR; ++s) acc[s] = 0.0;
    // word w of a block's message: value (row k, rhs s) half h at w = 2 (s DB + k) + h
    const int nj = LOWER ? bw : nblk - 1 - bw;  // blocks before this one in solve order
    auto jblk = [&](int t) { return LOWER ? t : nblk - 1 - t; };
    uint64_t x[WPT];
    auto issue = [&](int t) {
        const uint64_t* m = msg + (size_t)jblk(t) * TMW;
#pragma unroll
        for (int i = 0; i < WPT; ++i) {
            const int w = tid + 256 * i;
            x[i] = (w < 2 * DB * NR && w / (2 * DB) < nrhs) ? ld_relaxed_u64(m + w) : 0;
        }
    };
    // the 64 x 64 tile of block j: lane rr of group g holds F[row, j DB + 16 g + q], q < 16, in
    // the factor's own type (converted where it is used, so that no wait for the load is placed
    // early); tiles are loaded one block ahead of their use
    auto tload = [&](int t, T (&lv)[16]) {
        const int j = jblk(t);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int64_t col = (int64_t)j * DB + 16 * g + q;
            lv[q] = (row < n && col < n) ? F[col * ld + row] : T(0);
        }
    };
    // block t of the solve order: its message (issued earlier) until current, into Ys, the next
    // block's message issued, then acc += tile * x_j
    auto step = [&](int t, const T (&lv)[16]) {
        const int j = jblk(t), buf = t & 1;
        const uint64_t* m = msg + (size_t)j * TMW;
        uint32_t stale = 0;
#pragma unroll
        for (int i = 0; i < WPT; ++i) {
            const int w = tid + 256 * i;
            if (w < 2 * DB * NR && w / (2 * DB) < nrhs && (uint32_t)(x[i] >> 32) != seq) stale |= 1u << i;
        }
        if (stale) {
            const int64_t t0 = wall_clock64();
            while (stale) {
                __builtin_amdgcn_s_sleep(1);
#pragma unroll
                for (int i = 0; i < WPT; ++i)
                    if (stale & (1u << i)) x[i] = ld_relaxed_u64(m + tid + 256 * i);
#pragma unroll
                for (int i = 0; i < WPT; ++i)
                    if ((stale & (1u << i)) && (uint32_t)(x[i] >> 32) == seq) stale &= ~(1u << i);
                if (stale && wall_clock64() - t0 > timeout) {
                    atomicOr((unsigned long long*)err, 2ull);
                    break;
                }
                if (stale && ld_relaxed_u64((const uint64_t*)err) != 0) break;  // an earlier wait gave up
            }
        }
        // the two halves of a value sit in words w, w + 1 of threads tid, tid ^ 1
#pragma unroll
        for (int i = 0; i < WPT; ++i) {
            const int w = tid + 256 * i;
            const uint32_t mine = (uint32_t)x[i];
            const uint32_t other = (uint32_t)__shfl_xor((int)mine, 1);
            const int k = (w >> 1) & (DB - 1), s = w / (2 * DB);
            if (!(w & 1) && w < 2 * DB * NR && s < nrhs) Ys[buf][k][s] = __hiloint2double((int)other, (int)mine);
        }
        if (t + 1 < nj) issue(t + 1);  // the next block's words travel while this one is used
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q)
#pragma unroll
            for (int s = 0; s < NR; ++s) acc[s] = fma((double)lv[q], Ys[buf][16 * g + q][s], acc[s]);
        // Ys is double-buffered: buffer buf is written again two blocks on, after the next barrier
    };
    if (nj > 0) {
        issue(0);
        T la[16], lb[16];
        tload(0, la);
        for (int t = 0; t < nj; t += 2) {
            if (t + 1 < nj) tload(t + 1, lb);
            step(t, la);
            if (t + 1 >= nj) break;
            if (t + 2 < nj) tload(t + 2, la);
            step(t + 1, lb);
        }
    }
#pragma unroll
    for (int s = 0; s < NR; ++s) Ps[g][rr][s] = acc[s];
    __syncthreads();
    // right-hand side minus contributions, one (row, rhs) per thread-slot
    double (*Yb)[NR + 1] = Ys[nj & 1];
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int idx = tid + 256 * i, k = idx & (DB - 1), s = idx >> 6;
        if (idx < DB * NR) Yb[k][s] = bin[i] - (Ps[0][k][s] + Ps[1][k][s] + Ps[2][k][s] + Ps[3][k][s]);
    }
    __syncthreads();
    // y = Dinv Yb: group g forms the partial products over k in [16 g, 16 g + 16) (four short
    // chains instead of one of 64), reduced through Ps
#pragma unroll
    for (int s = 0; s < NR; ++s) {
        double y = 0.0;
#pragma unroll
        for (int q = 0; q < 16; ++q) y = fma(Ds[rr][16 * g + q], Yb[16 * g + q][s], y);
        Ps[g][rr][s] = y;
    }
    __syncthreads();
    uint64_t* mine = msg + (size_t)bw * TMW;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int s = g + 4 * q;
        if (s >= NR) continue;
        const double y = (Ps[0][rr][s] + Ps[1][rr][s]) + (Ps[2][rr][s] + Ps[3][rr][s]);
        if (s < nrhs) {
            if (row < n) Out[(int64_t)s * ldout + row] = y;
            st_relaxed_u64(mine + 2 * (s * DB + rr), ll_pack((uint32_t)__double2loint(y), seq));
            st_relaxed_u64(mine + 2 * (s * DB + rr) + 1, ll_pack((uint32_t)__double2hiint(y), seq));
        }
    }
}

// Transposed triangular solves (getrs with trans = 'T', A^T x = b with P A = L U:
// U^T w = b, L^T v = w, x = P^T v), one workgroup per DB-row block, one right-hand side.
//   FIRST:  Out = U^{-T} In         (U^T is lower triangular: blocks in forward order)
//   !FIRST: Out[perm[i]] = (L^{-T} In)[i]   (L^T unit upper: blocks in backward order)
// The tile of block row b and block column j of the transposed factor is the (j, b) tile of F
// read by columns: lane r of wave g reads F[(r0 + 16 g + q) ld + j DB + r] (coalesced) and keeps
// its own partial sums of the 16 output rows 16 g + q over all earlier blocks but the last one;
// those are reduced across the wave while the last block's solution travels.  The last block's
// tile is read by rows instead (four threads per output row, 16 values each, loaded before the
// wait), so after its solution arrives only a two-step shuffle remains; the product with the
// diagonal block's inverse (the transpose of the stored one) is split the same way.  Block solutions travel as in k_trsv (two 64-bit words per value, 32-bit halves
// each with the launch's sequence number).
template <typename T, bool FIRST>
__global__ __launch_bounds__(256) void k_trsv_t(const T* __restrict F, int64_t ld, int64_t n,
                                                const int* __restrict perm, const T* __restrict Dinv,
                                                const double* In, double* Out, uint64_t* msg, uint32_t seq,
                                                int64_t timeout, int64_t* err) {
    const int nblk = (int)((n + DB - 1) / DB);
    const int bw = FIRST ? (int)blockIdx.x : nblk - 1 - (int)blockIdx.x;
    const int64_t r0 = (int64_t)bw * DB;
    const int tid = threadIdx.x, r = tid & 63, g = tid >> 6;
    __shared__ double Ds[DB][DB + 1];  // Ds[i][k] = Dinv^T(i, k)
    __shared__ double Wb[2][DB];       // solution of the block being consumed (double-buffered)
    __shared__ double Red[DB];
    __shared__ double Yb[DB];
    __shared__ double Bb[DB];          // the block's right-hand side
    if (tid < DB) Bb[tid] = (r0 + tid < n) ? In[r0 + tid] : 0.0;
    for (int idx = tid; idx < DB * DB; idx += 256) {
        const int i = idx >> 6, k = idx & (DB - 1);
        Ds[i][k] = (double)Dinv[(int64_t)bw * DB * DB + (int64_t)i * DB + k];
    }
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0;
    const int nj = FIRST ? bw : nblk - 1 - bw;
    auto jblk = [&](int t) { return FIRST ? t : nblk - 1 - t; };
    uint64_t x = 0;
    auto issue = [&](int t) {
        if (tid < 2 * DB) x = ld_relaxed_u64(msg + (size_t)jblk(t) * TMW + tid);
    };
    // the solution of chain block j into Wb[buf] (bounded wait)
    auto receive = [&](int j, int buf) {
        if (tid < 2 * DB && (uint32_t)(x >> 32) != seq) {
            const uint64_t* m = msg + (size_t)j * TMW + tid;
            const int64_t t0 = wall_clock64();
            while (true) {
                __builtin_amdgcn_s_sleep(1);
                x = ld_relaxed_u64(m);
                if ((uint32_t)(x >> 32) == seq) break;
                if (wall_clock64() - t0 > timeout) {
                    atomicOr((unsigned long long*)err, 2ull);
                    break;
                }
                if (ld_relaxed_u64((const uint64_t*)err) != 0) break;  // an earlier wait gave up
            }
        }
        const uint32_t mine = (uint32_t)x;
        const uint32_t other = (uint32_t)__shfl_xor((int)mine, 1);
        if (tid < 2 * DB && !(tid & 1)) Wb[buf][tid >> 1] = __hiloint2double((int)other, (int)mine);
    };
    if (nj > 0) issue(0);
    // every block but the last one in the chain: wave lanes over the source rows
    for (int t = 0; t + 1 < nj; ++t) {
        const int j = jblk(t), buf = t & 1;
        const int64_t srow = (int64_t)j * DB + r;
        double lv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int64_t ocol = r0 + 16 * g + q;
            lv[q] = (ocol < n && srow < n) ? (double)F[ocol * ld + srow] : 0.0;
        }
        receive(j, buf);
        issue(t + 1);
        __syncthreads();
        const double wr = Wb[buf][r];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] = fma(lv[q], wr, acc[q]);
    }
    // their sums across the wave, while the last block's solution travels
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        double v = acc[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (r == 0) Red[16 * g + q] = v;
    }
    // the last block: four threads per output row (i = tid / 4), each with 16 values of column
    // r0 + i of F, loaded before the wait, and a two-step sum over the four
    double last = 0.0;
    if (nj > 0) {
        const int jl = jblk(nj - 1), buf = (nj - 1) & 1;
        const int i = tid >> 2, p4 = tid & 3;
        double f[16];
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            const int64_t srow = (int64_t)jl * DB + 16 * p4 + kk;
            f[kk] = (r0 + i < n && srow < n) ? (double)F[(r0 + i) * ld + srow] : 0.0;
        }
        receive(jl, buf);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) last = fma(f[kk], Wb[buf][16 * p4 + kk], last);
        last += __shfl_xor(last, 1);
        last += __shfl_xor(last, 2);
    }
    __syncthreads();
    if ((tid & 3) == 0) Yb[tid >> 2] = Bb[tid >> 2] - Red[tid >> 2] - last;
    __syncthreads();
    // y = Dinv^T Yb, again four threads per row, 16 terms each
    {
        const int i = tid >> 2, p4 = tid & 3;
        double y = 0.0;
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) y = fma(Ds[i][16 * p4 + kk], Yb[16 * p4 + kk], y);
        y += __shfl_xor(y, 1);
        y += __shfl_xor(y, 2);
        if (p4 == 0) {
            const int64_t row = r0 + i;
            if (row < n) Out[FIRST ? row : (int64_t)perm[row]] = y;
            uint64_t* mine = msg + (size_t)bw * TMW;
            st_relaxed_u64(mine + 2 * i, ll_pack((uint32_t)__double2loint(y), seq));
            st_relaxed_u64(mine + 2 * i + 1, ll_pack((uint32_t)__double2hiint(y), seq));
        }
    }
}

}  // namespace lsx
