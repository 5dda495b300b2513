// calu.hpp — block-column factorization of the mixed-precision LU with tournament pivoting
// (communication-avoiding LU, Grigori, Demmel and Xiang, SIAM J. Matrix Anal. Appl. 32, 2011).
//
// The mixed path's factors only precondition an FP64 refinement, so its pivots need not be
// LAPACK's; they only need partial pivoting's stability, which tournament pivoting has in
// practice.  What it removes is the exact panel's chain of one global argmax per column (one
// memory round trip between the panel's workgroups per column): here a 32-column inner panel
// costs two.
//
// One launch per outer block column F[jo:n, jo:jo+JB) (JB <= 128), P = ceil((n - jo) / 512)
// workgroups of 512 threads, thread t of workgroup w owning row jo + 512 w + t.  Rows stay in
// place; for each inner panel of 32 columns c0..c0+W:
//   1. every workgroup runs partial pivoting on its live rows' panel values (in registers, the
//      rotating layout of k_panel) and publishes the rows it picked, in order, as candidates:
//      each candidate's row label, its panel values and its values in the block's remaining
//      columns;
//   2. the workgroup that arrives last (an atomic counter) runs partial pivoting on all
//      P x 32 candidates (their values before this panel): its choices are the panel's pivot rows
//      and its result the packed L11 \ U11 of the panel's diagonal block; it forms
//      U12 = L11^{-1} A12 of the pivot rows over the block's remaining columns and publishes
//      L11 \ U11, U12 and the pivot rows;
//   3. every workgroup then writes the pivot rows it owns as U rows (L11 | U11 | U12), and for
//      every other live row solves l U11 = a (the row's multipliers) and updates its remaining
//      columns a -= l U12.
// Every element of the block is read and written only by the thread that owns its row, and
// everything that crosses workgroups travels as flag-in-word messages (32 payload bits and the
// panel's sequence number in one 64-bit word, read until the number matches), so the kernel needs
// no memory fences (an agent-scope fence writes back and invalidates the L2).
// The last panel's deciding workgroup turns the block's pivot rows into LAPACK's swap sequence
// ipiv; the interchanges are then applied to the block's own columns (k_laswp_trsm) like those
// of the columns right of it.  A zero pivot (all candidates zero in a column) takes the candidate
// with the smallest row index, its multipliers are left zero, and info = first such column.
#pragma once
#include <type_traits>

#include "lu_kernels.hpp"

namespace lsx {

constexpr int CBS = 512;        // threads (= rows) per workgroup
constexpr int CALU_PMAX = 32;   // workgroups: blocks of up to 16384 rows
constexpr int CW = 32;          // inner panel width
constexpr int CREM = NBO - CW;  // columns of a block right of its first inner panel

constexpr int CSL = CALU_PMAX * CW;           // candidate slots
constexpr int CREMW = (1 + CW) * CSL;         // first word of the candidates' remaining columns
constexpr int CCAND = CREMW + CSL * CREM;     // words of the candidate records
constexpr int CBC = CW * CW + CW * CREM + CW;   // words of the broadcast: L11\U11, U12, pivot rows

struct CaluArgs {
    float* A;
    int64_t ld;
    int64_t n;
    int64_t jo;
    int JB;
    int P;
    int64_t* ipiv;     // LAPACK swap sequence of the block, 1-based
    int64_t* dinfo;    // [0] info, [1] wait error flags
    uint64_t* cand;    // candidate records, CCAND words (label -1: none)
    uint64_t* bc;      // broadcast [CBC]
    uint32_t* ctr;     // arrival counter
    uint32_t cbase;    // counter value when this launch starts
    uint32_t seq;      // sequence number of this launch's first panel
    int64_t timeout;
};

// Packed pivot key of a candidate with index idx < 2^IB: |x|'s bit pattern (monotone in |x|, NaN
// counted as 0) with its IB low bits replaced by the complement of idx, plus one unit of the kept
// bits so that every candidate stays above 0 = none.  One unsigned max is then the argmax: the
// largest magnitude to a relative 2^(IB-23) (so every multiplier stays below 1 + 2^(IB-22) in
// magnitude) and, among equal kept bits, the smallest index.
template <int IB>
__device__ __forceinline__ uint32_t calu_key(float x, uint32_t idx) {
    const float t = fabsf(x);
    const uint32_t b = t == t ? __float_as_uint(t) : 0u;
    return (((b >> IB) + 1u) << IB) | ((1u << IB) - 1u - idx);
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
    uint32_t y;
    y = dpp32<0xB1>(x);  x = y > x ? y : x;   // lane ^ 1
    y = dpp32<0x4E>(x);  x = y > x ? y : x;   // lane ^ 2
    y = dpp32<0x141>(x); x = y > x ? y : x;   // the two quads of each 8
    y = dpp32<0x140>(x); x = y > x ? y : x;   // the two halves of each 16
    y = xrow16(x);       x = y > x ? y : x;
    y = xrow32(x);       x = y > x ? y : x;
    return x;
}
// f(integral_constant<int, I>) for I = 0 .. N-1, each call with its index a compile-time constant
// (a loop whose steps index register arrays, too long for the unroller).
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// Values C .. CW-1 of a 16-byte aligned row in LDS, read as the 16-byte words that hold them.
template <int C>
__device__ __forceinline__ void load_row(const float* row, float (&r)[CW]) {
    float4 t[CW / 4];
#pragma unroll
    for (int k = C / 4; k < CW / 4; ++k) t[k] = reinterpret_cast<const float4*>(row)[k];
#pragma unroll
    for (int k = 0; k < CW / 4; ++k) {
        if (k < C / 4) {
            r[4 * k] = r[4 * k + 1] = r[4 * k + 2] = r[4 * k + 3] = 0.f;
        } else {
            r[4 * k] = t[k].x;
            r[4 * k + 1] = t[k].y;
            r[4 * k + 2] = t[k].z;
            r[4 * k + 3] = t[k].w;
        }
    }
}

// The largest of the eight waves' keys (two 16-byte LDS reads, the same address in every lane).
__device__ __forceinline__ uint32_t max8_u32(const uint32_t* s) {
    const uint4 a = *reinterpret_cast<const uint4*>(s);
    const uint4 b = *reinterpret_cast<const uint4*>(s + 4);
    return max(max(max(a.x, a.y), max(a.z, a.w)), max(max(b.x, b.y), max(b.z, b.w)));
}

// Payloads of the flag-in-word messages addr(i) selected by mask (i < N <= 64) once they carry
// sequence number seq: all are read, then the stale ones again until none is left, the bound is
// reached or a wait of the launch has given up (*abort, shared by the workgroup, stops its later
// waits at once).  Returns false when it gave up.  The reads go in groups of 16, every read of a
// group issued before any is tested (a test right after each read would wait for the reads one
// at a time; all N at once would hold too many registers).
template <int N, typename F>
__device__ __forceinline__ bool calu_poll(F addr, uint64_t mask, uint32_t seq, uint32_t (&out)[N], const CaluArgs& a,
                                          int* abort) {
    constexpr int PB = 16;
    uint64_t stale = mask;
    int64_t t0 = 0;
    bool first = true;
    while (stale) {
#pragma unroll
        for (int i0 = 0; i0 < N; i0 += PB) {
            uint64_t x[PB];
#pragma unroll
            for (int j = 0; j < PB; ++j) {
                const int i = i0 + j;
                x[j] = (i < N && (stale & (1ull << i))) ? ld_relaxed_u64(addr(i < N ? i : 0)) : 0ull;
            }
#pragma unroll
            for (int j = 0; j < PB; ++j) {
                const int i = i0 + j;
                if (i < N && (stale & (1ull << i)) && (uint32_t)(x[j] >> 32) == seq) {
                    out[i] = (uint32_t)x[j];
                    stale &= ~(1ull << i);
                }
            }
        }
        if (!stale) break;
        if (*abort) return false;
        if (first) {
            t0 = wall_clock64();
            first = false;
        } else if (wall_clock64() - t0 > a.timeout) {
            atomicOr((unsigned long long*)&a.dinfo[1], 4ull);
            *abort = 1;
            return false;
        } else if (ld_relaxed_u64((const uint64_t*)&a.dinfo[1]) != 0) {
            *abort = 1;  // another wait of this launch gave up: its results are invalid anyway
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// The broadcast in LDS: packed L11\U11 of the panel and U12 of its pivot rows (namespace scope,
// so that the row update below, a function of its own, reads them as LDS).
// Rows of s_LU are 16-byte aligned, so that a row's values are read four at a time; s_dinv holds
// the reciprocals of U11's diagonal (0 for a zero pivot).
__shared__ __align__(16) float s_LU[CW][CW + 4];
__shared__ __align__(16) float s_U12[CW][CREM + 4];
__shared__ float s_dinv[CW];

// A live row that is not a pivot row: l U11 = a (row vector, U11 upper with its diagonal),
// column by column (times the diagonal's reciprocals), then a_rest -= l U12.  pc0: the panel's
// first column from the block's first row; rob: the row's byte offset in a column.  FULL: the panel has all CW columns (every panel
// but a last partial one); only then are the multipliers stored without a test per column (a
// test per column makes the compiler keep every address and value live).
template <bool FULL>
__device__ __forceinline__ void calu_row_update(float* pc0, int64_t ld, uint32_t rob, int W, int rem) {
    auto at = [&](int c) -> float& {
        return *reinterpret_cast<float*>(reinterpret_cast<char*>(pc0 + (int64_t)c * ld) + rob);
    };
    // the remaining columns eight at a time, each group's loads issued one group ahead of its
    // update (the first group's with the panel values, before the multipliers are formed)
    auto ld8 = [&](int j, float (&x)[8]) {
#pragma unroll
        for (int e = 0; e < 8; ++e) x[e] = at(W + j + e);
    };
    auto upd8 = [&](int j, float (&x)[8], const float (&l)[CW]) {
#pragma unroll
        for (int c = 0; c < CW; ++c) {  // two 16-byte LDS reads of U12 per multiplier
            const float4 u0 = *reinterpret_cast<const float4*>(&s_U12[c][j]);
            const float4 u1 = *reinterpret_cast<const float4*>(&s_U12[c][j + 4]);
            x[0] = fmaf(-l[c], u0.x, x[0]);
            x[1] = fmaf(-l[c], u0.y, x[1]);
            x[2] = fmaf(-l[c], u0.z, x[2]);
            x[3] = fmaf(-l[c], u0.w, x[3]);
            x[4] = fmaf(-l[c], u1.x, x[4]);
            x[5] = fmaf(-l[c], u1.y, x[5]);
            x[6] = fmaf(-l[c], u1.z, x[6]);
            x[7] = fmaf(-l[c], u1.w, x[7]);
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) at(W + j + e) = x[e];
    };
    const int ng = rem / 8;
    float l[CW], xa[8], xb[8];
#pragma unroll
    for (int c = 0; c < CW; ++c) {
        const float x = at(FULL ? c : min(c, W - 1));
        l[c] = (FULL || c < W) ? x : 0.f;
    }
    if (ng > 0) ld8(0, xa);
#pragma unroll
    for (int i = 0; i < CW; ++i) {
        l[i] = (FULL || i < W) ? l[i] * s_dinv[i] : 0.f;
#pragma unroll
        for (int c = i + 1; c < CW; ++c) l[c] = fmaf(-l[i], s_LU[i][c], l[c]);
    }
    if (FULL) {
#pragma unroll
        for (int c = 0; c < CW; ++c) at(c) = l[c];
    } else {
#pragma unroll 1
        for (int c = 0; c < W; ++c) at(c) = l[c];
    }
#pragma unroll 1
    for (int g = 0; g < ng; g += 2) {
        if (g + 1 < ng) ld8(8 * (g + 1), xb);
        upd8(8 * g, xa, l);
        if (g + 1 >= ng) break;
        if (g + 2 < ng) ld8(8 * (g + 2), xa);
        upd8(8 * (g + 1), xb, l);
    }
#pragma unroll 1
    for (int j = 8 * ng; j < rem; ++j) {
        float* pa = &at(W + j);
        float x = *pa;
#pragma unroll
        for (int c = 0; c < CW; ++c) x = fmaf(-l[c], s_U12[c][j], x);
        *pa = x;
    }
}

__global__ __launch_bounds__(CBS) void k_block_calu(CaluArgs a) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, w = blockIdx.x;
    const int64_t r = a.jo + (int64_t)w * CBS + tid;
    const bool valid = r < a.n;
    // 32-bit row offset from the block's first row: a strided access is then a uniform column
    // pointer (scalar registers) plus this offset, with no 64-bit address per access
    const uint32_t ro = valid ? (uint32_t)(r - a.jo) : 0u;
    const uint32_t rob = ro * 4u;  // in bytes (rows < 2^30)
    bool dead = false;  // a pivot row of an earlier inner panel of this block
    // one elimination step's exchange between the waves (two sets, used in turn, so that one
    // barrier per step suffices): each wave's largest key and that candidate's row and label
    __shared__ __align__(16) uint32_t s_key[2][CBS / 64];
    __shared__ __align__(16) float s_row[2][CBS / 64][CW];
    __shared__ int s_rlab[2][CBS / 64];
    __shared__ __align__(16) float s_zrow[CW];  // a zero column's pivot row (16-byte aligned)
    __shared__ int s_zaux[CBS / 64 + 2];
    __shared__ int s_lab[CW];
    __shared__ int s_slot[CW];       // deciding workgroup: candidate slot of each pivot
    __shared__ int s_pick[CW];       // this workgroup's candidates, in order: their threads
    // this workgroup's own candidate records (as published): the deciding workgroup, the last to
    // arrive, reads its own slots here instead of waiting for its stores to come back from memory
    __shared__ uint32_t s_own[1 + CW][CW];
    __shared__ float s_ownrem[CW][CREM];
    __shared__ int s_allab[NBO];     // the block's pivot rows, panel after panel
    __shared__ int s_final, s_abort;
    const int np = (a.JB + CW - 1) / CW;
    const int64_t ld = a.ld;
    float* A = a.A;
    if (tid == 0) s_abort = 0;

#pragma unroll 1
    for (int p = 0; p < np; ++p) {
        const int W = min(CW, a.JB - CW * p);
        const int rem = a.JB - CW * p - W;  // block columns right of this panel
        const int64_t c0 = a.jo + CW * p;
        const uint32_t seq = a.seq + (uint32_t)p;
        // the column stride and buffers, opaque to the compiler inside each panel: otherwise it
        // computes the 64-bit address of every strided access before the panel loop and keeps them
        int64_t ldp = ld;
        uint64_t* candp = a.cand;
        uint64_t* bcp = a.bc;
        asm volatile("" : "+s"(ldp), "+s"(candp), "+s"(bcp));
        float* const pc0 = A + c0 * ldp + a.jo;  // column c0 from the block's first row
        auto at = [&](int c) -> float& {         // this thread's element of column c0 + c
            return *reinterpret_cast<float*>(reinterpret_cast<char*>(pc0 + (int64_t)c * ldp) + rob);
        };

        // 1. local tournament: partial pivoting on this workgroup's live rows, then its picks.
        // The steps are unrolled (the column index is a constant in each), and only the multipliers'
        // choices matter here, so a row's panel values are updated in place with no multipliers kept.
        {
            float v[CW];
#pragma unroll
            for (int q = 0; q < CW; ++q) {
                const float x = at(min(q, W - 1));
                v[q] = (valid && q < W) ? x : 0.f;
            }
            bool live = valid && !dead;
            int myrank = -1;
            int cnt = W;  // candidates this workgroup offers (uniform)
            bool done = false;
            static_for<0, CW>([&](auto C) {
                constexpr int c = decltype(C)::value;
                if (done || c >= W) return;
                constexpr int par = c & 1;
                const uint32_t key = live ? calu_key<9>(v[c], (uint32_t)tid) : 0u;
                const uint32_t m = wave_max_u32(key);
                if (lane == 0) s_key[par][wave] = m;
                if (key == m && m != 0u) {
#pragma unroll
                    for (int q = 0; q < CW; q += 4)
                        *reinterpret_cast<float4*>(&s_row[par][wave][q]) = make_float4(v[q], v[q + 1], v[q + 2], v[q + 3]);
                }
                __syncthreads();
                const uint32_t M = max8_u32(s_key[par]);
                if (M == 0u) {  // no live row left (uniform)
                    cnt = c;
                    done = true;
                    return;
                }
                const int wt = (CBS - 1) - (int)(M & (CBS - 1));
                // the pivot row, 16 bytes per read, all reads issued before the first use
                float pr[CW];
                load_row<c>(s_row[par][wt >> 6], pr);
                if (tid == wt) {
                    myrank = c;
                    live = false;
                }
                const float pv = pr[c];
                const float l = (live && pv != 0.f) ? v[c] * __builtin_amdgcn_rcpf(pv) : 0.f;
#pragma unroll
                for (int q = c + 1; q < CW; ++q) v[q] = fmaf(-l, pr[q], v[q]);
            });
            // arrival first: the records are read by their sequence numbers, not by this count
            uint32_t arrived = 0;
            if (tid == 0) arrived = atomicAdd(a.ctr, 1u);
            if (myrank >= 0) s_pick[myrank] = tid;
            __syncthreads();
            // the candidates of slots w CW + k, copied by the whole workgroup (every load of a
            // thread in flight together): words q = 0 (row) and 1 .. CW (panel values before this
            // panel) at q CSL + slot, so that the deciding workgroup reads them a slot per lane;
            // values in the block's remaining columns at CREMW + slot CREM + j (label -1: none)
            {
                constexpr int NPW = ((1 + CW) * CW + CBS - 1) / CBS;
                float x[NPW];
#pragma unroll
                for (int i = 0; i < NPW; ++i) {
                    const int idx = tid + i * CBS, q = idx / CW, k = idx % CW;
                    const int pick = k < cnt ? s_pick[k] : 0;
                    const int qc = min(max(q - 1, 0), W - 1);
                    x[i] = (idx < (1 + CW) * CW && q >= 1 && q - 1 < W && k < cnt)
                               ? *reinterpret_cast<const float*>(reinterpret_cast<const char*>(pc0 + (int64_t)qc * ldp) + 4u * (uint32_t)pick + 4u * (uint32_t)(w * CBS))
                               : 0.f;
                }
#pragma unroll
                for (int i = 0; i < NPW; ++i) {
                    const int idx = tid + i * CBS, q = idx / CW, k = idx % CW;
                    if (idx < (1 + CW) * CW) {
                        const uint32_t pl = q == 0 ? (k < cnt ? (uint32_t)(a.jo + (int64_t)w * CBS + s_pick[k]) : 0xFFFFFFFFu)
                                                   : __float_as_uint(x[i]);
                        st_relaxed_u64(candp + (size_t)q * CSL + w * CW + k, ll_pack(pl, seq));
                        s_own[q][k] = pl;
                    }
                }
                constexpr int NRW = (CW * CREM + CBS - 1) / CBS;
                float y[NRW];
#pragma unroll
                for (int i = 0; i < NRW; ++i) {
                    const int idx = tid + i * CBS, k = idx / CREM, j = idx % CREM;
                    const int pick = k < cnt ? s_pick[k] : 0;
                    const int col = rem > 0 ? W + min(j, rem - 1) : W - 1;  // inside the block
                    y[i] = (k < cnt && j < rem)
                               ? *reinterpret_cast<const float*>(reinterpret_cast<const char*>(pc0 + (int64_t)col * ldp) + 4u * (uint32_t)pick + 4u * (uint32_t)(w * CBS))
                               : 0.f;
                }
#pragma unroll
                for (int i = 0; i < NRW; ++i) {
                    const int idx = tid + i * CBS, k = idx / CREM, j = idx % CREM;
                    if (k < cnt && j < rem) {
                        st_relaxed_u64(candp + CREMW + (size_t)(w * CW + k) * CREM + j, ll_pack(__float_as_uint(y[i]), seq));
                        s_ownrem[k][j] = y[i];
                    }
                }
            }
            if (tid == 0) s_final = (arrived == a.cbase + (uint32_t)(p * a.P + a.P - 1)) ? 1 : 0;
            __syncthreads();
        }

        // 2. the deciding workgroup: partial pivoting on all candidates, U12, broadcast
        const bool fin = s_final != 0;
        if (fin) {
            float f[2][CW];
            int fl[2];
            bool fa[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int sl = tid + h * CBS;
                fl[h] = -1;
#pragma unroll
                for (int q = 0; q < CW; ++q) f[h][q] = 0.f;
                if (sl / CW < a.P) {
                    const bool own = sl / CW == w;
                    uint32_t x[1 + CW];
                    if (calu_poll<1 + CW>([&](int i) { return candp + (size_t)i * CSL + sl; },
                                          own ? 0ull : (1ull << (1 + CW)) - 1, seq, x, a, &s_abort)) {
                        if (own) {
#pragma unroll
                            for (int i = 0; i < 1 + CW; ++i) x[i] = s_own[i][sl % CW];
                        }
                        fl[h] = (int)x[0];
#pragma unroll
                        for (int q = 0; q < CW; ++q) f[h][q] = __uint_as_float(x[1 + q]);
                    }
                }
                fa[h] = fl[h] >= 0;
            }
            // the steps unrolled as in the local tournament; here each candidate keeps its
            // multipliers in place (f[h][q], q < c), so the pivot row's values are its packed row of
            // L11 \\ U11 as they stand
            int zero_col = -1;  // uniform
            static_for<0, CW>([&](auto C) {
                constexpr int c = decltype(C)::value;
                if (c >= W) return;
                constexpr int par = c & 1;
                const uint32_t k0 = fa[0] ? calu_key<10>(f[0][c], (uint32_t)tid) : 0u;
                const uint32_t k1 = fa[1] ? calu_key<10>(f[1][c], (uint32_t)(tid + CBS)) : 0u;
                const uint32_t m = wave_max_u32(max(k0, k1));
                if (lane == 0) s_key[par][wave] = m;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if ((h ? k1 : k0) == m && m != 0u) {
#pragma unroll
                        for (int q = 0; q < CW; q += 4)
                            *reinterpret_cast<float4*>(&s_row[par][wave][q]) =
                                make_float4(f[h][q], f[h][q + 1], f[h][q + 2], f[h][q + 3]);
                        s_rlab[par][wave] = fl[h];
                    }
                }
                __syncthreads();
                const uint32_t M = max8_u32(s_key[par]);
                int slot = (2 * CBS - 1) - (int)(M & (2 * CBS - 1));
                const float* prow = s_row[par][(slot & (CBS - 1)) >> 6];
                int glab = s_rlab[par][(slot & (CBS - 1)) >> 6];
                if ((M >> 10) <= 1u) {
                    // every candidate left is zero (or NaN) in this column: the one with the smallest
                    // row index is the pivot, its multipliers stay zero (rare; two more barriers)
                    if (zero_col < 0) zero_col = c;
                    int lab = INT_MAX;
#pragma unroll
                    for (int h = 0; h < 2; ++h)
                        if (fa[h]) lab = min(lab, fl[h]);
                    lab = wave_min_i32(lab);
                    if (lane == 0) s_zaux[wave] = lab;
                    if (tid == 0) s_zaux[CBS / 64] = -1;  // stays so only when a wait gave up
                    __syncthreads();
                    int zl = INT_MAX;
#pragma unroll
                    for (int i = 0; i < CBS / 64; ++i) zl = min(zl, s_zaux[i]);
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        if (fa[h] && fl[h] == zl) {
#pragma unroll
                            for (int q = 0; q < CW; ++q) s_zrow[q] = f[h][q];
                            s_zaux[CBS / 64] = tid + h * CBS;
                        }
                    }
                    __syncthreads();
                    slot = s_zaux[CBS / 64];
                    prow = s_zrow;
                    glab = zl;
                }
                float pr[CW];
                load_row<c>(prow, pr);
                const float pv = pr[c];
                const float rinv = pv != 0.f ? 1.f / pv : 0.f;
                if (tid == 0) {
                    s_lab[c] = glab;
                    s_slot[c] = slot;
                    s_dinv[c] = rinv;
                }
                if (tid < CW) s_LU[c][tid] = prow[tid];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (fa[h] && tid + h * CBS == slot) fa[h] = false;
                    const float l = fa[h] ? f[h][c] * rinv : 0.f;
                    f[h][c] = l;
#pragma unroll
                    for (int q = c + 1; q < CW; ++q) f[h][q] = fmaf(-l, pr[q], f[h][q]);
                }
            });
            __syncthreads();
            // U12 = L11^{-1} A12 of the pivot rows (from their records), one column per thread
            if (tid < rem) {
                float u[CW];
                uint32_t x[CW];
                uint64_t m = 0;
#pragma unroll
                for (int c = 0; c < CW; ++c) {
                    x[c] = 0u;
                    const int sl = s_slot[c];
                    if (c < W && sl >= 0 && sl < a.P * CW && sl / CW != w) m |= 1ull << c;
                }
                (void)calu_poll<CW>([&](int c) { return candp + CREMW + (size_t)s_slot[c] * CREM + tid; }, m, seq, x, a,
                                    &s_abort);
#pragma unroll
                for (int c = 0; c < CW; ++c) {
                    const int sl = s_slot[c];
                    const bool own = c < W && sl >= 0 && sl / CW == w;
                    u[c] = (m >> c) & 1 ? __uint_as_float(x[c]) : (own ? s_ownrem[sl % CW][tid] : 0.f);
                }
#pragma unroll
                for (int c = 1; c < CW; ++c) {  // row c of L11 (contiguous in LDS) against u[0..c)
#pragma unroll
                    for (int i = 0; i < c; ++i) u[c] = fmaf(-s_LU[c][i], u[i], u[c]);
                }
#pragma unroll
                for (int c = 0; c < CW; ++c) {
                    s_U12[c][tid] = u[c];
                    st_relaxed_u64(bcp + CW * CW + c * CREM + tid, ll_pack(__float_as_uint(u[c]), seq));
                }
            }
            for (int i = tid; i < CW * CW; i += CBS)
                st_relaxed_u64(bcp + i, ll_pack(__float_as_uint(s_LU[i / CW][i % CW]), seq));
            if (tid < CW) st_relaxed_u64(bcp + CW * CW + CW * CREM + tid, ll_pack((uint32_t)(tid < W ? s_lab[tid] : -1), seq));
            if (tid == 0 && zero_col >= 0 && a.dinfo[0] == 0) a.dinfo[0] = c0 + zero_col + 1;
        } else {
            // 3a. the other workgroups: the broadcast into LDS (a share of the words per thread)
            constexpr int NB = (CBC + CBS - 1) / CBS;
            uint32_t x[NB];
            const int base = tid * NB;
            const int cntw = min(NB, CBC - base);
            // words of U12 beyond the block's remaining columns are not sent: not waited for
            uint64_t need = 0;
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int idx = base + i;
                x[i] = 0u;
                const bool used = i < cntw && (idx < CW * CW || idx >= CW * CW + CW * CREM ||
                                               (idx - CW * CW) % CREM < rem);
                if (used) need |= 1ull << i;
            }
            (void)calu_poll<NB>([&](int i) { return bcp + base + i; }, need, seq, x, a, &s_abort);
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int idx = base + i;
                if (i >= cntw) continue;
                if (idx < CW * CW) {
                    const float y = __uint_as_float(x[i]);
                    s_LU[idx / CW][idx % CW] = y;
                    if (idx / CW == idx % CW) s_dinv[idx / CW] = y != 0.f ? 1.f / y : 0.f;
                }
                else if (idx < CW * CW + CW * CREM) {
                    const int u = idx - CW * CW;
                    if (u % CREM < rem) s_U12[u / CREM][u % CREM] = __uint_as_float(x[i]);
                } else s_lab[idx - CW * CW - CW * CREM] = (int)x[i];
            }
        }
        __syncthreads();
        if (tid < W) s_allab[CW * p + tid] = s_lab[tid];
        // 3b. every workgroup: its rows
        if (valid && !dead && s_abort == 0) {
            int k = -1;
            for (int c = 0; c < W; ++c)
                if (s_lab[c] == (int)r) k = c;
            if (k >= 0) {
                // a pivot row: its U row (multipliers of the earlier columns stay as they are)
                for (int j = 0; j < W; ++j) at(j) = s_LU[k][j];
                for (int j = 0; j < rem; ++j) at(W + j) = s_U12[k][j];
                dead = true;
            } else {
                if (W == CW) calu_row_update<true>(pc0, ldp, rob, W, rem);
                else calu_row_update<false>(pc0, ldp, rob, W, rem);
            }
        }
        __syncthreads();
    }

    // the block's LAPACK swap sequence, by the last panel's deciding workgroup
    if (s_final && tid == 0) {
        __shared__ int who[NBO];   // row at position jo + i
        __shared__ int wtop[NBO];  // position of row jo + i
        const int JB = a.JB;
        const int jo = (int)a.jo;
        for (int i = 0; i < JB; ++i) {
            who[i] = jo + i;
            wtop[i] = jo + i;
        }
        for (int c = 0; c < JB; ++c) {
            int L = s_allab[c];
            if (L < jo || (int64_t)L >= a.n) L = jo + c;  // only after a wait gave up
            int pos = L < jo + JB ? wtop[L - jo] : L;
            if (pos < jo + c || (int64_t)pos >= a.n) pos = jo + c;
            a.ipiv[jo + c] = (int64_t)pos + 1;
            const int X = who[c];
            who[c] = L;
            if (L < jo + JB) wtop[L - jo] = jo + c;
            if (pos < jo + JB) who[pos - jo] = X;
            if (X < jo + JB && X >= jo) wtop[X - jo] = pos;
        }
    }
}

}  // namespace lsx
