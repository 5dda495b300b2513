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
#include "lu_kernels.hpp"

namespace lsx {

constexpr int CBS = 512;        // threads (= rows) per workgroup
constexpr int CALU_PMAX = 32;   // workgroups: blocks of up to 16384 rows
constexpr int CW = 32;          // inner panel width
constexpr int CREM = NBO - CW;  // columns of a block right of its first inner panel

constexpr int CREC = 1 + CW + CREM;           // words of a candidate record: label, panel, rest
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
    uint64_t* cand;    // candidate records [CALU_PMAX][CW][CREC] (label -1: none)
    uint64_t* bc;      // broadcast [CBC]
    uint32_t* ctr;     // arrival counter
    uint32_t cbase;    // counter value when this launch starts
    uint32_t seq;      // sequence number of this launch's first panel
    int64_t timeout;
};

// 32-bit pivot key of an FP32 value: |x| as its (monotone) bit pattern + 1, NaN counted as 0;
// 0 = not a candidate.  Half the shuffle work of the 64-bit key of the FP64 panel.
__device__ __forceinline__ uint32_t pivot_key32(float x) {
    const float t = fabsf(x);
    return (t == t ? __float_as_uint(t) : 0u) + 1u;
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
// Lane holding the best (key, lab) pair of the wave: largest key, then smallest label (uniform).
__device__ __forceinline__ int wave_argmax32(uint32_t u, int lab) {
    const uint32_t m = wave_max_u32(u);
    const uint64_t tie = __ballot(u == m);
    if (__popcll(tie) == 1 || m == 0) return __ffsll((unsigned long long)tie) - 1;
    const int mn = wave_min_i32(u == m ? lab : INT_MAX);
    return __ffsll((unsigned long long)__ballot(u == m && lab == mn)) - 1;
}

// Workgroup argmax of (key, label) over its 8 waves: each wave's best writes its key, label and
// NV values to s[par][wave]; returns the winning wave (uniform).  key 0 = no candidate.
// The values are those of v0, or of v1 where sel is set.
template <int NV>
__device__ __forceinline__ int calu_wg_argmax(uint32_t key, int lab, const float (&v0)[NV], const float (&v1)[NV],
                                              bool sel, float (*s)[4 + NV], int lane, int wave) {
    const int wl = wave_argmax32(key, lab);
    if (lane == wl) {
        float* d = s[wave];
        reinterpret_cast<uint32_t*>(d)[0] = key;
        reinterpret_cast<int*>(d)[2] = lab;
#pragma unroll
        for (int q = 0; q < NV; ++q) d[4 + q] = sel ? v1[q] : v0[q];
    }
    __syncthreads();
    const uint32_t cu = lane < CBS / 64 ? reinterpret_cast<const uint32_t*>(s[lane])[0] : 0u;
    const int cl = lane < CBS / 64 ? reinterpret_cast<const int*>(s[lane])[2] : INT_MAX;
    return wave_argmax32(cu, cl);
}

// The pivot row's values are all read first (unconditionally, so the reads overlap: pivot_row)
// and the update selects instead of branching: a test per slot around its read would make the
// compiler wait for each read in turn.
__device__ __forceinline__ void pivot_row(const float* prow, float (&pr)[CW - 1]) {
#pragma unroll
    for (int q = 0; q < CW - 1; ++q) pr[q] = prow[q + 1];
}
__device__ __forceinline__ void calu_step(float (&v)[CW], const float (&pr)[CW - 1], float pv, bool upd, int c) {
    const float l = (upd && pv != 0.f) ? v[0] * (1.f / pv) : 0.f;
    const int lim = upd ? CW - 1 - c : 0;
#pragma unroll
    for (int q = 0; q < CW - 1; ++q) {
        const float nv = v[q + 1];
        const float u = fmaf(-l, pr[q], nv);
        v[q] = q < lim ? u : nv;
    }
    v[CW - 1] = l;
}

// Payloads of the flag-in-word messages addr(i) selected by mask (i < N <= 64) once they carry
// sequence number seq: all are read, then the stale ones again until none is left, the bound is
// reached or a wait of the launch has given up (*abort, shared by the workgroup, stops its later
// waits at once).  Returns false when it gave up.
template <int N, typename F>
__device__ __forceinline__ bool calu_poll(F addr, uint64_t mask, uint32_t seq, uint32_t (&out)[N], const CaluArgs& a,
                                          int* abort) {
    uint64_t stale = mask;
    int64_t t0 = 0;
    bool first = true;
    while (stale) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            if (stale & (1ull << i)) {
                const uint64_t x = ld_relaxed_u64(addr(i));
                if ((uint32_t)(x >> 32) == seq) {
                    out[i] = (uint32_t)x;
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
__shared__ float s_LU[CW][CW + 1];
__shared__ __align__(16) float s_U12[CW][CREM + 4];

// A live row that is not a pivot row: l U11 = a (row vector, U11 upper with its diagonal),
// column by column, then a_rest -= l U12.  pc0: the panel's first column from the block's first
// row; rob: the row's byte offset in a column.  FULL: the panel has all CW columns (every panel
// but a last partial one); only then are the multipliers stored without a test per column (a
// test per column makes the compiler keep every address and value live).
template <bool FULL>
__device__ __forceinline__ void calu_row_update(float* pc0, int64_t ld, uint32_t rob, int W, int rem) {
    auto at = [&](int c) -> float& {
        return *reinterpret_cast<float*>(reinterpret_cast<char*>(pc0 + (int64_t)c * ld) + rob);
    };
    float l[CW];
#pragma unroll
    for (int c = 0; c < CW; ++c) {
        const float x = at(FULL ? c : min(c, W - 1));
        l[c] = (FULL || c < W) ? x : 0.f;
    }
#pragma unroll
    for (int i = 0; i < CW; ++i) {
        const float d = s_LU[i][i];
        l[i] = ((FULL || i < W) && d != 0.f) ? l[i] / d : 0.f;
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
    // eight columns at a time (eight loads in flight), two 16-byte LDS reads of U12 per multiplier
    int j = 0;
#pragma unroll 1
    for (; j + 8 <= rem; j += 8) {
        float x[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) x[e] = at(W + j + e);
#pragma unroll
        for (int c = 0; c < CW; ++c) {
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
    }
#pragma unroll 1
    for (; j < rem; ++j) {
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
    __shared__ __align__(16) float s_win[2][CBS / 64][4 + CW];
    __shared__ int s_lab[CW];
    __shared__ int s_slot[CW];       // deciding workgroup: candidate slot of each pivot
    __shared__ int s_allab[NBO];     // the block's pivot rows, panel after panel
    __shared__ int s_final, s_count, s_abort;
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

        // 1. local tournament: partial pivoting on this workgroup's live rows, then its picks
        {
            float v[CW];
#pragma unroll
            for (int q = 0; q < CW; ++q) {
                const float x = at(min(q, W - 1));
                v[q] = (valid && q < W) ? x : 0.f;
            }
            bool live = valid && !dead;
            int myrank = -1;
            if (tid == 0) s_count = 0;
#pragma unroll 1
            for (int c = 0; c < W; ++c) {
                const int par = c & 1;
                const uint32_t key = live ? pivot_key32(v[0]) : 0u;
                const int wi = calu_wg_argmax<CW>(key, (int)r, v, v, false, s_win[par], lane, wave);
                const float* win = s_win[par][wi];
                if (reinterpret_cast<const uint32_t*>(win)[0] == 0u) break;  // no live row left (uniform)
                const int glab = reinterpret_cast<const int*>(win)[2];
                const float pv = win[4];
                if (live && glab == (int)r) {
                    myrank = c;
                    live = false;
                }
                float pr[CW - 1];
                pivot_row(win + 4, pr);
                calu_step(v, pr, pv, live, c);
                if (tid == 0) s_count = c + 1;
            }
            __syncthreads();
            const int cnt = s_count;
            // record of slot k: [0] row, [1, 33) panel values before this panel, [33, 33 + rem)
            // the row's values in the block's remaining columns
            uint64_t* recs = candp + (size_t)w * CW * CREC;
            if (myrank >= 0) {
                uint64_t* rec = recs + (size_t)myrank * CREC;
                st_relaxed_u64(rec, ll_pack((uint32_t)r, seq));
#pragma unroll
                for (int q = 0; q < CW; ++q) {
                    const float x = at(min(q, W - 1));
                    st_relaxed_u64(rec + 1 + q, ll_pack(__float_as_uint(q < W ? x : 0.f), seq));
                }
                for (int j0 = 0; j0 < rem; j0 += 16) {  // 16 loads in flight, then their stores
                    float x[16];
#pragma unroll
                    for (int e = 0; e < 16; ++e) x[e] = at(W + min(j0 + e, rem - 1));
#pragma unroll
                    for (int e = 0; e < 16; ++e)
                        if (j0 + e < rem) st_relaxed_u64(rec + 1 + CW + j0 + e, ll_pack(__float_as_uint(x[e]), seq));
                }
            }
            if (tid >= cnt && tid < CW) {  // no candidate in this slot
                uint64_t* rec = recs + (size_t)tid * CREC;
                st_relaxed_u64(rec, ll_pack(0xFFFFFFFFu, seq));
#pragma unroll
                for (int q = 0; q < CW; ++q) st_relaxed_u64(rec + 1 + q, ll_pack(0u, seq));
            }
            if (tid == 0) {
                const uint32_t old = atomicAdd(a.ctr, 1u);
                s_final = (old == a.cbase + (uint32_t)(p * a.P + a.P - 1)) ? 1 : 0;
            }
            __syncthreads();
        }

        // 2. the deciding workgroup: partial pivoting on all candidates, U12, broadcast
        const bool fin = s_final != 0;
        if (fin) {
            float f[2][CW];
            int fl[2], fs[2];
            bool fa[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int sl = tid + h * CBS;
                fs[h] = sl;
                fl[h] = -1;
#pragma unroll
                for (int q = 0; q < CW; ++q) f[h][q] = 0.f;
                if (sl / CW < a.P) {
                    uint32_t x[1 + CW];
                    const uint64_t* rec = candp + (size_t)sl * CREC;
                    if (calu_poll<1 + CW>([&](int i) { return rec + i; }, (1ull << (1 + CW)) - 1, seq, x, a, &s_abort)) {
                        fl[h] = (int)x[0];
#pragma unroll
                        for (int q = 0; q < CW; ++q) f[h][q] = __uint_as_float(x[1 + q]);
                    }
                }
                fa[h] = fl[h] >= 0;
            }
            int zero_col = -1;
#pragma unroll 1
            for (int c = 0; c < W; ++c) {
                const int par = c & 1;
                const uint32_t k0 = fa[0] ? pivot_key32(f[0][0]) : 0u, k1 = fa[1] ? pivot_key32(f[1][0]) : 0u;
                const bool second = k1 > k0 || (k1 == k0 && k1 != 0 && fl[1] < fl[0]);
                const uint32_t key = second ? k1 : k0;
                const int lab = second ? fl[1] : (fa[0] ? fl[0] : INT_MAX);
                const int wl = wave_argmax32(key, lab);
                if (lane == wl) {  // the wave's best: key, row, slot, values
                    float* d = s_win[par][wave];
                    reinterpret_cast<uint32_t*>(d)[0] = key;
                    reinterpret_cast<int*>(d)[2] = lab;
                    reinterpret_cast<int*>(d)[3] = second ? fs[1] : fs[0];
#pragma unroll
                    for (int q = 0; q < CW; ++q) d[4 + q] = second ? f[1][q] : f[0][q];
                }
                __syncthreads();
                int wi;
                {
                    const uint32_t cu = lane < CBS / 64 ? reinterpret_cast<const uint32_t*>(s_win[par][lane])[0] : 0u;
                    const int cl = lane < CBS / 64 ? reinterpret_cast<const int*>(s_win[par][lane])[2] : INT_MAX;
                    wi = wave_argmax32(cu, cl);
                }
                const float* win = s_win[par][wi];
                const int glab = reinterpret_cast<const int*>(win)[2];
                const float pv = win[4];
                if (tid == 0) {
                    s_lab[c] = glab;
                    s_slot[c] = reinterpret_cast<const int*>(win)[3];
                    if (pv == 0.f && zero_col < 0) zero_col = c;
                }
                // the pivot row's packed values: L11 (c, 0..c) from the tail slots, U11 (c, c..W)
                if (tid < CW) s_LU[c][tid] = tid < c ? win[4 + CW - c + tid] : win[4 + tid - c];
                float pr[CW - 1];
                pivot_row(win + 4, pr);
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (fa[h] && fl[h] == glab) fa[h] = false;
                    calu_step(f[h], pr, pv, fa[h], c);
                }
            }
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
                    if (c < W && sl >= 0 && sl < a.P * CW) m |= 1ull << c;
                }
                (void)calu_poll<CW>([&](int c) { return candp + (size_t)s_slot[c] * CREC + 1 + CW + tid; }, m, seq, x, a,
                                    &s_abort);
#pragma unroll
                for (int c = 0; c < CW; ++c) u[c] = (m >> c) & 1 ? __uint_as_float(x[c]) : 0.f;
#pragma unroll
                for (int i = 0; i < CW - 1; ++i) {
#pragma unroll
                    for (int c = i + 1; c < CW; ++c) u[c] = fmaf(-s_LU[c][i], u[i], u[c]);
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
                if (idx < CW * CW) s_LU[idx / CW][idx % CW] = __uint_as_float(x[i]);
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
