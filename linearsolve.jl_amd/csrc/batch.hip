// batch.hip — thousands of independent small dense systems (n <= 512), each factored with
// partial pivoting and solved, in one persistent launch.
//
// Reference: the blockwise factorization of BlockDiagonalFactorization
// (ext/LinearSolveBlockDiagonalsExt.jl:101-114: lu!(B_i, pivot; check = false) per block,
// success = all(issuccess); :123-185 per-block ldiv!).  Per system the arithmetic of the FP64
// path is that of LAPACK dgetrf/dgetrs (pivot = first index of max |a|, NaN ignored; info = first
// zero pivot).  The FP32 path factors in FP32 and refines the solution in FP64
// (r = b - A x with the FP64 matrix, d = U \ L \ P r, x += d) until the normwise backward error
// ||r||_inf / (||A||_inf ||x||_inf + ||b||_inf) reaches the tolerance.
//
// Scheduling: systems are sorted by decreasing size on the host (longest first); a persistent
// grid of workgroups takes them from an atomic counter, one workgroup per system.
// Per system (BT threads): right-looking blocked LU with panels of BW columns factored in
// registers (pivot search on packed keys by DPP).  A panel's row interchanges are applied in
// memory to the columns right of it as U12 is formed (one thread per column moves the at most
// 2 BW rows involved), so the trailing update A22 -= L21 U12 reads and writes contiguous rows;
// it runs on MFMA: FP64 panels of 8 or 16 columns hold L21 in LDS and deal 64 x 32 tiles to the
// waves in turn; 32-column panels (and every FP32 panel) keep a 32-row block of L21 in each
// wave's registers and sweep the 32 x 32 tiles of its rows, U12 read from L2.  The L columns
// take the interchanges of later panels in one gather after the last panel (like dlaswp on the
// left, but once).  The panel width BW (8, 16 or 32) is chosen per system from a size-binned
// tile table (set by the autotuner through lsx_batch_set_tiles).  Triangular solves go by
// blocks of 64 (FP32: 32) rows: the diagonal block is staged in LDS and solved by one wave with
// shuffles, the rows beyond it are updated by all threads, sixteen loads in flight per row.
// The launch sequence of a handle (counter reset + persistent kernel) is captured once in a HIP
// graph and replayed by every factor_solve call (LSX_BATCH_NO_GRAPH=1 launches it directly).
#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include <limits>
#include <new>
#include <numeric>
#include <vector>

#include "batch.hpp"
#include "wave.hpp"

namespace lsx {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int BT = 256;            // threads per system
constexpr int BNMAX = 512;         // largest system
constexpr int RPT = (BNMAX + BT - 1) / BT;  // panel rows per thread
static_assert(RPT <= 2, "a pivot row is held by the first or the last of a thread's panel rows");

// Per factor type: systems (workgroups) per CU, the trailing-update block of a wave in 16 x 16
// tiles, whether L21 and the pivot row are held (L21 in LDS, the pivot row in registers) or read
// where they are used, and the row block SB of the triangular solves.
// FP64: two systems per CU (up to 256 registers, 80 KB of LDS each) with L21 in LDS and 64-row
// solve blocks.  FP32: three systems per CU (168 registers, 36 KB of LDS each: 17 KB of them
// stage U12 of 32-column panels), L21 read back from the factor in L2 and 32-row solve blocks;
// the extra systems hide the latency of its refinement passes.  (FP64 at four systems per CU
// spills registers and runs slower; FP32 at four spills 129 registers once U12 is staged.)
template <typename T> struct BatchCfg;
// Panels of 32 columns never hold L21 (it would take 136 KB of LDS); their trailing update
// (and every FP32 one) keeps a row block of L21 in registers and sweeps the column blocks of
// its rows, in RB2 x CB2 tiles.
// (Loading a wave's next tile while it forms the current one measured slower: the registers
// it takes cost spills and occupancy.)
template <> struct BatchCfg<double> {
    static constexpr int WG = 2, RB = 4, CB = 2, RB2 = 2, CB2 = 2, SB = 64;
    static constexpr bool HOLD = true;
};
template <> struct BatchCfg<float> {
    static constexpr int WG = 3, RB = 2, CB = 2, RB2 = 2, CB2 = 2, SB = 32;
    static constexpr bool HOLD = false;
};
constexpr int PMAX = 32;  // widest panel
constexpr int UCH_F = 128;  // FP32: columns of U12 staged in LDS at a time
// Whether a panel of width BW holds L21 in LDS and the pivot row in registers.
template <typename T, int BW> struct PanelHold { static constexpr bool value = BatchCfg<T>::HOLD && BW <= 16; };

// Panel-width index (0: 8, 1: 16, 2: 32 columns) of a system without a tile table (the widths
// the autotuner picks on MI355X for both factor types; the serving lane's batches use these).
__host__ __device__ inline int default_cfg(int n) { return n <= 64 ? 0 : (n <= 192 ? 1 : 2); }

struct BatchArgs {
    int64_t count;
    const int32_t* order;     // systems by decreasing n
    const int32_t* n;
    const uint8_t* cfg;       // per-system panel-width index (nullable: default by n)
    const int64_t* aoff;      // offset of A_i (n_i^2 elements)
    const int64_t* voff;      // offset of b_i / x_i / ipiv_i
    void* A;                  // workspace of the factor type: factored in place
    const double* A64;        // FP32 path: the FP64 matrices (refinement residuals)
    const double* b;
    double* x;
    int64_t* ipiv;            // LAPACK swap sequence (1-based), packed like b
    int32_t* perm;            // row i of P A is row perm[i] of A, packed like b
    int64_t* info;
    int32_t* iters;           // refinement steps per system (nullable)
    double* berr;             // FP32 path: final backward error per system (nullable)
    int maxref;
    double tol;
    uint32_t* counter;
};

template <typename T>
struct BatchShared {
    static constexpr int SB = BatchCfg<T>::SB;
    static_assert(BNMAX <= SB * (SB + 1), "the FP32 residual is kept in the solve staging block");
    static_assert(PMAX * (PMAX + 1) <= SB * (SB + 1), "L11 of the widest panel fits the staging block");
    union {
        double l21[BatchCfg<T>::HOLD ? BNMAX * 17 : 1];  // held L21 (row r of the panel at r * 17;
                                                         // its first rows are L11)
        double dblk[SB][SB + 1];   // staged diagonal block of the solves / FP32 path: the residual /
                                   // during the factorization, L11 of a panel without held L21
    };
    uint64_t wkey[BT / 64];   // pivot search: each wave's best key and row
    int wrow[BT / 64];
    double rowbuf[2][PMAX];    // pivot row and the row it displaces (panel columns)
    double red[2][BT / 64];
    int chg[PMAX];             // rows below the panel that its interchanges move
    int nchg;
    // FP32, 32-column panels: staged U12 of the trailing update (UCH_F columns, padded rows);
    // FP64 stages it in the held-L21 area instead
    T ustg[BatchCfg<T>::HOLD ? 1 : UCH_F * (PMAX + 1)];
    double xs[BNMAX];          // solve vector
    union {
        double xr[BNMAX];      // FP32 path: the refined solution (after the factorization)
        int piv[BNMAX];        // 0-based swap sequence (during the factorization)
    };
    int perm[BNMAX];
};

__device__ __forceinline__ int64_t uniform64(int64_t x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// MFMA of the factor type: the transposed 16 x 16 tile D = X Y (X(i, k) = -U12(k, col i),
// Y(k, j) = L21(row j, k)); lane l / slot q holds C(row rb + (l & 15), col cb + ccol(l, q)).
template <typename T> struct BMma;
template <> struct BMma<double> {
    typedef v4d acc_t;
    __device__ static acc_t mma(double x, double y, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0); }
    __device__ static int ccol(int lane, int q) { return 4 * q + (lane >> 4); }
};
template <> struct BMma<float> {
    typedef v4f acc_t;
    __device__ static acc_t mma(float x, float y, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0); }
    __device__ static int ccol(int lane, int q) { return 4 * (lane >> 4) + q; }
};

// Block-wide max of two non-negative values (all threads get both).
template <typename T>
__device__ __forceinline__ void block_max2(BatchShared<T>& sh, double& a, double& b) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a = fmax(a, __shfl_xor(a, off));
        b = fmax(b, __shfl_xor(b, off));
    }
    if (lane == 0) { sh.red[0][wave] = a; sh.red[1][wave] = b; }
    __syncthreads();
    a = sh.red[0][0];
    b = sh.red[1][0];
#pragma unroll
    for (int i = 1; i < BT / 64; ++i) { a = fmax(a, sh.red[0][i]); b = fmax(b, sh.red[1][i]); }
    __syncthreads();
}

// Column C of a panel held in registers (v: thread tid's logical rows tid + k BT): pivot
// search, the interchange of rows C and the pivot row through LDS, elimination below row C.
// Two barriers per column.  Instantiated for C = 0..BW-1 so that v is indexed by constants.
template <typename T, int BW, int C>
__device__ __forceinline__ void panel_col(BatchShared<T>& sh, T (&v)[RPT][BW], int m, int j0, int jb, int64_t vo,
                                          int64_t* ipiv, int64_t& info) {
    if constexpr (C < BW) {
        const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
        if (C < jb) {  // uniform: only the last panel is narrower
            // pivot: first row of max |a| among logical rows C..m-1 (NaN counts as 0), by packed
            // keys (0: no candidate): the thread's rows, the wave (DPP), then the four waves
            uint64_t key = 0;
            int krow = INT_MAX;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int r = tid + k * BT;
                if (r >= C && r < m) {
                    const uint64_t u = pivot_key(v[k][C]);
                    if (u > key) { key = u; krow = r; }  // rows increase with k: a tie keeps the first
                }
            }
            const int wl = wave_argmax(key, krow);
            if (lane == wl) { sh.wkey[wave] = key; sh.wrow[wave] = krow; }
            __syncthreads();
            uint64_t bk = sh.wkey[0];
            int pr = sh.wrow[0];
#pragma unroll
            for (int i = 1; i < BT / 64; ++i) {
                const uint64_t ki = sh.wkey[i];
                const int ri = sh.wrow[i];
                if (ki > bk || (ki == bk && ri < pr)) { bk = ki; pr = ri; }
            }
            // the pivot row and (when they differ) row C through LDS; row C is thread C's first row
            if (tid == (pr & (BT - 1))) {
#pragma unroll
                for (int q = 0; q < BW; ++q) {
                    T x = v[0][q];
                    if constexpr (RPT > 1) x = (pr >= BT) ? v[RPT - 1][q] : x;
                    sh.rowbuf[0][q] = (double)x;
                }
            }
            if (pr != C && tid == C) {
#pragma unroll
                for (int q = 0; q < BW; ++q) sh.rowbuf[1][q] = (double)v[0][q];
            }
            if (tid == 0) {
                const int k = j0 + C;
                sh.piv[k] = j0 + pr;
                if (ipiv) ipiv[vo + k] = (int64_t)(j0 + pr) + 1;
                if (pr != C) {
                    const int t = sh.perm[k];
                    sh.perm[k] = sh.perm[j0 + pr];
                    sh.perm[j0 + pr] = t;
                }
            }
            __syncthreads();
            // the pivot row: held in registers, or read from LDS where it is used
            T prh[BW];
#pragma unroll
            for (int q = 0; q < BW; ++q) prh[q] = PanelHold<T, BW>::value ? (T)sh.rowbuf[0][q] : T(0);
            auto prow = [&](int q) -> T { return PanelHold<T, BW>::value ? prh[q] : (T)sh.rowbuf[0][q]; };
            if (pr != C) {
                if (tid == C) {
#pragma unroll
                    for (int q = 0; q < BW; ++q) v[0][q] = prow(q);
                }
                if (tid == (pr & (BT - 1))) {
#pragma unroll
                    for (int q = 0; q < BW; ++q) {
                        const T t = (T)sh.rowbuf[1][q];
                        if constexpr (RPT > 1) {
                            if (pr >= BT) v[RPT - 1][q] = t;
                            else v[0][q] = t;
                        } else {
                            v[0][q] = t;
                        }
                    }
                }
            }
            const T pv = prow(C);
            if (pv == T(0) && info == 0) info = j0 + C + 1;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int r = tid + k * BT;
                if (r > C && r < m) {
                    T l = v[k][C];
                    if (pv != T(0)) {
                        if (fabs((double)pv) >= (double)std::numeric_limits<T>::min()) l *= (T(1) / pv);
                        else l /= pv;
                    }
                    v[k][C] = l;
#pragma unroll
                    for (int q = C + 1; q < BW; ++q) v[k][q] = fma(-l, prow(q), v[k][q]);
                }
            }
    }
        panel_col<T, BW, C + 1>(sh, v, m, j0, jb, vo, ipiv, info);
    }
}

// LU with partial pivoting of one n x n system (column-major, in place).  Each panel's row
// interchanges reach the columns right of it while U12 is formed, so the trailing update reads
// and writes rows in place (coalesced); the columns left of it (L) take all later interchanges in
// one gather at the end, which also leaves the system's permutation (logical row -> original row)
// in sh.perm.  On return A holds L and U in LAPACK's row order.
template <typename T, int BW>
__device__ __forceinline__ void factor_sys(BatchShared<T>& sh, T* A, int n, int64_t vo, int64_t* ipiv, int64_t& info) {
    static_assert(BW == 8 || BW == 16 || BW == PMAX, "panel width");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr bool HL = PanelHold<T, BW>::value;
    double (*L11)[PMAX + 1] = reinterpret_cast<double (*)[PMAX + 1]>(&sh.dblk[0][0]);
    for (int j0 = 0; j0 < n; j0 += BW) {
        const int jb = min(BW, n - j0);
        const int m = n - j0;
        // rows j0..n-1 of columns j0..n-1 are in place: the panel's permutation starts as identity
        for (int i = j0 + tid; i < n; i += BT) sh.perm[i] = i;
        // the panel in registers: thread tid holds logical rows tid + k BT, k < RPT (m <= RPT BT)
        T v[RPT][BW];
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = tid + k * BT;
#pragma unroll
            for (int q = 0; q < BW; ++q) v[k][q] = (r < m && q < jb) ? A[(j0 + q) * n + j0 + r] : T(0);
        }
        panel_col<T, BW, 0>(sh, v, m, j0, jb, vo, ipiv, info);
        // the factored panel, rows in LAPACK's order, to A; L11 also to LDS for U12, and L21 to
        // LDS when it is held (else the trailing update reads it back from A, out of L2)
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = tid + k * BT;
            if (r < m) {
#pragma unroll
                for (int q = 0; q < BW; ++q) {
                    if (q < jb) A[(j0 + q) * n + j0 + r] = v[k][q];
                    if constexpr (HL) sh.l21[r * 17 + q] = (double)v[k][q];
                    else if (k == 0 && r < BW) L11[r][q] = (double)v[k][q];
                }
            }
        }
        __syncthreads();
        const int c1 = j0 + jb;
        if (c1 < n) {
            // the rows below the panel that its swaps move: those it swaps with (at most jb), in
            // the order of the swaps, each once.  Lane k of wave 0 keeps swap k's row unless an
            // earlier swap names it; the kept rows are compacted by a ballot.
            if (wave == 0) {
                const int p = lane < jb ? sh.piv[j0 + lane] : -1;
                bool keep = p >= c1;
                for (int k = 0; k < jb; ++k) keep = keep && !(k < lane && sh.piv[j0 + k] == p);
                const unsigned long long kept = __ballot(keep);
                if (keep) sh.chg[__popcll(kept & ((1ull << lane) - 1ull))] = p;
                if (lane == 0) sh.nchg = __popcll(kept);
            }
            __syncthreads();
            const int nchg = sh.nchg;
            // here jb == BW (only the last panel is narrower, and it has no columns right of it).
            // Right of the panel, one thread per column: the interchanges, then U12 = L11^{-1} A12.
            // The new rows j0..c1-1 come from anywhere at or below j0, so they are read first.
            // A moved row below c1 takes the old content of a row in [j0, c1) (a row below the
            // panel only ever receives what a swap brings from the panel's own rows), which is
            // not written until U12 is, so the moves go in chunks once the U12 rows are read
            // (slots past nchg repeat the last move).  Offsets within a system fit 32 bits.
            for (int col = c1 + tid; col < n; col += BT) {
                const int co = col * n;
                T u[BW];
#pragma unroll
                for (int i = 0; i < BW; ++i) u[i] = A[co + sh.perm[j0 + i]];
#pragma unroll
                for (int t0 = 0; t0 < BW; t0 += 8) {
                    if (t0 < nchg) {
                        T w[8];
#pragma unroll
                        for (int t = 0; t < 8; ++t) w[t] = A[co + sh.perm[sh.chg[min(t0 + t, nchg - 1)]]];
#pragma unroll
                        for (int t = 0; t < 8; ++t) A[co + sh.chg[min(t0 + t, nchg - 1)]] = w[t];
                    }
                }
#pragma unroll
                for (int kk = 0; kk < BW; ++kk) {
#pragma unroll
                    for (int i = kk + 1; i < BW; ++i) {
                        double l;
                        if constexpr (HL) l = sh.l21[i * 17 + kk];
                        else l = L11[i][kk];
                        u[i] = fma(-(T)l, u[kk], u[i]);
                    }
                    __builtin_amdgcn_sched_barrier(0);  // L11 read column by column, not all at once
                }
#pragma unroll
                for (int i = 0; i < BW; ++i) A[co + j0 + i] = u[i];
            }
            __syncthreads();
            const int mr = n - c1;
            T* A22 = A + c1 * n + c1;
            const T* U12 = A + c1 * n + j0;
            const T* L21 = A + j0 * n + c1;
            using acc_t = typename BMma<T>::acc_t;
            constexpr int KS = BW / 4, NW = BT / 64;
            // C tile (R x C 16 x 16 blocks at rows rb, columns cb of A22) and its U12 operands
            // (lane l, k-step s: -U12(4 s + l / 16, column cb + 16 b + l % 16))
            auto load_c = [&](auto& c, int rb, int cb) {
                constexpr int R = sizeof(c) / sizeof(c[0]), C = sizeof(c[0]) / sizeof(c[0][0]);
#pragma unroll
                for (int a2 = 0; a2 < R; ++a2)
#pragma unroll
                    for (int b2 = 0; b2 < C; ++b2)
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int lr = rb + 16 * a2 + (lane & 15), lc = cb + 16 * b2 + BMma<T>::ccol(lane, q);
                            c[a2][b2][q] = (lr < mr && lc < mr) ? A22[lc * n + lr] : T(0);
                        }
            };
            auto load_u = [&](auto& x, int cb) {
                constexpr int C = sizeof(x) / sizeof(x[0]);
#pragma unroll
                for (int b2 = 0; b2 < C; ++b2) {
                    const int lc = cb + 16 * b2 + (lane & 15);
#pragma unroll
                    for (int k4 = 0; k4 < KS; ++k4) x[b2][k4] = (lc < mr) ? -U12[lc * n + k4 * 4 + (lane >> 4)] : T(0);
                }
            };
            auto store_c = [&](const auto& c, int rb, int cb) {
                constexpr int R = sizeof(c) / sizeof(c[0]), C = sizeof(c[0]) / sizeof(c[0][0]);
#pragma unroll
                for (int a2 = 0; a2 < R; ++a2)
#pragma unroll
                    for (int b2 = 0; b2 < C; ++b2)
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int lr = rb + 16 * a2 + (lane & 15), lc = cb + 16 * b2 + BMma<T>::ccol(lane, q);
                            if (lr < mr && lc < mr) A22[lc * n + lr] = c[a2][b2][q];
                        }
            };
            if constexpr (HL) {
                // A22 -= L21 U12 on MFMA: tiles of RBT x CBT 16 x 16 blocks, waves round-robin,
                // L21 from LDS
                constexpr int RBT = BatchCfg<T>::RB, CBT = BatchCfg<T>::CB;
                const int tbr = (mr + 16 * RBT - 1) / (16 * RBT), tbc = (mr + 16 * CBT - 1) / (16 * CBT);
                for (int t = wave; t < tbr * tbc; t += NW) {
                    const int rb = (t % tbr) * (16 * RBT), cb = (t / tbr) * (16 * CBT);
                    acc_t acc[RBT][CBT];
                    T xv[CBT][KS];
                    load_c(acc, rb, cb);
                    load_u(xv, cb);
#pragma unroll
                    for (int k4 = 0; k4 < KS; ++k4) {
                        const int kk = k4 * 4 + (lane >> 4);
                        T yv[RBT];
#pragma unroll
                        for (int a2 = 0; a2 < RBT; ++a2) {
                            const int lr = rb + 16 * a2 + (lane & 15);
                            yv[a2] = lr < mr ? (T)sh.l21[(BW + lr) * 17 + kk] : T(0);
                        }
#pragma unroll
                        for (int a2 = 0; a2 < RBT; ++a2)
#pragma unroll
                            for (int b2 = 0; b2 < CBT; ++b2) acc[a2][b2] = BMma<T>::mma(xv[b2][k4], yv[a2], acc[a2][b2]);
                    }
                    store_c(acc, rb, cb);
                }
            } else {
                // A22 -= L21 U12 on MFMA by row blocks of 16 R2 rows: a wave loads its row block
                // of L21 into registers once and sweeps the column blocks (16 C2 columns) of a
                // segment of A22's columns; with fewer row blocks than waves, the columns are cut
                // into segments so that every wave has work.  32-column panels stage U12 in LDS,
                // UCH columns at a time (FP64: in the held-L21 area; FP32: in a buffer of its
                // own), so the waves do not read it again from L2 or HBM for every row block;
                // narrower FP32 panels read it per tile (out of L2).
                constexpr int R2 = BatchCfg<T>::RB2, C2 = BatchCfg<T>::CB2;
                constexpr bool HOLDT = BatchCfg<T>::HOLD;
                constexpr bool STG = BW == PMAX;
                constexpr int UCH = !STG ? BNMAX : (HOLDT ? (BNMAX * 17) / (BW + 1) / (16 * C2) * (16 * C2) : UCH_F);
                static_assert(UCH >= 16 * C2 && UCH % (16 * C2) == 0, "a staged U12 chunk holds whole column blocks");
                auto ubuf = [&]() -> T* {
                    if constexpr (HOLDT) return reinterpret_cast<T*>(&sh.l21[0]);
                    else return &sh.ustg[0];
                };
                const int tbr = (mr + 16 * R2 - 1) / (16 * R2);
                for (int cc0 = 0; cc0 < mr; cc0 += UCH) {
                    const int ncc = min(UCH, mr - cc0);
                    if constexpr (STG) {
                        T* Us = ubuf();   // Us[c (BW + 1) + k] = -U12(k, column cc0 + c)
                        __syncthreads();  // the previous chunk's tiles are done with Us
                        for (int base = tid; base < ncc * BW; base += 8 * BT) {
                            T t8[8];
#pragma unroll
                            for (int u8 = 0; u8 < 8; ++u8) {
                                const int idx = base + u8 * BT, c = idx / BW, k = idx - c * BW;
                                t8[u8] = idx < ncc * BW ? U12[(cc0 + c) * n + k] : T(0);
                            }
#pragma unroll
                            for (int u8 = 0; u8 < 8; ++u8) {
                                const int idx = base + u8 * BT, c = idx / BW, k = idx - c * BW;
                                if (idx < ncc * BW) Us[c * (BW + 1) + k] = -t8[u8];
                            }
                        }
                        __syncthreads();
                    }
                    const int tbc = (ncc + 16 * C2 - 1) / (16 * C2);
                    const int nseg = min(tbc, (NW + tbr - 1) / tbr);
                    const int segw = (tbc + nseg - 1) / nseg;
                    for (int u = wave; u < tbr * nseg; u += NW) {
                        const int rb = (u % tbr) * (16 * R2), sg = u / tbr;
                        T yl[R2][KS];
#pragma unroll
                        for (int a2 = 0; a2 < R2; ++a2) {
                            const int lr = rb + 16 * a2 + (lane & 15);
#pragma unroll
                            for (int k4 = 0; k4 < KS; ++k4) yl[a2][k4] = lr < mr ? L21[(k4 * 4 + (lane >> 4)) * n + lr] : T(0);
                        }
                        const int cbs = sg * segw, cbe = min(tbc, (sg + 1) * segw);
                        for (int cbi = cbs; cbi < cbe; ++cbi) {
                            const int cb = cc0 + cbi * (16 * C2);
                            acc_t acc[R2][C2];
                            T xv[C2][KS];
                            load_c(acc, rb, cb);
                            if constexpr (STG) {
                                const T* Us = ubuf();
#pragma unroll
                                for (int b2 = 0; b2 < C2; ++b2) {
                                    const int lc = cb + 16 * b2 + (lane & 15);
#pragma unroll
                                    for (int k4 = 0; k4 < KS; ++k4)
                                        xv[b2][k4] = (lc < mr) ? Us[(lc - cc0) * (BW + 1) + k4 * 4 + (lane >> 4)] : T(0);
                                }
                            } else {
                                load_u(xv, cb);
                            }
#pragma unroll
                            for (int k4 = 0; k4 < KS; ++k4)
#pragma unroll
                                for (int a2 = 0; a2 < R2; ++a2)
#pragma unroll
                                    for (int b2 = 0; b2 < C2; ++b2)
                                        acc[a2][b2] = BMma<T>::mma(xv[b2][k4], yl[a2][k4], acc[a2][b2]);
                            store_c(acc, rb, cb);
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    // The columns of panel p hold their rows below the panel in the order after panel p; later
    // panels swapped those rows.  S (kept in sh.perm) is the composition of the swaps after panel
    // p, built from the last panel back (row i of the final order is row S(i) of that order).
    for (int i = tid; i < n; i += BT) sh.perm[i] = i;
    __syncthreads();
    bool moved = false;  // some later swap exchanges two different rows
    for (int p = (n - 1) / BW; p >= 0; --p) {
        const int j0 = p * BW, jb = min(BW, n - j0), c1 = j0 + jb, m2 = n - c1;
        if (moved) {
            // thread tid moves rows c1 + tid + k BT, k < RPT (m2 < RPT BT), of the panel's columns,
            // through registers, GC columns at a time: in each group every source is read before
            // any row is written
            constexpr int GC = BW < 16 ? BW : 16;
#pragma unroll
            for (int cg = 0; cg < BW; cg += GC) {
                T g[RPT][GC];
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const int r = tid + k * BT;
#pragma unroll
                    for (int c = 0; c < GC; ++c)
                        g[k][c] = (r < m2 && cg + c < jb) ? A[(j0 + cg + c) * n + sh.perm[c1 + r]] : T(0);
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const int r = tid + k * BT;
#pragma unroll
                    for (int c = 0; c < GC; ++c)
                        if (r < m2 && cg + c < jb) A[(j0 + cg + c) * n + c1 + r] = g[k][c];
                }
            }
        }
        // S := tau(j0) o ... o tau(c1 - 1) o S, tau(k) exchanging rows k and piv[k]
        int pk[BW];
#pragma unroll
        for (int k = 0; k < BW; ++k) pk[k] = sh.piv[min(j0 + k, c1 - 1)];
        for (int i = j0 + tid; i < n; i += BT) {
            int x = sh.perm[i];
#pragma unroll
            for (int k = BW - 1; k >= 0; --k)
                if (k < jb) x = (x == j0 + k) ? pk[k] : (x == pk[k] ? j0 + k : x);
            sh.perm[i] = x;
        }
        for (int k = j0; k < c1; ++k) moved = moved || sh.piv[k] != k;
        __syncthreads();
    }
}

// s - sum_{k0 <= k < k1} a[k * n] x[k] in the order of k (FP64), sixteen loads in flight.
template <typename T>
__device__ __forceinline__ double row_dot_sub(double s, const T* a, int n, const double* x, int k0, int k1) {
    int k = k0;
    for (; k + 16 <= k1; k += 16) {
        T av[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) av[u] = a[(int64_t)(k + u) * n];
#pragma unroll
        for (int u = 0; u < 16; ++u) s = fma(-(double)av[u], x[k + u], s);
    }
    for (; k < k1; ++k) s = fma(-(double)a[(int64_t)k * n], x[k], s);
    return s;
}

// The kb x kb diagonal block at (k0, k0) into D (FP64); a full block with all of a thread's
// loads in flight before its LDS stores.
template <typename T, int SB>
__device__ __forceinline__ void stage_block(double (*D)[SB + 1], const T* A, int n, int k0, int kb) {
    const int tid = threadIdx.x;
    if (kb == SB) {
        T v[SB * SB / BT];
#pragma unroll
        for (int u = 0; u < SB * SB / BT; ++u) {
            const int idx = tid + u * BT, r = idx % SB, c = idx / SB;
            v[u] = A[(int64_t)(k0 + c) * n + k0 + r];
        }
#pragma unroll
        for (int u = 0; u < SB * SB / BT; ++u) {
            const int idx = tid + u * BT;
            D[idx % SB][idx / SB] = (double)v[u];
        }
    } else {
        for (int idx = tid; idx < kb * kb; idx += BT) {
            const int r = idx % kb, c = idx / kb;
            D[r][c] = (double)A[(int64_t)(k0 + c) * n + k0 + r];
        }
    }
}

// sh.xs := U^{-1} L^{-1} sh.xs (sh.xs already holds P b), arithmetic in FP64.
template <typename T>
__device__ __forceinline__ void solve_sys(BatchShared<T>& sh, const T* A, int n) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int SB = BatchCfg<T>::SB;
    static_assert((SB * SB) % BT == 0, "a full diagonal block is whole loads per thread");
    double (*D)[SB + 1] = sh.dblk;
    // forward: unit lower L
    for (int k0 = 0; k0 < n; k0 += SB) {
        const int kb = min(SB, n - k0);
        stage_block<T, SB>(D, A, n, k0, kb);
        __syncthreads();
        if (wave == 0) {
            double xv = lane < kb ? sh.xs[k0 + lane] : 0.0;
            for (int c = 0; c < kb; ++c) {
                const double xc = __shfl(xv, c);
                if (lane > c && lane < kb) xv = fma(-D[lane][c], xc, xv);
            }
            if (lane < kb) sh.xs[k0 + lane] = xv;
        }
        __syncthreads();
        for (int i = k0 + kb + tid; i < n; i += BT) sh.xs[i] = row_dot_sub(sh.xs[i], A + i, n, sh.xs, k0, k0 + kb);
        __syncthreads();
    }
    // backward: U
    for (int k1 = n; k1 > 0; k1 -= SB) {
        const int k0 = max(0, k1 - SB), kb = k1 - k0;
        stage_block<T, SB>(D, A, n, k0, kb);
        __syncthreads();
        if (wave == 0) {
            double xv = lane < kb ? sh.xs[k0 + lane] : 0.0;
            for (int c = kb - 1; c >= 0; --c) {
                if (lane == c) xv = xv / D[c][c];
                const double xc = __shfl(xv, c);
                if (lane < c) xv = fma(-D[lane][c], xc, xv);
            }
            if (lane < kb) sh.xs[k0 + lane] = xv;
        }
        __syncthreads();
        for (int i = tid; i < k0; i += BT) sh.xs[i] = row_dot_sub(sh.xs[i], A + i, n, sh.xs, k0, k1);
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(BT, BatchCfg<T>::WG * BT / 256) void k_batch_lu(BatchArgs a) {  // min waves per SIMD
    __shared__ BatchShared<T> sh;
    __shared__ int s_item;
    const int tid = threadIdx.x;
    while (true) {
        if (tid == 0) s_item = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        // the system's scalars are wave-uniform (held in scalar registers)
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= a.count) break;
        const int sys = __builtin_amdgcn_readfirstlane(a.order[item]);
        const int n = __builtin_amdgcn_readfirstlane(a.n[sys]);
        const int64_t ao = uniform64(a.aoff[sys]);
        T* A = reinterpret_cast<T*>(a.A) + ao;
        const int64_t vo = uniform64(a.voff[sys]);
        int64_t info = 0;
        const double* A64 = a.A64 ? a.A64 + ao : nullptr;
        double an = 0.0;  // FP32 path: this thread's rows' largest sum |a_ij| (||A||_inf)
        if (A64) {  // FP32 path: working copy of the FP64 matrix, by rows, with the row sums
            for (int i = tid; i < n; i += BT) {
                double rs = 0.0;
                int j = 0;
                for (; j + 16 <= n; j += 16) {
                    double v[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u) v[u] = A64[(j + u) * n + i];
#pragma unroll
                    for (int u = 0; u < 16; ++u) {
                        A[(j + u) * n + i] = (T)v[u];
                        rs += fabs(v[u]);
                    }
                }
                for (; j < n; ++j) {
                    const double v = A64[j * n + i];
                    A[j * n + i] = (T)v;
                    rs += fabs(v);
                }
                an = fmax(an, rs);
            }
        }
        for (int i = tid; i < n; i += BT) sh.perm[i] = i;
        __syncthreads();
        const int cfg = a.cfg ? a.cfg[sys] : default_cfg(n);
        if (cfg == 0) factor_sys<T, 8>(sh, A, n, vo, a.ipiv, info);
        else if (cfg == 1) factor_sys<T, 16>(sh, A, n, vo, a.ipiv, info);
        else factor_sys<T, PMAX>(sh, A, n, vo, a.ipiv, info);
        // solve: x = U^{-1} L^{-1} P b
        for (int i = tid; i < n; i += BT) sh.xs[i] = a.b[vo + sh.perm[i]];
        __syncthreads();
        solve_sys<T>(sh, A, n);
        if (A64) {
            // FP64 refinement
            double bn = 0.0;
            for (int i = tid; i < n; i += BT) bn = fmax(bn, fabs(a.b[vo + i]));
            block_max2(sh, an, bn);
            for (int i = tid; i < n; i += BT) sh.xr[i] = sh.xs[i];
            __syncthreads();
            double* rr = &sh.dblk[0][0];  // residual by original row (n <= (SB + 1) SB)
            int it = 0;
            double be = 0.0;
            while (true) {
                double rn = 0.0, xn = 0.0;
                for (int i = tid; i < n; i += BT) {
                    const double s = row_dot_sub(a.b[vo + i], A64 + i, n, sh.xr, 0, n);
                    rr[i] = s;
                    rn = fmax(rn, fabs(s));
                    xn = fmax(xn, fabs(sh.xr[i]));
                }
                block_max2(sh, rn, xn);
                const double den = an * xn + bn;
                be = den > 0.0 ? rn / den : 0.0;
                if (!(be > a.tol) || it >= a.maxref || info != 0) break;
                for (int i = tid; i < n; i += BT) sh.xs[i] = rr[sh.perm[i]];
                __syncthreads();
                solve_sys<T>(sh, A, n);
                for (int i = tid; i < n; i += BT) sh.xr[i] += sh.xs[i];
                __syncthreads();
                ++it;
            }
            for (int i = tid; i < n; i += BT) a.x[vo + i] = sh.xr[i];
            if (tid == 0) {
                if (a.iters) a.iters[sys] = it;
                if (a.berr) a.berr[sys] = be;
            }
        } else {
            for (int i = tid; i < n; i += BT) a.x[vo + i] = sh.xs[i];
            if (tid == 0) {
                if (a.iters) a.iters[sys] = 0;
                if (a.berr) a.berr[sys] = -1.0;
            }
        }
        for (int i = tid; i < n; i += BT) a.perm[vo + i] = sh.perm[i];
        if (tid == 0) a.info[sys] = info;
        __syncthreads();
    }
}

__global__ void k_reset_counter(uint32_t* c) { *c = 0u; }

int batch_launch_raw(lsx_ctx* ctx, hipStream_t st, int64_t count, const int32_t* order, const int32_t* n,
                     const int64_t* aoff, const int64_t* voff, double* A, const double* b, double* x, int64_t* ipiv,
                     int32_t* perm, int64_t* info, uint32_t* counter) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, counter);
    BatchArgs a{};
    a.count = count;
    a.order = order;
    a.n = n;
    a.cfg = nullptr;
    a.aoff = aoff;
    a.voff = voff;
    a.A = A;
    a.b = b;
    a.x = x;
    a.ipiv = ipiv;
    a.perm = perm;
    a.info = info;
    a.counter = counter;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(count, (int64_t)ctx->num_cus * BatchCfg<double>::WG));
    hipLaunchKernelGGL(k_batch_lu<double>, dim3(grid), dim3(BT), 0, st, a);
    LSX_HIP(ctx, hipGetLastError());
    return LSX_OK;
}

// The launch sequence of a handle: counter reset + the persistent kernel on the handle's buffers.
static void batch_enqueue(lsx_batch* B) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, B->stream, B->counter);
    BatchArgs a{};
    a.count = B->count;
    a.order = B->order;
    a.n = B->n;
    a.cfg = B->cfg;
    a.aoff = B->aoff;
    a.voff = B->voff;
    a.b = B->b;
    a.x = B->x;
    a.ipiv = B->ipiv;
    a.perm = B->perm;
    a.info = B->info;
    a.iters = B->iters;
    a.berr = B->berr;
    a.maxref = B->maxref;
    a.tol = B->tol;
    a.counter = B->counter;
    auto grid = [&](int wg) { return (int)std::max<int64_t>(1, std::min<int64_t>(B->count, (int64_t)B->ctx->num_cus * wg)); };
    if (B->dtype == LSX_F64) {
        a.A = B->A;
        a.A64 = nullptr;
        hipLaunchKernelGGL(k_batch_lu<double>, dim3(grid(BatchCfg<double>::WG)), dim3(BT), 0, B->stream, a);
    } else {
        a.A = B->A32;
        a.A64 = B->A;
        hipLaunchKernelGGL(k_batch_lu<float>, dim3(grid(BatchCfg<float>::WG)), dim3(BT), 0, B->stream, a);
    }
}

int batch_run(lsx_batch* B, const double* dA_src, const double* db, double* dx, bool copy_a) {
    lsx_ctx* ctx = B->ctx;
    if (copy_a)
        LSX_HIP(ctx, hipMemcpyAsync(B->A, dA_src, sizeof(double) * B->total_a, hipMemcpyDefault, B->stream));
    if (db != B->b) LSX_HIP(ctx, hipMemcpyAsync(B->b, db, sizeof(double) * B->total_v, hipMemcpyDefault, B->stream));
    const int pi = B->prof.begin(B->stream, PK_OTHER, B->flops, sizeof(double) * 2.0 * (double)B->total_a);
    if (B->use_graph) {
        if (!B->graph_exec) {
            hipGraph_t g = nullptr;
            LSX_HIP(ctx, hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
            batch_enqueue(B);
            LSX_HIP(ctx, hipStreamEndCapture(B->stream, &g));
            const hipError_t e = hipGraphInstantiate(&B->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            LSX_HIP(ctx, e);
        }
        LSX_HIP(ctx, hipGraphLaunch(B->graph_exec, B->stream));
    } else {
        batch_enqueue(B);
    }
    B->prof.end(B->stream, pi);
    LSX_HIP(ctx, hipGetLastError());
    if (dx != B->x) LSX_HIP(ctx, hipMemcpyAsync(dx, B->x, sizeof(double) * B->total_v, hipMemcpyDefault, B->stream));
    return LSX_OK;
}

// Per-system panel-width indices from a size-binned table (first bin with n <= upper).
static int tiles_to_cfg(lsx_ctx* ctx, const std::vector<int32_t>& hn, int nbins, const int32_t* upper,
                        const int32_t* bw, std::vector<uint8_t>& cfg) {
    for (int k = 0; k < nbins; ++k)
        if (bw[k] != 8 && bw[k] != 16 && bw[k] != PMAX)
            return fail(ctx, LSX_EINVAL, "lsx_batch_set_tiles: panel width %d is not 8, 16 or 32", bw[k]);
    cfg.resize(hn.size());
    for (size_t i = 0; i < hn.size(); ++i) {
        const int n = hn[i];
        int c = default_cfg(n);
        for (int k = 0; k < nbins; ++k) {
            if (n <= upper[k]) {
                c = bw[k] == 8 ? 0 : (bw[k] == 16 ? 1 : 2);
                break;
            }
        }
        cfg[i] = (uint8_t)c;
    }
    return LSX_OK;
}

}  // namespace lsx

using namespace lsx;

static void batch_free(lsx_batch* B) {
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    if (B->graph_exec) (void)hipGraphExecDestroy(B->graph_exec);
    (void)hipFree(B->order);
    (void)hipFree(B->n);
    (void)hipFree(B->cfg);
    (void)hipFree(B->aoff);
    (void)hipFree(B->voff);
    (void)hipFree(B->A);
    (void)hipFree(B->A32);
    (void)hipFree(B->b);
    (void)hipFree(B->x);
    (void)hipFree(B->ipiv);
    (void)hipFree(B->perm);
    (void)hipFree(B->info);
    (void)hipFree(B->iters);
    (void)hipFree(B->berr);
    (void)hipFree(B->counter);
    if (B->hinfo) (void)hipHostFree(B->hinfo);
    if (B->stream) (void)hipStreamDestroy(B->stream);
}

extern "C" {

int lsx_batch_create(lsx_ctx* ctx, int64_t count, const int64_t* n, int dtype, lsx_batch** out) {
    if (!ctx || !out || !n || count < 1) return LSX_EINVAL;
    *out = nullptr;
    if (dtype != LSX_F64 && dtype != LSX_F32)
        return fail(ctx, LSX_EINVAL, "lsx_batch_create: dtype must be LSX_F64 or LSX_F32");
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
    std::vector<uint8_t> cfg(count);
    for (int64_t i = 0; i < count; ++i) cfg[i] = (uint8_t)default_cfg(hn[i]);
    DeviceGuard g(ctx);
    lsx_batch* B = new (std::nothrow) lsx_batch();
    if (!B) return LSX_ENOMEM;
    B->ctx = ctx;
    B->dtype = dtype;
    B->count = count;
    B->total_a = ta;
    B->total_v = tv;
    B->flops = flops;
    B->hn = hn;
    const char* ng = std::getenv("LSX_BATCH_NO_GRAPH");
    B->use_graph = !(ng && ng[0] == '1');
    auto bad = [&](int rc) {
        batch_free(B);
        delete B;
        return rc;
    };
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&B->order, sizeof(int32_t) * count) != hipSuccess ||
        hipMalloc(&B->n, sizeof(int32_t) * count) != hipSuccess ||
        hipMalloc(&B->cfg, sizeof(uint8_t) * count) != hipSuccess ||
        hipMalloc(&B->aoff, sizeof(int64_t) * count) != hipSuccess ||
        hipMalloc(&B->voff, sizeof(int64_t) * count) != hipSuccess ||
        hipMalloc(&B->A, sizeof(double) * ta) != hipSuccess ||
        (dtype == LSX_F32 && hipMalloc(&B->A32, sizeof(float) * ta) != hipSuccess) ||
        hipMalloc(&B->b, sizeof(double) * tv) != hipSuccess ||
        hipMalloc(&B->x, sizeof(double) * tv) != hipSuccess || hipMalloc(&B->ipiv, sizeof(int64_t) * tv) != hipSuccess ||
        hipMalloc(&B->perm, sizeof(int32_t) * tv) != hipSuccess ||
        hipMalloc(&B->info, sizeof(int64_t) * count) != hipSuccess ||
        hipMalloc(&B->iters, sizeof(int32_t) * count) != hipSuccess ||
        hipMalloc(&B->berr, sizeof(double) * count) != hipSuccess ||
        hipMalloc(&B->counter, 256) != hipSuccess ||
        hipHostMalloc(&B->hinfo, sizeof(int64_t) * count, hipHostMallocDefault) != hipSuccess)
        return bad(fail(ctx, LSX_ENOMEM, "lsx_batch_create: device allocation failed"));
    if (hipMemcpy(B->order, order.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->n, hn.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->cfg, cfg.data(), sizeof(uint8_t) * count, hipMemcpyHostToDevice) != hipSuccess ||
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

int lsx_batch_set_tiles(lsx_batch* B, int32_t nbins, const int32_t* upper, const int32_t* bw) {
    if (!B || nbins < 0 || (nbins > 0 && (!upper || !bw))) return LSX_EINVAL;
    lsx_ctx* ctx = B->ctx;
    DeviceGuard g(ctx);
    std::vector<uint8_t> cfg;
    const int rc = tiles_to_cfg(ctx, B->hn, nbins, upper, bw, cfg);
    if (rc) return rc;
    LSX_HIP(ctx, hipMemcpyAsync(B->cfg, cfg.data(), sizeof(uint8_t) * B->count, hipMemcpyHostToDevice, B->stream));
    LSX_HIP(ctx, hipStreamSynchronize(B->stream));
    return LSX_OK;
}

int lsx_batch_set_refinement(lsx_batch* B, double tol, int32_t max_iters) {
    if (!B || !(tol >= 0.0) || max_iters < 0) return LSX_EINVAL;
    B->tol = tol;
    B->maxref = max_iters;
    if (B->graph_exec) {  // the values are kernel arguments of the captured launch
        (void)hipGraphExecDestroy(B->graph_exec);
        B->graph_exec = nullptr;
    }
    return LSX_OK;
}

int lsx_batch_factor_solve(lsx_batch* B, const double* A, const double* b, double* x, int64_t* info,
                           int64_t* ipiv) {
    if (!B || !A || !b || !x) return LSX_EINVAL;
    lsx_ctx* ctx = B->ctx;
    DeviceGuard g(ctx);
    int rc = batch_run(B, A, b, x, true);
    if (rc) return rc;
    LSX_HIP(ctx, hipMemcpyAsync(B->hinfo, B->info, sizeof(int64_t) * B->count, hipMemcpyDeviceToHost, B->stream));
    if (info) LSX_HIP(ctx, hipMemcpyAsync(info, B->info, sizeof(int64_t) * B->count, hipMemcpyDefault, B->stream));
    if (ipiv) LSX_HIP(ctx, hipMemcpyAsync(ipiv, B->ipiv, sizeof(int64_t) * B->total_v, hipMemcpyDefault, B->stream));
    LSX_HIP(ctx, hipStreamSynchronize(B->stream));
    B->prof.collect();
    int nfail = 0;
    for (int64_t i = 0; i < B->count; ++i) nfail += B->hinfo[i] > 0;
    return nfail;
}

int lsx_batch_stats(lsx_batch* B, int32_t* iters, double* backward_error) {
    if (!B) return LSX_EINVAL;
    lsx_ctx* ctx = B->ctx;
    DeviceGuard g(ctx);
    if (iters) LSX_HIP(ctx, hipMemcpyAsync(iters, B->iters, sizeof(int32_t) * B->count, hipMemcpyDefault, B->stream));
    if (backward_error)
        LSX_HIP(ctx, hipMemcpyAsync(backward_error, B->berr, sizeof(double) * B->count, hipMemcpyDefault, B->stream));
    LSX_HIP(ctx, hipStreamSynchronize(B->stream));
    return LSX_OK;
}

}  // extern "C"
