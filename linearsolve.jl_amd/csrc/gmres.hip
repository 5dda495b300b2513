// gmres.hip — SimpleGMRES (restarted GMRES, modified Gram-Schmidt Arnoldi, symmetric Givens)
// on a CSR matrix, one persistent kernel per solve.
//
// Follows src/simplegmres.jl (scalar layout) step for step:
//   reinit / _init_cacheval (:110-129, :420-462): x = 0, r0 = Pl \ b, beta = ||r0||,
//     eps = abstol + reltol * beta;
//   outer restart loop (:500-601): on passes >= 2, w = b - A x, r0 = Pl \ w;
//   Arnoldi step (:521-574): q = Pl \ (A (Pr \ v_k)); MGS h_i = <v_i, q>, q -= h_i v_i (i = 1..k);
//     Hbis = ||q||; previous Givens applied to the new column, new Givens (_sym_givens
//     :140-162); zeta update; stopping tests rNorm <= eps, rNorm + 1 <= 1, Hbis <= eps^(3/4);
//     v_{k+1} = q / Hbis;
//   back-substitution R y = z with zero for |R_ii| <= btol (:576-585), x_r += sum y_i v_i,
//     x_r = Pr \ x_r, x += x_r (:587-595); retcodes (:603-606).
//
// Layout: workgroup w of P owns rows [w R, (w+1) R), R = GBS * RPT (GBS = 512 threads); thread
// t holds rows w R + t + i GBS (i < RPT) of q and of the current basis vector in LDS.  Every
// workgroup keeps its own copy of the Hessenberg/Givens scalars (R, c, s, z) in LDS and
// computes them identically from all-reduced values, so all workgroups take the same branches.
// All-reduces combine per-workgroup partial sums in a fixed order (deterministic, the same
// result in every workgroup) through flag-in-word messages; basis vectors written by one
// workgroup and read by the SpMV of another are ordered by a counter barrier with agent-scope
// release/acquire fences.
#include <algorithm>
#include <cfloat>
#include <new>
#include <vector>
#include <cmath>

#include "gmres.hpp"

namespace lsx {

constexpr int GBS = 512;            // threads per workgroup
constexpr int GNW = GBS / 64;
constexpr int GLEAD = 16;           // from this many workgroups on, all-reduces go through workgroup 0

struct GShared {
    double wsum[2][GNW];
    double lsum[2][GNW];        // workgroup 0 from GLEAD workgroups on: wave sums of the partials
    uint32_t rbuf[2][2 * 256];  // 2 words per workgroup, up to 256 workgroups
    int abort;                  // a wait of this solve gave up: leave the solve
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <int RPT>
struct GmresCtx {
    const GmresArgs& a;
    GShared& sh;
    int tid, lane, wave, w;
    int64_t row0;     // first row of this workgroup
    uint32_t nred = 0, nbar = 0;
    bool dead = false;  // a wait of this solve gave up (uniform in the workgroup): its results
                        // are invalid and every loop of the solve ends

    __device__ GmresCtx(const GmresArgs& a_, GShared& sh_, int w_) : a(a_), sh(sh_) {
        tid = threadIdx.x;
        lane = tid & 63;
        wave = tid >> 6;
        w = w_;
        row0 = (int64_t)w * GBS * RPT;
    }
    __device__ int64_t row(int i) const { return row0 + tid + (int64_t)i * GBS; }

    // Sum of v over all threads of all workgroups, identical in every thread.
    __device__ double allreduce(double v) {
        const int par = (int)(nred & 1);
        ++nred;
        v = wave_sum(v);
        if (lane == 0) sh.wsum[par][wave] = v;
        __syncthreads();
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < GNW; ++i) s += sh.wsum[par][i];
        if (a.P == 1) return s;
        const uint32_t seq = nred;
        uint64_t* slot = a.red + (size_t)par * a.P * 2;
        if (tid == 0) {
            st_relaxed_u64(slot + 2 * w, ll_pack((uint32_t)__double2loint(s), seq));
            st_relaxed_u64(slot + 2 * w + 1, ll_pack((uint32_t)__double2hiint(s), seq));
        }
        // Small P: every workgroup reads all P partials and sums them in order.  From GLEAD
        // workgroups on, workgroup 0 alone reads them (thread p reads partial p), sums them by a
        // fixed tree (wave sums, then the waves in order) and publishes the total, so the others
        // poll two words instead of 2P (the polling traffic grows as P^2 otherwise).
        const bool lead = a.P >= GLEAD;
        if (!lead) {
            if (tid < 2 * a.P) sh.rbuf[par][tid] = poll(slot + tid, seq);
            __syncthreads();
            dead = sh.abort != 0;
            double tot = 0.0;
            for (int p = 0; p < a.P; ++p)
                tot += __hiloint2double((int)sh.rbuf[par][2 * p + 1], (int)sh.rbuf[par][2 * p]);
            return tot;
        }
        if (w == 0) {
            double pv = 0.0;
            if (tid < a.P) {
                const uint32_t lo = poll(slot + 2 * tid, seq), hi = poll(slot + 2 * tid + 1, seq);
                pv = __hiloint2double((int)hi, (int)lo);
            }
            pv = wave_sum(pv);
            if (lane == 0) sh.lsum[par][wave] = pv;
            __syncthreads();
            dead = sh.abort != 0;
            double tot = 0.0;
#pragma unroll
            for (int i = 0; i < GNW; ++i) tot += sh.lsum[par][i];
            if (tid == 0) {
                uint64_t* res = a.red + 4 * (int64_t)a.P + 8 + 2 * par;
                st_relaxed_u64(res, ll_pack((uint32_t)__double2loint(tot), seq));
                st_relaxed_u64(res + 1, ll_pack((uint32_t)__double2hiint(tot), seq));
            }
            return tot;
        }
        if (tid < 2) sh.rbuf[par][tid] = poll(a.red + 4 * (int64_t)a.P + 8 + 2 * par + tid, seq);
        __syncthreads();
        dead = sh.abort != 0;
        return __hiloint2double((int)sh.rbuf[par][1], (int)sh.rbuf[par][0]);
    }

    // Payload of a flag-in-word message once it carries sequence number seq.  The wait ends at
    // the bound, or at once when any wait of this solve has given up (the solve's error word is
    // set); either way the workgroup's abort flag is raised, read after the caller's barrier.
    __device__ uint32_t poll(const uint64_t* word, uint32_t seq) const {
        uint64_t x = ld_relaxed_u64(word);
        if ((uint32_t)(x >> 32) != seq) {
            const int64_t t0 = wall_clock64();
            while (true) {
                x = ld_relaxed_u64(word);
                if ((uint32_t)(x >> 32) == seq) break;
                if (wall_clock64() - t0 > a.timeout) {
                    atomicOr((unsigned long long*)&a.out->err, 1ull);
                    sh.abort = 1;
                    break;
                }
                if (ld_relaxed_u64(&a.out->err) != 0) {
                    sh.abort = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);  // back off: many workgroups poll the same words
            }
        }
        return (uint32_t)x;
    }

    // Orders this workgroup's global writes before every workgroup's later reads.
    __device__ void barrier() {
        if (a.P == 1) {
            __syncthreads();
            return;
        }
        ++nbar;
        __threadfence();
        __syncthreads();
        if (tid == 0) {
            __hip_atomic_fetch_add(a.bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t target = nbar * (uint32_t)a.P;
            if (__hip_atomic_load(a.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                const int64_t t0 = wall_clock64();
                while (__hip_atomic_load(a.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                    if (wall_clock64() - t0 > a.timeout) {
                        atomicOr((unsigned long long*)&a.out->err, 2ull);
                        sh.abort = 1;
                        break;
                    }
                    if (ld_relaxed_u64(&a.out->err) != 0) {
                        sh.abort = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __atomic_thread_fence(__ATOMIC_ACQUIRE);
        }
        __syncthreads();
        dead = sh.abort != 0;
    }

    // y_r = (A (v ./ pr))_r, terms summed in ascending column order without fusion (the order
    // in which SparseArrays' CSC mul! accumulates a row).
    __device__ double spmv_row(int64_t r, const double* v) const {
        const int32_t k0 = a.rowptr[r], k1 = a.rowptr[r + 1];
        double s = 0.0;
        for (int32_t k = k0; k < k1; ++k) {
            const int32_t j = a.col[k];
            double xv = v[j];
            if (a.pr) xv = xv / a.pr[j];
            s = __dadd_rn(s, __dmul_rn(a.val[k], xv));
        }
        return s;
    }
};

__device__ void sym_givens(double a, double b, double& c, double& s, double& rho) {
    auto sgn = [](double t) { return t > 0.0 ? 1.0 : (t < 0.0 ? -1.0 : 0.0); };
    if (b == 0.0) {
        c = (a == 0.0) ? 1.0 : sgn(a);
        s = 0.0;
        rho = fabs(a);
    } else if (a == 0.0) {
        c = 0.0;
        s = sgn(b);
        rho = fabs(b);
    } else if (fabs(b) > fabs(a)) {
        const double t = a / b;
        s = sgn(b) / sqrt(1.0 + t * t);
        c = s * t;
        rho = b / s;
    } else {
        const double t = b / a;
        c = sgn(a) / sqrt(1.0 + t * t);
        s = c * t;
        rho = a / c;
    }
}

// The solve on workgroup w of a.P (sh, dyn: the workgroup's static and dynamic LDS).  RG: the
// bookkeeping is in HBM (a.Rg) instead of LDS; a compile-time choice, so that each pointer below
// is known to be LDS or global memory and gets that memory's instructions.
template <int RPT, bool RG>
__device__ __forceinline__ void gmres_body(const GmresArgs& a, int w, GShared& sh, double* dyn) {
    const int cap = a.cap;
    // the bookkeeping, read and written by thread 0 (z also by every thread after a barrier): in
    // LDS, or in this workgroup's HBM area when the Krylov dimension is too large for the LDS
    double* R = RG ? a.Rg + (size_t)w * (size_t)gmres_book_words(cap) : dyn;  // cap (cap + 1) / 2
    double* cs = R + ((size_t)cap * (cap + 1)) / 2;  // cap
    double* sn = cs + cap;                    // cap
    double* z = sn + cap;                     // cap + 1
    // q and the current basis slice, one LDS column per thread (each thread touches only its own
    // slots, so no synchronization is needed for them)
    double* qs = dyn + (RG ? 0 : gmres_book_words(cap));  // RPT x GBS
    double* vs = qs + RPT * GBS;              // RPT x GBS
    __shared__ int s_flags[4];
    GmresCtx<RPT> g(a, sh, w);
    if (g.tid == 0) sh.abort = 0;  // ordered before any poll by the first all-reduce's barrier
    const int64_t N = a.N;
    const int tid = g.tid;
    double* x = a.x;
    double* xr = a.restart ? a.xr : a.x;
#define q(i) qs[(i) * GBS + tid]
#define vk(i) vs[(i) * GBS + tid]

    // x = 0; r0 = Pl \ b
    double part = 0.0;
#pragma unroll 2
    for (int i = 0; i < RPT; ++i) {
        const int64_t r = g.row(i);
        q(i) = 0.0;
        if (r < N) {
            x[r] = 0.0;
            q(i) = a.pl ? a.b[r] / a.pl[r] : a.b[r];
            part += q(i) * q(i);
        }
    }
    const double beta0 = sqrt(g.allreduce(part));
    const double eps_ = a.abstol + a.reltol * beta0;
    double rNorm = beta0;
    int64_t iter = 0;
    int status = LSX_RET_DEFAULT;
    if (g.dead) {
        // a wait gave up: the error word tells the host, nothing else is computed
    } else if (beta0 == 0.0) {
        status = LSX_RET_SUCCESS;
    } else {
        int npass = 0;
        const double btol = pow(DBL_EPSILON, 0.75);
        bool breakdown = false, inconsistent = false;
        bool solved = rNorm <= eps_;
        int64_t inner_maxiters = a.maxiters;
        bool tired = iter >= a.maxiters;
        while (!(solved || tired || breakdown || g.dead)) {
            int64_t nr = 0;
            if (a.restart) {
#pragma unroll 2
                for (int i = 0; i < RPT; ++i) {
                    const int64_t r = g.row(i);
                    if (r < N) xr[r] = 0.0;
                }
                if (npass >= 1) {
                    g.barrier();  // x complete in every workgroup before the SpMV reads it
#pragma unroll 2
                    for (int i = 0; i < RPT; ++i) {
                        const int64_t r = g.row(i);
                        if (r < N) {
                            double Ax = 0.0;
                            const int32_t k0 = a.rowptr[r], k1 = a.rowptr[r + 1];
                            for (int32_t k = k0; k < k1; ++k) Ax = __dadd_rn(Ax, __dmul_rn(a.val[k], x[a.col[k]]));
                            const double wv = a.b[r] - Ax;
                            q(i) = a.pl ? wv / a.pl[r] : wv;
                        }
                    }
                }
            }
            // zeta_1 = ||r0||, v_1 = r0 / ||r0||
            part = 0.0;
#pragma unroll 2
            for (int i = 0; i < RPT; ++i) part += (g.row(i) < N) ? q(i) * q(i) : 0.0;
            const double b0 = sqrt(g.allreduce(part));
            if (tid == 0) z[0] = b0;
#pragma unroll 2
            for (int i = 0; i < RPT; ++i) {
                const int64_t r = g.row(i);
                if (r < N) a.V[r] = q(i) / b0;
            }
            g.barrier();
            ++npass;
            int k = 0;
            bool inner_tired = false;
            while (!(solved || inner_tired || breakdown || g.dead)) {
                ++k;
                const double* vcur = a.V + (int64_t)(k - 1) * N;
                // q = Pl \ (A (Pr \ v_k))
#pragma unroll 2
                for (int i = 0; i < RPT; ++i) {
                    const int64_t r = g.row(i);
                    q(i) = 0.0;
                    if (r < N) {
                        const double wv = g.spmv_row(r, vcur);
                        q(i) = a.pl ? wv / a.pl[r] : wv;
                    }
                }
                // modified Gram-Schmidt, the axpy of basis vector i fused with the dot of i + 1;
                // the rows of the next basis vector are loaded while the all-reduce travels
                double h = 0.0;
                double pre[RPT];
#pragma unroll
                for (int i = 0; i < RPT; ++i) {
                    const int64_t r = g.row(i);
                    pre[i] = (r < N) ? a.V[r] : 0.0;
                }
                for (int ii = 1; ii <= k + 1; ++ii) {
                    part = 0.0;
#pragma unroll
                    for (int i = 0; i < RPT; ++i) {
                        const int64_t r = g.row(i);
                        if (r < N) {
                            if (ii > 1) q(i) = q(i) - h * vk(i);
                            if (ii <= k) {
                                vk(i) = pre[i];
                                part += vk(i) * q(i);
                            } else {
                                part += q(i) * q(i);
                            }
                        }
                    }
                    if (ii < k) {
                        const double* vn = a.V + (int64_t)ii * N;
#pragma unroll
                        for (int i = 0; i < RPT; ++i) {
                            const int64_t r = g.row(i);
                            pre[i] = (r < N) ? vn[r] : 0.0;
                        }
                    }
                    const double red = g.allreduce(part);
                    if (g.dead) break;
                    if (ii <= k) {
                        h = red;
                        if (tid == 0) R[nr + ii - 1] = h;
                    } else {
                        h = sqrt(red);  // Hbis
                    }
                }
                const double Hbis = h;
                if (tid == 0) {
                    for (int i = 1; i <= k - 1; ++i) {
                        const double Rtmp = cs[i - 1] * R[nr + i - 1] + sn[i - 1] * R[nr + i];
                        R[nr + i] = sn[i - 1] * R[nr + i - 1] - cs[i - 1] * R[nr + i];
                        R[nr + i - 1] = Rtmp;
                    }
                    double c_, s_, rho;
                    sym_givens(R[nr + k - 1], Hbis, c_, s_, rho);
                    cs[k - 1] = c_;
                    sn[k - 1] = s_;
                    R[nr + k - 1] = rho;
                    const double zeta = s_ * z[k - 1];
                    z[k - 1] = c_ * z[k - 1];
                    z[k] = zeta;  // consumed only if the iteration continues
                }
                __syncthreads();
                rNorm = fabs(z[k]);
                nr += k;
                const bool mach = (rNorm + 1.0 <= 1.0);
                const bool lim = rNorm <= eps_;
                breakdown = Hbis <= btol;
                solved = lim || mach;
                inner_tired = a.restart ? (k >= (int64_t)min((int64_t)a.memory, inner_maxiters))
                                        : (k >= inner_maxiters);
                if (!(solved || inner_tired || breakdown)) {
                    double* vn = a.V + (int64_t)k * N;
#pragma unroll 2
                    for (int i = 0; i < RPT; ++i) {
                        const int64_t r = g.row(i);
                        if (r < N) vn[r] = q(i) / Hbis;
                    }
                    g.barrier();
                }
            }
            // back-substitution R y = z (y overwrites z)
            if (tid == 0) {
                int inc = 0;
                for (int i = k; i >= 1; --i) {
                    int64_t pos = nr + i - k;
                    for (int j = k; j >= i + 1; --j) {
                        z[i - 1] = z[i - 1] - R[pos - 1] * z[j - 1];
                        pos = pos - j + 1;
                    }
                    if (fabs(R[pos - 1]) <= btol) {
                        z[i - 1] = 0.0;
                        inc = 1;
                    } else {
                        z[i - 1] = z[i - 1] / R[pos - 1];
                    }
                }
                s_flags[0] = inc;
            }
            __syncthreads();
            inconsistent = inconsistent || s_flags[0];
            // x_r += sum_i y_i v_i; x_r = Pr \ x_r; x += x_r
#pragma unroll 2
            for (int i = 0; i < RPT; ++i) {
                const int64_t r = g.row(i);
                if (r < N) {
                    double acc = xr[r];
                    for (int ii = 0; ii < k; ++ii) acc = acc + z[ii] * a.V[(int64_t)ii * N + r];
                    if (a.pr) acc = acc / a.pr[r];
                    if (a.restart) x[r] = x[r] + acc;
                    else xr[r] = acc;
                }
            }
            inner_maxiters -= k;
            iter += k;
            tired = iter >= a.maxiters;
            __syncthreads();
        }
        if (tired) status = LSX_RET_MAXITERS;
        if (solved) status = LSX_RET_SUCCESS;
        if (inconsistent) status = LSX_RET_INFEASIBLE;
    }
#undef q
#undef vk
    if (g.w == 0 && tid == 0) {
        a.out->retcode = status;
        a.out->iters = iter;
        a.out->resid = rNorm;
    }
}

template <int RPT, bool RG>
__global__ __launch_bounds__(GBS) void k_gmres(GmresArgs a) {
    __shared__ GShared sh;
    extern __shared__ double dyn[];
    gmres_body<RPT, RG>(a, (int)blockIdx.x, sh, dyn);
}

// Several solves in one launch (gmres_launch_multi), on a pool of workgroups that may be smaller
// than the solves' total: a workgroup takes the next number from a counter, runs that rank of
// its solve, and takes another number when the solve ends.  Numbers go out in start order, so
// every solve but the one holding the highest number handed out so far has all its ranks
// running and finishes by itself; with at least as many workgroups as the widest solve, that
// last one fills up as the others end.
template <int RPT>
__global__ __launch_bounds__(GBS) void k_gmres_multi(const GmresArgs* __restrict items, const int* __restrict first,
                                                     int nitems, uint32_t* ticket) {
    __shared__ GShared sh;
    __shared__ int s_id;
    extern __shared__ double dyn[];
    const int total = first[nitems];
    while (true) {
        __syncthreads();  // the previous solve's LDS is no longer read
        if (threadIdx.x == 0) s_id = (int)atomicAdd(ticket, 1u);
        __syncthreads();
        const int id = __builtin_amdgcn_readfirstlane(s_id);
        if (id >= total) return;
        int lo = 0, hi = nitems - 1;  // last solve with first[i] <= id
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (first[mid] <= id) lo = mid;
            else hi = mid - 1;
        }
        const int item = __builtin_amdgcn_readfirstlane(lo);
        const GmresArgs a = items[item];
        gmres_body<RPT, false>(a, id - first[item], sh, dyn);
    }
}

// Zeroes the exchange area and result of every solve of a multi launch, and its ticket counter.
__global__ void k_gmres_clear(const GmresArgs* __restrict items, uint32_t* ticket) {
    const GmresArgs& a = items[blockIdx.x];
    const int64_t words = gmres_red_words(a.P);
    for (int64_t i = threadIdx.x; i < words; i += blockDim.x) a.red[i] = 0ull;
    if (threadIdx.x == 0) {
        a.out->retcode = 0;
        a.out->pad = 0;
        a.out->iters = 0;
        a.out->resid = 0.0;
        a.out->err = 0ull;
        if (blockIdx.x == 0) *ticket = 0u;
    }
}

// ------------------------------------------------------------------------------------------
__global__ void k_spmv(int64_t N, const int32_t* __restrict rowptr, const int32_t* __restrict col,
                       const double* __restrict val, const double* __restrict xv, double* __restrict y) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    double s = 0.0;
    for (int32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) s = __dadd_rn(s, __dmul_rn(val[k], xv[col[k]]));
    y[r] = s;
}

// Rows per thread for a matrix of N rows: a single workgroup up to 8192 rows, then 16 rows per
// thread over ceil(N / 8192) workgroups.
// Rows per thread and workgroups of a solve of N rows: beyond 8192 rows, the smallest of 4, 8, 16
// rows per thread that keeps the solve within GMRES_WIDE workgroups, so that more CUs stream each
// basis vector (measured: a 256^2 grid 105 -> 69 us per Arnoldi step with 32 workgroups of 2048
// rows instead of 8 of 8192).  The bound keeps every solve inside the serving lanes' share of the
// CUs, so a solve has the same shape, and the same bits, in every launch form.
constexpr int GMRES_WIDE = 128;
static int pick_rpt(int64_t N, int& P) {
    if (N > 16 * GBS) {
        for (int r : {4, 8}) {
            const int64_t p = cdiv(N, (int64_t)r * GBS);
            if (p <= GMRES_WIDE) {
                P = (int)p;
                return r;
            }
        }
    }
    if (N <= GBS) { P = 1; return 1; }
    if (N <= 4 * GBS) { P = 1; return 4; }
    if (N <= 16 * GBS) { P = 1; return 16; }
    P = (int)cdiv(N, 16 * GBS);
    return 16;
}

int csr_spmv_dev(lsx_csr* A, const double* dx, double* dy, hipStream_t st) {
    hipLaunchKernelGGL(k_spmv, dim3((unsigned)cdiv(A->N, 256)), dim3(256), 0, st, A->N, A->rowptr, A->col, A->val,
                       dx, dy);
    return hipGetLastError() == hipSuccess ? LSX_OK : LSX_EHIP;
}

// Dynamic LDS of a solve: q and the basis slice, and (rglob false) the packed triangle R, the
// Givens pairs and z.
static size_t gmres_dynlds(int64_t cap, int rpt, bool rglob) {
    return sizeof(double) * ((rglob ? 0 : (size_t)gmres_book_words(cap)) + 2 * (size_t)rpt * GBS);
}
constexpr size_t GMRES_LDS_MAX = 150 * 1024;

// The largest Krylov dimension of one solve: GMRES_MAX_CAP_G (bookkeeping in HBM from the LDS
// bound on); with hbm false (the multi-solve launches) the largest whose bookkeeping fits the LDS
// of a workgroup of a matrix of N rows.
int gmres_cap_limit(int64_t N, bool hbm) {
    if (hbm) return GMRES_MAX_CAP_G;
    int P = 1;
    const int rpt = pick_rpt(N, P);
    int cap = GMRES_MAX_CAP;
    while (cap > 1 && gmres_dynlds(cap, rpt, false) > GMRES_LDS_MAX) --cap;
    return cap;
}

int gmres_plan(lsx_ctx* ctx, int64_t N, const GmresParams& p, GmresPlan* out, bool hbm) {
    int P = 1;
    const int rpt = pick_rpt(N, P);
    if (P > PMAX_G) return fail(ctx, LSX_EINVAL, "gmres: N = %lld exceeds %d rows", (long long)N, PMAX_G * 16 * GBS);
    const int64_t memory = std::max<int64_t>(1, std::min<int64_t>(p.memory, p.maxiters));
    const int64_t cap = p.restart ? memory : std::max<int64_t>(memory, p.maxiters);
    const int lim = gmres_cap_limit(N, hbm);
    if (cap > lim)
        return fail(ctx, LSX_EINVAL, "gmres: Krylov dimension %lld exceeds %d (restart = %d, N = %lld)", (long long)cap,
                    lim, p.restart, (long long)N);
    const bool rglob = gmres_dynlds(cap, rpt, false) > GMRES_LDS_MAX;
    out->P = P;
    out->rpt = rpt;
    out->cap = (int)cap;
    out->rglob = rglob;
    out->dynlds = gmres_dynlds(cap, rpt, rglob);
    return LSX_OK;
}

int gmres_launch_multi(lsx_ctx* ctx, hipStream_t st, const GmresArgs* items, const int* first, int nitems,
                       int pool, int rpt, size_t dynlds, uint32_t* ticket) {
    if (nitems < 1 || pool < 1) return fail(ctx, LSX_EINVAL, "gmres: empty multi launch");
    hipLaunchKernelGGL(k_gmres_clear, dim3((unsigned)nitems), dim3(256), 0, st, items, ticket);
    switch (rpt) {
        case 1: hipLaunchKernelGGL(k_gmres_multi<1>, dim3(pool), dim3(GBS), dynlds, st, items, first, nitems, ticket); break;
        case 4: hipLaunchKernelGGL(k_gmres_multi<4>, dim3(pool), dim3(GBS), dynlds, st, items, first, nitems, ticket); break;
        case 8: hipLaunchKernelGGL(k_gmres_multi<8>, dim3(pool), dim3(GBS), dynlds, st, items, first, nitems, ticket); break;
        default: hipLaunchKernelGGL(k_gmres_multi<16>, dim3(pool), dim3(GBS), dynlds, st, items, first, nitems, ticket); break;
    }
    LSX_HIP(ctx, hipGetLastError());
    return LSX_OK;
}

int gmres_launch(lsx_csr* A, const GmresParams& p, const double* db, double* dx, const double* dpl,
                 const double* dpr, hipStream_t st) {
    lsx_ctx* ctx = A->ctx;
    const int64_t N = A->N;
    GmresPlan plan;
    int rc = gmres_plan(ctx, N, p, &plan, true);
    if (rc) return rc;
    const int P = plan.P, rpt = plan.rpt;
    const int64_t memory = std::max<int64_t>(1, std::min<int64_t>(p.memory, p.maxiters));
    const int64_t cap = plan.cap;
    // workspace: basis (cap + 1) x N, correction N
    const size_t need = sizeof(double) * (size_t)N * (size_t)(cap + 2);
    if (need > A->ws_bytes) {
        (void)hipFree(A->ws);
        A->ws = nullptr;
        A->ws_bytes = 0;
        if (hipMalloc(&A->ws, need) != hipSuccess)
            return fail(ctx, LSX_ENOMEM, "gmres: cannot allocate %zu bytes of Krylov workspace", need);
        A->ws_bytes = need;
    }
    if (plan.rglob) {
        const size_t rneed = sizeof(double) * (size_t)P * (size_t)gmres_book_words(cap);
        if (rneed > A->rg_bytes) {
            (void)hipFree(A->rg);
            A->rg = nullptr;
            A->rg_bytes = 0;
            if (hipMalloc(&A->rg, rneed) != hipSuccess)
                return fail(ctx, LSX_ENOMEM, "gmres: cannot allocate %zu bytes of Krylov bookkeeping", rneed);
            A->rg_bytes = rneed;
        }
    }
    LSX_HIP(ctx, hipMemsetAsync(A->red, 0, sizeof(uint64_t) * 4 * PMAX_G + 256, st));
    LSX_HIP(ctx, hipMemsetAsync(A->out, 0, sizeof(GmresOut), st));
    GmresArgs g;
    g.N = N;
    g.rowptr = A->rowptr;
    g.col = A->col;
    g.val = A->val;
    g.b = db;
    g.x = dx;
    g.pl = dpl;
    g.pr = dpr;
    g.V = A->ws;
    g.xr = A->ws + (size_t)N * (size_t)(cap + 1);
    g.memory = (int)memory;
    g.restart = p.restart;
    g.cap = (int)cap;
    g.Rg = plan.rglob ? A->rg : nullptr;
    g.maxiters = p.maxiters;
    g.abstol = p.abstol;
    g.reltol = p.reltol;
    g.P = P;
    g.red = A->red;
    g.bar = reinterpret_cast<uint32_t*>(A->red + 4 * (int64_t)P);
    g.out = A->out;
    g.timeout = wait_ticks(ctx, 20.0);
    const size_t dynlds = plan.dynlds;
    switch (rpt) {
        case 1:
            if (plan.rglob) hipLaunchKernelGGL((k_gmres<1, true>), dim3(P), dim3(GBS), dynlds, st, g);
            else hipLaunchKernelGGL((k_gmres<1, false>), dim3(P), dim3(GBS), dynlds, st, g);
            break;
        case 4:
            if (plan.rglob) hipLaunchKernelGGL((k_gmres<4, true>), dim3(P), dim3(GBS), dynlds, st, g);
            else hipLaunchKernelGGL((k_gmres<4, false>), dim3(P), dim3(GBS), dynlds, st, g);
            break;
        case 8:
            if (plan.rglob) hipLaunchKernelGGL((k_gmres<8, true>), dim3(P), dim3(GBS), dynlds, st, g);
            else hipLaunchKernelGGL((k_gmres<8, false>), dim3(P), dim3(GBS), dynlds, st, g);
            break;
        default:
            if (plan.rglob) hipLaunchKernelGGL((k_gmres<16, true>), dim3(P), dim3(GBS), dynlds, st, g);
            else hipLaunchKernelGGL((k_gmres<16, false>), dim3(P), dim3(GBS), dynlds, st, g);
            break;
    }
    LSX_HIP(ctx, hipGetLastError());
    return LSX_OK;
}

}  // namespace lsx

namespace lsx {
// CSR structure check: rowptr[0] = 0, rowptr[N] = nnz, rows non-decreasing within [0, nnz], column
// indices in [0, N); *bad is set otherwise.  A row's entries are read only once its bounds passed.
__global__ void k_csr_check(int64_t N, int64_t nnz, const int32_t* __restrict rowptr, const int32_t* __restrict col,
                            int* bad) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k0 = rowptr[r], k1 = rowptr[r + 1];
        if ((r == 0 && k0 != 0) || (r == N - 1 && k1 != nnz) || k0 < 0 || k1 < k0 || k1 > nnz) {
            *bad = 1;
            continue;
        }
        for (int64_t k = k0; k < k1; ++k)
            if (col[k] < 0 || col[k] >= N) *bad = 1;
    }
}

int csr_validate(hipStream_t st, int64_t N, int64_t nnz, const int32_t* rowptr, const int32_t* col, int* dflag) {
    if (hipMemsetAsync(dflag, 0, sizeof(int), st) != hipSuccess) return LSX_EHIP;
    hipLaunchKernelGGL(k_csr_check, dim3((unsigned)std::min<int64_t>(cdiv(N, 256), 4096)), dim3(256), 0, st, N, nnz,
                       rowptr, col, dflag);
    return hipGetLastError() == hipSuccess ? LSX_OK : LSX_EHIP;
}
}  // namespace lsx

// Column indices local to each row's block of `bs` rows (block-diagonal matrices); *bad is set
// when an entry lies outside its row's diagonal block.
__global__ void k_local_cols(int64_t N, int64_t bs, const int32_t* __restrict rowptr, const int32_t* __restrict col,
                             int32_t* __restrict cl, int* bad) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b0 = (r / bs) * bs;
        for (int32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const int64_t c = col[k];
            if (c < b0 || c >= b0 + bs) *bad = 1;
            cl[k] = (int32_t)(c - b0);
        }
    }
}

// ------------------------------------------------------------------------------------------
// C-ABI
using namespace lsx;

static void csr_free(lsx_csr* A) {
    if (A->stream) (void)hipStreamSynchronize(A->stream);
    (void)hipFree(A->col_local);
    (void)hipFree(A->bred);
    (void)hipFree(A->bitems);
    (void)hipFree(A->bfirst);
    (void)hipFree(A->bout);
    (void)hipFree(A->bticket);
    (void)hipFree(A->bflag);
    (void)hipFree(A->bpart);
    (void)hipFree(A->rowptr);
    (void)hipFree(A->col);
    (void)hipFree(A->val);
    (void)hipFree(A->ws);
    (void)hipFree(A->rg);
    (void)hipFree(A->red);
    (void)hipFree(A->out);
    if (A->hout) (void)hipHostFree(A->hout);
    (void)hipFree(A->db);
    (void)hipFree(A->dx);
    (void)hipFree(A->dpl);
    (void)hipFree(A->dpr);
    if (A->stream) (void)hipStreamDestroy(A->stream);
}

namespace lsx {

// Stopping tolerance of a batched solve, on the device: partial sums of (b / pl)^2 over a fixed
// grid of NPART blocks, then every block of k_batch_tol sums the partials in the same fixed order
// and writes eps = abstol + reltol sqrt(sum) into its share of the launch table (one value for all
// blocks, the same rounding on every call).
constexpr int NPART = 256;

__global__ __launch_bounds__(256) void k_sumsq_partial(const double* b, const double* pl, int64_t N, double* part) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (int64_t)NPART * 256) {
        const double r = pl ? b[i] / pl[i] : b[i];
        acc = fma(r, r, acc);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_batch_tol(const double* part, double abstol, double reltol, GmresArgs* items,
                                                   int64_t nitems) {
    __shared__ double s[NPART];
    s[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int w = NPART / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    const double eps = abstol + reltol * sqrt(s[0]);
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nitems; i += (int64_t)gridDim.x * 256)
        items[i].abstol = eps;
}

}  // namespace lsx

extern "C" {

int lsx_csr_create(lsx_ctx* ctx, int64_t N, int64_t nnz, const int32_t* rowptr, const int32_t* colind,
                   const double* vals, lsx_csr** out) {
    if (!ctx || !out || !rowptr || (nnz > 0 && (!colind || !vals))) return LSX_EINVAL;
    *out = nullptr;
    if (N < 1 || nnz < 0 || nnz > INT32_MAX)
        return fail(ctx, LSX_EINVAL, "lsx_csr_create: N = %lld, nnz = %lld out of range", (long long)N,
                    (long long)nnz);
    DeviceGuard g(ctx);
    lsx_csr* A = new (std::nothrow) lsx_csr();
    if (!A) return LSX_ENOMEM;
    A->ctx = ctx;
    A->N = N;
    A->nnz = nnz;
    auto bad = [&](int rc) {
        csr_free(A);
        delete A;
        return rc;
    };
    if (hipStreamCreateWithFlags(&A->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&A->rowptr, sizeof(int32_t) * (N + 1)) != hipSuccess ||
        hipMalloc(&A->col, sizeof(int32_t) * std::max<int64_t>(nnz, 1)) != hipSuccess ||
        hipMalloc(&A->val, sizeof(double) * std::max<int64_t>(nnz, 1)) != hipSuccess ||
        hipMalloc(&A->red, sizeof(uint64_t) * 4 * PMAX_G + 256) != hipSuccess ||
        hipMalloc(&A->out, sizeof(GmresOut)) != hipSuccess ||
        hipHostMalloc(&A->hout, sizeof(GmresOut), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&A->db, sizeof(double) * N) != hipSuccess || hipMalloc(&A->dx, sizeof(double) * N) != hipSuccess ||
        hipMalloc(&A->dpl, sizeof(double) * N) != hipSuccess || hipMalloc(&A->dpr, sizeof(double) * N) != hipSuccess ||
        hipMalloc(&A->bflag, sizeof(int)) != hipSuccess)
        return bad(fail(ctx, LSX_ENOMEM, "lsx_csr_create: device allocation failed (N = %lld)", (long long)N));
    if (hipMemcpyAsync(A->rowptr, rowptr, sizeof(int32_t) * (N + 1), hipMemcpyDefault, A->stream) != hipSuccess ||
        (nnz > 0 && hipMemcpyAsync(A->col, colind, sizeof(int32_t) * nnz, hipMemcpyDefault, A->stream) != hipSuccess) ||
        (nnz > 0 && hipMemcpyAsync(A->val, vals, sizeof(double) * nnz, hipMemcpyDefault, A->stream) != hipSuccess) ||
        hipStreamSynchronize(A->stream) != hipSuccess)
        return bad(fail(ctx, LSX_EHIP, "lsx_csr_create: upload failed"));
    int badcsr = 0;
    if (csr_validate(A->stream, N, nnz, A->rowptr, A->col, A->bflag) != LSX_OK ||
        hipMemcpyAsync(&badcsr, A->bflag, sizeof(int), hipMemcpyDeviceToHost, A->stream) != hipSuccess ||
        hipStreamSynchronize(A->stream) != hipSuccess)
        return bad(fail(ctx, LSX_EHIP, "lsx_csr_create: structure check failed to run"));
    if (badcsr)
        return bad(fail(ctx, LSX_EINVAL, "lsx_csr_create: not a CSR matrix of order %lld with %lld entries (row pointers "
                        "or column indices out of range)", (long long)N, (long long)nnz));
    *out = A;
    return LSX_OK;
}

int lsx_csr_destroy(lsx_csr* A) {
    if (!A) return LSX_EINVAL;
    DeviceGuard g(A->ctx);
    csr_free(A);
    delete A;
    return LSX_OK;
}

int lsx_csr_spmv(lsx_csr* A, const double* x, double* y) {
    if (!A || !x || !y) return LSX_EINVAL;
    lsx_ctx* ctx = A->ctx;
    DeviceGuard g(ctx);
    LSX_HIP(ctx, hipMemcpyAsync(A->dx, x, sizeof(double) * A->N, hipMemcpyDefault, A->stream));
    int rc = csr_spmv_dev(A, A->dx, A->db, A->stream);
    if (rc) return fail(ctx, rc, "lsx_csr_spmv: launch failed");
    LSX_HIP(ctx, hipMemcpyAsync(y, A->db, sizeof(double) * A->N, hipMemcpyDefault, A->stream));
    LSX_HIP(ctx, hipStreamSynchronize(A->stream));
    return LSX_OK;
}

int lsx_gmres_solve(lsx_csr* A, const double* b, double* x, int32_t memory, int32_t restart, double abstol,
                    double reltol, int64_t maxiters, const double* pl_diag, const double* pr_diag,
                    lsx_gmres_stats* stats) {
    if (!A || !b || !x) return LSX_EINVAL;
    lsx_ctx* ctx = A->ctx;
    if (memory < 1 || maxiters < 0) return fail(ctx, LSX_EINVAL, "lsx_gmres_solve: memory >= 1, maxiters >= 0");
    DeviceGuard g(ctx);
    const size_t vb = sizeof(double) * A->N;
    LSX_HIP(ctx, hipMemcpyAsync(A->db, b, vb, hipMemcpyDefault, A->stream));
    if (pl_diag) LSX_HIP(ctx, hipMemcpyAsync(A->dpl, pl_diag, vb, hipMemcpyDefault, A->stream));
    if (pr_diag) LSX_HIP(ctx, hipMemcpyAsync(A->dpr, pr_diag, vb, hipMemcpyDefault, A->stream));
    GmresParams p{memory, restart ? 1 : 0, maxiters, abstol, reltol};
    int rc = gmres_launch(A, p, A->db, A->dx, pl_diag ? A->dpl : nullptr, pr_diag ? A->dpr : nullptr, A->stream);
    if (rc) return rc;
    LSX_HIP(ctx, hipMemcpyAsync(x, A->dx, vb, hipMemcpyDefault, A->stream));
    LSX_HIP(ctx, hipMemcpyAsync(A->hout, A->out, sizeof(GmresOut), hipMemcpyDeviceToHost, A->stream));
    LSX_HIP(ctx, hipStreamSynchronize(A->stream));
    if (A->hout->err) return fail(ctx, LSX_ETIMEOUT, "lsx_gmres_solve: a device-side wait exceeded its bound");
    if (stats) {
        stats->retcode = A->hout->retcode;
        stats->pad = 0;
        stats->iters = A->hout->iters;
        stats->resid = A->hout->resid;
    }
    return LSX_OK;
}

int lsx_gmres_max_memory(lsx_csr* A, int32_t* memory) {
    if (!A || !memory) return LSX_EINVAL;
    *memory = gmres_cap_limit(A->N, true);
    return LSX_OK;
}

int lsx_gmres_solve_batched(lsx_csr* A, int64_t blocksize, const double* b, double* x, int32_t memory,
                            int32_t restart, double abstol, double reltol, int64_t maxiters, const double* pl_diag,
                            const double* pr_diag, lsx_gmres_stats* stats) {
    if (!A || !b || !x) return LSX_EINVAL;
    lsx_ctx* ctx = A->ctx;
    const int64_t N = A->N;
    if (blocksize < 1 || N % blocksize != 0)
        return fail(ctx, LSX_EINVAL, "lsx_gmres_solve_batched: blocksize %lld must divide N = %lld", (long long)blocksize,
                    (long long)N);
    if (memory < 1 || maxiters < 0)
        return fail(ctx, LSX_EINVAL, "lsx_gmres_solve_batched: memory >= 1, maxiters >= 0");
    DeviceGuard g(ctx);
    const int64_t nb = N / blocksize;
    GmresParams p{memory, restart ? 1 : 0, maxiters, 0.0, 0.0};
    GmresPlan plan;
    int rc = gmres_plan(ctx, blocksize, p, &plan);
    if (rc) return rc;
    if (plan.P != 1)
        return fail(ctx, LSX_EINVAL, "lsx_gmres_solve_batched: blocks of more than %d rows are not supported", 16 * GBS);
    const hipStream_t st = A->stream;
    // local columns (once per block size), after checking that A is block diagonal
    if (A->local_bs != blocksize) {
        if (!A->col_local) LSX_HIP(ctx, hipMalloc(&A->col_local, sizeof(int32_t) * std::max<int64_t>(A->nnz, 1)));
        LSX_HIP(ctx, hipMemsetAsync(A->bflag, 0, sizeof(int), st));
        hipLaunchKernelGGL(k_local_cols, dim3((unsigned)std::min<int64_t>(cdiv(N, 256), 4096)), dim3(256), 0, st, N,
                           blocksize, A->rowptr, A->col, A->col_local, A->bflag);
        int bad = 0;
        LSX_HIP(ctx, hipMemcpyAsync(&bad, A->bflag, sizeof(int), hipMemcpyDeviceToHost, st));
        LSX_HIP(ctx, hipStreamSynchronize(st));
        if (bad) return fail(ctx, LSX_EINVAL, "lsx_gmres_solve_batched: A has entries outside its %lld x %lld diagonal blocks",
                             (long long)blocksize, (long long)blocksize);
        A->local_bs = blocksize;
    }
    // buffers for nb blocks
    const int64_t rw = gmres_red_words(1);
    if (nb > A->bcap) {
        (void)hipFree(A->bred);
        (void)hipFree(A->bitems);
        (void)hipFree(A->bfirst);
        (void)hipFree(A->bout);
        A->bred = nullptr;
        A->bitems = nullptr;
        A->bfirst = nullptr;
        A->bout = nullptr;
        A->bcap = 0;
        LSX_HIP(ctx, hipMalloc(&A->bred, sizeof(uint64_t) * rw * nb));
        LSX_HIP(ctx, hipMalloc(&A->bitems, sizeof(GmresArgs) * nb));
        LSX_HIP(ctx, hipMalloc(&A->bfirst, sizeof(int) * (nb + 1)));
        LSX_HIP(ctx, hipMalloc(&A->bout, sizeof(GmresOut) * nb));
        A->bcap = nb;
        A->bkey[0] = 0;  // new table buffers: rebuild
    }
    if (!A->bticket) LSX_HIP(ctx, hipMalloc(&A->bticket, 256));
    const size_t need = sizeof(double) * (size_t)N * (size_t)(plan.cap + 2);
    if (need > A->ws_bytes) {
        (void)hipFree(A->ws);
        A->ws = nullptr;
        A->ws_bytes = 0;
        if (hipMalloc(&A->ws, need) != hipSuccess)
            return fail(ctx, LSX_ENOMEM, "gmres: cannot allocate %zu bytes of Krylov workspace", need);
        A->ws_bytes = need;
    }
    if (!A->bpart) LSX_HIP(ctx, hipMalloc(&A->bpart, sizeof(double) * NPART));
    const size_t vb = sizeof(double) * N;
    LSX_HIP(ctx, hipMemcpyAsync(A->db, b, vb, hipMemcpyDefault, st));
    if (pl_diag) LSX_HIP(ctx, hipMemcpyAsync(A->dpl, pl_diag, vb, hipMemcpyDefault, st));
    if (pr_diag) LSX_HIP(ctx, hipMemcpyAsync(A->dpr, pr_diag, vb, hipMemcpyDefault, st));
    // the launch table, rebuilt when what it holds changes; its tolerance is written on the device
    const int64_t memv = std::max<int64_t>(1, std::min<int64_t>(memory, maxiters));
    const int64_t wt = wait_ticks(ctx, 20.0);
    const int64_t key[9] = {blocksize, (pl_diag ? 1 : 0) | (pr_diag ? 2 : 0), memv, restart ? 1 : 0, maxiters,
                            (int64_t)(uintptr_t)A->ws, (int64_t)(uintptr_t)A->bred, plan.cap, wt};
    if (!std::equal(key, key + 9, A->bkey)) {
        std::vector<GmresArgs> items(nb);
        std::vector<int> first(nb + 1);
        for (int64_t i = 0; i < nb; ++i) {
            GmresArgs& a = items[i];
            const int64_t r0 = i * blocksize;
            a.N = blocksize;
            a.rowptr = A->rowptr + r0;
            a.col = A->col_local;
            a.val = A->val;
            a.b = A->db + r0;
            a.x = A->dx + r0;
            a.pl = pl_diag ? A->dpl + r0 : nullptr;
            a.pr = pr_diag ? A->dpr + r0 : nullptr;
            a.V = A->ws + r0 * (plan.cap + 2);
            a.xr = a.V + blocksize * (plan.cap + 1);
            a.memory = (int)memv;
            a.restart = restart ? 1 : 0;
            a.cap = plan.cap;
            a.maxiters = maxiters;
            a.abstol = 0.0;
            a.reltol = 0.0;
            a.P = 1;
            a.red = A->bred + i * rw;
            a.bar = reinterpret_cast<uint32_t*>(a.red + 4);
            a.out = A->bout + i;
            a.timeout = wt;
            first[i] = (int)i;
        }
        first[nb] = (int)nb;
        LSX_HIP(ctx, hipMemcpyAsync(A->bitems, items.data(), sizeof(GmresArgs) * nb, hipMemcpyHostToDevice, st));
        LSX_HIP(ctx, hipMemcpyAsync(A->bfirst, first.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, st));
        LSX_HIP(ctx, hipStreamSynchronize(st));  // the host tables are freed on return
        std::copy(key, key + 9, A->bkey);
    }
    // the stopping tolerance of the whole batch: eps = abstol + reltol ||Pl \ b||_2
    // (src/simplegmres.jl:110-129: beta is the norm of the whole vector)
    hipLaunchKernelGGL(k_sumsq_partial, dim3(NPART), dim3(256), 0, st, A->db, pl_diag ? A->dpl : nullptr, N, A->bpart);
    hipLaunchKernelGGL(k_batch_tol, dim3((unsigned)std::min<int64_t>(cdiv(nb, 256), 64)), dim3(256), 0, st, A->bpart,
                       abstol, reltol, A->bitems, nb);
    const int pool = (int)std::min<int64_t>(nb, 4 * (int64_t)ctx->num_cus);
    rc = gmres_launch_multi(ctx, st, A->bitems, A->bfirst, (int)nb, pool, plan.rpt, plan.dynlds, A->bticket);
    if (rc) return rc;
    std::vector<GmresOut> outs(nb);
    LSX_HIP(ctx, hipMemcpyAsync(outs.data(), A->bout, sizeof(GmresOut) * nb, hipMemcpyDeviceToHost, st));
    LSX_HIP(ctx, hipMemcpyAsync(x, A->dx, vb, hipMemcpyDefault, st));
    LSX_HIP(ctx, hipStreamSynchronize(st));
    // the batch's result: iterations and rNorm of its slowest block; Infeasible if any block is,
    // else Success if every block is solved, else MaxIters (src/simplegmres.jl:603-606)
    int64_t iters = 0;
    double resid = 0.0;
    bool any_inf = false, all_ok = true;
    for (const GmresOut& o : outs) {
        if (o.err) return fail(ctx, LSX_ETIMEOUT, "lsx_gmres_solve_batched: a device-side wait exceeded its bound");
        iters = std::max<int64_t>(iters, o.iters);
        resid = std::max(resid, o.resid);
        any_inf = any_inf || o.retcode == LSX_RET_INFEASIBLE;
        all_ok = all_ok && o.retcode == LSX_RET_SUCCESS;
    }
    if (stats) {
        stats->retcode = any_inf ? LSX_RET_INFEASIBLE : (all_ok ? LSX_RET_SUCCESS : LSX_RET_MAXITERS);
        stats->pad = 0;
        stats->iters = iters;
        stats->resid = resid;
    }
    return LSX_OK;
}

}  // extern "C"
