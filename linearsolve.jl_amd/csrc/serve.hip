// serve.hip — continuous-batching request loop for mixed dense / sparse solves on one GPU.
//
// Requests are submitted from any host thread into the queue of their lane, in arrival order.
// Admission: a request starts when its lane has a free worker, its device memory can be carved
// out of the server's heap (one allocation of the HBM budget made at creation, so admission is
// the allocation itself and no request calls hipMalloc/hipFree), and the workgroups that wait on
// each other fit the cooperative budget (below).  Memory and workgroups return when it completes.
//   batch lane   dense systems with n <= 512: every waiting system that fits the batch arena goes
//                into one persistent batched-LU launch, so requests that arrive while a launch
//                runs join the next one (continuous batching);
//   dense lanes  dense systems with n > 512, one at a time per lane on a reusable LU handle
//                (grown when a larger order arrives) whose factors live in the request's heap
//                block;
//   GMRES lanes  SimpleGMRES solves: every waiting solve of the same rows-per-thread class that
//                fits goes into one multi-solve launch (continuous batching again) on a pool of
//                persistent workgroups, so the GMRES lanes together hold at most 3/4 of the CUs.
// Cooperative budget: the panel kernel of a dense LU and every GMRES solve spread over
// workgroups that wait for each other, so a solve makes progress only once all its workgroups
// run.  Within a multi-solve launch workgroups take their numbers in start order, so at most one
// solve per launch is partly started (the pool is at least as wide as the widest solve, so the
// other solves' ends fill it up); a lane reserves the largest solve's workgroup count and
// the reservations of all lanes stay below the CU count minus 16.  Then the partly started
// solves can never hold every CU, and the other kernels always finish, so every solve starts
// completely in the end.
// Lanes run on their own streams and threads, so their work overlaps on the device.  Every
// kernel is deterministic and a request's arithmetic does not depend on what else is in flight:
// a result is the same alone and in the mix.  Latency = completion - submission on the serving
// clock (steady clock, seconds since the server was created).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <climits>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "gmres.hpp"
#include "lu.hpp"

namespace lsx {

// ------------------------------------------------------------------------------------------
// Device-side synthetic inputs (counter-based: value i of stream `seed` depends only on both).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t seed, uint64_t i) {  // (0, 1]
    return ((double)(mix64(seed * 0x100000001B3ull ^ mix64(i)) >> 11) + 1.0) * 0x1.0p-53;
}
__device__ __forceinline__ double gauss(uint64_t seed, uint64_t i) {
    const double a = u01(seed, 2 * i), b = u01(seed, 2 * i + 1);
    return sqrt(-2.0 * log(a)) * cos(6.283185307179586 * b);
}

struct GenItem {
    int64_t n, aoff, voff;
    uint64_t seed;
    int64_t ld;  // leading dimension of the system's A
};

// A = I - g J (J iid N(0,1)/sqrt(n), g ~ U[1e-3, 1e-1]) and b iid N(0,1), for a list of systems
// packed back to back; grid.y walks the systems.
__global__ void k_gen_dense(const GenItem* __restrict items, double* __restrict A, double* __restrict b) {
    const GenItem it = items[blockIdx.y];
    const int64_t n = it.n;
    const double g = 1e-3 + (1e-1 - 1e-3) * u01(it.seed, 0xFFFFFFFFull);
    const double sc = g / sqrt((double)n);
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n * n;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t col = idx / n, row = idx - col * n;
        A[it.aoff + col * it.ld + row] = (row == col ? 1.0 : 0.0) - sc * gauss(it.seed, (uint64_t)idx);
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[it.voff + i] = gauss(it.seed ^ 0xB5ull, (uint64_t)i);
}

// 5-point convection-diffusion CSR (row r = i + m j, columns ascending), b and the diagonal.
__global__ void k_gen_convdiff(int m, double pe, uint64_t seed, int32_t* __restrict rowptr, int32_t* __restrict col,
                               double* __restrict val, double* __restrict b, double* __restrict diag) {
    const int64_t N = (int64_t)m * m;
    const double h = 1.0 / (m + 1), lo = -1.0 - pe * h / 2.0, hi = -1.0 + pe * h / 2.0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = r / m, i = r - j * m;
        // entries before row r (closed form of the row-length prefix sum)
        int64_t off = j * (5 * (int64_t)m - 2) - (j > 0 ? m : 0) + 5 * i - (i > 0 ? 1 : 0) -
                      i * ((j == 0 ? 1 : 0) + (j == m - 1 ? 1 : 0));
        rowptr[r] = (int32_t)off;
        if (j > 0) { col[off] = (int32_t)(r - m); val[off++] = lo; }
        if (i > 0) { col[off] = (int32_t)(r - 1); val[off++] = lo; }
        col[off] = (int32_t)r;
        val[off++] = 4.0;
        if (i < m - 1) { col[off] = (int32_t)(r + 1); val[off++] = hi; }
        if (j < m - 1) { col[off] = (int32_t)(r + m); val[off++] = hi; }
        if (r == N - 1) rowptr[N] = (int32_t)off;
        b[r] = gauss(seed ^ 0xC3ull, (uint64_t)r);
        diag[r] = 4.0;
    }
}

static int64_t convdiff_nnz(int64_t m) { return 5 * m * m - 4 * m; }

// Batched LU on caller-provided device arrays (one launch, systems in the given order).
int batch_launch_raw(lsx_ctx* ctx, hipStream_t st, int64_t count, const int32_t* order, const int32_t* n,
                     const int64_t* aoff, const int64_t* voff, double* A, const double* b, double* x, int64_t* ipiv,
                     int32_t* perm, int64_t* info, uint32_t* counter);

}  // namespace lsx

using namespace lsx;
using Clock = std::chrono::steady_clock;

namespace {

enum Lane { LANE_BATCH = 0, LANE_DENSE = 1, LANE_GMRES = 2, NLANES = 3 };
constexpr int MAXG = 64;           // solves per GMRES launch
constexpr int COOP_MARGIN = 16;    // CUs never reserved by waiting workgroups
constexpr size_t HALIGN = 256;     // heap and carve alignment

inline size_t aup(size_t b) { return (b + HALIGN - 1) & ~(HALIGN - 1); }

// Consecutive arrays of one heap block (base 0: measures the block).
struct Carve {
    uintptr_t p;
    template <typename T>
    T* take(int64_t count) {
        T* r = reinterpret_cast<T*>(p);
        p += aup(sizeof(T) * (size_t)std::max<int64_t>(count, 1));
        return r;
    }
};

// First-fit allocator over one device allocation; free ranges kept sorted and coalesced.
struct Heap {
    uintptr_t base = 0;
    size_t size = 0, used = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes

    void init(void* b, size_t s) {
        base = reinterpret_cast<uintptr_t>(b);
        size = s;
        used = 0;
        free_.clear();
        if (s) free_[0] = s;
    }
    size_t alloc(size_t bytes) {  // offset, or SIZE_MAX when no free range is large enough
        bytes = aup(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            used += bytes;
            return off;
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes) {
        bytes = aup(bytes);
        used -= bytes;
        auto nx = free_.lower_bound(off);
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) {
                off = pv->first;
                bytes += pv->second;
                free_.erase(pv);
            }
        }
        if (nx != free_.end() && off + bytes == nx->first) {
            bytes += nx->second;
            free_.erase(nx);
        }
        free_[off] = bytes;
    }
};

struct Job {
    uint64_t ticket = 0;
    lsx_request req{};
    double* x_out = nullptr;
    int lane = 0;
    size_t footprint = 0;   // bytes: heap block (dense, GMRES) or batch-arena share (batch)
    size_t off = 0;         // heap offset
    int coop = 0;           // workgroups that wait on each other
    GmresPlan plan{};       // GMRES: launch shape
    int64_t N = 0, nnz = 0; // GMRES: rows and nonzeros
    bool done = false;
    lsx_result res{};
};

// Heap block of a GMRES request.
struct GmresBlock {
    int32_t* rowptr;
    int32_t* col;
    double* val;
    double* b;
    double* pl;
    double* x;
    double* V;
    double* xr;
    uint64_t* red;
    size_t bytes;
};

GmresBlock gmres_block(uintptr_t base, const Job& j) {
    Carve c{base};
    GmresBlock g;
    g.rowptr = c.take<int32_t>(j.N + 1);
    g.col = c.take<int32_t>(j.nnz);
    g.val = c.take<double>(j.nnz);
    g.b = c.take<double>(j.N);
    g.pl = c.take<double>(j.N);
    g.x = c.take<double>(j.N);
    g.V = c.take<double>((int64_t)(j.plan.cap + 1) * j.N);
    g.xr = c.take<double>(j.N);
    g.red = c.take<uint64_t>(gmres_red_words(j.plan.P));
    g.bytes = c.p - base;
    return g;
}

// Heap block of a dense request: the factors (leading dimension pad_ld(n)) and b, solved in place.
struct DenseBlock {
    double* F;
    double* x;
    size_t bytes;
};

DenseBlock dense_block(uintptr_t base, int64_t n) {
    Carve c{base};
    DenseBlock d;
    d.F = c.take<double>(pad_ld(n) * n);
    d.x = c.take<double>(n);
    d.bytes = c.p - base;
    return d;
}

struct Arena {  // batch lane
    double* A = nullptr;
    double* b = nullptr;
    double* x = nullptr;
    int64_t* ipiv = nullptr;
    int32_t* perm = nullptr;
    int64_t* info = nullptr;
    int32_t* meta32 = nullptr;   // order | n
    int64_t* meta64 = nullptr;   // aoff | voff
    GenItem* gen = nullptr;
    uint32_t* counter = nullptr;
    int64_t capA = 0, capV = 0, capN = 0;
    size_t bytes = 0;            // admission share of the budget
    // pinned staging
    int32_t* h32 = nullptr;
    int64_t* h64 = nullptr;
    GenItem* hgen = nullptr;
    int64_t* hinfo = nullptr;
    hipStream_t stream = nullptr;
};

struct DenseLane {
    lsx_lu* lu = nullptr;        // reusable handle (factors point into the request's block)
    GenItem* d_gen = nullptr;
    GenItem* h_gen = nullptr;    // pinned
};

struct GmresLane {
    hipStream_t stream = nullptr;
    GmresArgs* h_items = nullptr;  // pinned
    int* h_first = nullptr;        // pinned
    GmresOut* h_out = nullptr;     // pinned
    int* h_flags = nullptr;        // pinned: CSR checks of a launch's requests
    int* d_flags = nullptr;
    GmresArgs* d_items = nullptr;
    int* d_first = nullptr;
    GmresOut* d_out = nullptr;
    uint32_t* d_ticket = nullptr;
};

}  // namespace

struct lsx_serve {
    lsx_ctx* ctx = nullptr;
    size_t budget = 0, peak = 0, batch_used = 0;
    int coop_total = 0, coop_used = 0;
    int gmres_pool = 0;               // workgroups of one GMRES lane's launch (CUs it occupies)
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv;       // workers wait for lane work
    std::condition_variable cv_done;  // drain waits for completions
    std::deque<Job*> lanes[NLANES];   // waiting for their lane, arrival order
    std::unordered_map<uint64_t, Job*> jobs;
    uint64_t next_ticket = 1;
    bool stop = false;
    int64_t submitted = 0, completed = 0, batches = 0, batched_reqs = 0, inflight = 0;
    Heap heap;
    void* heap_mem = nullptr;
    Arena arena;
    std::vector<DenseLane> dense;
    std::vector<GmresLane> gmres;
    std::vector<std::thread> workers;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    void note_use() { peak = std::max(peak, heap.used + batch_used); }
};

namespace {

// Sizes, lane, footprint and cooperative workgroups of a request; LSX_OK or an error (message set).
int classify(lsx_serve* s, Job* j) {
    const lsx_request& r = j->req;
    if (r.kind == 0) {
        const int64_t n = r.n;
        if (n <= 512) {
            j->lane = LANE_BATCH;
            j->footprint = (size_t)(8 * n * n + 64 * n);
            return LSX_OK;
        }
        if (n > (int64_t)PMAX * 512)
            return fail(s->ctx, LSX_EINVAL, "lsx_serve_submit: dense n = %lld exceeds %d", (long long)n, PMAX * 512);
        j->lane = LANE_DENSE;
        j->footprint = dense_block(0, n).bytes;
        j->coop = (int)cdiv(n, 512) + 1;  // panel workgroups, plus the chained solve's running one
        return LSX_OK;
    }
    j->lane = LANE_GMRES;
    j->N = r.synth ? (int64_t)r.grid_m * r.grid_m : r.n;
    j->nnz = r.synth ? convdiff_nnz(r.grid_m) : r.nnz;
    if (j->nnz < 0 || j->nnz > INT32_MAX) return fail(s->ctx, LSX_EINVAL, "lsx_serve_submit: nnz out of range");
    const GmresParams p{r.memory, r.restart ? 1 : 0, r.maxiters, r.abstol, r.reltol};
    int rc = gmres_plan(s->ctx, j->N, p, &j->plan);
    if (rc) return rc;
    j->coop = j->plan.P;
    j->footprint = gmres_block(0, *j).bytes;
    return LSX_OK;
}

void finish(lsx_serve* s, Job* j) {
    std::lock_guard<std::mutex> g(s->mu);
    j->res.t_done = s->now();
    j->done = true;
    if (j->lane == LANE_BATCH) s->batch_used -= j->footprint;
    else s->heap.release(j->off, j->footprint);
    s->completed += 1;
    s->inflight -= 1;
    s->cv.notify_all();
    s->cv_done.notify_all();
}

void release_coop(lsx_serve* s, int k) {
    std::lock_guard<std::mutex> g(s->mu);
    s->coop_used -= k;
    s->cv.notify_all();
}

// ---- dense lanes: one LU at a time per lane (called with s->mu held)
Job* take_dense(lsx_serve* s) {
    auto& q = s->lanes[LANE_DENSE];
    if (q.empty()) return nullptr;
    Job* j = q.front();
    if (j->coop > s->coop_total - s->coop_used) return nullptr;
    const size_t off = s->heap.alloc(j->footprint);
    if (off == SIZE_MAX) return nullptr;
    q.pop_front();
    j->off = off;
    s->coop_used += j->coop;
    s->note_use();
    j->res.t_start = s->now();
    return j;
}

// The lane's handle, grown to order n when needed (its buffers other than the factors: a few MB).
int dense_handle(lsx_ctx* ctx, DenseLane& L, int64_t n) {
    if (L.lu && L.lu->cap >= n) return lu64_reshape(L.lu, n);
    if (L.lu) {
        lu64_free(L.lu);
        delete L.lu;
        L.lu = nullptr;
    }
    lsx_lu* lu = new (std::nothrow) lsx_lu();
    if (!lu) return LSX_ENOMEM;
    lu->ctx = ctx;
    lu->n = n;
    lu->cap = std::max<int64_t>(n, 4096);
    lu->dtype = LSX_F64;
    lu->own_F = false;
    lu->lookahead = false;  // one stream per lane: the lanes overlap with each other instead
    int rc = lu64_alloc(lu);
    if (rc == LSX_OK) rc = lu_ensure_work(lu, 1);
    if (rc != LSX_OK) {
        lu64_free(lu);
        delete lu;
        return rc;
    }
    L.lu = lu;
    return lu64_reshape(lu, n);
}

void run_dense(lsx_serve* s, DenseLane& L, Job* j) {
    lsx_ctx* ctx = s->ctx;
    const lsx_request& r = j->req;
    const int64_t n = r.n;
    int rc = dense_handle(ctx, L, n);
    if (rc == LSX_OK) {
        lsx_lu* lu = L.lu;
        const hipStream_t st = lu->stream;
        const DenseBlock d = dense_block(s->heap.base + j->off, n);
        lu->F = d.F;
        if (r.synth) {
            *L.h_gen = GenItem{n, 0, 0, r.seed, lu->ld};
            if (hipMemcpyAsync(L.d_gen, L.h_gen, sizeof(GenItem), hipMemcpyHostToDevice, st) != hipSuccess) rc = LSX_EHIP;
            else
                hipLaunchKernelGGL(k_gen_dense, dim3((unsigned)std::min<int64_t>(cdiv(n * n, 256), 2048), 1), dim3(256), 0,
                                   st, (const GenItem*)L.d_gen, d.F, d.x);
        } else if (hipMemcpy2DAsync(d.F, sizeof(double) * lu->ld, r.A, sizeof(double) * n, sizeof(double) * n, n,
                                    hipMemcpyDefault, st) != hipSuccess ||
                   hipMemcpyAsync(d.x, r.b, sizeof(double) * n, hipMemcpyDefault, st) != hipSuccess) {
            rc = LSX_EHIP;
        }
        if (rc == LSX_OK) {
            const int info = lu64_getrf(lu);  // waits for its stream
            if (info < 0) rc = info;
            else {
                j->res.info = info;
                lu->factored = true;
                if (info > 0) j->res.retcode = LSX_RET_FAILURE;
                else {
                    rc = lu64_getrs_dev(lu, 1, d.x, n, d.x, n);
                    j->res.retcode = LSX_RET_SUCCESS;
                    if (rc == LSX_OK &&
                        (hipMemcpyAsync(j->x_out, d.x, sizeof(double) * n, hipMemcpyDefault, st) != hipSuccess ||
                         hipMemcpyAsync(lu->hinfo + 1, lu->dinfo + 1, sizeof(int64_t), hipMemcpyDeviceToHost, st) !=
                             hipSuccess ||
                         hipStreamSynchronize(st) != hipSuccess))
                        rc = LSX_EHIP;
                    if (rc == LSX_OK && lu->hinfo[1]) rc = LSX_ETIMEOUT;
                }
            }
        }
        lu->F = nullptr;
    }
    j->res.status = rc;
    if (rc != LSX_OK) j->res.retcode = LSX_RET_FAILURE;  // no valid result
}

void dense_worker(lsx_serve* s, int lane) {
    (void)hipSetDevice(s->ctx->device);
    DenseLane& L = s->dense[lane];
    while (true) {
        Job* j = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || (j = take_dense(s)) != nullptr; });
            if (!j) return;
        }
        run_dense(s, L, j);
        release_coop(s, j->coop);
        finish(s, j);
    }
}

// ---- GMRES lanes: multi-solve launches (called with s->mu held).  The front request must fit;
// later requests of the same rows-per-thread class join while memory and workgroups allow.
int take_gmres(lsx_serve* s, std::vector<Job*>& group) {
    auto& q = s->lanes[LANE_GMRES];
    if (q.empty()) return 0;
    Job* f = q.front();
    const int avail = s->coop_total - s->coop_used;
    if (f->coop > avail) return 0;
    size_t off = s->heap.alloc(f->footprint);
    if (off == SIZE_MAX) return 0;
    q.pop_front();
    f->off = off;
    group.push_back(f);
    int reserve = f->coop;
    int64_t total = f->coop;
    bool multi = f->coop > 1;  // the group holds a solve of several workgroups
    int scanned = 0;
    for (auto it = q.begin(); it != q.end() && (int)group.size() < MAXG && scanned < 512; ++scanned) {
        Job* j = *it;
        if (j->plan.rpt != f->plan.rpt) {
            ++it;
            continue;
        }
        const int grow = std::max(reserve, j->coop) - reserve;
        if (reserve + grow > avail) break;
        // A solve of several workgroups must start with all of them: its first all-reduce waits
        // for every rank, and a rank that only starts when an earlier solve ends could wait past
        // the bound.  So a group that holds one never needs more workgroups than its launch's pool
        // (every work number is taken at once); groups of one-workgroup solves, which wait on
        // nobody, take about two pool turns of work, so later arrivals join the next launch soon.
        const bool m2 = multi || j->coop > 1;
        const int64_t pool = std::max<int64_t>(reserve + grow, s->gmres_pool);
        if (m2 ? total + j->coop > pool : total >= 2 * (int64_t)s->gmres_pool) break;
        multi = m2;
        off = s->heap.alloc(j->footprint);
        if (off == SIZE_MAX) break;
        j->off = off;
        reserve += grow;
        total += j->coop;
        group.push_back(j);
        it = q.erase(it);
    }
    const double t = s->now();
    for (Job* j : group) j->res.t_start = t;
    s->coop_used += reserve;
    s->note_use();
    return reserve;
}

void run_gmres_group(lsx_serve* s, GmresLane& L, std::vector<Job*>& group) {
    lsx_ctx* ctx = s->ctx;
    const hipStream_t st = L.stream;
    const int cnt = (int)group.size();
    int rc = LSX_OK;
    std::vector<GmresBlock> blk(cnt);
    std::vector<char> has_pl(cnt, 0), bad(cnt, 0);
    bool user = false;
    // inputs into each request's heap block; a caller's CSR is checked before it is used
    for (int i = 0; i < cnt && rc == LSX_OK; ++i) {
        Job* j = group[i];
        const lsx_request& r = j->req;
        const GmresBlock g = gmres_block(s->heap.base + j->off, *j);
        blk[i] = g;
        const int64_t N = j->N;
        has_pl[i] = r.pl_diag != nullptr;
        if (r.synth) {
            hipLaunchKernelGGL(k_gen_convdiff, dim3((unsigned)std::min<int64_t>(cdiv(N, 256), 4096)), dim3(256), 0, st,
                               (int)r.grid_m, r.pe, r.seed, g.rowptr, g.col, g.val, g.b, g.pl);
            has_pl[i] = 1;
        } else if (hipMemcpyAsync(g.rowptr, r.rowptr, sizeof(int32_t) * (N + 1), hipMemcpyDefault, st) != hipSuccess ||
                   (j->nnz > 0 && (hipMemcpyAsync(g.col, r.colind, sizeof(int32_t) * j->nnz, hipMemcpyDefault, st) != hipSuccess ||
                                   hipMemcpyAsync(g.val, r.vals, sizeof(double) * j->nnz, hipMemcpyDefault, st) != hipSuccess)) ||
                   hipMemcpyAsync(g.b, r.b, sizeof(double) * N, hipMemcpyDefault, st) != hipSuccess ||
                   (has_pl[i] && hipMemcpyAsync(g.pl, r.pl_diag, sizeof(double) * N, hipMemcpyDefault, st) != hipSuccess) ||
                   csr_validate(st, N, j->nnz, g.rowptr, g.col, L.d_flags + i) != LSX_OK) {
            rc = LSX_EHIP;
        } else {
            user = true;
        }
    }
    if (rc == LSX_OK && user) {
        if (hipMemcpyAsync(L.h_flags, L.d_flags, sizeof(int) * cnt, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = LSX_EHIP;
        for (int i = 0; i < cnt && rc == LSX_OK; ++i) bad[i] = !group[i]->req.synth && L.h_flags[i] != 0;
    }
    // the launch over the requests that passed
    std::vector<int> idx;
    int total = 0, widest = 0;
    size_t lds = 0;
    for (int i = 0; i < cnt && rc == LSX_OK; ++i) {
        if (bad[i]) continue;
        Job* j = group[i];
        const lsx_request& r = j->req;
        const GmresBlock& g = blk[i];
        const int m = (int)idx.size();
        idx.push_back(i);
        GmresArgs& a = L.h_items[m];
        a.N = j->N;
        a.rowptr = g.rowptr;
        a.col = g.col;
        a.val = g.val;
        a.b = g.b;
        a.x = g.x;
        a.pl = has_pl[i] ? g.pl : nullptr;
        a.pr = nullptr;
        a.V = g.V;
        a.xr = g.xr;
        a.memory = (int)std::max<int64_t>(1, std::min<int64_t>(r.memory, r.maxiters));
        a.restart = r.restart ? 1 : 0;
        a.cap = j->plan.cap;
        a.Rg = nullptr;  // bookkeeping in LDS (the plan is made without the HBM option)
        a.maxiters = r.maxiters;
        a.abstol = r.abstol;
        a.reltol = r.reltol;
        a.P = j->plan.P;
        a.red = g.red;
        a.bar = reinterpret_cast<uint32_t*>(g.red + 4 * (int64_t)j->plan.P);
        a.out = L.d_out + m;
        a.timeout = wait_ticks(ctx, 20.0);
        L.h_first[m] = total;
        total += j->plan.P;
        widest = std::max(widest, j->plan.P);
        lds = std::max(lds, j->plan.dynlds);
    }
    const int nl = (int)idx.size();
    if (rc == LSX_OK && nl > 0) {
        L.h_first[nl] = total;
        const int pool = std::min(total, std::max(widest, s->gmres_pool));
        if (hipMemcpyAsync(L.d_items, L.h_items, sizeof(GmresArgs) * nl, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(L.d_first, L.h_first, sizeof(int) * (nl + 1), hipMemcpyHostToDevice, st) != hipSuccess)
            rc = LSX_EHIP;
        if (rc == LSX_OK)
            rc = gmres_launch_multi(ctx, st, L.d_items, L.d_first, nl, pool, group[idx[0]]->plan.rpt, lds, L.d_ticket);
        if (rc == LSX_OK &&
            hipMemcpyAsync(L.h_out, L.d_out, sizeof(GmresOut) * nl, hipMemcpyDeviceToHost, st) != hipSuccess)
            rc = LSX_EHIP;
        for (int m = 0; m < nl && rc == LSX_OK; ++m) {
            Job* j = group[idx[m]];
            if (hipMemcpyAsync(j->x_out, blk[idx[m]].x, sizeof(double) * j->N, hipMemcpyDefault, st) != hipSuccess)
                rc = LSX_EHIP;
        }
    }
    if (hipStreamSynchronize(st) != hipSuccess && rc == LSX_OK) rc = LSX_EHIP;
    for (int i = 0; i < cnt; ++i) {
        group[i]->res.status = rc != LSX_OK ? rc : (bad[i] ? LSX_EINVAL : LSX_OK);
        group[i]->res.retcode = LSX_RET_FAILURE;  // until a valid result is recorded below
    }
    for (int m = 0; m < nl && rc == LSX_OK; ++m) {
        Job* j = group[idx[m]];
        j->res.iters = L.h_out[m].iters;
        j->res.resid = L.h_out[m].resid;
        if (L.h_out[m].err) j->res.status = LSX_ETIMEOUT;  // a wait gave up: the result is invalid
        else j->res.retcode = L.h_out[m].retcode;
    }
}

void gmres_worker(lsx_serve* s, int lane) {
    (void)hipSetDevice(s->ctx->device);
    GmresLane& L = s->gmres[lane];
    while (true) {
        std::vector<Job*> group;
        int reserve = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || (reserve = take_gmres(s, group)) > 0; });
            if (group.empty()) return;
        }
        run_gmres_group(s, L, group);
        release_coop(s, reserve);
        for (Job* j : group) finish(s, j);
    }
}

// ---- batch lane: continuous batching of small dense systems
void batch_worker(lsx_serve* s) {
    (void)hipSetDevice(s->ctx->device);
    Arena& a = s->arena;
    while (true) {
        std::vector<Job*> group;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || !s->lanes[LANE_BATCH].empty(); });
            if (s->stop) return;
            int64_t ua = 0, uv = 0;
            size_t ub = 0;
            const double t = s->now();
            while (!s->lanes[LANE_BATCH].empty()) {
                Job* j = s->lanes[LANE_BATCH].front();
                const int64_t n = j->req.n;
                if ((int64_t)group.size() + 1 > a.capN || ua + n * n > a.capA || uv + n > a.capV) break;
                if (!group.empty() && ub + j->footprint > a.bytes) break;
                s->lanes[LANE_BATCH].pop_front();
                j->res.t_start = t;
                group.push_back(j);
                ua += n * n;
                uv += n;
                ub += j->footprint;
            }
            s->batch_used += ub;
            s->note_use();
        }
        const int64_t cnt = (int64_t)group.size();
        // metadata: order by decreasing n (longest first), offsets in arrival order
        std::vector<int32_t> idx(cnt);
        for (int64_t i = 0; i < cnt; ++i) idx[i] = (int32_t)i;
        std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return group[x]->req.n > group[y]->req.n; });
        int64_t ao = 0, vo = 0, ngen = 0;
        int rc = LSX_OK;
        for (int64_t i = 0; i < cnt; ++i) {
            const lsx_request& r = group[i]->req;
            a.h32[i] = idx[i];
            a.h32[a.capN + i] = (int32_t)r.n;
            a.h64[i] = ao;
            a.h64[a.capN + i] = vo;
            if (r.synth) a.hgen[ngen++] = GenItem{r.n, ao, vo, r.seed, r.n};
            else if (hipMemcpyAsync(a.A + ao, r.A, sizeof(double) * r.n * r.n, hipMemcpyDefault, a.stream) != hipSuccess ||
                     hipMemcpyAsync(a.b + vo, r.b, sizeof(double) * r.n, hipMemcpyDefault, a.stream) != hipSuccess)
                rc = LSX_EHIP;
            ao += r.n * r.n;
            vo += r.n;
        }
        if (rc == LSX_OK &&
            (hipMemcpyAsync(a.meta32, a.h32, sizeof(int32_t) * 2 * a.capN, hipMemcpyHostToDevice, a.stream) != hipSuccess ||
             hipMemcpyAsync(a.meta64, a.h64, sizeof(int64_t) * 2 * a.capN, hipMemcpyHostToDevice, a.stream) != hipSuccess))
            rc = LSX_EHIP;
        if (rc == LSX_OK && ngen > 0) {
            if (hipMemcpyAsync(a.gen, a.hgen, sizeof(GenItem) * ngen, hipMemcpyHostToDevice, a.stream) != hipSuccess)
                rc = LSX_EHIP;
            else
                hipLaunchKernelGGL(k_gen_dense, dim3(16, (unsigned)ngen), dim3(256), 0, a.stream, (const GenItem*)a.gen, a.A,
                                   a.b);
        }
        if (rc == LSX_OK)
            rc = batch_launch_raw(s->ctx, a.stream, cnt, a.meta32, a.meta32 + a.capN, a.meta64, a.meta64 + a.capN, a.A, a.b,
                                  a.x, a.ipiv, a.perm, a.info, a.counter);
        if (rc == LSX_OK &&
            hipMemcpyAsync(a.hinfo, a.info, sizeof(int64_t) * cnt, hipMemcpyDeviceToHost, a.stream) != hipSuccess)
            rc = LSX_EHIP;
        for (int64_t i = 0; i < cnt && rc == LSX_OK; ++i)
            if (hipMemcpyAsync(group[i]->x_out, a.x + a.h64[a.capN + i], sizeof(double) * group[i]->req.n, hipMemcpyDefault,
                               a.stream) != hipSuccess)
                rc = LSX_EHIP;
        if (hipStreamSynchronize(a.stream) != hipSuccess) rc = LSX_EHIP;
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->batches += 1;
            s->batched_reqs += cnt;
        }
        for (int64_t i = 0; i < cnt; ++i) {
            Job* j = group[i];
            j->res.status = rc;
            j->res.info = rc == LSX_OK ? a.hinfo[i] : 0;
            j->res.retcode = (rc == LSX_OK && a.hinfo[i] == 0) ? LSX_RET_SUCCESS : LSX_RET_FAILURE;
            finish(s, j);
        }
    }
}

int arena_alloc(lsx_ctx* ctx, Arena& a, size_t bytes) {
    a.bytes = bytes;
    a.capA = std::max<int64_t>(512 * 512, (int64_t)(bytes / sizeof(double)));
    a.capV = 4 << 20;
    a.capN = 16384;
    if (hipStreamCreateWithFlags(&a.stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&a.A, sizeof(double) * a.capA) != hipSuccess || hipMalloc(&a.b, sizeof(double) * a.capV) != hipSuccess ||
        hipMalloc(&a.x, sizeof(double) * a.capV) != hipSuccess || hipMalloc(&a.ipiv, sizeof(int64_t) * a.capV) != hipSuccess ||
        hipMalloc(&a.perm, sizeof(int32_t) * a.capV) != hipSuccess || hipMalloc(&a.info, sizeof(int64_t) * a.capN) != hipSuccess ||
        hipMalloc(&a.meta32, sizeof(int32_t) * 2 * a.capN) != hipSuccess ||
        hipMalloc(&a.meta64, sizeof(int64_t) * 2 * a.capN) != hipSuccess ||
        hipMalloc(&a.gen, sizeof(GenItem) * a.capN) != hipSuccess || hipMalloc(&a.counter, 256) != hipSuccess ||
        hipHostMalloc(&a.h32, sizeof(int32_t) * 2 * a.capN, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&a.h64, sizeof(int64_t) * 2 * a.capN, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&a.hgen, sizeof(GenItem) * a.capN, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&a.hinfo, sizeof(int64_t) * a.capN, hipHostMallocDefault) != hipSuccess)
        return fail(ctx, LSX_ENOMEM, "lsx_serve_create: cannot allocate the batch arena");
    return LSX_OK;
}

void arena_free(Arena& a) {
    if (a.stream) (void)hipStreamSynchronize(a.stream);
    (void)hipFree(a.A);
    (void)hipFree(a.b);
    (void)hipFree(a.x);
    (void)hipFree(a.ipiv);
    (void)hipFree(a.perm);
    (void)hipFree(a.info);
    (void)hipFree(a.meta32);
    (void)hipFree(a.meta64);
    (void)hipFree(a.gen);
    (void)hipFree(a.counter);
    if (a.h32) (void)hipHostFree(a.h32);
    if (a.h64) (void)hipHostFree(a.h64);
    if (a.hgen) (void)hipHostFree(a.hgen);
    if (a.hinfo) (void)hipHostFree(a.hinfo);
    if (a.stream) (void)hipStreamDestroy(a.stream);
}

int lanes_alloc(lsx_serve* s, int nd, int ng) {
    s->dense.resize(nd);
    s->gmres.resize(ng);
    for (auto& L : s->dense)
        if (hipMalloc(&L.d_gen, sizeof(GenItem)) != hipSuccess ||
            hipHostMalloc(&L.h_gen, sizeof(GenItem), hipHostMallocDefault) != hipSuccess)
            return fail(s->ctx, LSX_ENOMEM, "lsx_serve_create: cannot allocate a dense lane");
    for (auto& L : s->gmres)
        if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess ||
            hipHostMalloc(&L.h_items, sizeof(GmresArgs) * MAXG, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc(&L.h_first, sizeof(int) * (MAXG + 1), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc(&L.h_out, sizeof(GmresOut) * MAXG, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc(&L.h_flags, sizeof(int) * MAXG, hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&L.d_flags, sizeof(int) * MAXG) != hipSuccess ||
            hipMalloc(&L.d_items, sizeof(GmresArgs) * MAXG) != hipSuccess ||
            hipMalloc(&L.d_first, sizeof(int) * (MAXG + 1)) != hipSuccess ||
            hipMalloc(&L.d_out, sizeof(GmresOut) * MAXG) != hipSuccess || hipMalloc(&L.d_ticket, 256) != hipSuccess)
            return fail(s->ctx, LSX_ENOMEM, "lsx_serve_create: cannot allocate a GMRES lane");
    return LSX_OK;
}

void lanes_free(lsx_serve* s) {
    for (auto& L : s->dense) {
        if (L.lu) {
            lu64_free(L.lu);
            delete L.lu;
        }
        (void)hipFree(L.d_gen);
        if (L.h_gen) (void)hipHostFree(L.h_gen);
    }
    for (auto& L : s->gmres) {
        if (L.stream) (void)hipStreamSynchronize(L.stream);
        if (L.h_items) (void)hipHostFree(L.h_items);
        if (L.h_first) (void)hipHostFree(L.h_first);
        if (L.h_out) (void)hipHostFree(L.h_out);
        if (L.h_flags) (void)hipHostFree(L.h_flags);
        (void)hipFree(L.d_flags);
        (void)hipFree(L.d_items);
        (void)hipFree(L.d_first);
        (void)hipFree(L.d_out);
        (void)hipFree(L.d_ticket);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    s->dense.clear();
    s->gmres.clear();
}

void server_free(lsx_serve* s) {
    lanes_free(s);
    arena_free(s->arena);
    (void)hipFree(s->heap_mem);
}

}  // namespace

extern "C" {

int lsx_serve_create(lsx_ctx* ctx, size_t hbm_budget, int32_t nstreams, lsx_serve** out) {
    if (!ctx || !out || hbm_budget < (64u << 20)) return LSX_EINVAL;
    *out = nullptr;
    DeviceGuard g(ctx);
    lsx_serve* s = new (std::nothrow) lsx_serve();
    if (!s) return LSX_ENOMEM;
    s->ctx = ctx;
    s->t0 = Clock::now();
    s->budget = hbm_budget;
    s->coop_total = std::max(64, ctx->num_cus - COOP_MARGIN);
    const size_t arena_bytes = std::min<size_t>(hbm_budget / 4, (size_t)2 << 30);
    const size_t heap_bytes = (hbm_budget - arena_bytes) & ~(HALIGN - 1);
    int rc = arena_alloc(ctx, s->arena, arena_bytes);
    if (rc == LSX_OK && hipMalloc(&s->heap_mem, heap_bytes) != hipSuccess)
        rc = fail(ctx, LSX_ENOMEM, "lsx_serve_create: cannot allocate a %zu-byte request heap", heap_bytes);
    const int lanes = std::max(2, std::min<int>(nstreams, 8));
    const int nd = lanes / 2, ng = lanes - nd;
    if (rc == LSX_OK) rc = lanes_alloc(s, nd, ng);
    // the GMRES lanes together occupy at most 3/4 of the CUs (one 135 KB-LDS workgroup each), so
    // the dense and batched lanes' kernels always find CUs
    s->gmres_pool = std::max(1, ctx->num_cus * 3 / 4 / ng);
    if (rc) {
        server_free(s);
        delete s;
        return rc;
    }
    s->heap.init(s->heap_mem, heap_bytes);
    s->workers.emplace_back(batch_worker, s);
    for (int i = 0; i < nd; ++i) s->workers.emplace_back(dense_worker, s, i);
    for (int i = 0; i < ng; ++i) s->workers.emplace_back(gmres_worker, s, i);
    *out = s;
    return LSX_OK;
}

int lsx_serve_destroy(lsx_serve* s) {
    if (!s) return LSX_EINVAL;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    for (auto& t : s->workers) t.join();
    {
        DeviceGuard g(s->ctx);
        server_free(s);
    }
    for (auto& kv : s->jobs) delete kv.second;  // every unpolled request, queued ones included
    delete s;
    return LSX_OK;
}

int lsx_serve_submit(lsx_serve* s, const lsx_request* req, double* x_out, uint64_t* ticket) {
    if (!s || !req || !x_out || !ticket) return LSX_EINVAL;
    if (req->kind == 0) {
        if (req->n < 1 || req->dtype != LSX_F64 || (!req->synth && (!req->A || !req->b)))
            return fail(s->ctx, LSX_EINVAL, "lsx_serve_submit: bad dense request");
    } else if (req->kind == 1) {
        if (req->synth ? req->grid_m < 2 : (req->n < 1 || !req->rowptr || !req->b || (req->nnz > 0 && (!req->colind || !req->vals))))
            return fail(s->ctx, LSX_EINVAL, "lsx_serve_submit: bad sparse request");
        if (req->memory < 1 || req->maxiters < 0)
            return fail(s->ctx, LSX_EINVAL, "lsx_serve_submit: memory >= 1 and maxiters >= 0");
    } else {
        return fail(s->ctx, LSX_EINVAL, "lsx_serve_submit: kind must be 0 or 1");
    }
    Job* j = new (std::nothrow) Job();
    if (!j) return LSX_ENOMEM;
    j->req = *req;
    j->x_out = x_out;
    int rc = classify(s, j);
    if (rc) {
        delete j;
        return rc;
    }
    std::lock_guard<std::mutex> g(s->mu);
    const size_t limit = j->lane == LANE_BATCH ? s->arena.bytes : s->heap.size;
    if (j->footprint > limit || j->coop > s->coop_total) {
        delete j;
        return fail(s->ctx, LSX_ENOMEM, "lsx_serve_submit: request needs more HBM (or workgroups) than the whole budget");
    }
    j->ticket = s->next_ticket++;
    j->res.status = LSX_EAGAIN;
    j->res.t_submit = s->now();
    s->jobs[j->ticket] = j;
    s->lanes[j->lane].push_back(j);
    s->submitted += 1;
    s->inflight += 1;
    s->cv.notify_all();
    *ticket = j->ticket;
    return LSX_OK;
}

int lsx_serve_poll(lsx_serve* s, uint64_t ticket, lsx_result* res) {
    if (!s || !res) return LSX_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return fail(s->ctx, LSX_EINVAL, "lsx_serve_poll: unknown ticket");
    Job* j = it->second;
    if (!j->done) return LSX_EAGAIN;
    *res = j->res;
    s->jobs.erase(it);
    delete j;
    return LSX_OK;
}

int lsx_serve_drain(lsx_serve* s) {
    if (!s) return LSX_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->inflight == 0; });
    return LSX_OK;
}

double lsx_serve_clock(lsx_serve* s) { return s ? s->now() : 0.0; }

int lsx_serve_stats(lsx_serve* s, int64_t* o) {
    if (!s || !o) return LSX_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    o[0] = s->submitted;
    o[1] = s->completed;
    o[2] = s->batches;
    o[3] = s->batched_reqs;
    o[4] = (int64_t)s->peak;
    int64_t waiting = 0;
    for (int l = 0; l < NLANES; ++l) waiting += (int64_t)s->lanes[l].size();
    o[5] = waiting;
    return LSX_OK;
}

}  // extern "C"
