// serve.hip — continuous-batching request loop for mixed dense / sparse solves on one GPU.
//
// Requests are submitted from any host thread into the queue of their lane, in arrival order.
// A lane dispatches its next request(s) when it is free and their estimated HBM footprint fits
// what the running requests leave of the budget (admission), and releases the memory when they
// complete:
//   batch lane        dense systems with n <= 512: every waiting system that fits goes into one
//                     persistent batched-LU launch, so requests that arrive while a launch runs
//                     join the next one (continuous batching);
//   dense lane        dense systems with n > 512 (panel workgroups wait on each other: at most
//                     32 of them, one per CU);
//   large GMRES lane  SimpleGMRES solves that need more than one workgroup (at most 128);
//   GMRES lanes       SimpleGMRES solves that fit one workgroup (N <= 8192), several at once.
// The two lanes with cooperating workgroups need at most 160 CUs together, so both always fit
// on the GPU's 256 once the other (always terminating) kernels drain.
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
#include <deque>
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
        A[it.aoff + idx] = (row == col ? 1.0 : 0.0) - sc * gauss(it.seed, (uint64_t)idx);
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

enum Lane { LANE_BATCH = 0, LANE_DENSE = 1, LANE_GMRES_BIG = 2, LANE_GMRES = 3, NLANES = 4 };

struct Job {
    uint64_t ticket = 0;
    lsx_request req{};
    double* x_out = nullptr;
    size_t footprint = 0;
    int lane = 0;
    bool done = false;
    lsx_result res{};
};

struct Arena {
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
    // pinned staging
    int32_t* h32 = nullptr;
    int64_t* h64 = nullptr;
    GenItem* hgen = nullptr;
    int64_t* hinfo = nullptr;
    hipStream_t stream = nullptr;
};

}  // namespace

struct lsx_serve {
    lsx_ctx* ctx = nullptr;
    size_t budget = 0, used = 0, peak = 0;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv;       // workers wait for lane work
    std::condition_variable cv_done;  // drain waits for completions
    std::deque<Job*> lanes[NLANES];   // waiting for their lane, arrival order
    std::unordered_map<uint64_t, Job*> jobs;
    uint64_t next_ticket = 1;
    bool stop = false;
    int64_t submitted = 0, completed = 0, batches = 0, batched_reqs = 0, inflight = 0;
    Arena arena;
    std::vector<std::thread> workers;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

size_t footprint(const lsx_request& r, int& lane) {
    if (r.kind == 0) {
        const size_t n = (size_t)r.n;
        if (r.n <= 512) {
            lane = LANE_BATCH;
            return 8 * n * n + 64 * n;
        }
        lane = LANE_DENSE;
        return 16 * n * n + 256 * n + (2 << 20);  // A staging + resident factors and workspaces
    }
    const int64_t N = r.synth ? (int64_t)r.grid_m * r.grid_m : r.n;
    const int64_t nnz = r.synth ? convdiff_nnz(r.grid_m) : r.nnz;
    const int64_t mem = std::max<int64_t>(1, std::min<int64_t>(r.memory, r.maxiters));
    lane = (N <= 8192) ? LANE_GMRES : LANE_GMRES_BIG;
    return (size_t)(24 * nnz + 8 * N * (mem + 8)) + (1 << 20);
}

bool fits(lsx_serve* s, size_t f) { return s->used == 0 || s->used + f <= s->budget; }

void reserve_locked(lsx_serve* s, size_t f) {
    s->used += f;
    s->peak = std::max(s->peak, s->used);
}

void finish(lsx_serve* s, Job* j) {
    std::lock_guard<std::mutex> g(s->mu);
    j->res.t_done = s->now();
    j->done = true;
    s->used -= j->footprint;
    s->completed += 1;
    s->inflight -= 1;
    s->cv.notify_all();
    s->cv_done.notify_all();
}

// Next request of a lane, once the lane is free and its footprint fits the budget.
Job* take(lsx_serve* s, int lane) {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait(lk, [&] {
        return s->stop || (!s->lanes[lane].empty() && fits(s, s->lanes[lane].front()->footprint));
    });
    if (s->stop) return nullptr;
    Job* j = s->lanes[lane].front();
    s->lanes[lane].pop_front();
    reserve_locked(s, j->footprint);
    j->res.t_start = s->now();
    return j;
}

// ---- coop and gmres lanes: one request at a time per worker, through the C-ABI entry points
void run_dense_single(lsx_serve* s, Job* j) {
    lsx_ctx* ctx = s->ctx;
    const lsx_request& r = j->req;
    const int64_t n = r.n;
    double *dA = nullptr, *db = nullptr, *dx = nullptr;
    lsx_lu* lu = nullptr;
    int rc = LSX_OK;
    if (hipMalloc(&dA, sizeof(double) * n * n) != hipSuccess || hipMalloc(&db, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&dx, sizeof(double) * n) != hipSuccess) {
        rc = LSX_ENOMEM;
    } else {
        if (r.synth) {
            GenItem it{n, 0, 0, r.seed};
            GenItem* dit = nullptr;
            if (hipMalloc(&dit, sizeof(GenItem)) != hipSuccess || hipMemcpy(dit, &it, sizeof it, hipMemcpyHostToDevice) != hipSuccess)
                rc = LSX_EHIP;
            else {
                hipLaunchKernelGGL(k_gen_dense, dim3((unsigned)std::min<int64_t>(cdiv(n * n, 256), 2048), 1), dim3(256), 0, 0,
                                   dit, dA, db);
                if (hipDeviceSynchronize() != hipSuccess) rc = LSX_EHIP;
            }
            (void)hipFree(dit);
        } else {
            if (hipMemcpy(dA, r.A, sizeof(double) * n * n, hipMemcpyDefault) != hipSuccess ||
                hipMemcpy(db, r.b, sizeof(double) * n, hipMemcpyDefault) != hipSuccess)
                rc = LSX_EHIP;
        }
        if (rc == LSX_OK) rc = lsx_lu_create(ctx, n, LSX_F64, &lu);
        if (rc == LSX_OK) {
            const int info = lsx_lu_factor(lu, dA, n);
            if (info < 0) rc = info;
            else {
                j->res.info = info;
                if (info > 0) j->res.retcode = LSX_RET_FAILURE;
                else {
                    rc = lsx_lu_solve(lu, 1, db, n, dx, n);
                    j->res.retcode = LSX_RET_SUCCESS;
                }
            }
        }
        if (rc == LSX_OK && hipMemcpy(j->x_out, dx, sizeof(double) * n, hipMemcpyDefault) != hipSuccess) rc = LSX_EHIP;
    }
    if (lu) lsx_lu_destroy(lu);
    (void)hipFree(dA);
    (void)hipFree(db);
    (void)hipFree(dx);
    j->res.status = rc;
}

void run_gmres(lsx_serve* s, Job* j) {
    lsx_ctx* ctx = s->ctx;
    const lsx_request& r = j->req;
    int rc = LSX_OK;
    lsx_csr* A = nullptr;
    int32_t *rp = nullptr, *ci = nullptr;
    double *va = nullptr, *db = nullptr, *dd = nullptr, *dx = nullptr;
    const double* b = r.b;
    const double* pl = r.pl_diag;
    int64_t N = r.n;
    if (r.synth) {
        const int64_t m = r.grid_m;
        N = m * m;
        const int64_t nnz = convdiff_nnz(m);
        if (hipMalloc(&rp, sizeof(int32_t) * (N + 1)) != hipSuccess || hipMalloc(&ci, sizeof(int32_t) * nnz) != hipSuccess ||
            hipMalloc(&va, sizeof(double) * nnz) != hipSuccess || hipMalloc(&db, sizeof(double) * N) != hipSuccess ||
            hipMalloc(&dd, sizeof(double) * N) != hipSuccess)
            rc = LSX_ENOMEM;
        else {
            hipLaunchKernelGGL(k_gen_convdiff, dim3((unsigned)std::min<int64_t>(cdiv(N, 256), 4096)), dim3(256), 0, 0,
                               (int)m, r.pe, r.seed, rp, ci, va, db, dd);
            if (hipDeviceSynchronize() != hipSuccess) rc = LSX_EHIP;
            if (rc == LSX_OK) rc = lsx_csr_create(ctx, N, nnz, rp, ci, va, &A);
            b = db;
            pl = dd;
        }
    } else {
        rc = lsx_csr_create(ctx, N, r.nnz, r.rowptr, r.colind, r.vals, &A);
    }
    if (rc == LSX_OK && hipMalloc(&dx, sizeof(double) * N) != hipSuccess) rc = LSX_ENOMEM;
    if (rc == LSX_OK) {
        lsx_gmres_stats st{};
        rc = lsx_gmres_solve(A, b, dx, r.memory, r.restart, r.abstol, r.reltol, r.maxiters, pl, nullptr, &st);
        if (rc == LSX_OK) {
            j->res.retcode = st.retcode;
            j->res.iters = st.iters;
            j->res.resid = st.resid;
            if (hipMemcpy(j->x_out, dx, sizeof(double) * N, hipMemcpyDefault) != hipSuccess) rc = LSX_EHIP;
        }
    }
    if (A) lsx_csr_destroy(A);
    (void)hipFree(rp);
    (void)hipFree(ci);
    (void)hipFree(va);
    (void)hipFree(db);
    (void)hipFree(dd);
    (void)hipFree(dx);
    j->res.status = rc;
}

void single_worker(lsx_serve* s, int lane) {
    (void)hipSetDevice(s->ctx->device);
    while (Job* j = take(s, lane)) {
        if (j->req.kind == 0) run_dense_single(s, j);
        else run_gmres(s, j);
        finish(s, j);
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
            s->cv.wait(lk, [&] {
                return s->stop || (!s->lanes[LANE_BATCH].empty() && fits(s, s->lanes[LANE_BATCH].front()->footprint));
            });
            if (s->stop) return;
            int64_t ua = 0, uv = 0;
            const double t = s->now();
            while (!s->lanes[LANE_BATCH].empty()) {
                Job* j = s->lanes[LANE_BATCH].front();
                const int64_t n = j->req.n;
                if ((int64_t)group.size() + 1 > a.capN || ua + n * n > a.capA || uv + n > a.capV) break;
                if (!group.empty() && !fits(s, j->footprint)) break;
                s->lanes[LANE_BATCH].pop_front();
                reserve_locked(s, j->footprint);
                j->res.t_start = t;
                group.push_back(j);
                ua += n * n;
                uv += n;
            }
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
            if (r.synth) a.hgen[ngen++] = GenItem{r.n, ao, vo, r.seed};
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
    const size_t arena_bytes = std::min<size_t>(hbm_budget / 4, (size_t)2 << 30);
    int rc = arena_alloc(ctx, s->arena, arena_bytes);
    if (rc) {
        arena_free(s->arena);
        delete s;
        return rc;
    }
    s->budget = hbm_budget - arena_bytes;
    const int ng = std::max(1, std::min<int>(nstreams, 8));
    s->workers.emplace_back(batch_worker, s);
    s->workers.emplace_back(single_worker, s, (int)LANE_DENSE);
    s->workers.emplace_back(single_worker, s, (int)LANE_GMRES_BIG);
    for (int i = 0; i < ng; ++i) s->workers.emplace_back(single_worker, s, (int)LANE_GMRES);
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
    arena_free(s->arena);
    for (auto& kv : s->jobs) delete kv.second;
    delete s;
    return LSX_OK;
}

int lsx_serve_submit(lsx_serve* s, const lsx_request* req, double* x_out, uint64_t* ticket) {
    if (!s || !req || !x_out || !ticket) return LSX_EINVAL;
    if (req->kind == 0) {
        if (req->n < 1 || req->dtype != LSX_F64 || (!req->synth && (!req->A || !req->b)))
            return fail(s->ctx, LSX_EINVAL, "lsx_serve_submit: bad dense request");
    } else if (req->kind == 1) {
        if (req->synth ? req->grid_m < 2 : (req->n < 1 || !req->rowptr || !req->b))
            return fail(s->ctx, LSX_EINVAL, "lsx_serve_submit: bad sparse request");
        if (req->memory < 1) return fail(s->ctx, LSX_EINVAL, "lsx_serve_submit: memory >= 1");
    } else {
        return fail(s->ctx, LSX_EINVAL, "lsx_serve_submit: kind must be 0 or 1");
    }
    Job* j = new (std::nothrow) Job();
    if (!j) return LSX_ENOMEM;
    j->req = *req;
    j->x_out = x_out;
    j->footprint = footprint(*req, j->lane);
    std::lock_guard<std::mutex> g(s->mu);
    if (j->footprint > s->budget) {
        delete j;
        return fail(s->ctx, LSX_ENOMEM, "lsx_serve_submit: request needs more HBM than the whole budget");
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
