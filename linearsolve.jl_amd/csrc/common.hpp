// common.hpp — shared definitions of the gfx950 dense-LU / SimpleGMRES library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "lsx.h"

struct lsx_ctx {
    int device = 0;
    int num_cus = 0;
    int64_t wall_khz = 100000;  // device wall clock rate (kHz), for bounded device-side waits
    double wait_bound_s = 0.0;  // > 0: bound of every later device-side wait (seconds); 0: defaults
    double* scratch = nullptr;  // this GPU's address of the no-allocation scratch array (api.hip)
    std::mutex mu;
    std::string err;
};

namespace lsx {

// Records an error message on the context and returns code.
int fail(lsx_ctx* ctx, int code, const char* fmt, ...);

// Makes the context's device current for the lifetime of the guard.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(const lsx_ctx* c) {
        (void)hipGetDevice(&prev);
        if (c && prev != c->device) (void)hipSetDevice(c->device);
    }
    ~DeviceGuard() {
        int cur = -1;
        (void)hipGetDevice(&cur);
        if (prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};

#define LSX_HIP(ctx, call)                                                                  \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return ::lsx::fail((ctx), LSX_EHIP, "%s failed: %s (%s:%d)", #call,            \
                               hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Device wall-clock ticks of a device-side wait bound: the context's bound when one is set,
// else dflt_s seconds.
inline int64_t wait_ticks(const lsx_ctx* c, double dflt_s) {
    const double s = c->wait_bound_s > 0.0 ? c->wait_bound_s : dflt_s;
    const double t = s * 1e3 * (double)c->wall_khz;
    return t < 1.0 ? 1 : (int64_t)t;
}
inline int64_t rup(int64_t a, int64_t b) { return cdiv(a, b) * b; }

// Leading dimension of resident matrices: a multiple of 32 elements, nudged off multiples of
// 512 so that column strides do not all map to the same HBM channel.
inline int64_t pad_ld(int64_t n) {
    int64_t ld = rup(n < 1 ? 1 : n, 32);
    if (ld % 512 == 0) ld += 32;
    return ld;
}

// Whether p is memory of this device (host memory, pinned or pageable, is not).
inline bool is_device_ptr(int device, const void* p) {
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeDevice && at.device == device;
}

// Device-side helpers ---------------------------------------------------------------------
__device__ __forceinline__ uint64_t ld_relaxed_u64(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed_u64(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_relaxed_u32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_coherent_f64(const double* p) {
    return __longlong_as_double(
        (long long)__hip_atomic_load((const uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
// Flag-in-word message (low 32 bits payload, high 32 bits sequence number): a reader that sees
// the expected sequence number in a 64-bit word also sees that word's payload.
__device__ __forceinline__ uint64_t ll_pack(uint32_t payload, uint32_t seq) {
    return ((uint64_t)seq << 32) | payload;
}

// Per-kernel-class device timing: HIP events around launches while enabled, read back after
// the owning stream has been synchronized.
enum { PK_PANEL = 0, PK_LASWP = 1, PK_GEMM = 2, PK_PERM = 3, PK_TRSV = 4, PK_REFINE = 5, PK_CONVERT = 6,
       PK_OTHER = 7, PK_GEMM_SMALL = 8 };
struct Profiler {
    bool on = false;
    lsx_profile acc{};
    struct Rec {
        hipEvent_t s, e;
        int cls;
        double flops, bytes;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    int begin(hipStream_t st, int cls, double flops, double bytes) {
        if (!on) return -1;
        Rec r{get(), get(), cls, flops, bytes};
        (void)hipEventRecord(r.s, st);
        pending.push_back(r);
        return (int)pending.size() - 1;
    }
    void end(hipStream_t st, int idx) {
        if (idx >= 0) (void)hipEventRecord(pending[idx].e, st);
    }
    void collect() {
        for (auto& r : pending) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, r.s, r.e) == hipSuccess) {
                acc.launches[r.cls] += 1;
                acc.ms[r.cls] += ms;
                acc.flops[r.cls] += r.flops;
                acc.bytes[r.cls] += r.bytes;
            }
            pool.push_back(r.s);
            pool.push_back(r.e);
        }
        pending.clear();
    }
    void reset() { acc = lsx_profile{}; }
    ~Profiler() {
        for (auto& r : pending) {
            (void)hipEventDestroy(r.s);
            (void)hipEventDestroy(r.e);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace lsx
