// wave.hpp — wave-wide (64-lane) reductions on DPP and permlane swaps, and the pivot key of
// partial pivoting, shared by the panel kernels of the blocked LU and the batched LU.
#pragma once
#include <climits>
#include <cstdint>

#include <hip/hip_runtime.h>

namespace lsx {

// ---- wave-wide reductions (64 lanes): DPP within rows of 16 lanes, then two cross-row steps
template <int CTRL> __device__ __forceinline__ uint32_t dpp32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ __forceinline__ uint64_t dpp64(uint64_t x) {
    return ((uint64_t)dpp32<CTRL>((uint32_t)(x >> 32)) << 32) | dpp32<CTRL>((uint32_t)x);
}
// Value of the lane 16 (32) apart across rows of 16 (halves of 32), by v_permlane16_swap /
// v_permlane32_swap of x with itself: the swap gives each lane its own value in one result and the
// partner's in the other.
__device__ __forceinline__ uint32_t xrow16(uint32_t x) {
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    return ((threadIdx.x >> 4) & 1) ? r[0] : r[1];
}
__device__ __forceinline__ uint32_t xrow32(uint32_t x) {
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return ((threadIdx.x >> 5) & 1) ? r[0] : r[1];
}
__device__ __forceinline__ uint64_t xrow16_64(uint64_t x) {
    return ((uint64_t)xrow16((uint32_t)(x >> 32)) << 32) | xrow16((uint32_t)x);
}
__device__ __forceinline__ uint64_t xrow32_64(uint64_t x) {
    return ((uint64_t)xrow32((uint32_t)(x >> 32)) << 32) | xrow32((uint32_t)x);
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t x) {
    uint64_t y;
    y = dpp64<0xB1>(x);  x = y > x ? y : x;   // quad_perm [1,0,3,2]: lane ^ 1
    y = dpp64<0x4E>(x);  x = y > x ? y : x;   // quad_perm [2,3,0,1]: lane ^ 2
    y = dpp64<0x141>(x); x = y > x ? y : x;   // row_half_mirror: pairs the two quads of each 8
    y = dpp64<0x140>(x); x = y > x ? y : x;   // row_mirror: pairs the two halves of each 16
    y = xrow16_64(x);    x = y > x ? y : x;
    y = xrow32_64(x);    x = y > x ? y : x;
    return x;
}
__device__ __forceinline__ int wave_min_i32(int x) {
    int y;
    y = (int)dpp32<0xB1>((uint32_t)x);  x = y < x ? y : x;
    y = (int)dpp32<0x4E>((uint32_t)x);  x = y < x ? y : x;
    y = (int)dpp32<0x141>((uint32_t)x); x = y < x ? y : x;
    y = (int)dpp32<0x140>((uint32_t)x); x = y < x ? y : x;
    y = (int)xrow16((uint32_t)x);       x = y < x ? y : x;
    y = (int)xrow32((uint32_t)x);       x = y < x ? y : x;
    return x;
}
// Lane holding the best (u, lr) pair of the wave: largest u, then smallest logical row lr.
// u = 0 marks "no candidate"; the result is uniform across the wave.
__device__ __forceinline__ int wave_argmax(uint64_t u, int lr) {
    const uint64_t m = wave_max_u64(u);
    const uint64_t tie = __ballot(u == m);
    if (__popcll(tie) == 1 || m == 0) return __ffsll((unsigned long long)tie) - 1;
    const int mn = wave_min_i32(u == m ? lr : INT_MAX);
    return __ffsll((unsigned long long)__ballot(u == m && lr == mn)) - 1;
}
// Pivot key: |a| as its (monotone) bit pattern + 1, NaN counted as 0; 0 = not a candidate.
template <typename T> __device__ __forceinline__ uint64_t pivot_key(T x) {
    double t = fabs((double)x);
    if (!(t == t)) t = 0.0;
    return (uint64_t)__double_as_longlong(t) + 1u;
}

}  // namespace lsx
