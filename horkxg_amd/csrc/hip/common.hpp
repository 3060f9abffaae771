// common.hpp -- device helpers shared by the gfx950 kernels.
//
// CDNA4 executes 64-lane wavefronts. Kernels here map a vertex to a "group" of
// G lanes (G in {4, 8, 16, 32, 64}) or to a whole workgroup; group reductions use
// width-limited shuffles, which lower to DPP / ds_swizzle row operations.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gpart_hip {

constexpr int kWave = 64;
constexpr std::int32_t kSclpInvalid = 0x7fffffff;
constexpr std::uint64_t kInvalidKey = ~std::uint64_t{0};

// The public "lowbias32" integer hash.
__device__ __forceinline__ std::uint32_t mix32(std::uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    return x ^ (x >> 16);
}

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x & (kWave - 1)); }

__device__ __forceinline__ std::uint64_t lanemask_lt() {
    const int lane = lane_id();
    return lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
}

template <typename T>
__device__ __forceinline__ T shfl(T v, int src, int width) {
    return __shfl(v, src, width);
}

__device__ __forceinline__ unsigned long long shfl_u64(unsigned long long v, int src, int width) {
    const unsigned lo = __shfl(static_cast<unsigned>(v), src, width);
    const unsigned hi = __shfl(static_cast<unsigned>(v >> 32), src, width);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int mask, int width) {
    const unsigned lo = __shfl_xor(static_cast<unsigned>(v), mask, width);
    const unsigned hi = __shfl_xor(static_cast<unsigned>(v >> 32), mask, width);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

__device__ __forceinline__ unsigned long long shfl_down_u64(unsigned long long v, int delta, int width) {
    const unsigned lo = __shfl_down(static_cast<unsigned>(v), delta, width);
    const unsigned hi = __shfl_down(static_cast<unsigned>(v >> 32), delta, width);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

__device__ __forceinline__ unsigned long long shfl_up_u64(unsigned long long v, int delta, int width) {
    const unsigned lo = __shfl_up(static_cast<unsigned>(v), delta, width);
    const unsigned hi = __shfl_up(static_cast<unsigned>(v >> 32), delta, width);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

__device__ __forceinline__ long long shfl_xor_i64(long long v, int mask, int width) {
    return static_cast<long long>(shfl_xor_u64(static_cast<unsigned long long>(v), mask, width));
}

template <int G>
__device__ __forceinline__ unsigned long long group_sum_u64(unsigned long long v) {
#pragma unroll
    for (int m = G / 2; m > 0; m /= 2) v += shfl_xor_u64(v, m, G);
    return v;
}

template <int G>
__device__ __forceinline__ long long group_sum_i64(long long v) {
#pragma unroll
    for (int m = G / 2; m > 0; m /= 2) v += shfl_xor_i64(v, m, G);
    return v;
}

template <int G>
__device__ __forceinline__ unsigned long long group_max_u64(unsigned long long v) {
#pragma unroll
    for (int m = G / 2; m > 0; m /= 2) {
        const auto o = shfl_xor_u64(v, m, G);
        v = o > v ? o : v;
    }
    return v;
}

template <int G>
__device__ __forceinline__ bool group_any(bool p) {
    int v = p ? 1 : 0;
#pragma unroll
    for (int m = G / 2; m > 0; m /= 2) v |= __shfl_xor(v, m, G);
    return v != 0;
}

// Width traits: Standard = int32 offsets / uint32 weights, Big = int64 / uint64.
template <int Width>
struct WidthTypes;
template <>
struct WidthTypes<0> {
    using O = std::int32_t;
    using W = std::uint32_t;
};
template <>
struct WidthTypes<1> {
    using O = std::int64_t;
    using W = std::uint64_t;
};

__device__ __forceinline__ void atomic_add_u64(unsigned long long* p, unsigned long long v) {
    atomicAdd(p, v);
}

inline int ceil_div_int(std::int64_t a, std::int64_t b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace gpart_hip
