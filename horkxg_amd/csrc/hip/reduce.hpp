// reduce.hpp -- contention-free device-wide reductions.
//
// A single global address updated by every wave serialises in L2 (a CAS loop
// per wave costs tens of milliseconds at 10^5 waves). Reductions here run a
// capped grid with grid-stride loops, combine inside each workgroup through
// LDS, write one partial per workgroup, and fold the partials in one final
// workgroup: no atomics at all, and a fixed combination order.
#pragma once

#include "internal.hpp"

namespace gpart_hip {

constexpr int kReduceThreads = 256;
constexpr int kReduceMaxBlocks = 1024;

// Op: T identity(); T operator()(T, T); T load(int64 i) via Load functor.
template <typename T, typename Op>
__device__ T block_reduce(T value, Op op, T* shared) {
    const int lane = lane_id(), wave = threadIdx.x / kWave;
#pragma unroll
    for (int m = kWave / 2; m > 0; m /= 2) {
        T other;
        static_assert(sizeof(T) == 8, "block_reduce handles 8-byte values");
        const auto bits = shfl_xor_u64(*reinterpret_cast<unsigned long long*>(&value), m, kWave);
        *reinterpret_cast<unsigned long long*>(&other) = bits;
        value = op(value, other);
    }
    if (lane == 0) shared[wave] = value;
    __syncthreads();
    T total = op.identity();
    if (threadIdx.x == 0)
        for (int w = 0; w < static_cast<int>(blockDim.x) / kWave; ++w) total = op(total, shared[w]);
    return total;  // valid in thread 0
}

template <typename T, typename Op, typename Load>
__global__ void __launch_bounds__(kReduceThreads) reduce_partials_kernel(Load load, std::int64_t n, Op op,
                                                                         T* partials) {
    __shared__ T shared[kReduceThreads / kWave];
    T acc = op.identity();
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<std::int64_t>(gridDim.x) * blockDim.x)
        acc = op(acc, load(i));
    const T total = block_reduce(acc, op, shared);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

template <typename T, typename Op>
__global__ void __launch_bounds__(kReduceThreads) reduce_finish_kernel(const T* partials, int count, Op op, T* out) {
    __shared__ T shared[kReduceThreads / kWave];
    T acc = op.identity();
    for (int i = threadIdx.x; i < count; i += blockDim.x) acc = op(acc, partials[i]);
    const T total = block_reduce(acc, op, shared);
    if (threadIdx.x == 0) *out = total;
}

// Reduce n loaded values into *out (device); out may be read back by the caller.
template <typename T, typename Op, typename Load>
void device_reduce(gpart_ctx* ctx, Load load, std::int64_t n, Op op, T* out) {
    const std::int64_t want = (n + kReduceThreads * 4 - 1) / (kReduceThreads * 4);
    const int blocks = static_cast<int>(want < 1 ? 1 : (want > kReduceMaxBlocks ? kReduceMaxBlocks : want));
    Scratch<T> partials(ctx, blocks);
    hipLaunchKernelGGL((reduce_partials_kernel<T, Op, Load>), dim3(blocks), dim3(kReduceThreads), 0, ctx->stream,
                       load, n, op, partials.get());
    hipLaunchKernelGGL((reduce_finish_kernel<T, Op>), dim3(1), dim3(kReduceThreads), 0, ctx->stream,
                       partials.get(), blocks, op, out);
    GPART_LAUNCH_CHECK();
}

struct MaxU64 {
    __device__ __forceinline__ unsigned long long identity() const { return 0; }
    __device__ __forceinline__ unsigned long long operator()(unsigned long long a, unsigned long long b) const {
        return a > b ? a : b;
    }
};
struct MinU64 {
    __device__ __forceinline__ unsigned long long identity() const { return ~0ULL; }
    __device__ __forceinline__ unsigned long long operator()(unsigned long long a, unsigned long long b) const {
        return a < b ? a : b;
    }
};
struct SumU64 {
    __device__ __forceinline__ unsigned long long identity() const { return 0; }
    __device__ __forceinline__ unsigned long long operator()(unsigned long long a, unsigned long long b) const {
        return a + b;
    }
};
// Saturating at UINT64_MAX (CheckedSum, src/initial_partition_gpu.cu:31-35).
struct CheckedSumU64 {
    __device__ __forceinline__ unsigned long long identity() const { return 0; }
    __device__ __forceinline__ unsigned long long operator()(unsigned long long a, unsigned long long b) const {
        return ~0ULL - a < b ? ~0ULL : a + b;
    }
};
// Saturating at INT64_MAX/2 + 1 (RefinementWeightSum, src/refine.cu:367-372).
struct LimitedSumU64 {
    static constexpr unsigned long long kLimit = 0x7fffffffffffffffULL / 2;
    __device__ __forceinline__ unsigned long long identity() const { return 0; }
    __device__ __forceinline__ unsigned long long operator()(unsigned long long a, unsigned long long b) const {
        return a > kLimit || b > kLimit || a > kLimit - b ? kLimit + 1 : a + b;
    }
};

template <typename W>
struct WidenLoad {
    const W* data;
    __device__ __forceinline__ unsigned long long operator()(std::int64_t i) const {
        return static_cast<unsigned long long>(data[i]);
    }
};

struct NonzeroLoad {
    const unsigned long long* data;
    __device__ __forceinline__ unsigned long long operator()(std::int64_t i) const { return data[i] != 0; }
};

}  // namespace gpart_hip
