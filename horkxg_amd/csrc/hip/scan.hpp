// scan.hpp -- generic device-wide scan (reduce-then-scan) for gfx950.
//
// A tile is 256 threads x 8 items in blocked arrangement. Pass 1 reduces each
// tile; pass 2 scans the tile totals (recursively); pass 3 rescans each tile
// with its carried-in prefix and hands (index, exclusive prefix, item) to a
// store functor. Up to 1024 tiles pass 3 reduces the totals below its tile
// itself and pass 2 is not launched. Load/store functors let callers fuse
// their producers and consumers (flags -> compaction, (value, head) ->
// segmented prefix) into the scan instead of materialising temporaries.
#pragma once

#include "internal.hpp"

namespace gpart_hip {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr std::int64_t kScanInlineTiles = 1024;  // up to here two launches (the tiles' totals reduced inline)

template <typename T>
struct SumOp {
    __device__ __forceinline__ T operator()(const T& a, const T& b) const { return a + b; }
    __device__ __forceinline__ T identity() const { return T(0); }
};

// Segmented sum: (value, head) pairs; a head restarts the sum.
struct SegValue {
    unsigned long long value;
    unsigned int head;
};
struct SegOp {
    __device__ __forceinline__ SegValue operator()(const SegValue& a, const SegValue& b) const {
        return b.head ? b : SegValue{a.value + b.value, a.head};
    }
    __device__ __forceinline__ SegValue identity() const { return SegValue{0, 0}; }
};

template <typename T>
__device__ __forceinline__ T shfl_up_any(const T& v, int delta) {
    static_assert(sizeof(T) % 4 == 0, "scan value must be a multiple of 4 bytes");
    T out;
    const auto* src = reinterpret_cast<const int*>(&v);
    auto* dst = reinterpret_cast<int*>(&out);
#pragma unroll
    for (int i = 0; i < static_cast<int>(sizeof(T) / 4); ++i) dst[i] = __shfl_up(src[i], delta, kWave);
    return out;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive prefix
// and writes the block total to *total.
template <typename T, typename Op>
__device__ T block_exclusive_scan(T value, Op op, T* total, T* shared_waves) {
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const int waves = blockDim.x / kWave;
    T inclusive = value;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const T other = shfl_up_any(inclusive, d);
        if (lane >= d) inclusive = op(other, inclusive);
    }
    if (lane == kWave - 1) shared_waves[wave] = inclusive;
    __syncthreads();
    if (threadIdx.x == 0) {
        T running = op.identity();
        for (int w = 0; w < waves; ++w) {
            const T t = shared_waves[w];
            shared_waves[w] = running;
            running = op(running, t);
        }
        shared_waves[waves] = running;
    }
    __syncthreads();
    const T wave_prefix = shared_waves[wave];
    T exclusive = shfl_up_any(inclusive, 1);
    if (lane == 0) exclusive = op.identity();
    exclusive = op(wave_prefix, exclusive);
    *total = shared_waves[waves];
    __syncthreads();
    return exclusive;
}

template <typename T, typename Op, typename Load>
__global__ void __launch_bounds__(kScanThreads) scan_reduce_kernel(Load load, std::int64_t n, Op op, T* partials) {
    __shared__ T shared_waves[kScanThreads / kWave + 1];
    const std::int64_t base = static_cast<std::int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    T acc = op.identity();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const std::int64_t idx = base + i;
        if (idx < n) acc = op(acc, load(idx));
    }
    T total;
    (void)block_exclusive_scan(acc, op, &total, shared_waves);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// partial_prefix: null (one tile), the tiles' exclusive prefixes, or (tile_totals)
// the tiles' totals, which each workgroup then reduces up to its own tile, in
// order, itself: at most kScanInlineTiles tiles, so at most that many L2 reads
// per workgroup, and the launch that scans the totals is saved.
template <typename T, typename Op, typename Load, typename Store>
__global__ void __launch_bounds__(kScanThreads) scan_downsweep_kernel(Load load, Store store, std::int64_t n, Op op,
                                                                      const T* partial_prefix, int tile_totals) {
    __shared__ T shared_waves[kScanThreads / kWave + 1];
    const std::int64_t base = static_cast<std::int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    T items[kScanItems];
    T acc = op.identity();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const std::int64_t idx = base + i;
        items[i] = idx < n ? load(idx) : op.identity();
        acc = op(acc, items[i]);
    }
    T carry = op.identity();
    if (partial_prefix && tile_totals) {
        // thread t reduces totals [t * per, (t + 1) * per) below this tile; the
        // block scan's total combines the threads in order
        const int below = static_cast<int>(blockIdx.x);
        const int per = (below + kScanThreads - 1) / kScanThreads;
        T part = op.identity();
        for (int i = 0; i < per; ++i) {
            const int idx = static_cast<int>(threadIdx.x) * per + i;
            if (idx < below) part = op(part, partial_prefix[idx]);
        }
        (void)block_exclusive_scan(part, op, &carry, shared_waves);
    } else if (partial_prefix) {
        carry = partial_prefix[blockIdx.x];
    }
    T total;
    T prefix = block_exclusive_scan(acc, op, &total, shared_waves);
    if (partial_prefix) prefix = op(carry, prefix);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const std::int64_t idx = base + i;
        if (idx < n) store(idx, prefix, items[i]);
        prefix = op(prefix, items[i]);
    }
}

template <typename T>
struct ArrayLoad {
    const T* data;
    __device__ __forceinline__ T operator()(std::int64_t i) const { return data[i]; }
};

template <typename T>
struct ExclusiveStore {
    T* out;
    __device__ __forceinline__ void operator()(std::int64_t i, const T& prefix, const T&) const { out[i] = prefix; }
};

template <typename T>
struct InclusiveStore {
    T* out;
    __device__ __forceinline__ void operator()(std::int64_t i, const T& prefix, const T& item) const {
        out[i] = prefix + item;
    }
};

// Host driver: scans n items of load() into store(); returns nothing. Temporary
// partial arrays come from the context allocator.
template <typename T, typename Op, typename Load, typename Store>
void device_scan(gpart_ctx* ctx, Load load, Store store, std::int64_t n, Op op) {
    if (n <= 0) return;
    const std::int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles == 1) {
        hipLaunchKernelGGL((scan_downsweep_kernel<T, Op, Load, Store>), dim3(1), dim3(kScanThreads), 0,
                           ctx->stream, load, store, n, op, static_cast<const T*>(nullptr), 0);
        GPART_LAUNCH_CHECK();
        return;
    }
    Scratch<T> partials(ctx, tiles);
    hipLaunchKernelGGL((scan_reduce_kernel<T, Op, Load>), dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0,
                       ctx->stream, load, n, op, partials.get());
    GPART_LAUNCH_CHECK();
    const bool inline_totals = tiles <= kScanInlineTiles;
    // Otherwise: exclusive scan of the tile totals, in place.
    if (!inline_totals)
        device_scan<T, Op>(ctx, ArrayLoad<T>{partials.get()}, ExclusiveStore<T>{partials.get()}, tiles, op);
    hipLaunchKernelGGL((scan_downsweep_kernel<T, Op, Load, Store>), dim3(static_cast<unsigned>(tiles)),
                       dim3(kScanThreads), 0, ctx->stream, load, store, n, op,
                       static_cast<const T*>(partials.get()), inline_totals ? 1 : 0);
    GPART_LAUNCH_CHECK();
}

}  // namespace gpart_hip
