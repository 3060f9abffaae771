// hubs.hpp -- hub rows split into fixed chunks of entries.
//
// A kernel that gives one workgroup to each vertex of the high degree class is
// bounded by its largest row: on power-law graphs a single hub of 10^5 entries
// keeps one workgroup busy long after the rest of the grid has drained. Rows of
// degree above a threshold are cut into chunks of kHubChunk entries instead;
// chunk c belongs to hub chunk_hub[c] (an index into the high list) and covers
// entries [start[i] .. ) of it, (c - start[i]) * kHubChunk onwards. The chunk
// total lives on the device (start[high_count]); kernels loop grid-stride over
// it, so building the map needs no synchronisation.
#pragma once

#include "internal.hpp"
#include "scan.hpp"

#include <algorithm>

namespace gpart_hip {
namespace {

constexpr std::int64_t kHubChunk = 4096;
constexpr int kHubLaunchCap = 2048;  // workgroups of a grid-stride chunk launch

template <typename O>
struct HubChunkLoad {
    const std::int32_t* high;
    const O* offsets;
    std::int64_t count;
    std::int64_t min_degree;
    __device__ __forceinline__ std::int64_t operator()(std::int64_t i) const {
        if (i >= count) return 0;
        const std::int32_t v = high[i];
        const std::int64_t d = static_cast<std::int64_t>(offsets[v + 1] - offsets[v]);
        return d > min_degree ? (d + kHubChunk - 1) / kHubChunk : 0;
    }
};

__global__ void hub_chunk_map_kernel(std::int64_t count, const std::int64_t* start, std::int32_t* chunk_hub) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    for (std::int64_t c = start[i]; c < start[i + 1]; ++c) chunk_hub[c] = static_cast<std::int32_t>(i);
}

// Chunk map of the high-list vertices with degree > min_degree.
struct HubChunks {
    Scratch<std::int64_t> start;   // high_count + 1: exclusive prefix of chunk counts
    Scratch<std::int32_t> hub;     // chunk -> index into the high list
    std::int64_t bound = 0;        // host upper bound on the chunk count
    std::int64_t count = 0;        // high_count
    const std::int64_t* total() const { return start.get() + count; }
    unsigned grid() const {
        return static_cast<unsigned>(std::max<std::int64_t>(1, std::min<std::int64_t>(bound, kHubLaunchCap)));
    }
};

// Work items of the listed vertices: load(i) items for entry i (load(count) = 0),
// at most `bound` in total.
template <typename Load>
void build_work_items(gpart_ctx* ctx, Load load, std::int64_t count, std::int64_t bound, HubChunks& out) {
    out.count = count;
    out.bound = bound;
    out.start.reset(ctx, count + 1);
    out.hub.reset(ctx, bound);
    device_scan<std::int64_t, SumOp<std::int64_t>>(ctx, load, ExclusiveStore<std::int64_t>{out.start.get()},
                                                   count + 1, SumOp<std::int64_t>{});
    if (count > 0)
        hipLaunchKernelGGL(hub_chunk_map_kernel, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0,
                           ctx->stream, count, out.start.get(), out.hub.get());
    GPART_LAUNCH_CHECK();
}

// high_edges bounds the entries of the listed vertices (their degree sum).
template <typename O>
void build_hub_chunks(gpart_ctx* ctx, const O* offsets, const std::int32_t* high, std::int64_t high_count,
                      std::int64_t high_edges, std::int64_t min_degree, HubChunks& out) {
    // sum of ceil(d / C) over d > min_degree: at most edges / C plus one per hub
    const std::int64_t bound =
        high_edges / kHubChunk + std::min<std::int64_t>(high_count, high_edges / (min_degree + 1)) + 1;
    build_work_items(ctx, HubChunkLoad<O>{high, offsets, high_count, min_degree}, high_count, bound, out);
}

// Entry range of chunk c: hub index *i, vertex *v, entries [*begin, *end).
template <typename O>
__device__ __forceinline__ void hub_chunk_range(const O* offsets, const std::int32_t* high, const std::int64_t* start,
                                                const std::int32_t* chunk_hub, std::int64_t c, std::int32_t* i,
                                                std::int32_t* v, std::int64_t* begin, std::int64_t* end) {
    *i = chunk_hub[c];
    *v = high[*i];
    const std::int64_t row = static_cast<std::int64_t>(offsets[*v]);
    const std::int64_t row_end = static_cast<std::int64_t>(offsets[*v + 1]);
    *begin = row + (c - start[*i]) * kHubChunk;
    *end = *begin + kHubChunk < row_end ? *begin + kHubChunk : row_end;
}

}  // namespace
}  // namespace gpart_hip
