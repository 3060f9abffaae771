// tiles.hpp -- edge-balanced traversal of CSR rows.
//
// The edge range [0, m) of a CSR is cut into tiles of kTileEdges entries and
// tile_row[t] names the row holding entry t * kTileEdges. A workgroup takes one
// tile; lane i takes the run of kTileRun consecutive entries starting at
// t * kTileEdges + i * kTileRun, finds the run's first row by a binary search
// between tile_row[t] and tile_row[t + 1] (a window of a few hundred bytes of
// offsets, L1-resident) and walks forward across row ends. Every entry costs the
// same whatever the degree of its row, so a hub of a million entries spreads over
// hundreds of workgroups instead of serialising one group of lanes.
#pragma once

#include "common.hpp"

namespace gpart_hip {
namespace {

constexpr int kTileLanes = 256;
constexpr int kTileRun = 16;
constexpr std::int64_t kTileEdges = static_cast<std::int64_t>(kTileLanes) * kTileRun;

__host__ __device__ inline std::int64_t tile_count(std::int64_t m) { return (m + kTileEdges - 1) / kTileEdges; }

// tile_row[t] = the row r with offsets[r] <= t * kTileEdges < offsets[r + 1]; every
// tile start lies in exactly one non-empty row, so each entry is written once.
template <typename O>
__global__ void tile_rows_kernel(std::int64_t n, const O* offsets, std::int32_t* tile_row) {
    const std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const std::int64_t lo = static_cast<std::int64_t>(offsets[v]);
    const std::int64_t hi = static_cast<std::int64_t>(offsets[v + 1]);
    for (std::int64_t t = (lo + kTileEdges - 1) / kTileEdges; t * kTileEdges < hi; ++t)
        tile_row[t] = static_cast<std::int32_t>(v);
}

// Largest r in [lo, hi] with offsets[r] <= e, given offsets[lo] <= e.
template <typename O>
__device__ __forceinline__ std::int32_t row_of_entry(const O* offsets, std::int32_t lo, std::int32_t hi,
                                                     std::int64_t e) {
    while (lo < hi) {
        const std::int32_t mid = lo + (hi - lo + 1) / 2;
        if (static_cast<std::int64_t>(offsets[mid]) <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The calling lane's run [*begin, *end) in tile t and the row of *begin; false
// when the run is empty (the last tile is partial).
template <typename O>
__device__ __forceinline__ bool tile_run(const O* offsets, const std::int32_t* tile_row, std::int64_t tiles,
                                         std::int64_t n, std::int64_t m, std::int64_t t, std::int64_t* begin,
                                         std::int64_t* end, std::int32_t* row) {
    const std::int64_t e = t * kTileEdges + static_cast<std::int64_t>(threadIdx.x) * kTileRun;
    const std::int64_t stop = e + kTileRun < m ? e + kTileRun : m;
    if (e >= stop) return false;
    const std::int32_t first = tile_row[t];
    const std::int32_t last = t + 1 < tiles ? tile_row[t + 1] : static_cast<std::int32_t>(n - 1);
    *row = row_of_entry(offsets, first, last, e);
    *begin = e;
    *end = stop;
    return true;
}

}  // namespace
}  // namespace gpart_hip
