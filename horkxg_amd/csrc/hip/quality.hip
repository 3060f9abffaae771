// quality.hip -- partition verification and quality counters on the device.
//
// The reference checks a labelling on the host: tools/verify_partition.py:22-65
// (label range, loads, directed cut, non-loop entries) and
// graph_untils/analyze_partition.cpp:219-254 (per-part vertices and internal /
// outgoing / incoming entries with their weights). Here both are two passes over
// HBM: a vertex pass for label range and per-part vertex counts, and an edge pass
// over edge tiles (tiles.hpp) for the entry counters. Per-part counters live in
// LDS for the workgroup's lifetime and reach global memory once per workgroup;
// the totals are integers, so the result does not depend on the order of the
// additions.
#include "internal.hpp"
#include "reduce.hpp"
#include "tiles.hpp"

#include <algorithm>
#include <limits>

namespace gpart_hip {
namespace {

constexpr int kBlock = 256;
constexpr int kLdsParts = 1024;     // per-part counters in LDS up to this many parts
constexpr int kQualityBlocks = 2048;
constexpr int kHeader = 8;          // u64 words before the per-part counters

inline unsigned grid_for(std::int64_t threads) { return static_cast<unsigned>((threads + kBlock - 1) / kBlock); }

// header: [0] first vertex with a label outside [0, parts) (~0 when none),
// [1] first entry with a neighbour outside [0, n) (~0 when none), [2] non-loop
// entries, [3] cross-part entries, [4] their weight, [5] smallest label + 2^31,
// [6] largest label + 2^31 (unsigned min / max of the biased value), [7] unused.
// counters: vertices, internal, outgoing, incoming, weighted internal, weighted
// outgoing, weighted incoming; parts entries each.
enum Counter { kVertices = 0, kInternal, kOutgoing, kIncoming, kWInternal, kWOutgoing, kWIncoming, kCounters };

__global__ void quality_init_kernel(unsigned long long* words, std::int64_t count) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    words[i] = (i == 0 || i == 1 || i == 5) ? ~0ULL : 0ULL;
}

__device__ __forceinline__ unsigned long long biased(std::int32_t label) {
    return static_cast<unsigned long long>(static_cast<std::int64_t>(label) + (std::int64_t{1} << 31));
}

template <int Threads>
__device__ unsigned long long block_min_u64(unsigned long long v, unsigned long long* shared) {
    for (int m = kWave / 2; m > 0; m /= 2) {
        const unsigned long long o = shfl_xor_u64(v, m, kWave);
        v = o < v ? o : v;
    }
    if (lane_id() == 0) shared[threadIdx.x / kWave] = v;
    __syncthreads();
    unsigned long long r = ~0ULL;
    if (threadIdx.x == 0)
        for (int w = 0; w < Threads / kWave; ++w) r = shared[w] < r ? shared[w] : r;
    __syncthreads();
    return r;  // valid in thread 0
}

template <int Threads>
__device__ unsigned long long block_max_u64(unsigned long long v, unsigned long long* shared) {
    for (int m = kWave / 2; m > 0; m /= 2) {
        const unsigned long long o = shfl_xor_u64(v, m, kWave);
        v = o > v ? o : v;
    }
    if (lane_id() == 0) shared[threadIdx.x / kWave] = v;
    __syncthreads();
    unsigned long long r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < Threads / kWave; ++w) r = shared[w] > r ? shared[w] : r;
    __syncthreads();
    return r;  // valid in thread 0
}

template <int Threads>
__device__ unsigned long long block_sum_u64(unsigned long long v, unsigned long long* shared) {
    v = group_sum_u64<kWave>(v);
    if (lane_id() == 0) shared[threadIdx.x / kWave] = v;
    __syncthreads();
    unsigned long long r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < Threads / kWave; ++w) r += shared[w];
    __syncthreads();
    return r;  // valid in thread 0
}

// Label range and per-part vertex counts. lds: parts words when parts <= kLdsParts.
template <bool Lds>
__global__ void __launch_bounds__(kBlock) quality_vertices_kernel(std::int64_t n, const std::int32_t* labels,
                                                                  int parts, unsigned long long* header,
                                                                  unsigned long long* vertices) {
    extern __shared__ unsigned long long lds[];
    __shared__ unsigned long long shared[kBlock / kWave];
    if (Lds) {
        for (int i = threadIdx.x; i < parts; i += blockDim.x) lds[i] = 0;
        __syncthreads();
    }
    unsigned long long bad = ~0ULL, lo = ~0ULL, hi = 0;
    for (std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<std::int64_t>(gridDim.x) * blockDim.x) {
        const std::int32_t l = labels[v];
        const unsigned long long b = biased(l);
        lo = b < lo ? b : lo;
        hi = b > hi ? b : hi;
        if (l < 0 || l >= parts) {
            const auto vv = static_cast<unsigned long long>(v);
            bad = vv < bad ? vv : bad;
        } else if (Lds) {
            atomicAdd(&lds[l], 1ULL);
        } else {
            atomicAdd(&vertices[l], 1ULL);
        }
    }
    bad = block_min_u64<kBlock>(bad, shared);
    lo = block_min_u64<kBlock>(lo, shared);
    hi = block_max_u64<kBlock>(hi, shared);
    if (threadIdx.x == 0) {
        if (bad != ~0ULL) atomicMin(&header[0], bad);
        if (lo != ~0ULL) atomicMin(&header[5], lo);
        if (hi != 0) atomicMax(&header[6], hi);
    }
    if (Lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < parts; i += blockDim.x)
            if (lds[i]) atomicAdd(&vertices[i], lds[i]);
    }
}

// Entry counters over edge tiles. counters: the six entry counters (internal ..
// weighted incoming), parts words each. lds: the same, when parts <= kLdsParts.
template <typename O, typename W, bool Lds>
__global__ void __launch_bounds__(kTileLanes) quality_edges_kernel(
    std::int64_t n, std::int64_t m, std::int64_t tiles, const O* offsets, const std::int32_t* tile_row,
    const std::int32_t* neighbors, const W* weights, const std::int32_t* labels, int parts,
    unsigned long long* header, unsigned long long* counters) {
    extern __shared__ unsigned long long lds[];
    __shared__ unsigned long long shared[kTileLanes / kWave];
    unsigned long long* acc = Lds ? lds : counters;
    const std::int64_t stride = parts;  // words per counter
    if (Lds) {
        for (int i = threadIdx.x; i < 6 * parts; i += blockDim.x) lds[i] = 0;
        __syncthreads();
    }
    unsigned long long nonloop = 0, cut = 0, cut_w = 0, bad = ~0ULL;
    for (std::int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        std::int64_t e, stop;
        std::int32_t r;
        if (!tile_run(offsets, tile_row, tiles, n, m, t, &e, &stop, &r)) continue;
        std::int32_t nb[kTileRun];
        W w[kTileRun];
#pragma unroll
        for (int j = 0; j < kTileRun; ++j) {
            const bool in = e + j < stop;
            nb[j] = in ? neighbors[e + j] : 0;
            w[j] = in ? weights[e + j] : W{0};
        }
        std::int32_t q[kTileRun];
#pragma unroll
        for (int j = 0; j < kTileRun; ++j)
            q[j] = (e + j < stop && nb[j] >= 0 && nb[j] < n) ? labels[nb[j]] : -1;
        std::int64_t next = static_cast<std::int64_t>(offsets[r + 1]);
        std::int32_t p = labels[r];
        // the source part's own counters stay in registers while the row does not change
        unsigned long long in_c = 0, in_w = 0, out_c = 0, out_w = 0;
        auto flush = [&]() {
            if (p >= 0 && p < parts && (in_c | out_c)) {
                atomicAdd(&acc[0 * stride + p], in_c);
                atomicAdd(&acc[1 * stride + p], out_c);
                atomicAdd(&acc[3 * stride + p], in_w);
                atomicAdd(&acc[4 * stride + p], out_w);
            }
            in_c = in_w = out_c = out_w = 0;
        };
#pragma unroll
        for (int j = 0; j < kTileRun; ++j) {
            if (e + j < stop) {
                if (e + j >= next) {
                    flush();
                    do next = static_cast<std::int64_t>(offsets[++r + 1]);
                    while (e + j >= next);
                    p = labels[r];
                }
                if (nb[j] < 0 || nb[j] >= n) {
                    const auto ee = static_cast<unsigned long long>(e + j);
                    bad = ee < bad ? ee : bad;
                    continue;
                }
                nonloop += nb[j] != r ? 1ULL : 0ULL;
                const bool valid = p >= 0 && p < parts && q[j] >= 0 && q[j] < parts;
                if (!valid) continue;
                const auto wj = static_cast<unsigned long long>(w[j]);
                if (q[j] == p) {
                    in_c += 1;
                    in_w += wj;
                } else {
                    out_c += 1;
                    out_w += wj;
                    cut += 1;
                    cut_w += wj;
                    atomicAdd(&acc[2 * stride + q[j]], 1ULL);
                    atomicAdd(&acc[5 * stride + q[j]], wj);
                }
            }
        }
        flush();
    }
    nonloop = block_sum_u64<kTileLanes>(nonloop, shared);
    cut = block_sum_u64<kTileLanes>(cut, shared);
    cut_w = block_sum_u64<kTileLanes>(cut_w, shared);
    bad = block_min_u64<kTileLanes>(bad, shared);
    if (threadIdx.x == 0) {
        if (nonloop) atomicAdd(&header[2], nonloop);
        if (cut) atomicAdd(&header[3], cut);
        if (cut_w) atomicAdd(&header[4], cut_w);
        if (bad != ~0ULL) atomicMin(&header[1], bad);
    }
    if (Lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < 6 * parts; i += blockDim.x)
            if (lds[i]) atomicAdd(&counters[i], lds[i]);
    }
}

template <typename O, typename W>
void part_counters(gpart_ctx* ctx, const gpart_csr_view* g, const std::int32_t* labels, int parts,
                   gpart_part_counters_t* out, std::uint64_t* per_part) {
    const std::int64_t words = kHeader + static_cast<std::int64_t>(kCounters) * parts;
    Scratch<unsigned long long> buf(ctx, words);
    unsigned long long* header = buf.get();
    unsigned long long* counts = header + kHeader;
    hipLaunchKernelGGL(quality_init_kernel, dim3(ceil_div_int(words, kBlock)), dim3(kBlock), 0, ctx->stream, header,
                       words);
    const bool lds = parts <= kLdsParts;
    if (g->n > 0) {
        const unsigned blocks = static_cast<unsigned>(std::min<std::int64_t>(grid_for(g->n), kQualityBlocks));
        const std::size_t bytes = lds ? static_cast<std::size_t>(parts) * 8 : 0;
        if (lds)
            hipLaunchKernelGGL(quality_vertices_kernel<true>, dim3(blocks), dim3(kBlock), bytes, ctx->stream, g->n,
                               labels, parts, header, counts);
        else
            hipLaunchKernelGGL(quality_vertices_kernel<false>, dim3(blocks), dim3(kBlock), 0, ctx->stream, g->n,
                               labels, parts, header, counts);
    }
    if (parts > 0 && g->n > 0 && g->m > 0) {
        const auto* off = static_cast<const O*>(g->offsets);
        const auto* w = static_cast<const W*>(g->edge_weights);
        const std::int64_t tiles = tile_count(g->m);
        Scratch<std::int32_t> tile_row(ctx, tiles);
        hipLaunchKernelGGL(tile_rows_kernel<O>, dim3(grid_for(g->n)), dim3(kBlock), 0, ctx->stream, g->n, off,
                           tile_row.get());
        const unsigned blocks = static_cast<unsigned>(std::min<std::int64_t>(tiles, kQualityBlocks));
        unsigned long long* entry_counts = counts + parts;  // after the vertex counts
        if (lds)
            hipLaunchKernelGGL((quality_edges_kernel<O, W, true>), dim3(blocks), dim3(kTileLanes),
                               static_cast<std::size_t>(6) * parts * 8, ctx->stream, g->n, g->m, tiles, off,
                               tile_row.get(), g->neighbors, w, labels, parts, header, entry_counts);
        else
            hipLaunchKernelGGL((quality_edges_kernel<O, W, false>), dim3(blocks), dim3(kTileLanes), 0, ctx->stream,
                               g->n, g->m, tiles, off, tile_row.get(), g->neighbors, w, labels, parts, header,
                               entry_counts);
    }
    GPART_LAUNCH_CHECK();
    std::vector<unsigned long long> host(static_cast<std::size_t>(words));
    read_back(ctx, host.data(), header, static_cast<std::size_t>(words) * 8);
    out->first_invalid_label = host[0] == ~0ULL ? -1 : static_cast<std::int64_t>(host[0]);
    out->first_invalid_neighbor = host[1] == ~0ULL ? -1 : static_cast<std::int64_t>(host[1]);
    out->nonloop_entries = host[2];
    out->directed_cut = host[3];
    out->weighted_directed_cut = host[4];
    if (g->n > 0) {
        out->min_label = static_cast<std::int32_t>(static_cast<std::int64_t>(host[5]) - (std::int64_t{1} << 31));
        out->max_label = static_cast<std::int32_t>(static_cast<std::int64_t>(host[6]) - (std::int64_t{1} << 31));
    } else {
        out->min_label = 0;
        out->max_label = -1;
    }
    if (per_part)
        for (std::int64_t i = 0; i < static_cast<std::int64_t>(kCounters) * parts; ++i)
            per_part[i] = host[static_cast<std::size_t>(kHeader + i)];
}

}  // namespace
}  // namespace gpart_hip

using namespace gpart_hip;

extern "C" {

int gpart_part_counters(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* labels, int parts,
                        gpart_part_counters_t* out, uint64_t* per_part) {
    return guarded([&] {
        if (!g || !out) throw Error(GPART_ERR_INVALID_ARGUMENT, "null graph view or output");
        if (parts < 0) throw Error(GPART_ERR_INVALID_ARGUMENT, "part count must be non-negative");
        if (g->n > 0 && !labels) throw Error(GPART_ERR_INVALID_ARGUMENT, "null labels");
        if (g->width == GPART_WIDTH_STANDARD)
            part_counters<std::int32_t, std::uint32_t>(ctx, g, labels, parts, out, per_part);
        else
            part_counters<std::int64_t, std::uint64_t>(ctx, g, labels, parts, out, per_part);
    });
}

}  // extern "C"
