// refine.hip -- incremental bucketed refinement for gfx950
// (src/refine.cu and src/refine_incremental.cuh of the reference).
//
// Per-vertex edge loops use the level's degree classes: 8 lanes for degree <= 8,
// a wave for <= 256 and a workgroup above. Row scans of the n x k connectivity
// cache use 1, 2 or 4 lanes per row, each owning 8 consecutive parts, so a wave
// has 16-64 rows in flight per pass.
//
// The reference ranks vertices inside a rebalance bucket by the return value of
// atomicAdd (src/refine_incremental.cuh:185,216): any order is legal. Here the
// rank is the weight of the bucket's lower-id members, so refinement is
// reproducible run to run.
//
// A round (LP or rebalance) keeps its counts on the device: kernels after a
// compaction read the length from a RoundState and loop grid-stride, and the
// host reads one RoundState back at the end of the round.
#include "internal.hpp"
#include "reduce.hpp"
#include "scan.hpp"
#include "tiles.hpp"
#include "hubs.hpp"
#include "refine_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <list>
#include <unordered_set>

namespace gpart_hip {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(std::int64_t threads) { return static_cast<unsigned>((threads + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------- round state

// Device record of one round, read back once at its end.
struct RoundState {
    long long light;       // movers with degree <= kHeavyDegree
    long long heavy;       // movers above
    long long delta;       // exact cut decrease
    long long candidates;  // rebalance: vertices of overloaded parts
    unsigned long long max_weight;
    long long edges;       // CSR entries of all movers (light + heavy)
    unsigned long long cut_directed;  // rebuild rounds: directed cut weight after the round
    unsigned long long pad[1];
    unsigned long long pw[32];
};

constexpr int kHeavyDegree = 64;
// Levels whose largest degree is at most this update their heavy movers' cache rows
// a wave per mover (one launch); above it the movers' entries are flattened into
// tiles first (a scan and a tile map, three more launches) so that hubs spread
// over many workgroups.
constexpr int kWaveHeavyMaxDegree = 4096;
constexpr int kRoundBlocks = 1024;  // cap of grid-stride launches whose count lives on the device
constexpr int kCacheBlocks = 4096;  // cap of the cache build launches (grid-stride)


// ---------------------------------------------------------------- cut / weights

// Directed cut weight over edge tiles (grid-stride over tiles); one partial per
// workgroup in directed[blockIdx.x]. A lane loads its run's neighbours and
// weights up front, then gathers their labels, so 16 loads are in flight per lane.
template <typename O, typename W>
__global__ void __launch_bounds__(kTileLanes) cut_kernel(std::int64_t n, std::int64_t m, std::int64_t tiles,
                                                         const O* offsets, const std::int32_t* tile_row,
                                                         const std::int32_t* neighbors, const W* weights,
                                                         const std::int32_t* labels, unsigned long long* directed) {
    __shared__ unsigned long long waves[kTileLanes / kWave];
    unsigned long long local = 0;
    for (std::int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        std::int64_t e, stop;
        std::int32_t r;
        if (!tile_run(offsets, tile_row, tiles, n, m, t, &e, &stop, &r)) continue;
        std::int32_t other[kTileRun];
        W w[kTileRun];
#pragma unroll
        for (int j = 0; j < kTileRun; ++j) {
            const bool in = e + j < stop;
            other[j] = in ? neighbors[e + j] : 0;
            w[j] = in ? weights[e + j] : W{0};
        }
#pragma unroll
        for (int j = 0; j < kTileRun; ++j) other[j] = labels[other[j]];
        std::int64_t next = static_cast<std::int64_t>(offsets[r + 1]);
        std::int32_t p = labels[r];
#pragma unroll
        for (int j = 0; j < kTileRun; ++j) {
            if (e + j < stop) {
                if (e + j >= next) {
                    do next = static_cast<std::int64_t>(offsets[++r + 1]);
                    while (e + j >= next);
                    p = labels[r];
                }
                if (other[j] != p) local += static_cast<unsigned long long>(w[j]);
            }
        }
    }
    local = group_sum_u64<kWave>(local);
    if (lane_id() == 0) waves[threadIdx.x / kWave] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long total = 0;
        for (int w = 0; w < kTileLanes / kWave; ++w) total += waves[w];
        directed[blockIdx.x] = total;
    }
}

template <typename W>
__global__ void part_weights_kernel(std::int64_t n, const std::int32_t* labels, const W* vertex_weights, int k,
                                    unsigned long long* out) {
    __shared__ unsigned long long local[32];
    if (threadIdx.x < 32) local[threadIdx.x] = 0;
    __syncthreads();
    for (std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<std::int64_t>(gridDim.x) * blockDim.x)
        atomicAdd(&local[labels[v]], static_cast<unsigned long long>(vertex_weights[v]));
    __syncthreads();
    if (threadIdx.x < k && local[threadIdx.x]) atomicAdd(out + threadIdx.x, local[threadIdx.x]);
}

__global__ void project_kernel(std::int64_t n, const std::int32_t* map, const std::int32_t* coarse,
                               std::int32_t* fine) {
    const std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v < n) fine[v] = coarse[map[v]];
}

// ---------------------------------------------------------------- cache

// Adds the workgroup's sum of v to *out with one atomic. Called by every thread
// of a kBlock-thread workgroup; waves holds kBlock / kWave entries.
__device__ __forceinline__ void block_add_u64(unsigned long long v, unsigned long long* waves,
                                              unsigned long long* out) {
    v = group_sum_u64<kWave>(v);
    if (lane_id() == 0) waves[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long total = 0;
        for (int w = 0; w < kBlock / kWave; ++w) total += waves[w];
        if (total) atomicAdd(out, total);
    }
}

template <typename O, typename W, int G, typename C>
__global__ void __launch_bounds__(kBlock) build_cache_kernel(std::int64_t count, const std::int32_t* list,
                                                             const O* offsets, const std::int32_t* neighbors,
                                                             const W* weights, const std::int32_t* labels, int k,
                                                             C* cache, const long long* gate, long long above,
                                                             unsigned long long* cut) {
    // gate: a rebuild inside a round runs only when the round's mover entries exceed
    // `above`; grid-stride over groups so that a gated-off launch is a small grid.
    // cut (optional) receives the directed cut weight of `labels`.
    if (gate && *gate <= above) return;
    constexpr int kGroups = kBlock / G;
    __shared__ unsigned long long rows[kGroups][32];
    __shared__ unsigned long long wave_cut[kBlock / kWave];
    const int local_group = threadIdx.x / G;
    const int sub = threadIdx.x & (G - 1);
    unsigned long long cut_local = 0;
    for (std::int64_t first = static_cast<std::int64_t>(blockIdx.x) * kGroups; first < count;
         first += static_cast<std::int64_t>(gridDim.x) * kGroups) {
        const std::int64_t group = first + local_group;
        for (int p = sub; p < 32; p += G) rows[local_group][p] = 0;
        __syncthreads();
        std::int32_t v = 0;
        bool own = group < count;
        if (own) {
            v = list ? list[group] : static_cast<std::int32_t>(group);
            // rows above kHubChunk entries are hubs, built by build_cache_hub_kernel
            own = G < kBlock || offsets[v + 1] - offsets[v] <= kHubChunk;
            if (own) {
                const std::int32_t lv = labels[v];
                for (std::int64_t e = static_cast<std::int64_t>(offsets[v]) + sub; e < offsets[v + 1]; e += G) {
                    const std::int32_t u = neighbors[e];
                    if (u == v) continue;
                    const std::int32_t lu = labels[u];
                    const auto w = static_cast<unsigned long long>(weights[e]);
                    atomicAdd(&rows[local_group][lu], w);
                    if (lu != lv) cut_local += w;
                }
            }
        }
        __syncthreads();
        if (own)
            for (int p = sub; p < k; p += G)
                cache[static_cast<std::int64_t>(v) * k + p] = static_cast<C>(rows[local_group][p]);
        __syncthreads();
    }
    if (cut) block_add_u64(cut_local, wave_cut, cut);
}

template <typename C>
__global__ void hub_rows_zero_kernel(std::int64_t count, const std::int64_t* chunk_start, const std::int32_t* list,
                                     int k, C* cache, const long long* gate, long long above) {
    if (gate && *gate <= above) return;
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count || chunk_start[i] == chunk_start[i + 1]) return;
    C* row = cache + static_cast<std::int64_t>(list[i]) * k;
    for (int p = 0; p < k; ++p) row[p] = 0;
}

// Hub rows (degree > kHubChunk), one chunk per workgroup pass: the chunk's part
// sums gather in LDS and reach the zeroed cache row as k atomic adds.
template <typename O, typename W, typename C>
__global__ void __launch_bounds__(kBlock) build_cache_hub_kernel(const std::int64_t* chunk_total,
                                                                 const std::int64_t* chunk_start,
                                                                 const std::int32_t* chunk_hub,
                                                                 const std::int32_t* list, const O* offsets,
                                                                 const std::int32_t* neighbors, const W* weights,
                                                                 const std::int32_t* labels, int k, C* cache,
                                                                 const long long* gate, long long above,
                                                                 unsigned long long* cut) {
    __shared__ unsigned long long row[32];
    __shared__ unsigned long long wave_cut[kBlock / kWave];
    if (gate && *gate <= above) return;
    unsigned long long cut_local = 0;
    const std::int64_t chunks = *chunk_total;
    for (std::int64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        if (threadIdx.x < 32) row[threadIdx.x] = 0;
        __syncthreads();
        std::int32_t i, v;
        std::int64_t begin, end;
        hub_chunk_range(offsets, list, chunk_start, chunk_hub, c, &i, &v, &begin, &end);
        const std::int32_t lv = labels[v];
        for (std::int64_t e = begin + threadIdx.x; e < end; e += kBlock) {
            const std::int32_t u = neighbors[e];
            if (u == v) continue;
            const std::int32_t lu = labels[u];
            const auto w = static_cast<unsigned long long>(weights[e]);
            atomicAdd(&row[lu], w);
            if (lu != lv) cut_local += w;
        }
        __syncthreads();
        if (threadIdx.x < k && row[threadIdx.x])
            atomicAdd(cache + static_cast<std::int64_t>(v) * k + threadIdx.x, static_cast<C>(row[threadIdx.x]));
        __syncthreads();
    }
    if (cut) block_add_u64(cut_local, wave_cut, cut);
}

// incremental_lp_proposals (:27-46): argmax over p != source of cache > 0,
// lowest index on ties; accepted if best >= current or within the filter.
template <int G, typename C>
__global__ void __launch_bounds__(kBlock) lp_proposals_kernel(std::int64_t n, const C* cache,
                                                              const std::int32_t* labels, int k, double filter,
                                                              std::int32_t* locks, std::int32_t* targets,
                                                              long long* gains, RoundState* st) {
    // the round's state starts at zero (the filter appends the movers to it)
    if (blockIdx.x == 0 && threadIdx.x < sizeof(RoundState) / 8)
        reinterpret_cast<unsigned long long*>(st)[threadIdx.x] = 0;
    const std::int64_t v = (static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
    const int j = threadIdx.x & (G - 1);
    if (v >= n) return;  // uniform per group
    const std::int32_t source = labels[v];
    const C* row = cache + v * k;
    unsigned long long best = 0, current = 0;
    int best_p = 0x7fffffff;
    unsigned long long span[kRowSpan];
    load_span(row, j, k, span);
#pragma unroll
    for (int e = 0; e < kRowSpan; ++e) {
        const int p = j * kRowSpan + e;
        const unsigned long long value = span[e];
        if (p == source) current = value;
        else if (value > best) {  // strict: the lowest index keeps a tie
            best = value;
            best_p = p;
        }
    }
#pragma unroll
    for (int m = G / 2; m > 0; m /= 2) {
        const auto ob = shfl_xor_u64(best, m, G);
        const int op = __shfl_xor(best_p, m, G);
        const auto oc = shfl_xor_u64(current, m, G);
        if (ob > best || (ob == best && op < best_p)) {
            best = ob;
            best_p = op;
        }
        current = oc > current ? oc : current;  // only the source's lane holds it
    }
    if (j != 0) return;
    if (locks[v] != 0) {
        locks[v] = 0;
        targets[v] = -1;
        gains[v] = static_cast<long long>(0x8000000000000000ULL);
        return;
    }
    std::int32_t target = -1;
    long long gain = static_cast<long long>(0x8000000000000000ULL);
    if (best > 0 &&
        (best >= current ||
         current - best < static_cast<unsigned long long>(filter * static_cast<double>(current)))) {
        target = best_p;
        gain = best >= current ? static_cast<long long>(best - current) : -static_cast<long long>(current - best);
    }
    targets[v] = target;
    gains[v] = gain;
}

// incremental_priority_filter (:48-67) over one class list with G lanes per row,
// by the workgroups numbered `block` of the list's range. Called by every thread
// of the workgroup.
template <typename O, typename W, int G>
__device__ __forceinline__ void priority_filter_rows(std::int64_t count, const std::int32_t* list, long long block,
                                                     const O* offsets, const std::int32_t* neighbors,
                                                     const W* weights, const std::int32_t* labels,
                                                     const std::int32_t* targets, const long long* gains,
                                                     std::int32_t* flags, long long* block_sum) {
    const std::int64_t group = (block * blockDim.x + threadIdx.x) / G;
    const int sub = threadIdx.x & (G - 1);
    const bool active = group < count;
    const std::int32_t v = active ? (list ? list[group] : static_cast<std::int32_t>(group)) : 0;
    const std::int32_t target = active ? targets[v] : -1;
    // rows above kHubChunk entries are hubs, done in chunks by priority_filter_hub_kernel
    if (G == kBlock && active && offsets[v + 1] - offsets[v] > kHubChunk) return;  // uniform per block
    long long correction = 0;
    if (active && target >= 0)
        correction = filter_correction(v, target, labels[v], gains[v], static_cast<std::int64_t>(offsets[v]) + sub,
                                       static_cast<std::int64_t>(offsets[v + 1]), G, neighbors, weights, labels,
                                       targets, gains);
    if constexpr (G <= kWave) {
        correction = group_sum_i64<G>(correction);
        if (active && sub == 0) flags[v] = target >= 0 && gains[v] + correction >= 0;
    } else {
        correction = group_sum_i64<kWave>(correction);
        if (lane_id() == 0) block_sum[threadIdx.x / kWave] = correction;
        __syncthreads();
        if (threadIdx.x == 0 && active) {
            long long total = 0;
            for (int w = 0; w < kBlock / kWave; ++w) total += block_sum[w];
            flags[v] = target >= 0 && gains[v] + total >= 0;
        }
    }
}

// The three degree classes in one launch: workgroups [0, low_blocks) take the low
// rows (8 lanes each), the next medium_blocks the medium rows (a wave each), the
// rest the high rows (a workgroup each).
template <typename O, typename W>
__global__ void __launch_bounds__(kBlock) priority_filter_kernel(
    std::int64_t low_count, const std::int32_t* low, unsigned low_blocks, std::int64_t medium_count,
    const std::int32_t* medium, unsigned medium_blocks, std::int64_t high_count, const std::int32_t* high,
    const O* offsets, const std::int32_t* neighbors, const W* weights, const std::int32_t* labels,
    const std::int32_t* targets, const long long* gains, std::int32_t* flags) {
    __shared__ long long block_sum[kBlock / kWave];
    const long long b = blockIdx.x;
    if (b < low_blocks)
        priority_filter_rows<O, W, 8>(low_count, low, b, offsets, neighbors, weights, labels, targets, gains, flags,
                                      block_sum);
    else if (b < low_blocks + medium_blocks)
        priority_filter_rows<O, W, 64>(medium_count, medium, b - low_blocks, offsets, neighbors, weights, labels,
                                       targets, gains, flags, block_sum);
    else
        priority_filter_rows<O, W, kBlock>(high_count, high, b - low_blocks - medium_blocks, offsets, neighbors,
                                           weights, labels, targets, gains, flags, block_sum);
}

// Hub rows (degree > kHubChunk): each chunk adds its share of the correction to
// sums[i]; priority_filter_hub_finish_kernel then sets the hub's flag.
template <typename O, typename W>
__global__ void __launch_bounds__(kBlock) priority_filter_hub_kernel(const std::int64_t* chunk_total,
                                                                     const std::int64_t* chunk_start,
                                                                     const std::int32_t* chunk_hub,
                                                                     const std::int32_t* list, const O* offsets,
                                                                     const std::int32_t* neighbors, const W* weights,
                                                                     const std::int32_t* labels,
                                                                     const std::int32_t* targets,
                                                                     const long long* gains, long long* sums) {
    __shared__ long long block_sum[kBlock / kWave];
    const std::int64_t chunks = *chunk_total;
    for (std::int64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        std::int32_t i, v;
        std::int64_t begin, end;
        hub_chunk_range(offsets, list, chunk_start, chunk_hub, c, &i, &v, &begin, &end);
        const std::int32_t target = targets[v];
        long long correction = 0;
        if (target >= 0)
            correction = filter_correction(v, target, labels[v], gains[v], begin + threadIdx.x, end, kBlock,
                                           neighbors, weights, labels, targets, gains);
        correction = group_sum_i64<kWave>(correction);
        if (lane_id() == 0) block_sum[threadIdx.x / kWave] = correction;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long total = 0;
            for (int w = 0; w < kBlock / kWave; ++w) total += block_sum[w];
            if (total) atomicAdd(reinterpret_cast<unsigned long long*>(sums + i), static_cast<unsigned long long>(total));
        }
        __syncthreads();
    }
}

__global__ void priority_filter_hub_finish_kernel(std::int64_t count, const std::int64_t* chunk_start,
                                                  const std::int32_t* list, const std::int32_t* targets,
                                                  const long long* gains, long long* sums, std::int32_t* flags) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count || chunk_start[i] == chunk_start[i + 1]) return;
    const std::int32_t v = list[i];
    flags[v] = targets[v] >= 0 && gains[v] + sums[i] >= 0;
    sums[i] = 0;
}

// incremental_update_cache_and_cut (:69-95) over a mover list whose length is on
// the device, G lanes per mover, by the workgroups [0, blocks) numbered `block`;
// the block's gain sum goes to *delta. Called by every thread of the workgroup.
template <typename O, typename W, int G, typename C>
__device__ __forceinline__ void update_cache_movers(long long count, const std::int32_t* movers, const O* offsets,
                                                    const std::int32_t* neighbors, const W* weights,
                                                    const std::int32_t* labels, const std::int32_t* targets,
                                                    const std::int32_t* flags, int k, C* cache, long long* delta,
                                                    long long block, long long blocks, long long* block_sum) {
    const int sub = threadIdx.x & (G - 1);
    const long long groups = blocks * (kBlock / G);
    long long gain = 0;
    for (long long group = (block * blockDim.x + threadIdx.x) / G; group < count; group += groups) {
        const std::int32_t v = movers[group];
        const std::int32_t source = labels[v];
        const std::int32_t target = targets[v];
        for (std::int64_t e = static_cast<std::int64_t>(offsets[v]) + sub; e < offsets[v + 1]; e += G) {
            const std::int32_t u = neighbors[e];
            if (u == v) continue;
            const auto w = static_cast<unsigned long long>(weights[e]);
            // modular arithmetic: entries fit C at every point of the round
            atomicAdd(cache + static_cast<std::int64_t>(u) * k + source, static_cast<C>(0ULL - w));
            atomicAdd(cache + static_cast<std::int64_t>(u) * k + target, static_cast<C>(w));
            const bool fu = flags[u] != 0;
            if (!fu || u > v) {
                const std::int32_t lu = labels[u];
                const std::int32_t nu = fu ? targets[u] : lu;
                gain += static_cast<long long>(w) * ((source != lu) - (target != nu));
            }
        }
    }
    gain = group_sum_i64<kWave>(gain);
    if (lane_id() == 0) block_sum[threadIdx.x / kWave] = gain;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long total = 0;
        for (int w = 0; w < kBlock / kWave; ++w) total += block_sum[w];
        if (total) atomicAdd(reinterpret_cast<unsigned long long*>(delta), static_cast<unsigned long long>(total));
    }
}

// The light movers (8 lanes each) on workgroups [0, light_blocks), and, when
// heavy is given, the heavy movers (a wave each) on the rest of the grid.
template <typename O, typename W, typename C>
__global__ void __launch_bounds__(kBlock) update_cache_kernel(const RoundState* st, const std::int32_t* light,
                                                              const std::int32_t* heavy, unsigned light_blocks,
                                                              const O* offsets, const std::int32_t* neighbors,
                                                              const W* weights, const std::int32_t* labels,
                                                              const std::int32_t* targets,
                                                              const std::int32_t* flags, int k,
                                                              C* cache, long long* delta, long long rebuild_above) {
    __shared__ long long block_sum[kBlock / kWave];
    // a rebuild round recomputes cache and cut from the CSR after the commit
    if (rebuild_above >= 0 && st->edges > rebuild_above) return;
    if (blockIdx.x < light_blocks)
        update_cache_movers<O, W, 8, C>(st->light, light, offsets, neighbors, weights, labels, targets, flags, k,
                                        cache, delta, blockIdx.x, light_blocks, block_sum);
    else
        update_cache_movers<O, W, kWave, C>(st->heavy, heavy, offsets, neighbors, weights, labels, targets, flags,
                                            k, cache, delta, blockIdx.x - light_blocks, gridDim.x - light_blocks,
                                            block_sum);
}

// Heavy movers (degree > kHeavyDegree) as one flattened entry range: heavy_start
// holds the exclusive prefix of their degrees (count + 1 entries, the last one
// the total), tiled as in tiles.hpp with the movers as rows.
template <typename O>
struct HeavyDegreeLoad {
    const RoundState* st;
    const std::int32_t* heavy;
    const O* offsets;
    __device__ __forceinline__ long long operator()(std::int64_t i) const {
        if (i >= st->heavy) return 0;
        const std::int32_t v = heavy[i];
        return static_cast<long long>(offsets[v + 1] - offsets[v]);
    }
};

__global__ void heavy_tiles_kernel(const RoundState* st, const long long* heavy_start, std::int32_t* tile_mover) {
    const long long count = st->heavy;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long lo = heavy_start[i], hi = heavy_start[i + 1];
        for (long long t = (lo + kTileEdges - 1) / kTileEdges; t * kTileEdges < hi; ++t)
            tile_mover[t] = static_cast<std::int32_t>(i);
    }
}

// incremental_update_cache_and_cut over the heavy movers' entries, kTileRun per lane.
template <typename O, typename W, typename C>
__global__ void __launch_bounds__(kTileLanes) update_cache_tiles_kernel(
    const RoundState* st, const long long* heavy_start, const std::int32_t* tile_mover, const std::int32_t* movers,
    const O* offsets, const std::int32_t* neighbors, const W* weights, const std::int32_t* labels,
    const std::int32_t* targets, const std::int32_t* flags, int k, C* cache, long long* delta,
    long long rebuild_above) {
    __shared__ long long block_sum[kTileLanes / kWave];
    if (rebuild_above >= 0 && st->edges > rebuild_above) return;  // a rebuild round
    const long long count = st->heavy;
    const long long total = heavy_start[count];
    const long long tiles = tile_count(total);
    long long gain = 0;
    for (long long t = blockIdx.x; t < tiles; t += gridDim.x) {
        std::int64_t x, stop;
        std::int32_t i;
        if (!tile_run(heavy_start, tile_mover, tiles, count, total, t, &x, &stop, &i)) continue;
        std::int32_t v = movers[i];
        std::int32_t source = labels[v], target = targets[v];
        long long next = heavy_start[i + 1];
        std::int64_t base = static_cast<std::int64_t>(offsets[v]) - heavy_start[i];
        for (; x < stop; ++x) {
            while (x >= next) {
                ++i;
                v = movers[i];
                source = labels[v];
                target = targets[v];
                base = static_cast<std::int64_t>(offsets[v]) - heavy_start[i];
                next = heavy_start[i + 1];
            }
            const std::int64_t e = base + x;
            const std::int32_t u = neighbors[e];
            if (u == v) continue;
            const auto w = static_cast<unsigned long long>(weights[e]);
            atomicAdd(cache + static_cast<std::int64_t>(u) * k + source, static_cast<C>(0ULL - w));
            atomicAdd(cache + static_cast<std::int64_t>(u) * k + target, static_cast<C>(w));
            const bool fu = flags[u] != 0;
            if (!fu || u > v) {
                const std::int32_t lu = labels[u];
                const std::int32_t nu = fu ? targets[u] : lu;
                gain += static_cast<long long>(w) * ((source != lu) - (target != nu));
            }
        }
    }
    gain = group_sum_i64<kWave>(gain);
    if (lane_id() == 0) block_sum[threadIdx.x / kWave] = gain;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long sum = 0;
        for (int w = 0; w < kTileLanes / kWave; ++w) sum += block_sum[w];
        if (sum) atomicAdd(reinterpret_cast<unsigned long long*>(delta), static_cast<unsigned long long>(sum));
    }
}

// incremental_commit (:97-109) over light then heavy movers; part weight changes
// are summed per workgroup before they reach the k global counters.
template <typename W>
__global__ void __launch_bounds__(kBlock) commit_kernel(const RoundState* st, const std::int32_t* light,
                                                        const std::int32_t* heavy, std::int32_t* labels,
                                                        const std::int32_t* targets, const W* vertex_weights,
                                                        unsigned long long* part_weights, std::int32_t* locks,
                                                        bool lp, int k) {
    __shared__ unsigned long long gained[32], lost[32];
    if (threadIdx.x < 32) gained[threadIdx.x] = lost[threadIdx.x] = 0;
    __syncthreads();
    const long long nl = st->light, total = nl + st->heavy;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const std::int32_t v = i < nl ? light[i] : heavy[i - nl];
        const std::int32_t source = labels[v];
        const std::int32_t target = targets[v];
        const auto w = static_cast<unsigned long long>(vertex_weights[v]);
        atomicAdd(&lost[source], w);
        atomicAdd(&gained[target], w);
        labels[v] = target;
        if (lp) locks[v] = 1;
    }
    __syncthreads();
    if (threadIdx.x < k) {
        const unsigned long long change = gained[threadIdx.x] - lost[threadIdx.x];
        if (change) atomicAdd(part_weights + threadIdx.x, change);
    }
}

// The round's status, also copied to the host-visible status words (word 0: seq,
// stored last, after a system-scope fence).
// One wave: lane p < k holds part p's weight; every lane writes one status word.
__global__ void __launch_bounds__(kWave) round_status_kernel(RoundState* st, const unsigned long long* part_weights,
                                                             int k, unsigned long long* status,
                                                             unsigned long long* counter) {
    constexpr int kHead = static_cast<int>(offsetof(RoundState, pw) / 8);  // words before pw
    const int lane = threadIdx.x;
    const unsigned long long w = lane < k ? part_weights[lane] : 0;
    unsigned long long maximum = w;
    for (int m = kWave / 2; m > 0; m /= 2) {
        const unsigned long long o = shfl_xor_u64(maximum, m, kWave);
        maximum = o > maximum ? o : maximum;
    }
    if (lane < 32) st->pw[lane] = w;
    if (lane == 0) st->max_weight = maximum;
    __syncthreads();
    // words of the head (with max_weight) and pw[0, k): a store per lane, in parallel
    const auto* words = reinterpret_cast<const unsigned long long*>(st);
    if (lane < kHead + k) status[1 + lane] = words[lane];
    __threadfence_system();
    __syncthreads();
    if (lane == 0) {
        const unsigned long long seq = *counter + 1;
        *counter = seq;
        __hip_atomic_store(status, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}
static_assert(sizeof(RoundState) % 8 == 0 && sizeof(RoundState) / 8 + 1 <= kStatusWords, "status words");
static_assert(offsetof(RoundState, pw) / 8 + 32 <= static_cast<std::size_t>(kWave), "one wave writes the status");

// Movers are split by degree so the cache update can give heavy rows a workgroup.
struct MoverPair {
    long long light, heavy, edges;
};
struct MoverPairOp {
    __device__ __forceinline__ MoverPair identity() const { return MoverPair{0, 0, 0}; }
    __device__ __forceinline__ MoverPair operator()(const MoverPair& a, const MoverPair& b) const {
        return MoverPair{a.light + b.light, a.heavy + b.heavy, a.edges + b.edges};
    }
};
template <typename O>
struct MoverLoad {
    const std::int32_t* flags;
    const O* offsets;
    __device__ __forceinline__ MoverPair operator()(std::int64_t v) const {
        if (!flags[v]) return MoverPair{0, 0, 0};
        const long long degree = static_cast<long long>(offsets[v + 1] - offsets[v]);
        const bool heavy = degree > kHeavyDegree;
        return MoverPair{heavy ? 0 : 1, heavy ? 1 : 0, degree};
    }
};
struct MoverStore {
    std::int32_t* light;
    std::int32_t* heavy;
    std::int64_t n;
    RoundState* st;
    __device__ __forceinline__ void operator()(std::int64_t v, const MoverPair& prefix, const MoverPair& item) const {
        if (item.light) light[prefix.light] = static_cast<std::int32_t>(v);
        if (item.heavy) heavy[prefix.heavy] = static_cast<std::int32_t>(v);
        if (v == n - 1) {
            st->light = prefix.light + item.light;
            st->heavy = prefix.heavy + item.heavy;
            st->edges = prefix.edges + item.edges;
        }
    }
};

// ---------------------------------------------------------------- rebalance
//
// One iteration of the reference's two-sided bucketed rebalance:
//   proposals   every vertex of an overloaded part picks a destination and lands
//               in a source bucket (part, gain class, section);
//   sources     a vertex leaves when the weight ranked before it in its part's
//               bucket order is below the part's overload;
//   destinations it is admitted when the weight ranked before it plus its own
//               fits the destination's slack.
// Buckets are visited in a fixed order, so within one part every bucket is
// either wholly taken, wholly refused, or the single boundary bucket that
// straddles the part's threshold. Only boundary members need a rank; they get
// the weight of the lower-id members of their bucket (a blocked multisplit),
// one legal outcome of the reference's atomicAdd ranking.

// Appends the taking threads' vertices to the light/heavy mover lists (one
// global atomic per list per workgroup pass) and flags them. Called by every
// thread of the workgroup.
template <typename O>
__device__ void append_movers(bool take, std::int32_t v, const O* offsets, std::int32_t* light,
                              std::int32_t* heavy, std::int32_t* flags, RoundState* st, unsigned* counts,
                              unsigned long long* base) {
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    const long long degree = take ? static_cast<long long>(offsets[v + 1] - offsets[v]) : 0;
    const bool is_heavy = degree > kHeavyDegree;
    const long long edges = group_sum_i64<kWave>(degree);
    if (lane == 0 && edges) atomicAdd(reinterpret_cast<unsigned long long*>(&st->edges),
                                      static_cast<unsigned long long>(edges));
    const bool is_light = take && !is_heavy;
    const unsigned long long bl = __ballot(is_light), bh = __ballot(is_heavy);
    if (lane == 0) {
        counts[wave] = static_cast<unsigned>(__popcll(bl));
        counts[kBlock / kWave + wave] = static_cast<unsigned>(__popcll(bh));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned tl = 0, th = 0;
        for (int q = 0; q < kBlock / kWave; ++q) {
            tl += counts[q];
            th += counts[kBlock / kWave + q];
        }
        base[0] = tl ? atomicAdd(reinterpret_cast<unsigned long long*>(&st->light), tl) : 0;
        base[1] = th ? atomicAdd(reinterpret_cast<unsigned long long*>(&st->heavy), th) : 0;
    }
    __syncthreads();
    if (take) {
        unsigned before = 0;
        const unsigned* row = is_heavy ? counts + kBlock / kWave : counts;
        for (int q = 0; q < wave; ++q) before += row[q];
        const unsigned long long mask = is_heavy ? bh : bl;
        before += static_cast<unsigned>(__popcll(mask & lanemask_lt()));
        (is_heavy ? heavy : light)[(is_heavy ? base[1] : base[0]) + before] = v;
        flags[v] = 1;
    }
    __syncthreads();
}

// Candidates: vertices of overloaded parts, in id order. Every flag is cleared
// on the way (movers are flagged later in the iteration).
struct OverloadedLoad {
    const std::int32_t* labels;
    const unsigned long long* part_weights;
    std::uint64_t capacity;
    // 32-bit counts: vertex ids, and so candidate counts, stay below 2^31
    __device__ __forceinline__ int operator()(std::int64_t v) const {
        return part_weights[labels[v]] > capacity ? 1 : 0;
    }
};
struct CandidateStore {
    std::int32_t* candidates;
    std::int32_t* flags;
    std::int64_t n;
    RoundState* st;
    __device__ __forceinline__ void operator()(std::int64_t v, int prefix, int item) const {
        flags[v] = 0;
        if (item) candidates[prefix] = static_cast<std::int32_t>(v);
        if (v == n - 1) st->candidates = prefix + item;
    }
};

// incremental_rebalance_proposals (:111-157) + balance_source_histogram (:174-186).
// One group of G lanes per candidate, lane j owning parts [8j, 8j + 8).
template <typename W, int G, typename C>
__global__ void __launch_bounds__(kBlock) balance_proposals_kernel(RoundState* st, const std::int32_t* cand,
                                                                   const C* cache,
                                                                   const std::int32_t* labels, const W* vertex_weights,
                                                                   int k, const unsigned long long* part_weights,
                                                                   std::uint64_t capacity, PowerTable powers,
                                                                   std::int32_t* targets, std::int32_t* cbucket,
                                                                   unsigned long long* replicas) {
    __shared__ unsigned long long pw[32];
    __shared__ BucketTable table;
    // the round state past the candidate count starts at zero
    constexpr unsigned kCandidateWord = offsetof(RoundState, candidates) / 8;
    if (blockIdx.x == 0 && threadIdx.x < sizeof(RoundState) / 8 && threadIdx.x != kCandidateWord)
        reinterpret_cast<unsigned long long*>(st)[threadIdx.x] = 0;
    if (threadIdx.x < 32) pw[threadIdx.x] = threadIdx.x < k ? part_weights[threadIdx.x] : 0;
    table_clear(table);
    __syncthreads();
    unsigned long long* counts = replicas + static_cast<long long>(blockIdx.x % kReplicas) * k * kBalanceWidth;
    const long long count = st->candidates;
    const int j = threadIdx.x & (G - 1);
    const long long groups = static_cast<long long>(gridDim.x) * (kBlock / G);
    for (long long index = (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x) / G; index < count;
         index += groups) {
        const std::int64_t v = cand[index];
        const std::int32_t source = labels[v];
        const std::uint64_t weight = static_cast<std::uint64_t>(vertex_weights[v]);
        const C* row = cache + v * k;
        unsigned long long slack[kRowSpan];
        unsigned long long conn = 0, best_slack = 0, current = 0;
        int bp = 0x7fffffff;
        unsigned long long span[kRowSpan];
        load_span(row, j, k, span);
        // lane-local best by (conn, slack), lowest index on ties (ascending p)
#pragma unroll
        for (int e = 0; e < kRowSpan; ++e) {
            const int p = j * kRowSpan + e;
            const unsigned long long value = span[e];
            if (p == source) current = value;
            const bool ok = p < k && p != source && pw[p] < capacity && weight <= capacity - pw[p];
            slack[e] = ok ? capacity - pw[p] : 0;
            if (ok && (bp == 0x7fffffff || value > conn || (value == conn && slack[e] > best_slack))) {
                conn = value;
                best_slack = slack[e];
                bp = p;
            }
        }
#pragma unroll
        for (int m = G / 2; m > 0; m /= 2) {
            const auto oc = shfl_xor_u64(conn, m, G);
            const auto os = shfl_xor_u64(best_slack, m, G);
            const int op = __shfl_xor(bp, m, G);
            const auto ocur = shfl_xor_u64(current, m, G);
            const bool ob = op != 0x7fffffff;
            const bool mine = bp != 0x7fffffff;
            if (ob && (!mine || oc > conn || (oc == conn && (os > best_slack || (os == best_slack && op < bp))))) {
                conn = oc;
                best_slack = os;
                bp = op;
            }
            current = ocur > current ? ocur : current;  // only the source's lane holds it
        }
        int best = bp == 0x7fffffff ? -1 : bp;
        if (best >= 0 && conn == 0) {
            // No affinity anywhere: a ticket mix32(v) mod total slack walks the
            // eligible parts in index order. Saturating sums of non-negative terms do
            // not depend on the order of addition, so the running total before part p
            // is the lane's running prefix plus the saturated total of lower lanes;
            // the walk stops at the first part whose running total exceeds the ticket.
            unsigned long long running[kRowSpan];
            unsigned long long lane_total = 0;
#pragma unroll
            for (int e = 0; e < kRowSpan; ++e) {
                lane_total = ~0ULL - lane_total < slack[e] ? ~0ULL : lane_total + slack[e];
                running[e] = lane_total;
            }
            unsigned long long inclusive = lane_total;
#pragma unroll
            for (int d = 1; d < G; d <<= 1) {
                const auto other = shfl_up_u64(inclusive, d, G);
                if (j >= d) inclusive = ~0ULL - inclusive < other ? ~0ULL : inclusive + other;
            }
            const unsigned long long before = G > 1 ? shfl_up_u64(inclusive, 1, G) : 0;
            const unsigned long long lower = j > 0 ? before : 0;
            const unsigned long long total = shfl_u64(inclusive, G - 1, G);  // > 0: some part is eligible
            const std::uint32_t hash = mix32(static_cast<std::uint32_t>(v));
            const unsigned long long ticket =
                total > 0xffffffffULL ? hash : hash % static_cast<std::uint32_t>(total);
            int hit = 0x7fffffff;
#pragma unroll
            for (int e = kRowSpan - 1; e >= 0; --e) {
                const unsigned long long upto = ~0ULL - lower < running[e] ? ~0ULL : lower + running[e];
                if (upto > ticket) hit = j * kRowSpan + e;
            }
#pragma unroll
            for (int m = G / 2; m > 0; m /= 2) {
                const int other = __shfl_xor(hit, m, G);
                hit = other < hit ? other : hit;
            }
            best = hit;
        }
        if (j == 0) {
            int bucket = -1;
            if (best >= 0) {
                const double gain =
                    conn >= current ? static_cast<double>(conn - current) : -static_cast<double>(current - conn);
                const double per_weight = gain / static_cast<double>(weight);
                bucket = source * kBalanceWidth + gain_bucket(per_weight, powers) * kBalanceSections +
                         static_cast<int>(mix32(static_cast<std::uint32_t>(v)) % kBalanceSections);
                table_add(table, static_cast<unsigned>(bucket), weight, counts);
            }
            targets[v] = best;
            cbucket[index] = bucket;
        }
    }
    __syncthreads();
    table_flush(table, counts);
}

// balance_scan_buckets (:188-202): bucket totals (the kReplicas copies summed),
// the exclusive sum over the 800 buckets of each part, and the part's threshold:
// its overload (sources) or slack (destinations).
// One workgroup per part, a thread per bucket: the 16 replica loads of a thread
// are independent and all in flight at once.
constexpr int kBucketScanThreads = 1024;
static_assert(kBucketScanThreads >= kBalanceWidth, "a thread per bucket");
static_assert(kRoundBlocks % kBucketScanThreads == 0, "column scan items per thread");
__global__ void __launch_bounds__(kBucketScanThreads) balance_scan_kernel(unsigned long long* replicas, int k,
                                                                    unsigned long long* counts,
                                                                    unsigned long long* offsets,
                                                                    const unsigned long long* part_weights,
                                                                    std::uint64_t capacity, bool destinations,
                                                                    unsigned long long* thresholds) {
    __shared__ unsigned long long waves[kBucketScanThreads / kWave + 1];
    constexpr int kPer = 1;
    const long long width = static_cast<long long>(k) * kBalanceWidth;
    unsigned long long items[kPer];
    unsigned long long local = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int b = threadIdx.x * kPer + i;
        unsigned long long total = 0;
        if (b < kBalanceWidth) {
            const long long bucket = static_cast<long long>(blockIdx.x) * kBalanceWidth + b;
#pragma unroll
            for (int r = 0; r < kReplicas; ++r) total += replicas[r * width + bucket];
#pragma unroll
            for (int r = 0; r < kReplicas; ++r) replicas[r * width + bucket] = 0;  // zero for the next round
            counts[bucket] = total;
        }
        items[i] = total;
        local += items[i];
    }
    unsigned long long total;
    unsigned long long prefix = block_exclusive_scan(local, SumOp<unsigned long long>{}, &total, waves);
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int b = threadIdx.x * kPer + i;
        if (b < kBalanceWidth) offsets[static_cast<std::int64_t>(blockIdx.x) * kBalanceWidth + b] = prefix;
        prefix += items[i];
    }
    if (threadIdx.x == 0) {
        const unsigned long long w = part_weights[blockIdx.x];
        thresholds[blockIdx.x] =
            destinations ? (w < capacity ? capacity - w : 0) : (w > capacity ? w - capacity : 0);
    }
}

// balance_select_sources (:204-217) for non-boundary buckets; boundary members
// are summed per part and workgroup for the multisplit.
template <typename W>
__global__ void __launch_bounds__(kBlock) balance_sources_kernel(const RoundState* st, const std::int32_t* cand,
                                                                 const std::int32_t* cbucket, std::int32_t* dbucket,
                                                                 const std::int32_t* targets,
                                                                 const W* vertex_weights,
                                                                 const unsigned long long* counts,
                                                                 const unsigned long long* offsets,
                                                                 const unsigned long long* thresholds,
                                                                 unsigned long long* block_sums,
                                                                 unsigned long long* dst_replicas, long long width) {
    __shared__ BucketTable table;
    __shared__ unsigned long long sums[32];
    unsigned long long* dst_counts = dst_replicas + static_cast<long long>(blockIdx.x % kReplicas) * width;
    table_clear(table);
    if (threadIdx.x < 32) sums[threadIdx.x] = 0;
    __syncthreads();
    long long begin, end;
    block_slice(st->candidates, begin, end);
    for (long long i = begin + threadIdx.x; i < end; i += blockDim.x) {
        const int b = cbucket[i];
        int d = kNoBucket;
        if (b >= 0) {
            const int s = b / kBalanceWidth;
            const std::int32_t v = cand[i];
            const auto w = static_cast<unsigned long long>(vertex_weights[v]);
            const BucketSide side = source_side(offsets[b], counts[b], thresholds[s]);
            if (side == kTaken) {
                d = targets[v] * kBalanceWidth + b % kBalanceWidth;
                table_add(table, static_cast<unsigned>(d), w, dst_counts);
            } else if (side == kBoundary) {
                d = kSourcePending;
                atomicAdd(&sums[s], w);
            }
        }
        dbucket[i] = d;
    }
    __syncthreads();
    table_flush(table, dst_counts);
    if (threadIdx.x < 32) block_sums[static_cast<long long>(blockIdx.x) * 32 + threadIdx.x] = sums[threadIdx.x];
}

// Exclusive scan of each part's column of per-workgroup sums (blocks <= 1024).
__global__ void __launch_bounds__(kBucketScanThreads) column_scan_kernel(const unsigned long long* sums, int blocks,
                                                                   unsigned long long* prefix) {
    __shared__ unsigned long long waves[kBucketScanThreads / kWave + 1];
    constexpr int kPer = kRoundBlocks / kBucketScanThreads;
    const int column = blockIdx.x;
    unsigned long long items[kPer];
    unsigned long long local = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int b = threadIdx.x * kPer + i;
        items[i] = b < blocks ? sums[static_cast<long long>(b) * 32 + column] : 0;
        local += items[i];
    }
    unsigned long long total;
    unsigned long long running = block_exclusive_scan(local, SumOp<unsigned long long>{}, &total, waves);
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int b = threadIdx.x * kPer + i;
        if (b < blocks) prefix[static_cast<long long>(b) * 32 + column] = running;
        running += items[i];
    }
}

// The exclusive prefix over the workgroups before this one of each of the 32
// columns of block_sums (a row of 32 per workgroup) into running[]: read from
// column_prefix when the column scan ran, else summed here from the rows, which
// grids of at most kInlineColumnBlocks workgroups do instead of that launch
// (thread t sums column t % 32 over the rows t / 32, t / 32 + 8, ...).
constexpr int kInlineColumnBlocks = 256;
__device__ __forceinline__ void column_running(const unsigned long long* column_prefix,
                                               const unsigned long long* block_sums, unsigned long long* running,
                                               unsigned long long (*groups)[32]) {
    static_assert(kBlock == 256, "8 row groups of 32 columns");
    if (column_prefix) {
        if (threadIdx.x < 32)
            running[threadIdx.x] = column_prefix[static_cast<long long>(blockIdx.x) * 32 + threadIdx.x];
    } else {
        const int column = threadIdx.x & 31, group = threadIdx.x >> 5;
        unsigned long long sum = 0;
        for (int row = group; row < static_cast<int>(blockIdx.x); row += kBlock / 32)
            sum += block_sums[static_cast<long long>(row) * 32 + column];
        groups[group][column] = sum;
        __syncthreads();
        if (threadIdx.x < 32) {
            unsigned long long total = 0;
#pragma unroll
            for (int g = 0; g < kBlock / 32; ++g) total += groups[g][threadIdx.x];
            running[threadIdx.x] = total;
        }
    }
}

// Source boundary members: taken while offset + rank < overload.
template <typename W>
__global__ void __launch_bounds__(kBlock) balance_source_ranks_kernel(const RoundState* st, const std::int32_t* cand,
                                                                      const std::int32_t* cbucket,
                                                                      std::int32_t* dbucket,
                                                                      const std::int32_t* targets,
                                                                      const W* vertex_weights,
                                                                      const unsigned long long* offsets,
                                                                      const unsigned long long* thresholds,
                                                                      const unsigned long long* column_prefix,
                                                                      const unsigned long long* block_sums,
                                                                      unsigned long long* dst_replicas,
                                                                      long long width) {
    __shared__ BucketTable table;
    unsigned long long* dst_counts = dst_replicas + static_cast<long long>(blockIdx.x % kReplicas) * width;
    __shared__ unsigned long long wave_total[kBlock / kWave][32];
    __shared__ unsigned long long running[32];
    __shared__ unsigned long long groups[kBlock / 32][32];
    table_clear(table);
    column_running(column_prefix, block_sums, running, groups);
    __syncthreads();
    long long begin, end;
    block_slice(st->candidates, begin, end);
    for (long long base = begin; base < end; base += blockDim.x) {
        const long long i = base + threadIdx.x;
        const bool pending = i < end && dbucket[i] == kSourcePending;
        if (!__syncthreads_or(pending)) continue;
        int b = 0, s = 0;
        std::int32_t v = 0;
        unsigned long long w = 0;
        if (pending) {
            b = cbucket[i];
            s = b / kBalanceWidth;
            v = cand[i];
            w = static_cast<unsigned long long>(vertex_weights[v]);
        }
        const unsigned long long rank = block_key_prefix<kBlock>(pending, s, w, wave_total, running);
        if (pending) {
            if (offsets[b] + rank < thresholds[s]) {
                const int d = targets[v] * kBalanceWidth + b % kBalanceWidth;
                dbucket[i] = d;
                table_add(table, static_cast<unsigned>(d), w, dst_counts);
            } else {
                dbucket[i] = kNoBucket;
            }
        }
    }
    __syncthreads();
    table_flush(table, dst_counts);
}

// balance_select_destinations (:219-232) for non-boundary buckets: taken ones
// become movers; boundary members are summed for the multisplit.
template <typename O, typename W>
__global__ void __launch_bounds__(kBlock) balance_destinations_kernel(
    RoundState* st, const std::int32_t* cand, std::int32_t* dbucket, const W* vertex_weights, const O* offsets_csr,
    const unsigned long long* counts, const unsigned long long* offsets, const unsigned long long* thresholds,
    unsigned long long* block_sums, std::int32_t* light, std::int32_t* heavy, std::int32_t* flags) {
    __shared__ unsigned long long sums[32];
    __shared__ unsigned counts_lds[2 * (kBlock / kWave)];
    __shared__ unsigned long long base_lds[2];
    if (threadIdx.x < 32) sums[threadIdx.x] = 0;
    __syncthreads();
    long long begin, end;
    block_slice(st->candidates, begin, end);
    for (long long base = begin; base < end; base += blockDim.x) {
        const long long i = base + threadIdx.x;
        bool take = false;
        std::int32_t v = 0;
        if (i < end) {
            const int d = dbucket[i];
            if (d >= 0) {
                const int part = d / kBalanceWidth;
                v = cand[i];
                const BucketSide side = destination_side(offsets[d], counts[d], thresholds[part]);
                if (side == kTaken) {
                    take = true;
                } else if (side == kBoundary) {
                    dbucket[i] = -3 - d;
                    atomicAdd(&sums[part], static_cast<unsigned long long>(vertex_weights[v]));
                }
            }
        }
        append_movers(take, v, offsets_csr, light, heavy, flags, st, counts_lds, base_lds);
    }
    if (threadIdx.x < 32) block_sums[static_cast<long long>(blockIdx.x) * 32 + threadIdx.x] = sums[threadIdx.x];
}

// Destination boundary members: admitted while offset + rank + weight <= slack.
template <typename O, typename W>
__global__ void __launch_bounds__(kBlock) balance_destination_ranks_kernel(
    RoundState* st, const std::int32_t* cand, const std::int32_t* dbucket, const W* vertex_weights,
    const O* offsets_csr, const unsigned long long* offsets, const unsigned long long* thresholds,
    const unsigned long long* column_prefix, const unsigned long long* block_sums, std::int32_t* light,
    std::int32_t* heavy, std::int32_t* flags) {
    __shared__ unsigned long long wave_total[kBlock / kWave][32];
    __shared__ unsigned long long running[32];
    __shared__ unsigned long long groups[kBlock / 32][32];
    __shared__ unsigned counts_lds[2 * (kBlock / kWave)];
    __shared__ unsigned long long base_lds[2];
    column_running(column_prefix, block_sums, running, groups);
    __syncthreads();
    long long begin, end;
    block_slice(st->candidates, begin, end);
    for (long long base = begin; base < end; base += blockDim.x) {
        const long long i = base + threadIdx.x;
        const int code = i < end ? dbucket[i] : kNoBucket;
        const bool pending = code <= -3;
        if (!__syncthreads_or(pending)) continue;
        int d = 0, part = 0;
        std::int32_t v = 0;
        unsigned long long w = 0;
        if (pending) {
            d = -3 - code;
            part = d / kBalanceWidth;
            v = cand[i];
            w = static_cast<unsigned long long>(vertex_weights[v]);
        }
        const unsigned long long rank = block_key_prefix<kBlock>(pending, part, w, wave_total, running);
        const bool take = pending && offsets[d] + rank + w <= thresholds[part];
        append_movers(take, v, offsets_csr, light, heavy, flags, st, counts_lds, base_lds);
    }
}

// ---------------------------------------------------------------- round graphs
//
// A refinement round (an LP round or a rebalance iteration) is a fixed sequence
// of launches for a given level and set of arrays, repeated tens of times per
// level, each followed by a host wait for its status. The first call of a
// sequence captures its launches into hipGraphs; later calls with the same key
// (every pointer, size and setting the launches depend on) replay them, which
// submits a round at once instead of launch by launch (on MI355X 2.0-2.6 us per
// kernel instead of 2.8-4.3 us with a host wait every 20 kernels). The round's
// scratch comes from its plan, whose blocks keep their addresses from call to
// call. A family timed with events runs outside the graphs (events recorded in a
// capture carry no timing), which splits the sequence into segments.
constexpr std::size_t kMaxRoundPlans = 64;
constexpr std::int64_t kMaxGraphVertices = std::int64_t{1} << 23;  // larger levels are not launch-bound

BlockCache* plan_cache(gpart_ctx* ctx) {
    if (!ctx->plan_cache) {
        SegmentSource source;
        source.allocate = [](std::size_t bytes) -> void* {
            void* q = nullptr;
            const hipError_t e = hipMalloc(&q, bytes);
            if (e == hipErrorOutOfMemory) {
                (void)hipGetLastError();
                return nullptr;
            }
            GPART_HIP_CHECK(e);
            return q;
        };
        source.release = [](void* q) { GPART_HIP_CHECK(hipFree(q)); };
        source.drain = [ctx] { GPART_HIP_CHECK(hipStreamSynchronize(ctx->stream)); };
        ctx->plan_cache = std::make_unique<BlockCache>(std::move(source));
    }
    return ctx->plan_cache.get();
}

struct RoundPlan final : ScratchRecorder {
    explicit RoundPlan(gpart_ctx* c, std::vector<std::uint64_t> k) : ctx(c), key(std::move(k)) {}
    ~RoundPlan() override {
        for (auto g : segments) (void)hipGraphExecDestroy(g);
        for (auto& list : blocks)
            for (auto& b : list) ctx->plan_cache->free(b.first);
    }
    void* take(gpart_ctx*, std::size_t bytes) override {
        if (blocks.size() <= static_cast<std::size_t>(phase)) {
            blocks.resize(phase + 1);
            cursor.resize(phase + 1, 0);
        }
        auto& list = blocks[phase];
        std::size_t& at = cursor[phase];
        if (!recorded) {
            void* p = nullptr;
            try {
                p = plan_cache(ctx)->allocate(bytes);
            } catch (const std::runtime_error& e) {
                if (dynamic_cast<const Error*>(&e)) throw;
                throw Error(GPART_ERR_RUNTIME, e.what());
            }
            list.emplace_back(p, bytes);
            owned.insert(p);
            ++at;
            return p;
        }
        if (at >= list.size() || list[at].second != bytes)
            throw Error(GPART_ERR_RUNTIME, "a recorded refinement round asked for different scratch");
        return list[at++].first;
    }
    bool owns(const void* p) const override { return owned.count(p) != 0; }

    gpart_ctx* ctx;
    std::vector<std::uint64_t> key;
    std::vector<std::vector<std::pair<void*, std::size_t>>> blocks;  // per phase; 0: the round's body
    std::vector<std::size_t> cursor;
    std::unordered_set<const void*> owned;
    std::vector<hipGraphExec_t> segments;
    bool recorded = false;  // a first call has captured the segments
    bool broken = false;    // a capture failed: the plan is dropped
    int phase = 0;
    long long calls = 0, captures = 0, replays = 0;
};

struct RoundPlans {
    gpart_ctx* ctx = nullptr;
    std::list<std::unique_ptr<RoundPlan>> plans;  // most recently used first
    long long captures = 0, replays = 0;
    ~RoundPlans() {
        if (ctx && !plans.empty()) (void)hipStreamSynchronize(ctx->stream);
    }
};

RoundPlans& round_plans(gpart_ctx* ctx) {
    if (!ctx->round_plans) {
        auto* set = new RoundPlans;
        set->ctx = ctx;
        ctx->round_plans = std::shared_ptr<void>(set, [](void* q) { delete static_cast<RoundPlans*>(q); });
    }
    return *static_cast<RoundPlans*>(ctx->round_plans.get());
}

// The plan of a round with this key, or nullptr (launch directly) when the
// context does not use round graphs, the level is large, or the key is new: a
// round is recorded on its second call, so a sequence that runs once never pays
// for a capture.
RoundPlan* round_plan(gpart_ctx* ctx, std::int64_t n, std::vector<std::uint64_t> key) {
    if (!ctx->round_graphs || n > kMaxGraphVertices) return nullptr;
    for (int f = 0; f < GPART_PROFILE_FAMILIES; ++f) key.push_back(ctx->profile[f].enabled ? 1 : 0);
    auto& set = round_plans(ctx);
    for (auto it = set.plans.begin(); it != set.plans.end(); ++it) {
        if ((*it)->key != key) continue;
        if ((*it)->broken) {
            set.plans.erase(it);
            break;
        }
        set.plans.splice(set.plans.begin(), set.plans, it);
        RoundPlan* plan = set.plans.front().get();
        return ++plan->calls >= 2 ? plan : nullptr;
    }
    if (set.plans.size() >= kMaxRoundPlans) {
        GPART_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // its graphs may still run
        set.plans.pop_back();
    }
    set.plans.push_front(std::make_unique<RoundPlan>(ctx, std::move(key)));
    set.plans.front()->calls = 1;
    return nullptr;
}

// One call of a round: phases run their launches into the plan's graphs (first
// call) or replay them (later calls). Scratch of the round's body is taken
// before, between or after phases in the same order every call.
class RoundRun {
public:
    RoundRun(gpart_ctx* ctx, RoundPlan* plan) : ctx_(ctx), plan_(plan) {
        if (!plan_) return;
        ctx_->recorder = plan_;
        plan_->phase = 0;
        std::fill(plan_->cursor.begin(), plan_->cursor.end(), 0);
        if (plan_->recorded) {
            ++plan_->replays;
            ++round_plans(ctx_).replays;
        } else {
            ++plan_->captures;
            ++round_plans(ctx_).captures;
        }
    }
    ~RoundRun() {
        if (!plan_) return;
        if (capturing_) {  // an exception left a capture open
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(ctx_->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            plan_->broken = true;
        }
        ctx_->recorder = nullptr;
    }
    RoundRun(const RoundRun&) = delete;
    RoundRun& operator=(const RoundRun&) = delete;

    // The launches of one phase; family >= 0: the profile family they report to
    // (run outside the graphs while that family is timed).
    template <typename F>
    void phase(int family, F&& launches) {
        if (!plan_) {
            launches();
            return;
        }
        plan_->phase = ++phases_;
        if (plan_->cursor.size() > static_cast<std::size_t>(phases_)) plan_->cursor[phases_] = 0;
        const bool direct = family >= 0 && ctx_->profile[family].enabled;
        if (direct) {
            close_segment();
            in_segment_ = false;
            launches();
        } else if (!plan_->recorded) {
            if (!capturing_) {
                GPART_HIP_CHECK(hipStreamBeginCapture(ctx_->stream, hipStreamCaptureModeRelaxed));
                capturing_ = true;
            }
            launches();
        } else if (!in_segment_) {
            if (segment_ >= plan_->segments.size())
                throw Error(GPART_ERR_RUNTIME, "a recorded refinement round has fewer segments than its calls");
            GPART_HIP_CHECK(hipGraphLaunch(plan_->segments[segment_++], ctx_->stream));
            in_segment_ = true;
        }
        plan_->phase = 0;
    }
    // After the last phase.
    void end() {
        if (!plan_) return;
        close_segment();
        plan_->recorded = true;
    }

private:
    // First call: the open capture becomes a segment and runs now.
    void close_segment() {
        if (!capturing_) return;
        hipGraph_t graph = nullptr;
        capturing_ = false;
        GPART_HIP_CHECK(hipStreamEndCapture(ctx_->stream, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            plan_->broken = true;
            GPART_HIP_CHECK(e);
        }
        plan_->segments.push_back(exec);
        GPART_HIP_CHECK(hipGraphLaunch(exec, ctx_->stream));
    }

    gpart_ctx* ctx_;
    RoundPlan* plan_;
    int phases_ = 0;
    bool capturing_ = false;
    bool in_segment_ = false;
    std::size_t segment_ = 0;
};

inline std::uint64_t key_of(const void* p) { return reinterpret_cast<std::uintptr_t>(p); }
inline std::uint64_t key_of(double v) {
    std::uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    return b;
}

// ---------------------------------------------------------------- drivers

template <typename O, typename W, typename C = unsigned long long>
struct Ops {
    static void cut(gpart_ctx* ctx, const gpart_csr_view* g, const std::int32_t* labels, std::uint64_t* out) {
        const std::int64_t tiles = g->n > 0 && g->m > 0 ? tile_count(g->m) : 0;
        Scratch<unsigned long long> d(ctx, 1);
        Scratch<std::int32_t> tile_row(ctx, tiles);
        Scratch<unsigned long long> partials(ctx, std::min<std::int64_t>(tiles, kReduceMaxBlocks));
        cut_with(ctx, g, labels, d.get(), tile_row.get(), partials.get(), out);
    }

    // The cut over caller-provided device scratch: total (1 word), tile_row
    // (tile_count(m) entries) and partials (min(tiles, kReduceMaxBlocks) words).
    static void cut_with(gpart_ctx* ctx, const gpart_csr_view* g, const std::int32_t* labels,
                         unsigned long long* total, std::int32_t* tile_row, unsigned long long* partials,
                         std::uint64_t* out) {
        fill_bytes(ctx, total, 0, 8);
        const auto* off = static_cast<const O*>(g->offsets);
        const auto* w = static_cast<const W*>(g->edge_weights);
        if (g->n > 0 && g->m > 0) {
            // edge tiles, capped grid, partials folded by one workgroup
            const std::int64_t tiles = tile_count(g->m);
            hipLaunchKernelGGL(tile_rows_kernel<O>, dim3(grid_for(g->n)), dim3(kBlock), 0, ctx->stream, g->n, off,
                               tile_row);
            const int blocks = static_cast<int>(std::min<std::int64_t>(tiles, kReduceMaxBlocks));
            hipLaunchKernelGGL((cut_kernel<O, W>), dim3(blocks), dim3(kTileLanes), 0, ctx->stream, g->n, g->m, tiles,
                               off, tile_row, g->neighbors, w, labels, partials);
            hipLaunchKernelGGL((reduce_finish_kernel<unsigned long long, SumU64>), dim3(1), dim3(kReduceThreads), 0,
                               ctx->stream, partials, blocks, SumU64{}, total);
            GPART_LAUNCH_CHECK();
        }
        const auto directed = read_scalar(ctx, total);
        if (directed & 1ULL) throw Error(GPART_ERR_RUNTIME, "refinement cut is not symmetric");
        *out = directed / 2;
    }

    // Writes every row of the cache. With a gate, the kernels run only when
    // *gate > above (decided on the device, inside a round); cut, when given,
    // receives the directed cut weight of `labels`.
    static void build_cache(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* c,
                            const std::int32_t* labels, int k, C* cache, const long long* gate = nullptr,
                            long long above = 0, unsigned long long* cut = nullptr) {
        const auto* off = static_cast<const O*>(g->offsets);
        const auto* w = static_cast<const W*>(g->edge_weights);
        if (!gate) profile_begin(ctx, GPART_PROFILE_REFINE_CACHE);
        auto grid = [](std::int64_t threads) {
            return static_cast<unsigned>(std::min<std::int64_t>(grid_for(threads), kCacheBlocks));
        };
        if (c->low_count)
            hipLaunchKernelGGL((build_cache_kernel<O, W, 8, C>), dim3(grid(c->low_count * 8)), dim3(kBlock), 0,
                               ctx->stream, c->low_count, c->low, off, g->neighbors, w, labels, k, cache, gate, above,
                               cut);
        if (c->medium_count)
            hipLaunchKernelGGL((build_cache_kernel<O, W, 64, C>), dim3(grid(c->medium_count * 64)), dim3(kBlock), 0,
                               ctx->stream, c->medium_count, c->medium, off, g->neighbors, w, labels, k, cache, gate,
                               above, cut);
        if (c->high_count)
            hipLaunchKernelGGL((build_cache_kernel<O, W, kBlock, C>), dim3(grid(c->high_count * kBlock)),
                               dim3(kBlock), 0, ctx->stream, c->high_count, c->high, off, g->neighbors, w, labels, k,
                               cache, gate, above, cut);
        if (c->max_degree > kHubChunk) {
            HubChunks chunks;
            build_hub_chunks(ctx, off, c->high, c->high_count, c->high_edges, kHubChunk, chunks);
            hipLaunchKernelGGL(hub_rows_zero_kernel<C>, dim3(grid_for(c->high_count)), dim3(kBlock), 0, ctx->stream,
                               c->high_count, chunks.start.get(), c->high, k, cache, gate, above);
            hipLaunchKernelGGL((build_cache_hub_kernel<O, W, C>), dim3(chunks.grid()), dim3(kBlock), 0, ctx->stream,
                               chunks.total(), chunks.start.get(), chunks.hub.get(), c->high, off, g->neighbors, w,
                               labels, k, cache, gate, above, cut);
        }
        GPART_LAUNCH_CHECK();
        if (gate) return;
        const double n = static_cast<double>(g->n), m = static_cast<double>(g->m);
        profile_end(ctx, GPART_PROFILE_REFINE_CACHE,
                    (n + 1) * sizeof(O) + m * (4 + sizeof(W)) + 4 * m + static_cast<double>(sizeof(C)) * n * k);
    }

    static void filter_movers(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* c,
                              const std::int32_t* labels, const std::int32_t* targets, const long long* gains,
                              std::int32_t* flags) {
        const auto* off = static_cast<const O*>(g->offsets);
        const auto* w = static_cast<const W*>(g->edge_weights);
        profile_begin(ctx, GPART_PROFILE_REFINE_FILTER);
        const unsigned low_blocks = grid_for(c->low_count * 8), medium_blocks = grid_for(c->medium_count * 64);
        const auto high_blocks = static_cast<unsigned>(c->high_count);
        if (low_blocks + medium_blocks + high_blocks > 0)
            hipLaunchKernelGGL((priority_filter_kernel<O, W>), dim3(low_blocks + medium_blocks + high_blocks),
                               dim3(kBlock), 0, ctx->stream, c->low_count, c->low, low_blocks, c->medium_count,
                               c->medium, medium_blocks, c->high_count, c->high, off, g->neighbors, w, labels,
                               targets, gains, flags);
        if (c->max_degree > kHubChunk) {
            HubChunks chunks;
            build_hub_chunks(ctx, off, c->high, c->high_count, c->high_edges, kHubChunk, chunks);
            Scratch<long long> sums(ctx, c->high_count);
            fill_bytes(ctx, sums.get(), 0, static_cast<std::size_t>(c->high_count) * 8);
            hipLaunchKernelGGL((priority_filter_hub_kernel<O, W>), dim3(chunks.grid()), dim3(kBlock), 0, ctx->stream,
                               chunks.total(), chunks.start.get(), chunks.hub.get(), c->high, off, g->neighbors, w,
                               labels, targets, gains, sums.get());
            hipLaunchKernelGGL(priority_filter_hub_finish_kernel, dim3(grid_for(c->high_count)), dim3(kBlock), 0,
                               ctx->stream, c->high_count, chunks.start.get(), c->high, targets, gains, sums.get(),
                               flags);
        }
        GPART_LAUNCH_CHECK();
        const double nd = static_cast<double>(g->n), m = static_cast<double>(g->m);
        profile_end(ctx, GPART_PROFILE_REFINE_FILTER, (nd + 1) * sizeof(O) + m * (4 + sizeof(W)) + 16 * m + 20 * nd);
    }

    // incremental_update_cache_and_cut + incremental_commit over the round's movers.
    // On levels of at least ctx->rebuild_min_entries entries the device chooses,
    // from the round's mover entries, between atomic updates (two per mover entry)
    // and a full rebuild of the cache after the commit (one gather per entry, no
    // atomics, same cache); the cut delta is summed from the movers' entries
    // either way.
    static void apply(gpart_ctx* ctx, RoundRun& run, const gpart_csr_view* g, const gpart_degree_classes* c,
                      RoundState* st, const std::int32_t* light, const std::int32_t* heavy, std::int32_t* labels,
                      const std::int32_t* targets, const std::int32_t* flags, int k, C* cache,
                      unsigned long long* part_weights, std::int32_t* locks, bool lp) {
        const std::int64_t n = g->n;
        const auto* off = static_cast<const O*>(g->offsets);
        const auto* w = static_cast<const W*>(g->edge_weights);
        const long long rebuild_above = rebuild_threshold(ctx, g);
        const auto light_blocks = static_cast<unsigned>(std::min<std::int64_t>(grid_for(n * 8), 4 * kRoundBlocks));
        run.phase(GPART_PROFILE_CACHE_UPDATE, [&] {
        profile_begin(ctx, GPART_PROFILE_CACHE_UPDATE);
        // levels whose degrees stay at most kWaveHeavyMaxDegree update their heavy
        // movers a wave each, in the same launch
        const bool wave_heavy = c->max_degree > kHeavyDegree && c->max_degree <= kWaveHeavyMaxDegree;
        const std::int64_t heavy_bound = c->medium_count + c->high_count;
        const unsigned heavy_blocks =
            wave_heavy ? static_cast<unsigned>(std::max<std::int64_t>(
                             1, std::min<std::int64_t>(grid_for(heavy_bound * kWave), 4 * kRoundBlocks)))
                       : 0;
        hipLaunchKernelGGL((update_cache_kernel<O, W, C>), dim3(light_blocks + heavy_blocks), dim3(kBlock), 0,
                           ctx->stream, st, light, wave_heavy ? heavy : nullptr, light_blocks, off, g->neighbors, w,
                           labels, targets, flags, k, cache, &st->delta, rebuild_above);
        if (!wave_heavy && c->max_degree > kHeavyDegree) {
            Scratch<long long> heavy_start(ctx, heavy_bound + 1);
            Scratch<std::int32_t> tile_mover(ctx, tile_count(g->m) + 1);
            device_scan<long long, SumOp<long long>>(ctx, HeavyDegreeLoad<O>{st, heavy, off},
                                                     ExclusiveStore<long long>{heavy_start.get()}, heavy_bound + 1,
                                                     SumOp<long long>{});
            hipLaunchKernelGGL(heavy_tiles_kernel,
                               dim3(static_cast<unsigned>(std::min<std::int64_t>(grid_for(heavy_bound), kRoundBlocks))),
                               dim3(kBlock), 0, ctx->stream, st, heavy_start.get(), tile_mover.get());
            const auto tiles = static_cast<unsigned>(std::max<std::int64_t>(
                1, std::min<std::int64_t>(tile_count(g->m), 2 * kRoundBlocks)));
            hipLaunchKernelGGL((update_cache_tiles_kernel<O, W, C>), dim3(tiles), dim3(kTileLanes), 0, ctx->stream,
                               st, heavy_start.get(), tile_mover.get(), heavy, off, g->neighbors, w, labels, targets,
                               flags, k, cache, &st->delta, rebuild_above);
        }
        profile_end(ctx, GPART_PROFILE_CACHE_UPDATE, 0.0);
        });
        run.phase(-1, [&] {
            hipLaunchKernelGGL(commit_kernel<W>,
                               dim3(static_cast<unsigned>(std::min<std::int64_t>(grid_for(n), kRoundBlocks))),
                               dim3(kBlock), 0, ctx->stream, st, light, heavy, labels, targets,
                               static_cast<const W*>(g->vertex_weights), part_weights, locks, lp, k);
            GPART_LAUNCH_CHECK();
            if (rebuild_above >= 0)
                build_cache(ctx, g, c, labels, k, cache, &st->edges, rebuild_above, &st->cut_directed);
        });
    }

    // Mover entries above which a round rebuilds the cache; -1: never.
    static long long rebuild_threshold(const gpart_ctx* ctx, const gpart_csr_view* g) {
        return ctx->rebuild_share > 0 && g->m >= ctx->rebuild_min_entries ? g->m / ctx->rebuild_share : -1;
    }

    // The round's status into the fine-grained status words (the last phase), then
    // the host's wait and reading of it.
    static void finish(gpart_ctx* ctx, RoundRun& run, const gpart_csr_view* g, RoundState* st,
                       const unsigned long long* part_weights, int k, gpart_refine_status* out) {
        run.phase(-1, [&] {
            hipLaunchKernelGGL(round_status_kernel, dim3(1), dim3(kWave), 0, ctx->stream, st, part_weights, k,
                               reinterpret_cast<unsigned long long*>(ctx->status_device), ctx->status_counter);
            GPART_LAUNCH_CHECK();
        });
        run.end();
        const std::uint64_t seq = ++ctx->status_seq;
        wait_status(ctx, seq);
        RoundState h{};
        std::memcpy(&h, ctx->status_host + 1, offsetof(RoundState, pw) + 8 * static_cast<std::size_t>(k));
        out->movers = h.light + h.heavy;
        out->cut_delta = h.delta;
        out->candidates = h.candidates;
        out->max_part_weight = h.max_weight;
        for (int p = 0; p < 32; ++p) out->part_weights[p] = p < k ? h.pw[p] : 0;
        out->mover_entries = h.edges;
        const long long rebuild_above = rebuild_threshold(ctx, g);
        out->rebuilt = rebuild_above >= 0 && h.edges > rebuild_above ? 1 : 0;
        out->cut_after = 0;
        if (!out->rebuilt) {
            // the movers' rows and ids, per entry a neighbour, a weight, two cache
            // atomics (read and write each) and the label, target and flag gathers
            const double movers = static_cast<double>(out->movers), entries = static_cast<double>(h.edges);
            profile_add_bytes(ctx, GPART_PROFILE_CACHE_UPDATE,
                              movers * (2.0 * sizeof(O) + 16.0) + entries * (4.0 + sizeof(W) + 4.0 * sizeof(C) + 12.0));
            profile_add_units(ctx, GPART_PROFILE_CACHE_UPDATE, 2.0 * entries);
        }
        if (out->rebuilt) {
            if (h.cut_directed & 1ULL) throw Error(GPART_ERR_RUNTIME, "refinement cut is not symmetric");
            out->cut_after = h.cut_directed / 2;
            out->cut_delta = 0;
        }
    }

    // One LP round: proposals, priority filter, movers, apply; one synchronisation.
    static void lp_round(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* c,
                         C* cache, std::int32_t* labels, int k, double filter,
                         std::int32_t* locks, std::int32_t* targets, long long* gains, std::int32_t* flags,
                         unsigned long long* part_weights, gpart_refine_status* out) {
        const std::int64_t n = g->n;
        const auto* off = static_cast<const O*>(g->offsets);
        RoundRun run(ctx, round_plan(ctx, n, round_key(ctx, g, c, cache, labels, k, targets, flags, part_weights,
                                                       {0, key_of(filter), key_of(locks), key_of(gains)})));
        // the proposals kernel zeroes the round state
        Scratch<RoundState> st(ctx, 1);
        Scratch<std::int32_t> light(ctx, n), heavy(ctx, n);
        const int G = group_for_k(k);
        run.phase(GPART_PROFILE_REFINE_PROPOSALS, [&] {
            profile_begin(ctx, GPART_PROFILE_REFINE_PROPOSALS);
            const unsigned blocks = std::max<unsigned>(1, grid_for(n * G));
            switch (G) {
                case 1: hipLaunchKernelGGL((lp_proposals_kernel<1, C>), dim3(blocks), dim3(kBlock), 0, ctx->stream, n, cache, labels, k, filter, locks, targets, gains, st.get()); break;
                case 2: hipLaunchKernelGGL((lp_proposals_kernel<2, C>), dim3(blocks), dim3(kBlock), 0, ctx->stream, n, cache, labels, k, filter, locks, targets, gains, st.get()); break;
                default: hipLaunchKernelGGL((lp_proposals_kernel<4, C>), dim3(blocks), dim3(kBlock), 0, ctx->stream, n, cache, labels, k, filter, locks, targets, gains, st.get()); break;
            }
            GPART_LAUNCH_CHECK();
            profile_end(ctx, GPART_PROFILE_REFINE_PROPOSALS, static_cast<double>(n) * (20.0 + sizeof(C) * k));
        });
        run.phase(GPART_PROFILE_REFINE_FILTER, [&] { filter_movers(ctx, g, c, labels, targets, gains, flags); });
        run.phase(-1, [&] {
            device_scan<MoverPair, MoverPairOp>(ctx, MoverLoad<O>{flags, off},
                                                MoverStore{light.get(), heavy.get(), n, st.get()}, n, MoverPairOp{});
        });
        apply(ctx, run, g, c, st.get(), light.get(), heavy.get(), labels, targets, flags, k, cache, part_weights,
              locks, true);
        finish(ctx, run, g, st.get(), part_weights, k, out);
    }

    // Everything a round's launches depend on besides its own scratch.
    static std::vector<std::uint64_t> round_key(const gpart_ctx* ctx, const gpart_csr_view* g,
                                                const gpart_degree_classes* c, const C* cache,
                                                const std::int32_t* labels, int k, const std::int32_t* targets,
                                                const std::int32_t* flags, const unsigned long long* part_weights,
                                                std::initializer_list<std::uint64_t> extra) {
        std::vector<std::uint64_t> key{sizeof(O), sizeof(C), static_cast<std::uint64_t>(g->n),
                                       static_cast<std::uint64_t>(g->m), key_of(g->offsets), key_of(g->neighbors),
                                       key_of(g->edge_weights), key_of(g->vertex_weights), key_of(c->low),
                                       key_of(c->medium), key_of(c->high), static_cast<std::uint64_t>(c->low_count),
                                       static_cast<std::uint64_t>(c->medium_count),
                                       static_cast<std::uint64_t>(c->high_count),
                                       static_cast<std::uint64_t>(c->high_edges),
                                       static_cast<std::uint64_t>(c->max_degree), key_of(cache), key_of(labels),
                                       static_cast<std::uint64_t>(k), key_of(targets), key_of(flags),
                                       key_of(part_weights), static_cast<std::uint64_t>(rebuild_threshold(ctx, g)),
                                       key_of(ctx->status_device), key_of(ctx->status_counter)};
        key.insert(key.end(), extra.begin(), extra.end());
        return key;
    }

    // One rebalance iteration: proposals, source and destination bucket quotas,
    // apply; one synchronisation.
    static void rebalance_round(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* c,
                                C* cache, std::int32_t* labels, int k,
                                unsigned long long* part_weights, std::uint64_t capacity, std::int32_t* targets,
                                std::int32_t* flags, gpart_refine_status* out) {
        const std::int64_t n = g->n;
        const auto* off = static_cast<const O*>(g->offsets);
        const auto* vw = static_cast<const W*>(g->vertex_weights);
        static const PowerTable powers = make_power_table();
        const std::int64_t width = static_cast<std::int64_t>(k) * kBalanceWidth;
        // Both replica arrays live in the context between rounds and are zero at rest
        // (the bucket scans clear what they read); the proposals kernel zeroes the
        // round state after the candidate count.
        const std::size_t counter_bytes = static_cast<std::size_t>(2 * kReplicas * width) * 8;
        if (ctx->rebalance_counter_bytes < counter_bytes) {
            if (ctx->rebalance_counters) ctx_free(ctx, ctx->rebalance_counters);
            ctx->rebalance_counters = nullptr;
            ctx->rebalance_counter_bytes = 0;
            ctx->rebalance_counters = ctx_alloc(ctx, counter_bytes);
            ctx->rebalance_counter_bytes = counter_bytes;
            ctx->rebalance_counters_zero = false;
        }
        if (!ctx->rebalance_counters_zero) fill_bytes(ctx, ctx->rebalance_counters, 0, ctx->rebalance_counter_bytes);
        ctx->rebalance_counters_zero = false;  // until this round completes
        RoundRun run(ctx, round_plan(ctx, n, round_key(ctx, g, c, cache, labels, k, targets, flags, part_weights,
                                                       {1, capacity, key_of(ctx->rebalance_counters)})));
        Scratch<RoundState> st(ctx, 1);
        Scratch<unsigned long long> counts(ctx, 2 * width), offsets(ctx, 2 * width), thresholds(ctx, 64);
        auto* st_ptr = st.get();
        auto* src_replicas = static_cast<unsigned long long*>(ctx->rebalance_counters);
        unsigned long long* dst_replicas = src_replicas + kReplicas * width;
        unsigned long long* src_counts = counts.get();
        unsigned long long* dst_counts = counts.get() + width;
        unsigned long long* src_offsets = offsets.get();
        unsigned long long* dst_offsets = offsets.get() + width;
        unsigned long long* src_thresholds = thresholds.get();
        unsigned long long* dst_thresholds = thresholds.get() + 32;
        Scratch<std::int32_t> cand(ctx, n), cbucket(ctx, n), dbucket(ctx, n), light(ctx, n), heavy(ctx, n);
        const int slices = static_cast<int>(std::max<std::int64_t>(1, std::min<std::int64_t>(grid_for(n), kRoundBlocks)));
        Scratch<unsigned long long> sums(ctx, static_cast<std::int64_t>(slices) * 32),
            prefix(ctx, static_cast<std::int64_t>(slices) * 32);
        // small grids sum the per-workgroup boundary weights in the ranks kernels
        const bool inline_columns = slices <= kInlineColumnBlocks;
        const unsigned long long* column_prefix = inline_columns ? nullptr : prefix.get();

        run.phase(-1, [&] {
            device_scan<int, SumOp<int>>(ctx, OverloadedLoad{labels, part_weights, capacity},
                                         CandidateStore{cand.get(), flags, n, st_ptr}, n, SumOp<int>{});
        });
        const int G = group_for_k(k);
        const auto blocks = static_cast<unsigned>(std::min<std::int64_t>(grid_for(n * G), ctx->balance_blocks));
        run.phase(GPART_PROFILE_REBALANCE, [&] {
        profile_begin(ctx, GPART_PROFILE_REBALANCE);
#define GPART_BALANCE_PROPOSALS(GG)                                                                              \
    hipLaunchKernelGGL((balance_proposals_kernel<W, GG, C>), dim3(blocks), dim3(kBlock), 0, ctx->stream, st_ptr,   \
                       cand.get(), cache, labels, vw, k, part_weights, capacity, powers, targets, cbucket.get(), \
                       src_replicas)
        switch (G) {
            case 1: GPART_BALANCE_PROPOSALS(1); break;
            case 2: GPART_BALANCE_PROPOSALS(2); break;
            default: GPART_BALANCE_PROPOSALS(4); break;
        }
#undef GPART_BALANCE_PROPOSALS
        GPART_LAUNCH_CHECK();
        profile_end(ctx, GPART_PROFILE_REBALANCE, 0.0);
        });
        run.phase(-1, [&] {
        hipLaunchKernelGGL(balance_scan_kernel, dim3(k), dim3(kBucketScanThreads), 0, ctx->stream, src_replicas, k,
                           src_counts, src_offsets, part_weights, capacity, false, src_thresholds);
        hipLaunchKernelGGL(balance_sources_kernel<W>, dim3(slices), dim3(kBlock), 0, ctx->stream, st_ptr,
                           cand.get(), cbucket.get(), dbucket.get(), targets, vw, src_counts, src_offsets,
                           src_thresholds, sums.get(), dst_replicas, width);
        if (!inline_columns)
            hipLaunchKernelGGL(column_scan_kernel, dim3(k), dim3(kBucketScanThreads), 0, ctx->stream, sums.get(),
                               slices, prefix.get());
        hipLaunchKernelGGL(balance_source_ranks_kernel<W>, dim3(slices), dim3(kBlock), 0, ctx->stream, st_ptr,
                           cand.get(), cbucket.get(), dbucket.get(), targets, vw, src_offsets, src_thresholds,
                           column_prefix, sums.get(), dst_replicas, width);
        hipLaunchKernelGGL(balance_scan_kernel, dim3(k), dim3(kBucketScanThreads), 0, ctx->stream, dst_replicas, k,
                           dst_counts, dst_offsets, part_weights, capacity, true, dst_thresholds);
        hipLaunchKernelGGL((balance_destinations_kernel<O, W>), dim3(slices), dim3(kBlock), 0, ctx->stream,
                           st_ptr, cand.get(), dbucket.get(), vw, off, dst_counts, dst_offsets, dst_thresholds,
                           sums.get(), light.get(), heavy.get(), flags);
        if (!inline_columns)
            hipLaunchKernelGGL(column_scan_kernel, dim3(k), dim3(kBucketScanThreads), 0, ctx->stream, sums.get(),
                               slices, prefix.get());
        hipLaunchKernelGGL((balance_destination_ranks_kernel<O, W>), dim3(slices), dim3(kBlock), 0, ctx->stream,
                           st_ptr, cand.get(), dbucket.get(), vw, off, dst_offsets, dst_thresholds, column_prefix,
                           sums.get(), light.get(), heavy.get(), flags);
        GPART_LAUNCH_CHECK();
        });
        apply(ctx, run, g, c, st_ptr, light.get(), heavy.get(), labels, targets, flags, k, cache, part_weights,
              nullptr, false);
        finish(ctx, run, g, st_ptr, part_weights, k, out);
        ctx->rebalance_counters_zero = true;
        // per candidate: id, label, weight, cache row, target and bucket writes
        profile_add_bytes(ctx, GPART_PROFILE_REBALANCE,
                          static_cast<double>(out->candidates) * (16.0 + sizeof(W) + sizeof(C) * k));
        // one bucket-histogram add per candidate (balance_source_histogram, :173-186)
        profile_add_units(ctx, GPART_PROFILE_REBALANCE, static_cast<double>(out->candidates));
    }
};

}  // namespace
}  // namespace gpart_hip

// Device scratch of gpart_cut_noalloc, one copy per GPU for the life of the
// process: the directed total, the per-workgroup partials and the tile rows.
// Host code never passes this name to a kernel; it passes the device address
// that cut_scratch_setup looks up for each device.
namespace gpart_hip {
namespace {
constexpr std::int64_t kCutScratchTiles = std::int64_t{1} << 20;  // CSRs up to 2^20 * 4096 entries
constexpr std::size_t kCutScratchTotalOffset = 0;
constexpr std::size_t kCutScratchPartialsOffset = 256;
constexpr std::size_t kCutScratchTilesOffset = kCutScratchPartialsOffset + 8 * kReduceMaxBlocks;
constexpr std::size_t kCutScratchBytes = kCutScratchTilesOffset + 4 * kCutScratchTiles;  // 4 MiB + 8.25 KiB
constexpr int kCutScratchDevices = 64;

std::mutex& cut_scratch_mutex() {
    static std::mutex mutex;
    return mutex;
}
void* cut_scratch_address[kCutScratchDevices] = {};
}  // namespace
}  // namespace gpart_hip

static __device__ __align__(256) unsigned char scratch[gpart_hip::kCutScratchBytes];

namespace gpart_hip {
// Called once per context with its device current (gpart_ctx_create); the first
// call on a device resolves that device's copy of the array.
void cut_scratch_setup(int device) {
    if (device < 0 || device >= kCutScratchDevices)
        throw Error(GPART_ERR_RUNTIME, "device ordinal beyond the no-allocation scratch table");
    std::lock_guard<std::mutex> lock(cut_scratch_mutex());
    if (cut_scratch_address[device]) return;
    void* address = nullptr;
    GPART_HIP_CHECK(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)));
    cut_scratch_address[device] = address;
}
}  // namespace gpart_hip

using namespace gpart_hip;

#define GPART_WIDTH_DISPATCH(width, call)                            \
    do {                                                             \
        if ((width) == GPART_WIDTH_STANDARD) {                        \
            using OpsT = Ops<std::int32_t, std::uint32_t>;           \
            call;                                                    \
        } else {                                                     \
            using OpsT = Ops<std::int64_t, std::uint64_t>;           \
            call;                                                    \
        }                                                            \
    } while (0)

// Width and cache entry type (4 or 8 bytes) of the refinement entry points.
#define GPART_CACHE_DISPATCH(width, bytes, call)                                                      \
    do {                                                                                              \
        if ((bytes) != 4 && (bytes) != 8) throw Error(GPART_ERR_INVALID_ARGUMENT, "cache entries are 4 or 8 bytes"); \
        if ((width) == GPART_WIDTH_STANDARD && (bytes) == 4) {                                        \
            using OpsT = Ops<std::int32_t, std::uint32_t, unsigned>;                                  \
            using CacheT = unsigned;                                                                  \
            call;                                                                                     \
        } else if ((width) == GPART_WIDTH_STANDARD) {                                                 \
            using OpsT = Ops<std::int32_t, std::uint32_t, unsigned long long>;                        \
            using CacheT = unsigned long long;                                                        \
            call;                                                                                     \
        } else if ((bytes) == 4) {                                                                    \
            using OpsT = Ops<std::int64_t, std::uint64_t, unsigned>;                                  \
            using CacheT = unsigned;                                                                  \
            call;                                                                                     \
        } else {                                                                                      \
            using OpsT = Ops<std::int64_t, std::uint64_t, unsigned long long>;                        \
            using CacheT = unsigned long long;                                                        \
            call;                                                                                     \
        }                                                                                             \
    } while (0)

extern "C" {

int gpart_refine_set_rebuild(gpart_ctx* ctx, int64_t min_entries, int64_t share) {
    return guarded([&] {
        if (!ctx) throw Error(GPART_ERR_INVALID_ARGUMENT, "null context");
        ctx->rebuild_min_entries = min_entries < 0 ? 0 : min_entries;
        ctx->rebuild_share = share;
    });
}

int gpart_cut(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* labels, uint64_t* cut) {
    return guarded([&] { GPART_WIDTH_DISPATCH(g->width, OpsT::cut(ctx, g, labels, cut)); });
}

int gpart_cut_noalloc(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* labels, uint64_t* cut) {
    return guarded([&] {
        if (!ctx || !g || !cut) throw Error(GPART_ERR_INVALID_ARGUMENT, "null context, graph or output");
        if (g->m > gpart_cut_noalloc_max_entries())
            throw Error(GPART_ERR_INVALID_ARGUMENT,
                        "gpart_cut_noalloc: " + std::to_string(g->m) + " entries exceed its limit of " +
                            std::to_string(gpart_cut_noalloc_max_entries()));
        void* base = nullptr;
        if (ctx->device >= 0 && ctx->device < kCutScratchDevices) {
            std::lock_guard<std::mutex> lock(cut_scratch_mutex());
            base = cut_scratch_address[ctx->device];
        }
        if (!base) throw Error(GPART_ERR_RUNTIME, "gpart_cut_noalloc: scratch of this device was not set up");
        auto* bytes = static_cast<unsigned char*>(base);
        auto* total = reinterpret_cast<unsigned long long*>(bytes + kCutScratchTotalOffset);
        auto* partials = reinterpret_cast<unsigned long long*>(bytes + kCutScratchPartialsOffset);
        auto* tile_row = reinterpret_cast<std::int32_t*>(bytes + kCutScratchTilesOffset);
        GPART_WIDTH_DISPATCH(g->width, OpsT::cut_with(ctx, g, labels, total, tile_row, partials, cut));
    });
}

int64_t gpart_cut_noalloc_max_entries(void) { return kCutScratchTiles * kTileEdges; }

int gpart_part_weights(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* labels, int k,
                       uint64_t* part_weights) {
    return guarded([&] {
        if (k < 1 || k > 32) throw Error(GPART_ERR_INVALID_ARGUMENT, "part count must be in [1, 32]");
        fill_bytes(ctx, part_weights, 0, static_cast<std::size_t>(k) * 8);
        if (g->n <= 0) return;
        const unsigned blocks = std::min<unsigned>(grid_for(g->n), 1024);
        auto* pw = reinterpret_cast<unsigned long long*>(part_weights);
        if (g->width == GPART_WIDTH_STANDARD)
            hipLaunchKernelGGL(part_weights_kernel<std::uint32_t>, dim3(blocks), dim3(kBlock), 0, ctx->stream, g->n,
                               labels, static_cast<const std::uint32_t*>(g->vertex_weights), k, pw);
        else
            hipLaunchKernelGGL(part_weights_kernel<std::uint64_t>, dim3(blocks), dim3(kBlock), 0, ctx->stream, g->n,
                               labels, static_cast<const std::uint64_t*>(g->vertex_weights), k, pw);
        GPART_LAUNCH_CHECK();
    });
}

int gpart_project(gpart_ctx* ctx, int64_t n, const int32_t* map, const int32_t* coarse_labels, int32_t* fine_labels) {
    return guarded([&] {
        if (n <= 0) return;
        hipLaunchKernelGGL(project_kernel, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, map, coarse_labels,
                           fine_labels);
        GPART_LAUNCH_CHECK();
    });
}

int gpart_refine_build_cache(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* classes,
                             const int32_t* labels, int k, void* cache, int cache_bytes) {
    return guarded([&] {
        GPART_CACHE_DISPATCH(g->width, cache_bytes,
                             OpsT::build_cache(ctx, g, classes, labels, k, static_cast<CacheT*>(cache)));
    });
}

int gpart_refine_lp_round(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* classes,
                          void* cache, int cache_bytes, int32_t* labels, int k, double filter, int32_t* locks,
                          int32_t* targets, int64_t* gains, int32_t* flags, uint64_t* part_weights,
                          gpart_refine_status* status) {
    return guarded([&] {
        if (k < 1 || k > 32) throw Error(GPART_ERR_INVALID_ARGUMENT, "part count must be in [1, 32]");
        if (g->n <= 0) throw Error(GPART_ERR_INVALID_ARGUMENT, "refinement round needs vertices");
        GPART_CACHE_DISPATCH(g->width, cache_bytes,
                             OpsT::lp_round(ctx, g, classes, static_cast<CacheT*>(cache), labels, k, filter, locks,
                                            targets, reinterpret_cast<long long*>(gains), flags,
                                            reinterpret_cast<unsigned long long*>(part_weights), status));
    });
}

int gpart_refine_set_balance_blocks(gpart_ctx* ctx, int64_t blocks) {
    return guarded([&] {
        if (blocks < 1 || blocks > (int64_t{1} << 20)) throw Error(GPART_ERR_INVALID_ARGUMENT, "block count out of range");
        ctx->balance_blocks = blocks;
    });
}

int gpart_refine_set_graphs(gpart_ctx* ctx, int enabled) {
    return guarded([&] {
        GPART_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->round_plans.reset();
        ctx->round_graphs = enabled != 0;
    });
}

int gpart_refine_graphs_enabled(gpart_ctx* ctx) { return ctx && ctx->round_graphs ? 1 : 0; }

int gpart_refine_graph_stats(gpart_ctx* ctx, int64_t* plans, int64_t* captures, int64_t* replays) {
    return guarded([&] {
        const RoundPlans* set = ctx->round_plans ? static_cast<const RoundPlans*>(ctx->round_plans.get()) : nullptr;
        if (plans) *plans = set ? static_cast<int64_t>(set->plans.size()) : 0;
        if (captures) *captures = set ? set->captures : 0;
        if (replays) *replays = set ? set->replays : 0;
    });
}

int gpart_refine_rebalance_round(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* classes,
                                 void* cache, int cache_bytes, int32_t* labels, int k, uint64_t* part_weights,
                                 uint64_t capacity, int32_t* targets, int32_t* flags, gpart_refine_status* status) {
    return guarded([&] {
        if (k < 1 || k > 32) throw Error(GPART_ERR_INVALID_ARGUMENT, "part count must be in [1, 32]");
        if (g->n <= 0) throw Error(GPART_ERR_INVALID_ARGUMENT, "refinement round needs vertices");
        GPART_CACHE_DISPATCH(g->width, cache_bytes,
                             OpsT::rebalance_round(ctx, g, classes, static_cast<CacheT*>(cache), labels, k,
                                                   reinterpret_cast<unsigned long long*>(part_weights), capacity,
                                                   targets, flags, status));
    });
}

int gpart_max_edge_weight(gpart_ctx* ctx, const gpart_csr_view* g, uint64_t* maximum) {
    return guarded([&] {
        Scratch<unsigned long long> out(ctx, 1);
        fill_bytes(ctx, out.get(), 0, 8);
        if (g->m > 0) {
            if (g->width == GPART_WIDTH_STANDARD)
                device_reduce<unsigned long long>(
                    ctx, WidenLoad<std::uint32_t>{static_cast<const std::uint32_t*>(g->edge_weights)}, g->m, MaxU64{},
                    out.get());
            else
                device_reduce<unsigned long long>(
                    ctx, WidenLoad<std::uint64_t>{static_cast<const std::uint64_t*>(g->edge_weights)}, g->m, MaxU64{},
                    out.get());
        }
        *maximum = read_scalar(ctx, out.get());
    });
}

int gpart_edge_weight_sum(gpart_ctx* ctx, const gpart_csr_view* g, uint64_t* checked, uint64_t* limited) {
    return guarded([&] {
        Scratch<unsigned long long> out(ctx, 2);
        fill_bytes(ctx, out.get(), 0, 16);
        if (g->m > 0) {
            if (g->width == GPART_WIDTH_STANDARD) {
                const WidenLoad<std::uint32_t> load{static_cast<const std::uint32_t*>(g->edge_weights)};
                device_reduce<unsigned long long>(ctx, load, g->m, CheckedSumU64{}, out.get());
                device_reduce<unsigned long long>(ctx, load, g->m, LimitedSumU64{}, out.get() + 1);
            } else {
                const WidenLoad<std::uint64_t> load{static_cast<const std::uint64_t*>(g->edge_weights)};
                device_reduce<unsigned long long>(ctx, load, g->m, CheckedSumU64{}, out.get());
                device_reduce<unsigned long long>(ctx, load, g->m, LimitedSumU64{}, out.get() + 1);
            }
        }
        unsigned long long h[2];
        read_back(ctx, h, out.get(), sizeof(h));
        *checked = h[0];
        *limited = h[1];
    });
}

int gpart_vertex_weight_stats(gpart_ctx* ctx, const gpart_csr_view* g, uint64_t* total, uint64_t* maximum,
                              uint64_t* minimum) {
    return guarded([&] {
        Scratch<unsigned long long> out(ctx, 3);
        fill_bytes(ctx, out.get(), 0, 24);
        if (g->n > 0) {
            if (g->width == GPART_WIDTH_STANDARD) {
                const WidenLoad<std::uint32_t> load{static_cast<const std::uint32_t*>(g->vertex_weights)};
                device_reduce<unsigned long long>(ctx, load, g->n, CheckedSumU64{}, out.get());
                device_reduce<unsigned long long>(ctx, load, g->n, MaxU64{}, out.get() + 1);
                device_reduce<unsigned long long>(ctx, load, g->n, MinU64{}, out.get() + 2);
            } else {
                const WidenLoad<std::uint64_t> load{static_cast<const std::uint64_t*>(g->vertex_weights)};
                device_reduce<unsigned long long>(ctx, load, g->n, CheckedSumU64{}, out.get());
                device_reduce<unsigned long long>(ctx, load, g->n, MaxU64{}, out.get() + 1);
                device_reduce<unsigned long long>(ctx, load, g->n, MinU64{}, out.get() + 2);
            }
        }
        unsigned long long h[3];
        read_back(ctx, h, out.get(), sizeof(h));
        *total = h[0];
        *maximum = h[1];
        *minimum = h[2];
    });
}

}  // extern "C"
