// coarsen.hip -- SCLP coarsening kernels for gfx950 (src/coarsen.cu of the reference).
//
// Per LP round every vertex needs, over its neighbour clusters t != cluster(v):
//   conn[t] = sum of w(v,u), u != v, cluster(u) == t
//   the lexicographic maximum of (conn[t], tie(v,t)) among role-admissible t
// plus current = conn[cluster(v)]. These are order-independent, so the kernels
// below choose their own thread mapping per degree class:
//   low    (deg <= 8)    8 lanes per vertex, duplicates merged with 8 shuffles
//   medium (deg <= 256)  one wave per vertex, 512-slot LDS open-addressing table
//   high   (deg > 256)   one workgroup per vertex, 4096-slot LDS table, or a
//                        per-hub table in global memory above 2048 entries
// and fuse the reference's decode step (sclp_decode_gain_kernel) into the
// reduction, writing proposal/gain/tie directly.
#include "internal.hpp"
#include "reduce.hpp"
#include "scan.hpp"
#include "hubs.hpp"

#include <algorithm>

namespace gpart_hip {
namespace {

constexpr int kLowDegree = 8;
constexpr int kMediumDegree = 256;
constexpr int kMediumSlots = 512;
constexpr int kHighLdsDegree = 1024;
constexpr int kHighSlots = 2 * kHighLdsDegree;
constexpr int kBlock = 256;

struct RoundParams {
    std::uint32_t role_salt, tie_salt, threshold;
    bool filter;
};

__device__ __forceinline__ bool is_mover(std::int32_t cluster, const RoundParams& r) {
    return mix32(static_cast<std::uint32_t>(cluster) ^ r.role_salt) < r.threshold;
}

__device__ __forceinline__ unsigned long long tie_key(std::int32_t v, std::int32_t t, std::uint32_t salt) {
    const auto hash = mix32(static_cast<std::uint32_t>(v) ^ mix32(static_cast<std::uint32_t>(t)) ^ salt);
    return (static_cast<unsigned long long>(hash) << 32) | (0xffffffffULL - static_cast<std::uint32_t>(t));
}

// sclp_better_candidate: affinity 0 never wins (src/coarsen.cu:224-229)
__device__ __forceinline__ bool better(unsigned long long a, unsigned long long at, unsigned long long b,
                                      unsigned long long bt) {
    return a > b || (a == b && a != 0 && at > bt);
}

template <int G>
__device__ __forceinline__ void group_best(unsigned long long& conn, unsigned long long& tie) {
#pragma unroll
    for (int m = G / 2; m > 0; m /= 2) {
        const auto oc = shfl_xor_u64(conn, m, G);
        const auto ot = shfl_xor_u64(tie, m, G);
        if (better(oc, ot, conn, tie)) {
            conn = oc;
            tie = ot;
        }
    }
}

// Candidate of one distinct neighbour cluster t with connection conn.
__device__ __forceinline__ void offer(std::int32_t v, std::int32_t t, std::int32_t source, unsigned long long conn,
                                      const RoundParams& r, unsigned long long& current,
                                      unsigned long long& best, unsigned long long& best_tie) {
    if (t == source) {
        current = conn > current ? conn : current;
        return;
    }
    if (conn == 0 || (r.filter && is_mover(t, r))) return;
    const auto tie = tie_key(v, t, r.tie_salt);
    if (better(conn, tie, best, best_tie)) {
        best = conn;
        best_tie = tie;
    }
}

// sclp_decode_gain_kernel (src/coarsen.cu:562-598)
__device__ __forceinline__ void decode(std::int32_t v, std::int32_t source, unsigned long long current,
                                       unsigned long long best, unsigned long long best_tie, const RoundParams& r,
                                       std::int32_t* proposals, unsigned long long* gains, std::uint32_t* ties) {
    const bool can_move = !r.filter || is_mover(source, r);
    std::int32_t p = kSclpInvalid;
    unsigned long long g = 0;
    std::uint32_t t32 = 0;
    if (can_move && best > current) {
        const auto target = static_cast<std::int32_t>(0xffffffffULL - (best_tie & 0xffffffffULL));
        if (target >= 0 && target != source) {
            p = target;
            g = best - current;
            t32 = static_cast<std::uint32_t>(best_tie >> 32);
        }
    }
    proposals[v] = p;
    gains[v] = g;
    ties[v] = t32;
}

// ---------------------------------------------------------------- classes

// Three-way stable split of vertex ids by degree class, as one scan that also
// carries the level's isolated-vertex count, the high rows' entries and the
// largest degree (prefix sums and a prefix maximum).
struct DegreeScan {
    long long c[3];
    long long isolated;
    long long high_edges;
    long long max_degree;
};
struct DegreeScanOp {
    __device__ __forceinline__ DegreeScan identity() const { return DegreeScan{{0, 0, 0}, 0, 0, 0}; }
    __device__ __forceinline__ DegreeScan operator()(const DegreeScan& a, const DegreeScan& b) const {
        return DegreeScan{{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]},
                          a.isolated + b.isolated,
                          a.high_edges + b.high_edges,
                          a.max_degree > b.max_degree ? a.max_degree : b.max_degree};
    }
};

template <typename O>
struct ClassLoad {
    const O* offsets;
    __device__ __forceinline__ DegreeScan operator()(std::int64_t v) const {
        const auto deg = static_cast<long long>(offsets[v + 1] - offsets[v]);
        DegreeScan r{{0, 0, 0}, deg == 0 ? 1 : 0, deg > kMediumDegree ? deg : 0, deg};
        r.c[deg <= kLowDegree ? 0 : deg <= kMediumDegree ? 1 : 2] = 1;
        return r;
    }
};

// counts: low, medium, high, high entries, largest degree, isolated vertices.
template <typename O>
struct ClassStore {
    const O* offsets;
    std::int32_t* low;
    std::int32_t* medium;
    std::int32_t* high;
    std::int64_t n;
    unsigned long long* counts;
    __device__ __forceinline__ void operator()(std::int64_t v, const DegreeScan& prefix,
                                               const DegreeScan& item) const {
        if (item.c[0]) low[prefix.c[0]] = static_cast<std::int32_t>(v);
        else if (item.c[1]) medium[prefix.c[1]] = static_cast<std::int32_t>(v);
        else high[prefix.c[2]] = static_cast<std::int32_t>(v);
        if (v == n - 1) {
            const DegreeScan total = DegreeScanOp{}(prefix, item);
            for (int j = 0; j < 3; ++j) counts[j] = static_cast<unsigned long long>(total.c[j]);
            counts[3] = static_cast<unsigned long long>(total.high_edges);
            counts[4] = static_cast<unsigned long long>(total.max_degree);
            counts[5] = static_cast<unsigned long long>(total.isolated);
        }
    }
};

// ---------------------------------------------------------------- affinity

template <typename O, typename W, bool Singleton>
__global__ void __launch_bounds__(kBlock) sclp_low_kernel(std::int64_t count, const std::int32_t* list, const O* offsets,
                                                          const std::int32_t* neighbors, const W* weights,
                                                          const std::int32_t* clusters, RoundParams r,
                                                          std::int32_t* proposals, unsigned long long* gains,
                                                          std::uint32_t* ties) {
    constexpr int G = kLowDegree;
    const std::int64_t group = (static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
    const int sub = threadIdx.x & (G - 1);
    if (group >= count) return;  // uniform per group
    const std::int32_t v = list[group];
    const std::int64_t begin = offsets[v];
    const int deg = static_cast<int>(offsets[v + 1] - begin);
    bool valid = sub < deg;
    std::int32_t t = -1;
    unsigned long long w = 0;
    if (valid) {
        const std::int32_t u = neighbors[begin + sub];
        valid = u != v;
        if (valid) {
            t = Singleton ? u : clusters[u];
            w = static_cast<unsigned long long>(weights[begin + sub]);
        }
    }
    unsigned long long conn = w;
    if constexpr (!Singleton) {
        conn = 0;
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const std::int32_t tj = __shfl(t, j, G);
            const auto wj = shfl_u64(w, j, G);
            if (valid && tj == t) conn += wj;
        }
    }
    const std::int32_t source = clusters[v];
    unsigned long long current = 0, best = 0, best_tie = 0;
    if (valid) offer(v, t, source, conn, r, current, best, best_tie);
    current = group_max_u64<G>(current);
    group_best<G>(best, best_tie);
    if (sub == 0) decode(v, source, current, best, best_tie, r, proposals, gains, ties);
}

template <typename O, typename W, bool Singleton>
__global__ void __launch_bounds__(kBlock) sclp_medium_kernel(std::int64_t count, const std::int32_t* list, const O* offsets,
                                                             const std::int32_t* neighbors, const W* weights,
                                                             const std::int32_t* clusters, RoundParams r,
                                                             std::int32_t* proposals, unsigned long long* gains,
                                                             std::uint32_t* ties) {
    constexpr int kWavesPerBlock = kBlock / kWave;
    __shared__ std::int32_t table_keys[Singleton ? 1 : kWavesPerBlock * kMediumSlots];
    __shared__ unsigned long long table_values[Singleton ? 1 : kWavesPerBlock * kMediumSlots];
    const int wave = threadIdx.x / kWave;
    const int lane = lane_id();
    const std::int64_t index = static_cast<std::int64_t>(blockIdx.x) * kWavesPerBlock + wave;
    const bool active = index < count;
    const std::int32_t v = active ? list[index] : 0;
    const std::int64_t begin = active ? static_cast<std::int64_t>(offsets[v]) : 0;
    const std::int64_t end = active ? static_cast<std::int64_t>(offsets[v + 1]) : 0;
    const std::int32_t source = active ? clusters[v] : 0;
    unsigned long long current = 0, best = 0, best_tie = 0;
    if constexpr (Singleton) {
        for (std::int64_t e = begin + lane; e < end; e += kWave) {
            const std::int32_t u = neighbors[e];
            if (u == v) continue;
            offer(v, u, source, static_cast<unsigned long long>(weights[e]), r, current, best, best_tie);
        }
    } else {
        std::int32_t* keys = table_keys + wave * kMediumSlots;
        unsigned long long* values = table_values + wave * kMediumSlots;
        for (int s = lane; s < kMediumSlots; s += kWave) {
            keys[s] = -1;
            values[s] = 0;
        }
        __syncthreads();
        for (std::int64_t e = begin + lane; e < end; e += kWave) {
            const std::int32_t u = neighbors[e];
            if (u == v) continue;
            const std::int32_t t = clusters[u];
            int slot = static_cast<int>(mix32(static_cast<std::uint32_t>(t))) & (kMediumSlots - 1);
            while (true) {
                const std::int32_t found = atomicCAS(keys + slot, -1, t);
                if (found == -1 || found == t) {
                    atomicAdd(values + slot, static_cast<unsigned long long>(weights[e]));
                    break;
                }
                slot = (slot + 1) & (kMediumSlots - 1);
            }
        }
        __syncthreads();
        if (active) {
            for (int s = lane; s < kMediumSlots; s += kWave) {
                const std::int32_t t = keys[s];
                if (t < 0) continue;
                offer(v, t, source, values[s], r, current, best, best_tie);
            }
        }
    }
    current = group_max_u64<kWave>(current);
    group_best<kWave>(best, best_tie);
    if (active && lane == 0) decode(v, source, current, best, best_tie, r, proposals, gains, ties);
}

// Folds each wave's (current, best, tie) into thread 0 of a Threads-wide workgroup.
template <int Threads = kBlock>
__device__ __forceinline__ void block_best(unsigned long long& current, unsigned long long& best,
                                           unsigned long long& best_tie) {
    __shared__ unsigned long long red_current[Threads / kWave], red_best[Threads / kWave], red_tie[Threads / kWave];
    current = group_max_u64<kWave>(current);
    group_best<kWave>(best, best_tie);
    const int wave = threadIdx.x / kWave;
    if (lane_id() == 0) {
        red_current[wave] = current;
        red_best[wave] = best;
        red_tie[wave] = best_tie;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < Threads / kWave; ++w) {
            current = red_current[w] > current ? red_current[w] : current;
            if (better(red_best[w], red_tie[w], best, best_tie)) {
                best = red_best[w];
                best_tie = red_tie[w];
            }
        }
    }
    __syncthreads();
}

// Workgroup per high vertex of degree <= kHighLdsDegree, with an LDS table of
// the smallest power of two >= 2 * degree slots; larger rows are hubs, done in
// chunks by the kernels below.
template <typename O, typename W, bool Singleton>
__global__ void __launch_bounds__(kBlock) sclp_high_kernel(std::int64_t count, const std::int32_t* list, const O* offsets,
                                                           const std::int32_t* neighbors, const W* weights,
                                                           const std::int32_t* clusters, RoundParams r,
                                                           std::int32_t* proposals, unsigned long long* gains,
                                                           std::uint32_t* ties) {
    __shared__ std::int32_t lds_keys[Singleton ? 1 : kHighSlots];
    __shared__ unsigned long long lds_values[Singleton ? 1 : kHighSlots];
    const std::int64_t index = blockIdx.x;
    if (index >= count) return;  // uniform per block
    const std::int32_t v = list[index];
    const std::int64_t begin = offsets[v];
    const std::int64_t end = offsets[v + 1];
    if (end - begin > kHighLdsDegree) return;  // a hub
    const std::int32_t source = clusters[v];
    unsigned long long current = 0, best = 0, best_tie = 0;
    if constexpr (Singleton) {
        for (std::int64_t e = begin + threadIdx.x; e < end; e += kBlock) {
            const std::int32_t u = neighbors[e];
            if (u == v) continue;
            offer(v, u, source, static_cast<unsigned long long>(weights[e]), r, current, best, best_tie);
        }
    } else {
        int slots = 1;
        while (slots < 2 * (end - begin)) slots <<= 1;
        const int slot_mask = slots - 1;
        for (int s = threadIdx.x; s < slots; s += kBlock) {
            lds_keys[s] = -1;
            lds_values[s] = 0;
        }
        __syncthreads();
        for (std::int64_t e = begin + threadIdx.x; e < end; e += kBlock) {
            const std::int32_t u = neighbors[e];
            if (u == v) continue;
            const std::int32_t t = clusters[u];
            int slot = static_cast<int>(mix32(static_cast<std::uint32_t>(t))) & slot_mask;
            while (true) {
                const std::int32_t found = atomicCAS(lds_keys + slot, -1, t);
                if (found == -1 || found == t) {
                    atomicAdd(lds_values + slot, static_cast<unsigned long long>(weights[e]));
                    break;
                }
                slot = (slot + 1) & slot_mask;
            }
        }
        __syncthreads();
        for (int s = threadIdx.x; s < slots; s += kBlock) {
            const std::int32_t t = lds_keys[s];
            if (t < 0) continue;
            offer(v, t, source, lds_values[s], r, current, best, best_tie);
        }
    }
    block_best(current, best, best_tie);
    if (threadIdx.x == 0) decode(v, source, current, best, best_tie, r, proposals, gains, ties);
}

// Inserts (t, w) into an open-addressing table whose keys never change once set:
// a key already present costs a load and an add, no compare-and-swap.
template <typename Index>
__device__ __forceinline__ void table_add(std::int32_t* keys, unsigned long long* values, Index slot, Index mask,
                                          std::int32_t t, unsigned long long w) {
    while (true) {
        std::int32_t found = __hip_atomic_load(keys + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (found == -1) found = atomicCAS(keys + slot, -1, t);
        if (found == -1 || found == t) {
            atomicAdd(values + slot, w);
            return;
        }
        slot = (slot + 1) & mask;
    }
}

// Hubs (degree > kHighLdsDegree), one chunk of kHubChunk entries per workgroup
// pass. Singleton rounds offer every entry directly and leave a partial
// (current, best, tie) per chunk. Otherwise each half chunk is summed per
// cluster in an LDS table first, whose entries then go into the hub's global
// table (table_offset[i] .. table_offset[i + 1], a power of two >= 2 * degree):
// a cluster that holds many of the hub's neighbours costs one global add per
// half chunk instead of one per entry.
constexpr int kHubStage = kHubChunk / 2;
constexpr int kHubStageSlots = 2 * kHubStage;

template <typename O, typename W, bool Singleton>
__global__ void __launch_bounds__(kBlock) sclp_hub_insert_kernel(const std::int64_t* chunk_total,
                                                                 const std::int64_t* chunk_start,
                                                                 const std::int32_t* chunk_hub,
                                                                 const std::int32_t* list, const O* offsets,
                                                                 const std::int32_t* neighbors, const W* weights,
                                                                 const std::int32_t* clusters, RoundParams r,
                                                                 const std::int64_t* table_offset,
                                                                 std::int32_t* global_keys,
                                                                 unsigned long long* global_values,
                                                                 unsigned long long* partial_current,
                                                                 unsigned long long* partial_best,
                                                                 unsigned long long* partial_tie) {
    __shared__ std::int32_t stage_keys[Singleton ? 1 : kHubStageSlots];
    __shared__ unsigned long long stage_values[Singleton ? 1 : kHubStageSlots];
    const std::int64_t chunks = *chunk_total;
    for (std::int64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        std::int32_t i, v;
        std::int64_t begin, end;
        hub_chunk_range(offsets, list, chunk_start, chunk_hub, c, &i, &v, &begin, &end);
        if constexpr (Singleton) {
            const std::int32_t source = clusters[v];
            unsigned long long current = 0, best = 0, best_tie = 0;
            for (std::int64_t e = begin + threadIdx.x; e < end; e += kBlock) {
                const std::int32_t u = neighbors[e];
                if (u == v) continue;
                offer(v, u, source, static_cast<unsigned long long>(weights[e]), r, current, best, best_tie);
            }
            block_best(current, best, best_tie);
            if (threadIdx.x == 0) {
                partial_current[c] = current;
                partial_best[c] = best;
                partial_tie[c] = best_tie;
            }
        } else {
            const std::int64_t off = table_offset[i];
            const std::int64_t global_mask = table_offset[i + 1] - off - 1;
            std::int32_t* keys = global_keys + off;
            unsigned long long* values = global_values + off;
            for (std::int64_t stage = begin; stage < end; stage += kHubStage) {
                const std::int64_t stop = stage + kHubStage < end ? stage + kHubStage : end;
                for (int s = threadIdx.x; s < kHubStageSlots; s += kBlock) {
                    stage_keys[s] = -1;
                    stage_values[s] = 0;
                }
                __syncthreads();
                for (std::int64_t e = stage + threadIdx.x; e < stop; e += kBlock) {
                    const std::int32_t u = neighbors[e];
                    if (u == v) continue;
                    const std::int32_t t = clusters[u];
                    int slot = static_cast<int>(mix32(static_cast<std::uint32_t>(t))) & (kHubStageSlots - 1);
                    while (true) {
                        const std::int32_t found = atomicCAS(stage_keys + slot, -1, t);
                        if (found == -1 || found == t) {
                            atomicAdd(stage_values + slot, static_cast<unsigned long long>(weights[e]));
                            break;
                        }
                        slot = (slot + 1) & (kHubStageSlots - 1);
                    }
                }
                __syncthreads();
                for (int s = threadIdx.x; s < kHubStageSlots; s += kBlock) {
                    const std::int32_t t = stage_keys[s];
                    if (t < 0) continue;
                    const std::int64_t slot =
                        static_cast<std::int64_t>(mix32(static_cast<std::uint32_t>(t))) & global_mask;
                    table_add<std::int64_t>(keys, values, slot, global_mask, t, stage_values[s]);
                }
                __syncthreads();
            }
        }
    }
}

// Partial (current, best, tie) of each hub chunk over a strided share of the
// hub's table: chunk j of a hub with h chunks reads slots j*256 + t + q*256*h.
template <typename O>
__global__ void __launch_bounds__(kBlock) sclp_hub_scan_kernel(const std::int64_t* chunk_total,
                                                               const std::int64_t* chunk_start,
                                                               const std::int32_t* chunk_hub,
                                                               const std::int32_t* list, const std::int32_t* clusters,
                                                               RoundParams r, const std::int64_t* table_offset,
                                                               const std::int32_t* global_keys,
                                                               const unsigned long long* global_values,
                                                               unsigned long long* partial_current,
                                                               unsigned long long* partial_best,
                                                               unsigned long long* partial_tie) {
    const std::int64_t chunks = *chunk_total;
    for (std::int64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        const std::int32_t i = chunk_hub[c];
        const std::int32_t v = list[i];
        const std::int32_t source = clusters[v];
        const std::int64_t first = chunk_start[i];
        const std::int64_t stride = (chunk_start[i + 1] - first) * kBlock;
        const std::int64_t off = table_offset[i];
        const std::int64_t slots = table_offset[i + 1] - off;
        unsigned long long current = 0, best = 0, best_tie = 0;
        for (std::int64_t s = (c - first) * kBlock + threadIdx.x; s < slots; s += stride) {
            const std::int32_t t = global_keys[off + s];
            if (t < 0) continue;
            offer(v, t, source, global_values[off + s], r, current, best, best_tie);
        }
        block_best(current, best, best_tie);
        if (threadIdx.x == 0) {
            partial_current[c] = current;
            partial_best[c] = best;
            partial_tie[c] = best_tie;
        }
    }
}

// One thread per high vertex: folds the chunk partials of a hub and decodes.
__global__ void sclp_hub_decode_kernel(std::int64_t count, const std::int64_t* chunk_start, const std::int32_t* list,
                                       const std::int32_t* clusters, RoundParams r,
                                       const unsigned long long* partial_current,
                                       const unsigned long long* partial_best, const unsigned long long* partial_tie,
                                       std::int32_t* proposals, unsigned long long* gains, std::uint32_t* ties) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const std::int64_t first = chunk_start[i], last = chunk_start[i + 1];
    if (first == last) return;
    unsigned long long current = 0, best = 0, best_tie = 0;
    for (std::int64_t c = first; c < last; ++c) {
        current = partial_current[c] > current ? partial_current[c] : current;
        if (better(partial_best[c], partial_tie[c], best, best_tie)) {
            best = partial_best[c];
            best_tie = partial_tie[c];
        }
    }
    const std::int32_t v = list[i];
    decode(v, clusters[v], current, best, best_tie, r, proposals, gains, ties);
}

// Rows of kHighLdsDegree < degree <= kPartMaxDegree outside singleton rounds:
// work item (v, p) of P = 2^j >= degree / kPartKeys items sums, in an LDS table,
// the clusters t of v's neighbours with (mix32(t) & (P - 1)) == p, and leaves a
// partial (current, best, tie). mix32 is a bijection, so the clusters of a row
// split evenly; a pass that still meets more than kPartCap distinct clusters is
// redone as 2, 4, ... finer passes (a finer split always separates distinct keys).
// No global table, fill or atomic: each item reads the row once per pass.
constexpr int kPartThreads = 512;
constexpr int kPartSlots = 8192;
constexpr std::int64_t kPartKeys = 4096;
constexpr std::int64_t kPartMaxDegree = 524288;
constexpr int kPartCap = kPartSlots - kPartThreads - 1;  // keys held + one in flight per thread < slots
constexpr int kPartLaunchCap = 1024;
constexpr int kPartUnroll = 4;

template <typename O>
struct PartLoad {
    const std::int32_t* high;
    const O* offsets;
    std::int64_t count;
    __device__ __forceinline__ std::int64_t operator()(std::int64_t i) const {
        if (i >= count) return 0;
        const std::int32_t v = high[i];
        const std::int64_t d = static_cast<std::int64_t>(offsets[v + 1] - offsets[v]);
        if (d <= kHighLdsDegree || d > kPartMaxDegree) return 0;
        std::int64_t parts = 1;
        while (parts * kPartKeys < d) parts <<= 1;
        return parts;
    }
};

// Rows of more than one item (kPartKeys < degree <= kPartMaxDegree) are grouped
// once per round, so that item p reads only its own group of the row instead of
// the whole row with the cluster gather of every entry. The row's region of
// entries starts at the exclusive prefix of SplitLoad over the high list.
//
// The split itself runs as P chunk items per row (split item j = chunk p of row
// i, entries [begin + p * len, ...) with len = ceil(degree / P)), so no workgroup
// walks a whole row: sclp_split_count_kernel stores each entry's cluster in row
// order (-1 for a self loop) and counts, per chunk, the entries of each group
// q = mix32(cluster) & (P - 1); sclp_split_offsets_kernel turns the counts into
// offsets and sclp_split_scatter_kernel places chunk p's entries of group q
// after those of the groups before q and of the chunks before p. Ranks within a chunk come from one LDS atomic per entry (measured faster
// than a ballot per group and wave, from 2 groups up); the LDS tables that read a
// group sum exactly, in any order.
constexpr int kMaxParts = static_cast<int>(kPartMaxDegree / kPartKeys);

template <typename O>
struct SplitLoad {
    const std::int32_t* high;
    const O* offsets;
    std::int64_t count;
    __device__ __forceinline__ std::int64_t operator()(std::int64_t i) const {
        if (i >= count) return 0;
        const std::int32_t v = high[i];
        const std::int64_t d = static_cast<std::int64_t>(offsets[v + 1] - offsets[v]);
        return d > kPartKeys && d <= kPartMaxDegree ? d : 0;
    }
};

// Chunk items of a split row: its item count; 0 for rows of one item.
template <typename O>
struct SplitItemLoad {
    PartLoad<O> parts;
    __device__ __forceinline__ std::int64_t operator()(std::int64_t i) const {
        const std::int64_t p = parts(i);
        return p > 1 ? p : 0;
    }
};

struct SplitChunk {
    std::int32_t i, v, p, parts;
    std::int64_t row_begin, begin, end;
};

template <typename O>
__device__ __forceinline__ SplitChunk split_chunk(std::int64_t j, const std::int32_t* split_row,
                                                  const std::int64_t* split_start, const std::int32_t* list,
                                                  const O* offsets) {
    SplitChunk c;
    c.i = split_row[j];
    const std::int64_t first = split_start[c.i];
    c.parts = static_cast<std::int32_t>(split_start[c.i + 1] - first);
    c.p = static_cast<std::int32_t>(j - first);
    c.v = list[c.i];
    c.row_begin = static_cast<std::int64_t>(offsets[c.v]);
    const std::int64_t d = static_cast<std::int64_t>(offsets[c.v + 1]) - c.row_begin;
    const std::int64_t len = (d + c.parts - 1) / c.parts;
    c.begin = c.row_begin + (c.p * len < d ? c.p * len : d);
    c.end = c.row_begin + ((c.p + 1) * len < d ? (c.p + 1) * len : d);
    return c;
}

template <typename O>
__global__ void __launch_bounds__(kPartThreads) sclp_split_count_kernel(
    const std::int64_t* split_total, const std::int32_t* split_row, const std::int64_t* split_start,
    const std::int32_t* list, const O* offsets, const std::int32_t* neighbors, const std::int32_t* clusters,
    const std::int64_t* region, std::int32_t* row_clusters, std::uint32_t* chunk_counts) {
    __shared__ unsigned int counts[kMaxParts];
    const std::int64_t total = *split_total;
    for (std::int64_t j = blockIdx.x; j < total; j += gridDim.x) {
        const SplitChunk c = split_chunk(j, split_row, split_start, list, offsets);
        const std::int64_t base = region[c.i] - c.row_begin;
        const auto pmask = static_cast<std::uint32_t>(c.parts - 1);
        if (threadIdx.x < kMaxParts) counts[threadIdx.x] = 0;
        __syncthreads();
        for (std::int64_t e0 = c.begin; e0 < c.end; e0 += kPartThreads * kPartUnroll) {  // same trip count in every lane
            std::int32_t t[kPartUnroll];
#pragma unroll
            for (int k = 0; k < kPartUnroll; ++k) {
                const std::int64_t e = e0 + threadIdx.x + k * kPartThreads;
                t[k] = -1;
                if (e < c.end) {
                    const std::int32_t u = neighbors[e];
                    if (u != c.v) t[k] = u;
                }
            }
#pragma unroll
            for (int k = 0; k < kPartUnroll; ++k) {
                const std::int64_t e = e0 + threadIdx.x + k * kPartThreads;
                if (t[k] >= 0) t[k] = clusters[t[k]];
                if (e < c.end) row_clusters[base + e] = t[k];
            }
#pragma unroll
            for (int k = 0; k < kPartUnroll; ++k) {
                const int q = t[k] >= 0 ? static_cast<int>(mix32(static_cast<std::uint32_t>(t[k])) & pmask) : -1;
                if (q >= 0) atomicAdd(&counts[q], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x < c.parts) chunk_counts[j * kMaxParts + threadIdx.x] = counts[threadIdx.x];
        __syncthreads();  // counts is reused by the next chunk
    }
}

// Once per split row (in the workgroup that meets its chunk 0): each group's
// size over the row's chunks, the groups' places in the row's region
// (group_begin / group_end) and, in place of the chunk counts, the offset in the
// region at which each chunk's entries of each group start (the groups before
// it, then the same group's entries of the chunks before). Thread g walks the
// chunks of group g: P * P loads per row, where every chunk item redoing the sums
// would take P * P per item.
__global__ void __launch_bounds__(kPartThreads) sclp_split_offsets_kernel(
    const std::int64_t* split_total, const std::int32_t* split_row, const std::int64_t* split_start,
    const std::int64_t* region, const std::int64_t* item_start, std::uint32_t* chunk_counts,
    std::int64_t* group_begin, std::int64_t* group_end) {
    __shared__ unsigned int sizes[kMaxParts];
    const std::int64_t total = *split_total;
    for (std::int64_t j = blockIdx.x; j < total; j += gridDim.x) {
        const std::int32_t i = split_row[j];
        const std::int64_t first = split_start[i];
        if (j != first) continue;  // the same j in every thread of the workgroup
        const int parts = static_cast<int>(split_start[i + 1] - first);
        std::uint32_t* column = chunk_counts + first * kMaxParts + threadIdx.x;  // group g's count in chunk 0
        if (static_cast<int>(threadIdx.x) < parts) {
            unsigned int size = 0;
            for (int p = 0; p < parts; ++p) size += column[static_cast<std::int64_t>(p) * kMaxParts];
            sizes[threadIdx.x] = size;
        }
        __syncthreads();
        if (static_cast<int>(threadIdx.x) < parts) {
            unsigned int at = 0;
            for (int g = 0; g < static_cast<int>(threadIdx.x); ++g) at += sizes[g];
            const std::int64_t base = region[i];
            group_begin[item_start[i] + threadIdx.x] = base + at;
            group_end[item_start[i] + threadIdx.x] = base + at + sizes[threadIdx.x];
            // eight chunks' counts loaded before any is overwritten, so the loads overlap
            constexpr int kBatch = 8;
            for (int p0 = 0; p0 < parts; p0 += kBatch) {
                unsigned int n[kBatch];
#pragma unroll
                for (int q = 0; q < kBatch; ++q)
                    n[q] = p0 + q < parts ? column[static_cast<std::int64_t>(p0 + q) * kMaxParts] : 0;
#pragma unroll
                for (int q = 0; q < kBatch; ++q) {
                    if (p0 + q < parts) column[static_cast<std::int64_t>(p0 + q) * kMaxParts] = at;
                    at += n[q];
                }
            }
        }
        __syncthreads();  // sizes is reused by the next row
    }
}

template <typename O, typename W>
__global__ void __launch_bounds__(kPartThreads) sclp_split_scatter_kernel(
    const std::int64_t* split_total, const std::int32_t* split_row, const std::int64_t* split_start,
    const std::int32_t* list, const O* offsets, const W* weights, const std::int64_t* region,
    const std::uint32_t* chunk_offsets, const std::int32_t* row_clusters, std::int32_t* group_keys,
    W* group_weights) {
    __shared__ unsigned long long cursor[kMaxParts];
    const std::int64_t total = *split_total;
    for (std::int64_t j = blockIdx.x; j < total; j += gridDim.x) {
        const SplitChunk c = split_chunk(j, split_row, split_start, list, offsets);
        const std::int64_t base = region[c.i];
        const auto pmask = static_cast<std::uint32_t>(c.parts - 1);
        if (static_cast<int>(threadIdx.x) < c.parts)
            cursor[threadIdx.x] = static_cast<unsigned long long>(base + chunk_offsets[j * kMaxParts + threadIdx.x]);
        __syncthreads();
        for (std::int64_t e0 = c.begin; e0 < c.end; e0 += kPartThreads) {
            const std::int64_t e = e0 + threadIdx.x;
            int q = -1;
            std::int32_t t = -1;
            if (e < c.end) {
                t = row_clusters[base + (e - c.row_begin)];
                if (t >= 0) q = static_cast<int>(mix32(static_cast<std::uint32_t>(t)) & pmask);
            }
            if (q >= 0) {
                const auto pos = static_cast<std::int64_t>(atomicAdd(&cursor[q], 1ULL));
                group_keys[pos] = t;
                group_weights[pos] = weights[e];
            }
        }
        __syncthreads();  // cursor is reused by the next chunk
    }
}

// Items of one-item rows read the row and gather its clusters; items of split
// rows read their group. Each new key is counted once per wave (a ballot), so at
// most kPartThreads insertions are uncounted at any time and a table holding
// fewer than kPartCap counted keys always has a free slot.
template <typename O, typename W, typename V>
__global__ void __launch_bounds__(kPartThreads) sclp_part_kernel(const std::int64_t* item_total,
                                                                 const std::int64_t* item_start,
                                                                 const std::int32_t* item_vertex,
                                                                 const std::int32_t* list, const O* offsets,
                                                                 const std::int32_t* neighbors, const W* weights,
                                                                 const std::int32_t* clusters, RoundParams r,
                                                                 const std::int32_t* group_keys,
                                                                 const W* group_weights,
                                                                 const std::int64_t* group_begin,
                                                                 const std::int64_t* group_end,
                                                                 unsigned long long* partial_current,
                                                                 unsigned long long* partial_best,
                                                                 unsigned long long* partial_tie) {
    __shared__ std::int32_t keys[kPartSlots];
    __shared__ V values[kPartSlots];
    __shared__ int held;
    const std::int64_t items = *item_total;
    const int lane = lane_id();
    for (std::int64_t w = blockIdx.x; w < items; w += gridDim.x) {
        const std::int32_t i = item_vertex[w];
        const std::int32_t v = list[i];
        const std::int64_t first = item_start[i];
        const auto parts = static_cast<std::uint32_t>(item_start[i + 1] - first);
        const auto p = static_cast<std::uint32_t>(w - first);
        const bool grouped = parts > 1;
        const std::int64_t begin = grouped ? group_begin[w] : static_cast<std::int64_t>(offsets[v]);
        const std::int64_t end = grouped ? group_end[w] : static_cast<std::int64_t>(offsets[v + 1]);
        const std::int32_t source = clusters[v];
        int slots = kPartSlots;
        if (parts == 1) {  // 2 * degree slots (at most kPartSlots) for a single pass
            slots = 1;
            while (slots < 2 * (end - begin) && slots < kPartSlots) slots <<= 1;
        }
        const int slot_mask = slots - 1;
        unsigned long long current = 0, best = 0, best_tie = 0;
        for (std::uint32_t split = 1;; split <<= 1) {
            current = best = best_tie = 0;
            bool overflow = false;
            const std::uint32_t mask = parts * split - 1;
            for (std::uint32_t q = 0; q < split; ++q) {
                for (int s = threadIdx.x; s < slots; s += kPartThreads) {
                    keys[s] = -1;
                    values[s] = 0;
                }
                if (threadIdx.x == 0) held = 0;
                __syncthreads();
                const std::uint32_t select = p + parts * q;
                // kPartUnroll entries per lane and trip: their loads are issued together
                for (std::int64_t e0 = begin; e0 < end; e0 += kPartThreads * kPartUnroll) {  // same trip count in every lane
                    std::int32_t t[kPartUnroll];
                    V wt[kPartUnroll];
#pragma unroll
                    for (int k = 0; k < kPartUnroll; ++k) {
                        const std::int64_t e = e0 + threadIdx.x + k * kPartThreads;
                        t[k] = -1;
                        wt[k] = 0;
                        if (e < end) {
                            if (grouped) {
                                t[k] = group_keys[e];
                                wt[k] = static_cast<V>(group_weights[e]);
                            } else {
                                const std::int32_t u = neighbors[e];
                                if (u != v) t[k] = u;
                                wt[k] = static_cast<V>(weights[e]);
                            }
                        }
                    }
                    if (!grouped) {
#pragma unroll
                        for (int k = 0; k < kPartUnroll; ++k)
                            if (t[k] >= 0) t[k] = clusters[t[k]];
                    }
#pragma unroll
                    for (int k = 0; k < kPartUnroll; ++k) {
                        bool fresh = false;
                        const std::uint32_t h = mix32(static_cast<std::uint32_t>(t[k]));
                        // a full table: this pass is redone finer
                        if (t[k] >= 0 && (h & mask) == select &&
                            __hip_atomic_load(&held, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < kPartCap) {
                            int slot = static_cast<int>(mix32(h ^ 0x9e3779b9U)) & slot_mask;
                            while (true) {
                                const std::int32_t found = atomicCAS(keys + slot, -1, t[k]);
                                if (found == -1 || found == t[k]) {
                                    fresh = found == -1;
                                    atomicAdd(values + slot, wt[k]);
                                    break;
                                }
                                slot = (slot + 1) & slot_mask;
                            }
                        }
                        // counted before the next insertion: at most one uncounted key per lane
                        const std::uint64_t added = __ballot(fresh);
                        if (lane == 0 && added) atomicAdd(&held, static_cast<int>(__popcll(added)));
                    }
                }
                __syncthreads();
                overflow = held >= kPartCap;
                if (!overflow)
                    for (int s = threadIdx.x; s < slots; s += kPartThreads) {
                        const std::int32_t t = keys[s];
                        if (t < 0) continue;
                        offer(v, t, source, static_cast<unsigned long long>(values[s]), r, current, best, best_tie);
                    }
                __syncthreads();
                if (overflow) break;
            }
            if (!overflow) break;
        }
        block_best<kPartThreads>(current, best, best_tie);
        if (threadIdx.x == 0) {
            partial_current[w] = current;
            partial_best[w] = best;
            partial_tie[w] = best_tie;
        }
    }
}

template <typename O>
__global__ void hub_table_size_kernel(std::int64_t count, const std::int32_t* list, const O* offsets,
                                      std::int64_t min_degree, std::int64_t* sizes) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i > count) return;
    std::int64_t size = 0;
    if (i < count) {
        const std::int32_t v = list[i];
        const std::int64_t deg = offsets[v + 1] - offsets[v];
        if (deg > min_degree) {
            size = 1;
            while (size < 2 * deg) size <<= 1;
        }
    }
    sizes[i] = size;
}

// ---------------------------------------------------------------- admission

// The vertices with a valid proposal (or favourite), compacted in id order
// straight from the proposal array.
struct ValidLoad {
    const std::int32_t* targets;
    __device__ __forceinline__ std::int64_t operator()(std::int64_t v) const {
        return targets[v] != kSclpInvalid ? 1 : 0;
    }
};
struct IdStore {
    std::int32_t* out;
    std::int64_t n;
    std::int64_t* count;
    __device__ __forceinline__ void operator()(std::int64_t v, std::int64_t prefix, std::int64_t item) const {
        if (item) out[prefix] = static_cast<std::int32_t>(v);
        if (v == n - 1) *count = prefix + item;
    }
};
void compact_valid(gpart_ctx* ctx, std::int64_t n, const std::int32_t* targets, std::int32_t* out,
                   std::int64_t* d_count) {
    if (n <= 0) {
        fill_bytes(ctx, d_count, 0, sizeof(std::int64_t));
        return;
    }
    device_scan<std::int64_t, SumOp<std::int64_t>>(ctx, ValidLoad{targets}, IdStore{out, n, d_count}, n,
                                                   SumOp<std::int64_t>{});
}

// Key components of the admission order (target asc, gain desc, tie desc, id asc).
enum KeyStage { kTieOnly = 0, kGainTie = 1, kGainOnly = 2, kTargetOnly = 3, kAll = 4 };

__global__ void admission_key_kernel(std::int64_t p, const std::int32_t* order, const std::int32_t* targets,
                                     const unsigned long long* gains, const std::uint32_t* ties, int stage,
                                     int gain_bits, std::uint64_t* keys) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= p) return;
    const std::int32_t v = order[i];
    const std::uint64_t gmask = gain_bits >= 64 ? ~0ULL : ((1ULL << gain_bits) - 1);
    const std::uint64_t rgain = gmask - (gains[v] & gmask);  // descending gain
    const std::uint64_t rtie = static_cast<std::uint32_t>(~ties[v]);
    const std::uint64_t target = static_cast<std::uint32_t>(targets[v]);
    std::uint64_t key = 0;
    switch (stage) {
        case kTieOnly: key = rtie; break;
        case kGainTie: key = (rgain << 32) | rtie; break;
        case kGainOnly: key = rgain; break;
        case kTargetOnly: key = target; break;
        default: key = (((target << gain_bits) | rgain) << 32) | rtie; break;
    }
    keys[i] = key;
}

// zero (optional): words that the pass clears for the kernels after it.
__global__ void gather_order_kernel(std::int64_t p, const std::int32_t* order, const std::int32_t* targets,
                                    const std::uint64_t* vertex_weights, std::int32_t* ordered_targets,
                                    std::uint64_t* ordered_weights, unsigned long long* zero, int zero_words) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (zero && i < zero_words) zero[i] = 0;
    if (i >= p) return;
    const std::int32_t v = order[i];
    ordered_targets[i] = targets[v];
    ordered_weights[i] = vertex_weights[v];
}

struct SegWeightLoad {
    const std::int32_t* targets;
    const std::uint64_t* weights;
    __device__ __forceinline__ SegValue operator()(std::int64_t i) const {
        return SegValue{weights[i], static_cast<unsigned>(i == 0 || targets[i] != targets[i - 1])};
    }
};
struct SegOneLoad {
    const std::int32_t* targets;
    __device__ __forceinline__ SegValue operator()(std::int64_t i) const {
        return SegValue{1, static_cast<unsigned>(i == 0 || targets[i] != targets[i - 1])};
    }
};
// Exclusive prefix per item, and the total after the last item.
struct ExclusiveTotalStore {
    std::int64_t* out;
    std::int64_t n;
    std::int64_t* total;
    __device__ __forceinline__ void operator()(std::int64_t i, std::int64_t prefix, std::int64_t item) const {
        out[i] = prefix;
        if (i == n - 1) *total = prefix + item;
    }
};
struct SegInclusiveStore {
    std::uint64_t* out;
    __device__ __forceinline__ void operator()(std::int64_t i, const SegValue& prefix, const SegValue& item) const {
        out[i] = SegOp{}(prefix, item).value;
    }
};

// sclp_commit_kernel (src/coarsen.cu:652-680), split into a decision pass that
// reads the pre-commit cluster weights and an apply pass with atomics.
__global__ void admit_decide_kernel(std::int64_t p, const std::int32_t* ordered_targets,
                                    const std::uint64_t* prefix, const unsigned long long* cluster_weights,
                                    std::uint64_t capacity, std::int32_t* accept) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= p) return;
    const std::uint64_t base = cluster_weights[ordered_targets[i]];
    accept[i] = base <= capacity && prefix[i] <= capacity - base;
}

// sclp_commit_kernel (:652-680): totals[0] counts accepted moves, totals[1] sums
// their gains (the diagnostics' predicted gain); one atomic per workgroup each.
__global__ void __launch_bounds__(kBlock) admit_apply_kernel(std::int64_t p, const std::int32_t* order,
                                                             const std::int32_t* ordered_targets,
                                                             const std::uint64_t* ordered_weights,
                                                             const std::int32_t* accept,
                                                             const unsigned long long* gains,
                                                             std::int32_t* clusters,
                                                             unsigned long long* cluster_weights,
                                                             unsigned long long* totals) {
    __shared__ unsigned long long partial[2][kBlock / kWave];
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool ok = false;
    unsigned long long gain = 0;
    if (i < p && accept[i]) {
        ok = true;
        const std::int32_t v = order[i];
        const std::int32_t t = ordered_targets[i];
        const unsigned long long w = ordered_weights[i];
        gain = gains[v];
        atomicAdd(cluster_weights + clusters[v], 0ULL - w);
        atomicAdd(cluster_weights + t, w);
        clusters[v] = t;
    }
    const unsigned long long c = static_cast<unsigned long long>(__popcll(__ballot(ok)));
    gain = group_sum_u64<kWave>(gain);
    if (lane_id() == 0) {
        partial[0][threadIdx.x / kWave] = c;
        partial[1][threadIdx.x / kWave] = gain;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        unsigned long long total = 0;
        for (int w = 0; w < kBlock / kWave; ++w) total += partial[threadIdx.x][w];
        if (total) atomicAdd(totals + threadIdx.x, total);
    }
}

// ---------------------------------------------------------------- isolated / two-hop

template <typename O>
__global__ void zero_degree_flags_kernel(std::int64_t n, const O* offsets, std::int32_t* flags) {
    const std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v < n) flags[v] = offsets[v] == offsets[v + 1];
}

struct GatherWeightLoad {
    const std::int32_t* list;
    const std::uint64_t* weights;
    __device__ __forceinline__ std::uint64_t operator()(std::int64_t i) const { return weights[list[i]]; }
};

// isolated_representative_kernel / isolated_commit_kernel (src/coarsen.cu:741-773)
// The representative of a bin is the least id among the vertices lying wholly in
// it (the reference's atomicMin). The list is in id order and the prefix does not
// decrease along it (after any leading zero-weight vertices, which lie in no bin
// wholly), so a bin's whole vertices form one run of the list and the least id is
// the run's first: it alone writes, without atomics (a run of thousands of
// vertices on one address serialised the atomics).
__device__ __forceinline__ bool whole_bin(std::uint64_t prefix, std::uint64_t weight, std::uint64_t capacity,
                                          std::uint64_t* bin) {
    const std::uint64_t first = (prefix - weight) / capacity;
    const std::uint64_t last = (prefix - 1) / capacity;
    *bin = first;
    return first == last;
}

__global__ void isolated_rep_kernel(std::int64_t count, const std::int32_t* vertices, const std::uint64_t* prefix,
                                    const std::uint64_t* weights, std::uint64_t capacity, std::int32_t* reps,
                                    std::uint64_t* run_base) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const std::int32_t v = vertices[i];
    std::uint64_t bin;
    if (!whole_bin(prefix[i], weights[v], capacity, &bin)) return;
    if (i > 0) {
        std::uint64_t before;
        if (whole_bin(prefix[i - 1], weights[vertices[i - 1]], capacity, &before) && before == bin) return;
    }
    reps[bin] = v;
    run_base[bin] = prefix[i] - weights[v];  // the weight before the run
}

__global__ void isolated_commit_kernel(std::int64_t count, const std::int32_t* vertices, const std::uint64_t* prefix,
                                       const std::uint64_t* weights, std::uint64_t capacity,
                                       const std::int32_t* reps, const std::uint64_t* run_base,
                                       std::int32_t* clusters, unsigned long long* cluster_weights) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const std::int32_t v = vertices[i];
    std::uint64_t bin;
    if (!whole_bin(prefix[i], weights[v], capacity, &bin)) return;
    const std::int32_t target = reps[bin];
    if (target != v) {
        clusters[v] = target;
        cluster_weights[v] = 0;
    }
    // the run's last vertex sets the representative's weight: the run's total
    // (the representative was a singleton, weighing its own weight)
    std::uint64_t after = ~0ULL;
    if (i + 1 < count && whole_bin(prefix[i + 1], weights[vertices[i + 1]], capacity, &after) && after == bin) return;
    cluster_weights[target] = static_cast<unsigned long long>(prefix[i] - run_base[bin]);
}

__global__ void fill_i32_kernel(std::int32_t* out, std::int32_t value, std::int64_t n) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = value;
}

// sclp_singleton_favorite_kernel (src/coarsen.cu:682-692)
__global__ void favorite_kernel(std::int64_t n, const std::uint64_t* vertex_weights,
                                const unsigned long long* cluster_weights, const std::int32_t* clusters,
                                const std::int32_t* proposals, std::int32_t* favorites) {
    const std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    favorites[v] = cluster_weights[clusters[v]] == vertex_weights[v] ? proposals[v] : kSclpInvalid;
}

// sclp_pair_flags_kernel (:694-704)
__global__ void pair_flags_kernel(std::int64_t count, const std::uint64_t* ordered_weights,
                                  const std::uint64_t* ranks, std::uint64_t capacity, std::int64_t* pair_flags) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const bool second = (ranks[i] & 1ULL) == 0ULL;
    pair_flags[i] = second && i > 0 && ordered_weights[i - 1] <= capacity - ordered_weights[i] ? 1 : 0;
}

// sclp_two_hop_pair_commit_kernel (:706-724)
__global__ void two_hop_commit_kernel(std::int64_t count, const std::int32_t* order, const std::int64_t* pair_flags,
                                      const std::int64_t* positions, std::uint64_t budget,
                                      const std::uint64_t* ordered_weights, std::int32_t* clusters,
                                      unsigned long long* cluster_weights, unsigned long long* merged) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool did = false;
    if (i < count && pair_flags[i] && static_cast<std::uint64_t>(positions[i]) < budget) {
        const std::int32_t v = order[i];
        const std::int32_t source = clusters[v];
        const std::int32_t target = clusters[order[i - 1]];
        if (source != target) {
            const unsigned long long w = ordered_weights[i];
            atomicAdd(cluster_weights + source, 0ULL - w);
            atomicAdd(cluster_weights + target, w);
            clusters[v] = target;
            did = true;
        }
    }
    const unsigned long long c = static_cast<unsigned long long>(__popcll(__ballot(did)));
    if (lane_id() == 0 && c) atomicAdd(merged, c);
}

// ---------------------------------------------------------------- compaction

struct NonemptyLoad {
    const unsigned long long* weights;
    __device__ __forceinline__ std::int64_t operator()(std::int64_t i) const { return weights[i] > 0 ? 1 : 0; }
};
struct CompactIdStore {
    std::int32_t* ids;
    __device__ __forceinline__ void operator()(std::int64_t i, std::int64_t prefix, std::int64_t) const {
        ids[i] = static_cast<std::int32_t>(prefix);
    }
};

template <typename W>
__global__ void map_and_weights_kernel(std::int64_t n, const std::int32_t* clusters, const std::int32_t* compact,
                                       const W* vertex_weights, std::int32_t* map, W* coarse_weights) {
    const std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const std::int32_t c = compact[clusters[v]];
    map[v] = c;
    if constexpr (sizeof(W) == 4)
        atomicAdd(reinterpret_cast<unsigned int*>(coarse_weights + c), static_cast<unsigned int>(vertex_weights[v]));
    else
        atomicAdd(reinterpret_cast<unsigned long long*>(coarse_weights + c),
                  static_cast<unsigned long long>(vertex_weights[v]));
}

// ---------------------------------------------------------------- record bounds

template <typename O, typename W, int G>
__global__ void record_bound_kernel(std::int64_t count, const std::int32_t* list, const O* offsets,
                                    const std::int32_t* neighbors, const W* weights,
                                    unsigned long long* row_bounds, unsigned long long* meta) {
    const std::int64_t group = (static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
    const int sub = threadIdx.x & (G - 1);
    if (group >= count) return;
    const std::int32_t v = list[group];
    const std::int64_t begin = offsets[v], end = offsets[v + 1];
    unsigned long long degree = 0;
    bool unsorted = false;
    for (std::int64_t e = begin + sub; e < end; e += G) {
        const auto w = static_cast<unsigned long long>(weights[e]);
        degree = ~0ULL - degree < w ? ~0ULL : degree + w;
        unsorted = unsorted || (e > begin && neighbors[e] <= neighbors[e - 1]);
    }
#pragma unroll
    for (int m = G / 2; m > 0; m /= 2) {
        const auto other = shfl_xor_u64(degree, m, G);
        degree = ~0ULL - degree < other ? ~0ULL : degree + other;
    }
    unsorted = group_any<G>(unsorted);
    if (sub == 0) {
        row_bounds[v] = degree;
        if (unsorted) atomicOr(meta + 1, 1ULL);  // rare: only unsorted inputs write it
    }
}

struct U64Load {
    const unsigned long long* data;
    __device__ __forceinline__ unsigned long long operator()(std::int64_t i) const { return data[i]; }
};

// contract_record_bounds (src/coarsen.cu:72-85): saturating add through the map.
// Wrapping adds first: a wrap sets *wrapped (the add that crosses 2^64 always
// sees it) and the host then redoes the level with the saturating kernel below.
// Zero bounds (isolated vertices, often packed into a few clusters) are skipped.
__global__ void contract_bounds_add_kernel(std::int64_t n, const std::int32_t* map, const unsigned long long* fine,
                                           unsigned long long* coarse, unsigned long long* wrapped) {
    const std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const unsigned long long amount = fine[v];
    if (amount == 0) return;
    const unsigned long long previous = atomicAdd(coarse + map[v], amount);
    if (previous > ~0ULL - amount) *wrapped = 1;
}

__global__ void contract_bounds_kernel(std::int64_t n, const std::int32_t* map, const unsigned long long* fine,
                                       unsigned long long* coarse) {
    const std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    unsigned long long* target = coarse + map[v];
    const unsigned long long amount = fine[v];
    if (amount == 0) return;
    unsigned long long previous = __hip_atomic_load(target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (true) {
        const unsigned long long sum = ~0ULL - previous < amount ? ~0ULL : previous + amount;
        const unsigned long long observed = atomicCAS(target, previous, sum);
        if (observed == previous) break;
        previous = observed;
    }
}

__global__ void identity_list_kernel(std::int64_t n, std::int32_t* list) {
    const std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v < n) list[v] = static_cast<std::int32_t>(v);
}

inline unsigned grid_for(std::int64_t threads) { return static_cast<unsigned>((threads + kBlock - 1) / kBlock); }

// Sort the valid-proposal list `order` (p entries, vertex ids ascending) into
// admission order. Returns the buffer holding the sorted order (order or alt).
std::int32_t* admission_sort(gpart_ctx* ctx, std::int64_t n, std::int64_t p, std::int32_t* order,
                             std::int32_t* order_alt, const std::int32_t* targets,
                             const unsigned long long* gains, const std::uint32_t* ties,
                             std::uint64_t gain_bound) {
    if (p <= 1) return order;
    const int target_bits = std::max(1, bit_width_u64(static_cast<std::uint64_t>(n - 1)));
    const int gain_bits = std::max(1, bit_width_u64(gain_bound));
    Scratch<std::uint64_t> keys(ctx, p), keys_alt(ctx, p);
    std::int32_t* cur = order;
    std::int32_t* alt = order_alt;
    auto stage = [&](int kind, int bits) {
        hipLaunchKernelGGL(admission_key_kernel, dim3(grid_for(p)), dim3(kBlock), 0, ctx->stream, p, cur, targets,
                           gains, ties, kind, gain_bits, keys.get());
        GPART_LAUNCH_CHECK();
        const bool swapped = radix_sort_pairs(ctx, keys.get(), keys_alt.get(), reinterpret_cast<std::uint32_t*>(cur),
                                              reinterpret_cast<std::uint32_t*>(alt), p, 0, bits);
        if (swapped) std::swap(cur, alt);
    };
    if (target_bits + gain_bits + 32 <= 64) {
        stage(kAll, target_bits + gain_bits + 32);
    } else if (gain_bits + 32 <= 64) {
        stage(kGainTie, gain_bits + 32);
        stage(kTargetOnly, target_bits);
    } else {
        stage(kTieOnly, 32);
        stage(kGainOnly, gain_bits);
        stage(kTargetOnly, target_bits);
    }
    return cur;
}

unsigned long long count_nonempty(gpart_ctx* ctx, std::int64_t n, const unsigned long long* weights) {
    Scratch<unsigned long long> c(ctx, 1);
    device_reduce<unsigned long long>(ctx, NonzeroLoad{weights}, n, SumU64{}, c.get());
    return read_scalar(ctx, c.get());
}

template <typename O, typename W>
void launch_proposals(gpart_ctx* ctx, const gpart_csr_view* g, const std::int32_t* clusters,
                      const gpart_degree_classes* c, const gpart_sclp_round* round, std::int32_t* proposals,
                      unsigned long long* gains, std::uint32_t* ties) {
    const auto* off = static_cast<const O*>(g->offsets);
    const auto* w = static_cast<const W*>(g->edge_weights);
    RoundParams r{round->role_salt, round->tie_salt, round->mover_threshold, round->filter_roles != 0};
    const bool single = round->singleton != 0;
    // Rows above kHighLdsDegree: singleton rounds offer every entry, in chunks.
    // Otherwise rows up to kPartMaxDegree go to the partitioned LDS tables and
    // larger hubs, in chunks, to per-hub global tables.
    const bool parted = !single && c->high_count > 0 && c->max_degree > kHighLdsDegree;
    const std::int64_t hub_degree = single ? kHighLdsDegree : kPartMaxDegree;
    const bool hubs = c->high_count > 0 && c->max_degree > hub_degree;
    HubChunks chunks, items, split_items;
    Scratch<std::int64_t> table_offset;
    Scratch<std::int32_t> gkeys;
    Scratch<unsigned long long> gvalues, pcurrent, pbest, ptie, icurrent, ibest, itie;
    Scratch<std::int64_t> region, group_begin, group_end;
    Scratch<std::int32_t> row_clusters, group_keys;
    Scratch<W> group_weights;
    Scratch<std::uint32_t> chunk_counts;
    if (parted) {
        // sum over rows of 2^ceil(log2(ceil(d / kPartKeys))) <= 2 * (entries / kPartKeys + rows)
        const std::int64_t bound = 2 * (c->high_edges / kPartKeys + 1) + 2 * c->high_count;
        build_work_items(ctx, PartLoad<O>{c->high, off, c->high_count}, c->high_count, bound, items);
        icurrent.reset(ctx, items.bound);
        ibest.reset(ctx, items.bound);
        itie.reset(ctx, items.bound);
        if (c->max_degree > kPartKeys) {  // some row takes more than one item
            region.reset(ctx, c->high_count + 1);
            device_scan<std::int64_t, SumOp<std::int64_t>>(ctx, SplitLoad<O>{c->high, off, c->high_count},
                                                           ExclusiveStore<std::int64_t>{region.get()},
                                                           c->high_count + 1, SumOp<std::int64_t>{});
            // at most high_edges / (kPartKeys + 1) rows exceed kPartKeys entries
            // a split row of d entries has fewer than 2 d / kPartKeys items
            build_work_items(ctx, SplitItemLoad<O>{PartLoad<O>{c->high, off, c->high_count}}, c->high_count,
                             std::min<std::int64_t>(items.bound, 2 * (c->high_edges / kPartKeys) + 1), split_items);
            chunk_counts.reset(ctx, split_items.bound * kMaxParts);
            const std::int64_t split_edges = read_scalar(ctx, region.get() + c->high_count);
            row_clusters.reset(ctx, split_edges);
            group_keys.reset(ctx, split_edges);
            group_weights.reset(ctx, split_edges);
            group_begin.reset(ctx, items.bound);
            group_end.reset(ctx, items.bound);
        }
    }
    if (hubs) {
        build_hub_chunks(ctx, off, c->high, c->high_count, c->high_edges, hub_degree, chunks);
        pcurrent.reset(ctx, chunks.bound);
        pbest.reset(ctx, chunks.bound);
        ptie.reset(ctx, chunks.bound);
    }
    if (hubs && !single) {
        table_offset.reset(ctx, c->high_count + 1);
        hipLaunchKernelGGL(hub_table_size_kernel<O>, dim3(grid_for(c->high_count + 1)), dim3(kBlock), 0, ctx->stream,
                           c->high_count, c->high, off, hub_degree, table_offset.get());
        GPART_LAUNCH_CHECK();
        exclusive_scan<std::int64_t>(ctx, table_offset.get(), table_offset.get(), c->high_count + 1);
        // the last entry (size 0 appended by the scan of count+1 values) holds the total
        const std::int64_t total = read_scalar(ctx, table_offset.get() + c->high_count);
        gkeys.reset(ctx, total);
        gvalues.reset(ctx, total);
        fill_bytes(ctx, gkeys.get(), 0xff, static_cast<std::size_t>(total) * 4);
        fill_bytes(ctx, gvalues.get(), 0, static_cast<std::size_t>(total) * 8);
    }
    profile_begin(ctx, GPART_PROFILE_SCLP_AFFINITY);
    if (c->low_count > 0) {
        const unsigned blocks = grid_for(c->low_count * kLowDegree);
        if (single)
            hipLaunchKernelGGL((sclp_low_kernel<O, W, true>), dim3(blocks), dim3(kBlock), 0, ctx->stream, c->low_count,
                               c->low, off, g->neighbors, w, clusters, r, proposals, gains, ties);
        else
            hipLaunchKernelGGL((sclp_low_kernel<O, W, false>), dim3(blocks), dim3(kBlock), 0, ctx->stream,
                               c->low_count, c->low, off, g->neighbors, w, clusters, r, proposals, gains, ties);
        GPART_LAUNCH_CHECK();
    }
    if (c->medium_count > 0) {
        const unsigned blocks = static_cast<unsigned>((c->medium_count + 3) / 4);
        if (single)
            hipLaunchKernelGGL((sclp_medium_kernel<O, W, true>), dim3(blocks), dim3(kBlock), 0, ctx->stream,
                               c->medium_count, c->medium, off, g->neighbors, w, clusters, r, proposals, gains, ties);
        else
            hipLaunchKernelGGL((sclp_medium_kernel<O, W, false>), dim3(blocks), dim3(kBlock), 0, ctx->stream,
                               c->medium_count, c->medium, off, g->neighbors, w, clusters, r, proposals, gains, ties);
        GPART_LAUNCH_CHECK();
    }
    if (c->high_count > 0) {
        const unsigned blocks = static_cast<unsigned>(c->high_count);
        if (single)
            hipLaunchKernelGGL((sclp_high_kernel<O, W, true>), dim3(blocks), dim3(kBlock), 0, ctx->stream,
                               c->high_count, c->high, off, g->neighbors, w, clusters, r, proposals, gains, ties);
        else
            hipLaunchKernelGGL((sclp_high_kernel<O, W, false>), dim3(blocks), dim3(kBlock), 0, ctx->stream,
                               c->high_count, c->high, off, g->neighbors, w, clusters, r, proposals, gains, ties);
        GPART_LAUNCH_CHECK();
    }
    if (parted) {
        if (row_clusters.get()) {
            const unsigned split_grid = static_cast<unsigned>(std::min<std::int64_t>(split_items.bound, kPartLaunchCap));
            hipLaunchKernelGGL((sclp_split_count_kernel<O>), dim3(split_grid), dim3(kPartThreads), 0, ctx->stream,
                               split_items.total(), split_items.hub.get(), split_items.start.get(), c->high, off,
                               g->neighbors, clusters, region.get(), row_clusters.get(), chunk_counts.get());
            hipLaunchKernelGGL(sclp_split_offsets_kernel, dim3(split_grid), dim3(kPartThreads), 0, ctx->stream,
                               split_items.total(), split_items.hub.get(), split_items.start.get(), region.get(),
                               items.start.get(), chunk_counts.get(), group_begin.get(), group_end.get());
            hipLaunchKernelGGL((sclp_split_scatter_kernel<O, W>), dim3(split_grid), dim3(kPartThreads), 0,
                               ctx->stream, split_items.total(), split_items.hub.get(), split_items.start.get(),
                               c->high, off, w, region.get(), chunk_counts.get(), row_clusters.get(),
                               group_keys.get(), group_weights.get());
        }
        const unsigned grid = static_cast<unsigned>(std::min<std::int64_t>(items.bound, kPartLaunchCap));
        // connection sums fit 32 bits when every row weight does
        if (round->value_bound != 0 && round->value_bound <= 0xffffffffULL)
            hipLaunchKernelGGL((sclp_part_kernel<O, W, unsigned>), dim3(grid), dim3(kPartThreads), 0, ctx->stream,
                               items.total(), items.start.get(), items.hub.get(), c->high, off, g->neighbors, w,
                               clusters, r, group_keys.get(), group_weights.get(), group_begin.get(), group_end.get(),
                               icurrent.get(), ibest.get(), itie.get());
        else
            hipLaunchKernelGGL((sclp_part_kernel<O, W, unsigned long long>), dim3(grid), dim3(kPartThreads), 0,
                               ctx->stream, items.total(), items.start.get(), items.hub.get(), c->high, off,
                               g->neighbors, w, clusters, r, group_keys.get(), group_weights.get(),
                               group_begin.get(), group_end.get(), icurrent.get(), ibest.get(), itie.get());
        hipLaunchKernelGGL(sclp_hub_decode_kernel, dim3(grid_for(c->high_count)), dim3(kBlock), 0, ctx->stream,
                           c->high_count, items.start.get(), c->high, clusters, r, icurrent.get(), ibest.get(),
                           itie.get(), proposals, gains, ties);
        GPART_LAUNCH_CHECK();
    }
    if (hubs) {
        if (single) {
            hipLaunchKernelGGL((sclp_hub_insert_kernel<O, W, true>), dim3(chunks.grid()), dim3(kBlock), 0,
                               ctx->stream, chunks.total(), chunks.start.get(), chunks.hub.get(), c->high, off,
                               g->neighbors, w, clusters, r, nullptr, nullptr, nullptr, pcurrent.get(), pbest.get(),
                               ptie.get());
        } else {
            hipLaunchKernelGGL((sclp_hub_insert_kernel<O, W, false>), dim3(chunks.grid()), dim3(kBlock), 0,
                               ctx->stream, chunks.total(), chunks.start.get(), chunks.hub.get(), c->high, off,
                               g->neighbors, w, clusters, r, table_offset.get(), gkeys.get(), gvalues.get(), nullptr,
                               nullptr, nullptr);
            hipLaunchKernelGGL(sclp_hub_scan_kernel<O>, dim3(chunks.grid()), dim3(kBlock), 0, ctx->stream,
                               chunks.total(), chunks.start.get(), chunks.hub.get(), c->high, clusters, r,
                               table_offset.get(), gkeys.get(), gvalues.get(), pcurrent.get(), pbest.get(),
                               ptie.get());
        }
        hipLaunchKernelGGL(sclp_hub_decode_kernel, dim3(grid_for(c->high_count)), dim3(kBlock), 0, ctx->stream,
                           c->high_count, chunks.start.get(), c->high, clusters, r, pcurrent.get(), pbest.get(),
                           ptie.get(), proposals, gains, ties);
        GPART_LAUNCH_CHECK();
    }
    const double n = static_cast<double>(g->n), m = static_cast<double>(g->m);
    const double csr = (n + 1) * sizeof(O) + m * (4 + sizeof(W));
    profile_end(ctx, GPART_PROFILE_SCLP_AFFINITY, csr + (single ? 0.0 : 4 * m) + 20 * n);
}

template <typename O, typename W>
void build_classes(gpart_ctx* ctx, const gpart_csr_view* g, gpart_degree_classes* c) {
    Scratch<unsigned long long> counters(ctx, 6);
    fill_bytes(ctx, counters.get(), 0, 48);
    if (g->n > 0) {
        const auto* off = static_cast<const O*>(g->offsets);
        device_scan<DegreeScan, DegreeScanOp>(
            ctx, ClassLoad<O>{off}, ClassStore<O>{off, c->low, c->medium, c->high, g->n, counters.get()}, g->n,
            DegreeScanOp{});
    }
    unsigned long long h[6];
    read_back(ctx, h, counters.get(), sizeof(h));
    c->low_count = static_cast<std::int64_t>(h[0]);
    c->medium_count = static_cast<std::int64_t>(h[1]);
    c->high_count = static_cast<std::int64_t>(h[2]);
    c->high_edges = static_cast<std::int64_t>(h[3]);
    c->max_degree = static_cast<std::int64_t>(h[4]);
    c->isolated_count = static_cast<std::int64_t>(h[5]);
}

template <typename O, typename W>
void record_bounds_impl(gpart_ctx* ctx, const gpart_csr_view* g, std::uint64_t* row_bounds,
                        std::uint64_t* maximum, int* unsorted) {
    Scratch<unsigned long long> meta(ctx, 2);
    fill_bytes(ctx, meta.get(), 0, 16);
    if (g->n > 0) {
        Scratch<std::int32_t> low(ctx, g->n), medium(ctx, g->n), high(ctx, g->n);
        gpart_degree_classes c{low.get(), medium.get(), high.get(), 0, 0, 0, 0, 0, 0};
        build_classes<O, W>(ctx, g, &c);
        const auto* off = static_cast<const O*>(g->offsets);
        const auto* w = static_cast<const W*>(g->edge_weights);
        auto* rb = reinterpret_cast<unsigned long long*>(row_bounds);
        if (c.low_count)
            hipLaunchKernelGGL((record_bound_kernel<O, W, 8>), dim3(grid_for(c.low_count * 8)), dim3(kBlock), 0,
                               ctx->stream, c.low_count, c.low, off, g->neighbors, w, rb, meta.get());
        if (c.medium_count)
            hipLaunchKernelGGL((record_bound_kernel<O, W, 64>), dim3(grid_for(c.medium_count * 64)), dim3(kBlock), 0,
                               ctx->stream, c.medium_count, c.medium, off, g->neighbors, w, rb, meta.get());
        if (c.high_count)
            hipLaunchKernelGGL((record_bound_kernel<O, W, 64>), dim3(grid_for(c.high_count * 64)), dim3(kBlock), 0,
                               ctx->stream, c.high_count, c.high, off, g->neighbors, w, rb, meta.get());
        GPART_LAUNCH_CHECK();
        device_reduce<unsigned long long>(ctx, U64Load{rb}, g->n, MaxU64{}, meta.get());
    }
    unsigned long long h[2];
    read_back(ctx, h, meta.get(), sizeof(h));
    *maximum = h[0];
    *unsorted = h[1] != 0;
}

}  // namespace
}  // namespace gpart_hip

using namespace gpart_hip;

extern "C" {

int gpart_degree_classes_build(gpart_ctx* ctx, const gpart_csr_view* g, gpart_degree_classes* classes) {
    return guarded([&] {
        if (g->width == GPART_WIDTH_STANDARD) build_classes<std::int32_t, std::uint32_t>(ctx, g, classes);
        else build_classes<std::int64_t, std::uint64_t>(ctx, g, classes);
    });
}

int gpart_record_bounds(gpart_ctx* ctx, const gpart_csr_view* g, uint64_t* row_bounds, uint64_t* maximum,
                        int* unsorted) {
    return guarded([&] {
        if (g->width == GPART_WIDTH_STANDARD)
            record_bounds_impl<std::int32_t, std::uint32_t>(ctx, g, row_bounds, maximum, unsorted);
        else
            record_bounds_impl<std::int64_t, std::uint64_t>(ctx, g, row_bounds, maximum, unsorted);
    });
}

int gpart_contract_record_bounds(gpart_ctx* ctx, int64_t n, const int32_t* map, const uint64_t* fine_bounds,
                                 int64_t n_coarse, uint64_t* coarse_bounds, uint64_t* maximum) {
    return guarded([&] {
        fill_bytes(ctx, coarse_bounds, 0, static_cast<std::size_t>(n_coarse) * 8);
        auto* cb = reinterpret_cast<unsigned long long*>(coarse_bounds);
        const auto* fb = reinterpret_cast<const unsigned long long*>(fine_bounds);
        Scratch<unsigned long long> meta(ctx, 2);  // maximum, wrapped
        fill_bytes(ctx, meta.get(), 0, 16);
        if (n > 0) {
            hipLaunchKernelGGL(contract_bounds_add_kernel, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, map,
                               fb, cb, meta.get() + 1);
            GPART_LAUNCH_CHECK();
        }
        if (n_coarse > 0) device_reduce<unsigned long long>(ctx, U64Load{cb}, n_coarse, MaxU64{}, meta.get());
        unsigned long long h[2];
        read_back(ctx, h, meta.get(), sizeof(h));
        if (h[1]) {
            fill_bytes(ctx, coarse_bounds, 0, static_cast<std::size_t>(n_coarse) * 8);
            fill_bytes(ctx, meta.get(), 0, 8);
            hipLaunchKernelGGL(contract_bounds_kernel, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, map, fb,
                               cb);
            GPART_LAUNCH_CHECK();
            if (n_coarse > 0) device_reduce<unsigned long long>(ctx, U64Load{cb}, n_coarse, MaxU64{}, meta.get());
            h[0] = read_scalar(ctx, meta.get());
        }
        *maximum = h[0];
    });
}

int gpart_sclp_proposals(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* clusters,
                         const gpart_degree_classes* classes, const gpart_sclp_round* round, int32_t* proposals,
                         uint64_t* gains, uint32_t* ties) {
    return guarded([&] {
        auto* gn = reinterpret_cast<unsigned long long*>(gains);
        if (g->width == GPART_WIDTH_STANDARD)
            launch_proposals<std::int32_t, std::uint32_t>(ctx, g, clusters, classes, round, proposals, gn, ties);
        else
            launch_proposals<std::int64_t, std::uint64_t>(ctx, g, clusters, classes, round, proposals, gn, ties);
    });
}

int gpart_sclp_admit(gpart_ctx* ctx, int64_t n, const int32_t* proposals, const uint64_t* gains,
                     const uint32_t* ties, const uint64_t* vertex_weights, uint64_t capacity, uint64_t gain_bound,
                     int32_t* clusters, uint64_t* cluster_weights, int64_t* valid, int64_t* accepted,
                     int64_t* nonempty, uint64_t* predicted_gain) {
    return guarded([&] {
        auto* cw = reinterpret_cast<unsigned long long*>(cluster_weights);
        const auto* gn = reinterpret_cast<const unsigned long long*>(gains);
        Scratch<std::int32_t> order(ctx, n), order_alt(ctx, n);
        Scratch<std::int64_t> d_count(ctx, 1);
        compact_valid(ctx, n, proposals, order.get(), d_count.get());
        const std::int64_t p = read_scalar(ctx, d_count.get());
        *valid = p;
        *accepted = 0;
        if (p > 0) {
            std::int32_t* sorted = admission_sort(ctx, n, p, order.get(), order_alt.get(), proposals, gn, ties,
                                                  gain_bound);
            Scratch<std::int32_t> ordered_targets(ctx, p), accept(ctx, p);
            Scratch<std::uint64_t> ordered_weights(ctx, p), prefix(ctx, p);
            Scratch<unsigned long long> acc(ctx, 3);  // accepted, predicted gain, non-empty clusters
            // the gather also clears the commit's two counters
            hipLaunchKernelGGL(gather_order_kernel, dim3(grid_for(p)), dim3(kBlock), 0, ctx->stream, p, sorted,
                               proposals, vertex_weights, ordered_targets.get(), ordered_weights.get(), acc.get(), 2);
            GPART_LAUNCH_CHECK();
            device_scan<SegValue, SegOp>(ctx, SegWeightLoad{ordered_targets.get(), ordered_weights.get()},
                                         SegInclusiveStore{prefix.get()}, p, SegOp{});
            hipLaunchKernelGGL(admit_decide_kernel, dim3(grid_for(p)), dim3(kBlock), 0, ctx->stream, p,
                               ordered_targets.get(), prefix.get(), cw, capacity, accept.get());
            GPART_LAUNCH_CHECK();
            profile_begin(ctx, GPART_PROFILE_SCLP_COMMIT);
            hipLaunchKernelGGL(admit_apply_kernel, dim3(grid_for(p)), dim3(kBlock), 0, ctx->stream, p, sorted,
                               ordered_targets.get(), ordered_weights.get(), accept.get(),
                               reinterpret_cast<const unsigned long long*>(gains), clusters, cw, acc.get());
            GPART_LAUNCH_CHECK();
            profile_end(ctx, GPART_PROFILE_SCLP_COMMIT, 0.0);
            // the non-empty cluster count comes back with the commit's counters
            device_reduce<unsigned long long>(ctx, NonzeroLoad{cw}, n, SumU64{}, acc.get() + 2);
            unsigned long long h[3];
            read_back(ctx, h, acc.get(), sizeof(h));
            *accepted = static_cast<std::int64_t>(h[0]);
            *nonempty = static_cast<std::int64_t>(h[2]);
            // per proposal its order, target, weight and decision; per accepted move
            // its gain, cluster read and write and two cluster-weight atomics
            profile_add_bytes(ctx, GPART_PROFILE_SCLP_COMMIT,
                              20.0 * static_cast<double>(p) + 48.0 * static_cast<double>(h[0]));
            profile_add_units(ctx, GPART_PROFILE_SCLP_COMMIT, 2.0 * static_cast<double>(h[0]));
            if (predicted_gain) *predicted_gain = h[1];
        } else {
            *nonempty = static_cast<std::int64_t>(count_nonempty(ctx, n, cw));
        }
    });
}

int gpart_sclp_isolated(gpart_ctx* ctx, const gpart_csr_view* g, const uint64_t* vertex_weights, uint64_t capacity,
                        int32_t* clusters, uint64_t* cluster_weights, int64_t* isolated, int64_t* nonempty) {
    return guarded([&] {
        const auto n = g->n;
        auto* cw = reinterpret_cast<unsigned long long*>(cluster_weights);
        Scratch<std::int32_t> flags(ctx, n), list(ctx, n);
        Scratch<std::int64_t> d_count(ctx, 1);
        if (g->width == GPART_WIDTH_STANDARD)
            hipLaunchKernelGGL(zero_degree_flags_kernel<std::int32_t>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream,
                               n, static_cast<const std::int32_t*>(g->offsets), flags.get());
        else
            hipLaunchKernelGGL(zero_degree_flags_kernel<std::int64_t>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream,
                               n, static_cast<const std::int64_t*>(g->offsets), flags.get());
        GPART_LAUNCH_CHECK();
        compact_indices(ctx, flags.get(), n, list.get(), d_count.get());
        const std::int64_t count = read_scalar(ctx, d_count.get());
        *isolated = count;
        if (count > 1) {
            Scratch<std::uint64_t> prefix(ctx, count);
            device_scan<std::uint64_t, SumOp<std::uint64_t>>(ctx, GatherWeightLoad{list.get(), vertex_weights},
                                                              InclusiveStore<std::uint64_t>{prefix.get()}, count,
                                                              SumOp<std::uint64_t>{});
            // every bin holds a vertex, so count bounds the bins (no read-back of the total)
            const std::int64_t bins = count;
            Scratch<std::int32_t> reps(ctx, bins);
            Scratch<std::uint64_t> run_base(ctx, bins);
            hipLaunchKernelGGL(fill_i32_kernel, dim3(grid_for(bins)), dim3(kBlock), 0, ctx->stream, reps.get(),
                               0x7fffffff, bins);
            hipLaunchKernelGGL(isolated_rep_kernel, dim3(grid_for(count)), dim3(kBlock), 0, ctx->stream, count,
                               list.get(), prefix.get(), vertex_weights, capacity, reps.get(), run_base.get());
            hipLaunchKernelGGL(isolated_commit_kernel, dim3(grid_for(count)), dim3(kBlock), 0, ctx->stream, count,
                               list.get(), prefix.get(), vertex_weights, capacity, reps.get(), run_base.get(),
                               clusters, cw);
            GPART_LAUNCH_CHECK();
        }
        *nonempty = static_cast<std::int64_t>(count_nonempty(ctx, n, cw));
    });
}

int gpart_sclp_two_hop(gpart_ctx* ctx, int64_t n, const int32_t* proposals, const uint64_t* gains,
                       const uint32_t* ties, const uint64_t* vertex_weights, uint64_t capacity, uint64_t gain_bound,
                       int64_t merge_budget, int32_t* clusters, uint64_t* cluster_weights, int64_t* candidates,
                       int64_t* pairable, int64_t* merged, int64_t* nonempty) {
    return guarded([&] {
        auto* cw = reinterpret_cast<unsigned long long*>(cluster_weights);
        const auto* gn = reinterpret_cast<const unsigned long long*>(gains);
        Scratch<std::int32_t> favorites(ctx, n), order(ctx, n), order_alt(ctx, n);
        Scratch<std::int64_t> d_count(ctx, 1);
        hipLaunchKernelGGL(favorite_kernel, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, vertex_weights, cw,
                           clusters, proposals, favorites.get());
        GPART_LAUNCH_CHECK();
        compact_valid(ctx, n, favorites.get(), order.get(), d_count.get());
        const std::int64_t p = read_scalar(ctx, d_count.get());
        *candidates = p;
        *pairable = 0;
        *merged = 0;
        if (p > 0) {
            std::int32_t* sorted = admission_sort(ctx, n, p, order.get(), order_alt.get(), favorites.get(), gn, ties,
                                                  gain_bound);
            Scratch<std::int32_t> ordered_targets(ctx, p);
            Scratch<std::uint64_t> ordered_weights(ctx, p), ranks(ctx, p);
            Scratch<std::int64_t> pair_flags(ctx, p), positions(ctx, p + 1);
            hipLaunchKernelGGL(gather_order_kernel, dim3(grid_for(p)), dim3(kBlock), 0, ctx->stream, p, sorted,
                               favorites.get(), vertex_weights, ordered_targets.get(), ordered_weights.get(),
                               nullptr, 0);
            GPART_LAUNCH_CHECK();
            device_scan<SegValue, SegOp>(ctx, SegOneLoad{ordered_targets.get()}, SegInclusiveStore{ranks.get()}, p,
                                         SegOp{});
            hipLaunchKernelGGL(pair_flags_kernel, dim3(grid_for(p)), dim3(kBlock), 0, ctx->stream, p,
                               ordered_weights.get(), ranks.get(), capacity, pair_flags.get());
            GPART_LAUNCH_CHECK();
            // positions and, in one read-back, the pair count
            device_scan<std::int64_t, SumOp<std::int64_t>>(ctx, ArrayLoad<std::int64_t>{pair_flags.get()},
                                                           ExclusiveTotalStore{positions.get(), p, d_count.get()},
                                                           p, SumOp<std::int64_t>{});
            const std::int64_t pair_count = read_scalar(ctx, d_count.get());
            *pairable = pair_count;
            const auto budget = static_cast<std::uint64_t>(std::max<std::int64_t>(0, merge_budget));
            if (pair_count > 0 && budget > 0) {
                Scratch<unsigned long long> m(ctx, 2);  // merged pairs, non-empty clusters
                fill_bytes(ctx, m.get(), 0, 16);
                hipLaunchKernelGGL(two_hop_commit_kernel, dim3(grid_for(p)), dim3(kBlock), 0, ctx->stream, p, sorted,
                                   pair_flags.get(), positions.get(), budget, ordered_weights.get(), clusters, cw,
                                   m.get());
                GPART_LAUNCH_CHECK();
                device_reduce<unsigned long long>(ctx, NonzeroLoad{cw}, n, SumU64{}, m.get() + 1);
                unsigned long long h[2];
                read_back(ctx, h, m.get(), sizeof(h));
                *merged = static_cast<std::int64_t>(h[0]);
                *nonempty = static_cast<std::int64_t>(h[1]);
                return;
            }
        }
        *nonempty = static_cast<std::int64_t>(count_nonempty(ctx, n, cw));
    });
}

int gpart_sclp_compact(gpart_ctx* ctx, int64_t n, const int32_t* clusters, const uint64_t* cluster_weights,
                       const gpart_csr_view* g, int32_t* map, int64_t n_coarse, void* coarse_vertex_weights,
                       uint64_t* maximum_weight) {
    return guarded([&] {
        const auto* cw = reinterpret_cast<const unsigned long long*>(cluster_weights);
        Scratch<std::int32_t> compact(ctx, n);
        device_scan<std::int64_t, SumOp<std::int64_t>>(ctx, NonemptyLoad{cw}, CompactIdStore{compact.get()}, n,
                                                       SumOp<std::int64_t>{});
        const std::size_t wbytes = g->width == GPART_WIDTH_STANDARD ? 4 : 8;
        fill_bytes(ctx, coarse_vertex_weights, 0, static_cast<std::size_t>(n_coarse) * wbytes);
        Scratch<unsigned long long> mx(ctx, 1);
        fill_bytes(ctx, mx.get(), 0, 8);
        if (g->width == GPART_WIDTH_STANDARD) {
            hipLaunchKernelGGL(map_and_weights_kernel<std::uint32_t>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream,
                               n, clusters, compact.get(), static_cast<const std::uint32_t*>(g->vertex_weights), map,
                               static_cast<std::uint32_t*>(coarse_vertex_weights));
            device_reduce<unsigned long long>(
                ctx, WidenLoad<std::uint32_t>{static_cast<const std::uint32_t*>(coarse_vertex_weights)}, n_coarse,
                MaxU64{}, mx.get());
        } else {
            hipLaunchKernelGGL(map_and_weights_kernel<std::uint64_t>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream,
                               n, clusters, compact.get(), static_cast<const std::uint64_t*>(g->vertex_weights), map,
                               static_cast<std::uint64_t*>(coarse_vertex_weights));
            device_reduce<unsigned long long>(
                ctx, WidenLoad<std::uint64_t>{static_cast<const std::uint64_t*>(coarse_vertex_weights)}, n_coarse,
                MaxU64{}, mx.get());
        }
        GPART_LAUNCH_CHECK();
        *maximum_weight = read_scalar(ctx, mx.get());
    });
}

}  // extern "C"
