// refine.hip -- incremental bucketed refinement for gfx950
// (src/refine.cu and src/refine_incremental.cuh of the reference).
//
// Per-vertex edge loops use the level's degree classes: 8 lanes for degree <= 8,
// a wave for <= 256 and a workgroup above. Row scans of the n x k connectivity
// cache use a group of next_pow2(k) lanes per vertex so every row is one
// contiguous, coalesced read.
//
// The reference ranks vertices inside a rebalance bucket by the return value of
// atomicAdd (src/refine_incremental.cuh:185,216): any order is legal. Here the
// rank is the weight of the bucket's lower-id members (a stable sort by bucket
// and a segmented scan), so refinement is reproducible run to run.
#include "internal.hpp"
#include "reduce.hpp"
#include "scan.hpp"

#include <algorithm>
#include <cmath>

namespace gpart_hip {
namespace {

constexpr int kBlock = 256;
constexpr int kBalanceBuckets = 50;
constexpr int kBalanceSections = 16;
constexpr int kBalanceWidth = kBalanceBuckets * kBalanceSections;

inline unsigned grid_for(std::int64_t threads) { return static_cast<unsigned>((threads + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------- cut / weights

template <typename O, typename W, int G>
__global__ void __launch_bounds__(kBlock) cut_kernel(std::int64_t count, const std::int32_t* list, const O* offsets,
                                                     const std::int32_t* neighbors, const W* weights,
                                                     const std::int32_t* labels, unsigned long long* directed) {
    // grid-stride over vertex groups; one partial per workgroup in directed[blockIdx.x]
    __shared__ unsigned long long waves[kBlock / kWave];
    const int sub = threadIdx.x & (G - 1);
    const std::int64_t groups_per_grid = static_cast<std::int64_t>(gridDim.x) * (blockDim.x / G);
    unsigned long long local = 0;
    for (std::int64_t group = (static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; group < count;
         group += groups_per_grid) {
        const std::int32_t v = list ? list[group] : static_cast<std::int32_t>(group);
        const std::int32_t p = labels[v];
        for (std::int64_t e = static_cast<std::int64_t>(offsets[v]) + sub; e < offsets[v + 1]; e += G)
            if (labels[neighbors[e]] != p) local += static_cast<unsigned long long>(weights[e]);
    }
    local = group_sum_u64<kWave>(local);
    if (lane_id() == 0) waves[threadIdx.x / kWave] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long total = 0;
        for (int w = 0; w < kBlock / kWave; ++w) total += waves[w];
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

template <typename O, typename W, int G>
__global__ void __launch_bounds__(kBlock) build_cache_kernel(std::int64_t count, const std::int32_t* list,
                                                             const O* offsets, const std::int32_t* neighbors,
                                                             const W* weights, const std::int32_t* labels, int k,
                                                             unsigned long long* cache) {
    constexpr int kGroups = kBlock / G;
    __shared__ unsigned long long rows[kGroups][32];
    const int local_group = threadIdx.x / G;
    const int sub = threadIdx.x & (G - 1);
    const std::int64_t group = static_cast<std::int64_t>(blockIdx.x) * kGroups + local_group;
    for (int p = sub; p < 32; p += G) rows[local_group][p] = 0;
    __syncthreads();
    std::int32_t v = 0;
    if (group < count) {
        v = list ? list[group] : static_cast<std::int32_t>(group);
        for (std::int64_t e = static_cast<std::int64_t>(offsets[v]) + sub; e < offsets[v + 1]; e += G) {
            const std::int32_t u = neighbors[e];
            if (u != v) atomicAdd(&rows[local_group][labels[u]], static_cast<unsigned long long>(weights[e]));
        }
    }
    __syncthreads();
    if (group < count)
        for (int p = sub; p < k; p += G) cache[static_cast<std::int64_t>(v) * k + p] = rows[local_group][p];
}

// incremental_lp_proposals (:27-46), one group of G >= k lanes per vertex.
template <int G>
__global__ void __launch_bounds__(kBlock) lp_proposals_kernel(std::int64_t n, const unsigned long long* cache,
                                                              const std::int32_t* labels, int k, double filter,
                                                              std::int32_t* locks, std::int32_t* targets,
                                                              long long* gains) {
    const std::int64_t v = (static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
    const int p = threadIdx.x & (G - 1);
    if (v >= n) return;  // uniform per group
    const bool locked = locks[v] != 0;
    const std::int32_t source = labels[v];
    const unsigned long long value = p < k ? cache[v * k + p] : 0;
    // argmax over p != source of value > 0, lowest index on ties
    unsigned long long best = (p < k && p != source) ? value : 0;
    int best_p = best > 0 ? p : 0x7fffffff;
#pragma unroll
    for (int m = G / 2; m > 0; m /= 2) {
        const auto ob = shfl_xor_u64(best, m, G);
        const int op = __shfl_xor(best_p, m, G);
        if (ob > best || (ob == best && op < best_p)) {
            best = ob;
            best_p = op;
        }
    }
    const unsigned long long current = shfl_u64(value, source, G);
    if (p != 0) return;
    if (locked) {
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

// incremental_priority_filter (:48-67)
template <typename O, typename W, int G>
__global__ void __launch_bounds__(kBlock) priority_filter_kernel(std::int64_t count, const std::int32_t* list,
                                                                 const O* offsets, const std::int32_t* neighbors,
                                                                 const W* weights, const std::int32_t* labels,
                                                                 const std::int32_t* targets, const long long* gains,
                                                                 std::int32_t* flags) {
    __shared__ long long block_sum[kBlock / kWave];
    const std::int64_t group = (static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
    const int sub = threadIdx.x & (G - 1);
    const bool active = group < count;
    const std::int32_t v = active ? (list ? list[group] : static_cast<std::int32_t>(group)) : 0;
    const std::int32_t target = active ? targets[v] : -1;
    long long correction = 0;
    if (active && target >= 0) {
        const std::int32_t source = labels[v];
        const long long gain = gains[v];
        for (std::int64_t e = static_cast<std::int64_t>(offsets[v]) + sub; e < offsets[v + 1]; e += G) {
            const std::int32_t u = neighbors[e];
            if (u == v) continue;
            const std::int32_t tu = targets[u];
            if (tu < 0) continue;
            const long long gu = gains[u];
            if (!(gu > gain || (gu == gain && u < v))) continue;
            const std::int32_t lu = labels[u];
            const long long w = static_cast<long long>(weights[e]);
            correction += w * ((tu == target) - (tu == source) - (lu == target) + (lu == source));
        }
    }
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

// incremental_update_cache_and_cut (:69-95)
template <typename O, typename W, int G>
__global__ void __launch_bounds__(kBlock) update_cache_kernel(std::int64_t count, const std::int32_t* movers,
                                                              const O* offsets, const std::int32_t* neighbors,
                                                              const W* weights, const std::int32_t* labels,
                                                              const std::int32_t* targets,
                                                              const std::int32_t* flags, int k,
                                                              unsigned long long* cache, long long* delta) {
    __shared__ long long block_sum[kBlock / kWave];
    const std::int64_t group = (static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
    const int sub = threadIdx.x & (G - 1);
    long long gain = 0;
    if (group < count) {
        const std::int32_t v = movers[group];
        const std::int32_t source = labels[v];
        const std::int32_t target = targets[v];
        for (std::int64_t e = static_cast<std::int64_t>(offsets[v]) + sub; e < offsets[v + 1]; e += G) {
            const std::int32_t u = neighbors[e];
            if (u == v) continue;
            const auto w = static_cast<unsigned long long>(weights[e]);
            atomicAdd(cache + static_cast<std::int64_t>(u) * k + source, 0ULL - w);
            atomicAdd(cache + static_cast<std::int64_t>(u) * k + target, w);
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

// incremental_commit (:97-109)
template <typename W>
__global__ void commit_kernel(std::int64_t count, const std::int32_t* movers, std::int32_t* labels,
                              const std::int32_t* targets, const W* vertex_weights, unsigned long long* part_weights,
                              std::int32_t* locks, bool lp) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const std::int32_t v = movers[i];
    const std::int32_t source = labels[v];
    const std::int32_t target = targets[v];
    const auto w = static_cast<unsigned long long>(vertex_weights[v]);
    atomicAdd(part_weights + source, 0ULL - w);
    atomicAdd(part_weights + target, w);
    labels[v] = target;
    if (lp) locks[v] = 1;
}

template <typename O>
__global__ void split_movers_kernel(std::int64_t count, const std::int32_t* movers, const O* offsets,
                                    std::int32_t* light, std::int32_t* heavy, unsigned long long* counters) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = i < count;
    const std::int32_t v = valid ? movers[i] : 0;
    const bool heavy_row = valid && (offsets[v + 1] - offsets[v]) > 64;
    const std::uint64_t below = lanemask_lt();
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const bool mine = valid && (c == 1) == heavy_row;
        const std::uint64_t mask = __ballot(mine);
        if (!mask) continue;
        const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        unsigned long long base = 0;
        if (lane_id() == leader) base = atomicAdd(counters + c, static_cast<unsigned long long>(__popcll(mask)));
        base = shfl_u64(base, leader, kWave);
        if (mine) (c ? heavy : light)[base + __popcll(mask & below)] = v;
    }
}

// ---------------------------------------------------------------- rebalance

struct PowerTable {
    double value[48];  // correctly rounded 1.5^j, j = -23..24
};

// balance_gain_bucket (:167-171): clamp(25 + ceil(log_1.5(-g)), 2, 49) for g < 0,
// evaluated as the smallest j with -g <= 1.5^j.
__device__ __forceinline__ int gain_bucket(double gain, const PowerTable& t) {
    if (gain > 0) return 0;
    if (gain == 0) return 1;
    const double x = -gain;
    int lo = 0, hi = 48;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (x <= t.value[mid]) hi = mid;
        else lo = mid + 1;
    }
    const int j = lo == 48 ? 25 : lo - 23;
    return max(2, min(49, 25 + j));
}

// incremental_rebalance_proposals (:111-157) and the source bucket of each proposal.
// Vertices of overloaded parts are the only possible proposers; everyone else is
// reset here and skipped by the row scan below.
__global__ void overloaded_flags_kernel(std::int64_t n, const std::int32_t* labels,
                                        const unsigned long long* part_weights, int k, std::uint64_t capacity,
                                        std::int32_t* flags, std::int32_t* targets, std::int32_t* buckets) {
    __shared__ unsigned long long pw[32];
    if (threadIdx.x < 32) pw[threadIdx.x] = threadIdx.x < k ? part_weights[threadIdx.x] : 0;
    __syncthreads();
    const std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    flags[v] = pw[labels[v]] > capacity;
    targets[v] = -1;
    buckets[v] = -1;
}

template <typename W, int G>
__global__ void __launch_bounds__(kBlock) rebalance_proposals_kernel(std::int64_t count, const std::int32_t* list,
                                                                     const unsigned long long* cache,
                                                                     const std::int32_t* labels, const W* vertex_weights,
                                                                     int k, const unsigned long long* part_weights,
                                                                     std::uint64_t capacity, PowerTable powers,
                                                                     std::int32_t* targets, std::int32_t* buckets) {
    __shared__ unsigned long long pw[32];
    if (threadIdx.x < 32) pw[threadIdx.x] = threadIdx.x < k ? part_weights[threadIdx.x] : 0;
    __syncthreads();
    const std::int64_t index = (static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
    const int p = threadIdx.x & (G - 1);
    if (index >= count) return;  // uniform per group
    const std::int64_t v = list[index];
    const std::int32_t source = labels[v];
    const std::uint64_t weight = static_cast<std::uint64_t>(vertex_weights[v]);
    const bool ok = p < k && p != source && pw[p] < capacity && weight <= capacity - pw[p];
    unsigned long long conn = ok ? cache[v * k + p] : 0;
    unsigned long long slack = ok ? capacity - pw[p] : 0;
    int bp = ok ? p : 0x7fffffff;
    // (conn, slack) lexicographic max, lowest index on ties
#pragma unroll
    for (int m = G / 2; m > 0; m /= 2) {
        const auto oc = shfl_xor_u64(conn, m, G);
        const auto os = shfl_xor_u64(slack, m, G);
        const int op = __shfl_xor(bp, m, G);
        const bool ob = op != 0x7fffffff;
        const bool mine = bp != 0x7fffffff;
        if (ob && (!mine || oc > conn || (oc == conn && (os > slack || (os == slack && op < bp))))) {
            conn = oc;
            slack = os;
            bp = op;
        }
    }
    const unsigned long long current = shfl_u64(p < k ? cache[v * k + p] : 0, source, G);
    if (p != 0) return;
    int best = bp == 0x7fffffff ? -1 : bp;
    if (best >= 0 && conn == 0) {
        std::uint64_t total_slack = 0;
        for (int q = 0; q < k; ++q)
            if (q != source && pw[q] < capacity && weight <= capacity - pw[q])
                total_slack = ~0ULL - total_slack < capacity - pw[q] ? ~0ULL : total_slack + capacity - pw[q];
        std::uint64_t ticket = static_cast<std::uint64_t>(mix32(static_cast<std::uint32_t>(v))) % total_slack;
        for (int q = 0; q < k; ++q) {
            if (q == source || pw[q] >= capacity || weight > capacity - pw[q]) continue;
            const std::uint64_t s = capacity - pw[q];
            if (ticket < s) {
                best = q;
                break;
            }
            ticket -= s;
        }
    }
    if (best >= 0) {
        const double gain = conn >= current ? static_cast<double>(conn - current) : -static_cast<double>(current - conn);
        const double per_weight = gain / static_cast<double>(weight);
        buckets[v] = source * kBalanceWidth + gain_bucket(per_weight, powers) * kBalanceSections +
                     static_cast<int>(mix32(static_cast<std::uint32_t>(v)) % kBalanceSections);
        targets[v] = best;
    } else {
        buckets[v] = -1;
        targets[v] = -1;
    }
}

__global__ void bucket_flags_kernel(std::int64_t n, const std::int32_t* buckets, std::int32_t* flags) {
    const std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v < n) flags[v] = buckets[v] >= 0;
}

template <typename W>
__global__ void bucket_keys_kernel(std::int64_t p, const std::int32_t* list, const std::int32_t* buckets,
                                   std::uint64_t* keys) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < p) keys[i] = static_cast<std::uint64_t>(buckets[list[i]]);
}

// Segmented exclusive rank of weights by bucket; bucket totals at segment ends.
template <typename W>
struct RankLoad {
    const std::int32_t* order;
    const std::int32_t* buckets;
    const W* weights;
    __device__ __forceinline__ SegValue operator()(std::int64_t i) const {
        const std::int32_t v = order[i];
        const bool head = i == 0 || buckets[order[i - 1]] != buckets[v];
        return SegValue{static_cast<unsigned long long>(weights[v]), head ? 1U : 0U};
    }
};
template <typename W>
struct RankStore {
    const std::int32_t* order;
    const std::int32_t* buckets;
    const W* weights;
    std::int64_t n;
    unsigned long long* ranks;   // by vertex
    unsigned long long* counts;  // by bucket
    __device__ __forceinline__ void operator()(std::int64_t i, const SegValue& prefix, const SegValue& item) const {
        const SegValue incl = SegOp{}(prefix, item);
        const std::int32_t v = order[i];
        ranks[v] = incl.value - static_cast<unsigned long long>(weights[v]);
        const bool last = i == n - 1 || buckets[order[i + 1]] != buckets[v];
        if (last) counts[buckets[v]] = incl.value;
    }
};

// balance_scan_buckets (:188-202): exclusive sum over the 800 buckets of each part.
__global__ void __launch_bounds__(kBlock) scan_buckets_kernel(const unsigned long long* counts,
                                                              unsigned long long* offsets) {
    __shared__ unsigned long long waves[kBlock / kWave + 1];
    constexpr int kPer = (kBalanceWidth + kBlock - 1) / kBlock;
    unsigned long long items[kPer];
    unsigned long long local = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int b = threadIdx.x * kPer + i;
        items[i] = b < kBalanceWidth ? counts[static_cast<std::int64_t>(blockIdx.x) * kBalanceWidth + b] : 0;
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
}

// balance_select_sources (:204-217)
__global__ void select_sources_kernel(std::int64_t p, const std::int32_t* list, std::int32_t* buckets,
                                      const unsigned long long* ranks, const unsigned long long* offsets,
                                      const unsigned long long* part_weights, std::uint64_t capacity,
                                      const std::int32_t* targets) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= p) return;
    const std::int32_t v = list[i];
    const int b = buckets[v];
    const int source = b / kBalanceWidth;
    if (offsets[b] + ranks[v] >= part_weights[source] - capacity) {
        buckets[v] = -1;
        return;
    }
    buckets[v] = targets[v] * kBalanceWidth + b % kBalanceWidth;
}

// balance_select_destinations (:219-232)
template <typename W>
__global__ void select_destinations_kernel(std::int64_t n, const std::int32_t* buckets,
                                           const unsigned long long* ranks, const unsigned long long* offsets,
                                           const unsigned long long* part_weights, std::uint64_t capacity,
                                           const W* weights, std::int32_t* flags) {
    const std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int b = buckets[v];
    std::int32_t f = 0;
    if (b >= 0) {
        const unsigned long long before = offsets[b] + ranks[v];
        const unsigned long long slack = capacity - part_weights[b / kBalanceWidth];
        f = before <= slack && static_cast<unsigned long long>(weights[v]) <= slack - before;
    }
    flags[v] = f;
}

// ---------------------------------------------------------------- drivers

int group_for_k(int k) {
    int g = 2;
    while (g < k) g <<= 1;
    return g;
}

PowerTable make_power_table() {
    PowerTable t;
    for (int j = -23; j <= 24; ++j) {
        double p;
        if (j >= 0) {
            double three = 1.0;
            for (int i = 0; i < j; ++i) three *= 3.0;
            p = std::ldexp(three, -j);
        } else {
            double three = 1.0;
            for (int i = 0; i < -j; ++i) three *= 3.0;
            p = std::ldexp(1.0, -j) / three;
        }
        t.value[j + 23] = p;
    }
    return t;
}

template <typename O, typename W>
struct Ops {
    static void cut(gpart_ctx* ctx, const gpart_csr_view* g, const std::int32_t* labels, std::uint64_t* out) {
        Scratch<unsigned long long> d(ctx, 1);
        fill_bytes(ctx, d.get(), 0, 8);
        const auto* off = static_cast<const O*>(g->offsets);
        const auto* w = static_cast<const W*>(g->edge_weights);
        if (g->n > 0) {
            // 16 lanes per vertex, capped grid, partials folded by one workgroup
            const int blocks = static_cast<int>(std::min<std::int64_t>(grid_for(g->n * 16), kReduceMaxBlocks));
            Scratch<unsigned long long> partials(ctx, blocks);
            hipLaunchKernelGGL((cut_kernel<O, W, 16>), dim3(blocks), dim3(kBlock), 0, ctx->stream, g->n, nullptr, off,
                               g->neighbors, w, labels, partials.get());
            hipLaunchKernelGGL((reduce_finish_kernel<unsigned long long, SumU64>), dim3(1), dim3(kReduceThreads), 0,
                               ctx->stream, partials.get(), blocks, SumU64{}, d.get());
            GPART_LAUNCH_CHECK();
        }
        const auto directed = read_scalar(ctx, d.get());
        if (directed & 1ULL) throw Error(GPART_ERR_RUNTIME, "refinement cut is not symmetric");
        *out = directed / 2;
    }

    static void build_cache(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* c,
                            const std::int32_t* labels, int k, unsigned long long* cache) {
        const auto* off = static_cast<const O*>(g->offsets);
        const auto* w = static_cast<const W*>(g->edge_weights);
        profile_begin(ctx, GPART_PROFILE_REFINE_CACHE);
        if (c->low_count)
            hipLaunchKernelGGL((build_cache_kernel<O, W, 8>), dim3(grid_for(c->low_count * 8)), dim3(kBlock), 0,
                               ctx->stream, c->low_count, c->low, off, g->neighbors, w, labels, k, cache);
        if (c->medium_count)
            hipLaunchKernelGGL((build_cache_kernel<O, W, 64>), dim3(grid_for(c->medium_count * 64)), dim3(kBlock), 0,
                               ctx->stream, c->medium_count, c->medium, off, g->neighbors, w, labels, k, cache);
        if (c->high_count)
            hipLaunchKernelGGL((build_cache_kernel<O, W, kBlock>), dim3(static_cast<unsigned>(c->high_count)),
                               dim3(kBlock), 0, ctx->stream, c->high_count, c->high, off, g->neighbors, w, labels, k,
                               cache);
        GPART_LAUNCH_CHECK();
        const double n = static_cast<double>(g->n), m = static_cast<double>(g->m);
        profile_end(ctx, GPART_PROFILE_REFINE_CACHE,
                    (n + 1) * sizeof(O) + m * (4 + sizeof(W)) + 4 * m + 8.0 * n * k);
    }

    static void lp_phase(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* c,
                         const unsigned long long* cache, const std::int32_t* labels, int k, double filter,
                         std::int32_t* locks, std::int32_t* targets, long long* gains, std::int32_t* flags,
                         std::int32_t* movers, std::int64_t* count) {
        const std::int64_t n = g->n;
        const auto* off = static_cast<const O*>(g->offsets);
        const auto* w = static_cast<const W*>(g->edge_weights);
        const int G = group_for_k(k);
        profile_begin(ctx, GPART_PROFILE_REFINE_PROPOSALS);
        const unsigned blocks = grid_for(n * G);
        switch (G) {
            case 2: hipLaunchKernelGGL(lp_proposals_kernel<2>, dim3(blocks), dim3(kBlock), 0, ctx->stream, n, cache, labels, k, filter, locks, targets, gains); break;
            case 4: hipLaunchKernelGGL(lp_proposals_kernel<4>, dim3(blocks), dim3(kBlock), 0, ctx->stream, n, cache, labels, k, filter, locks, targets, gains); break;
            case 8: hipLaunchKernelGGL(lp_proposals_kernel<8>, dim3(blocks), dim3(kBlock), 0, ctx->stream, n, cache, labels, k, filter, locks, targets, gains); break;
            case 16: hipLaunchKernelGGL(lp_proposals_kernel<16>, dim3(blocks), dim3(kBlock), 0, ctx->stream, n, cache, labels, k, filter, locks, targets, gains); break;
            default: hipLaunchKernelGGL(lp_proposals_kernel<32>, dim3(blocks), dim3(kBlock), 0, ctx->stream, n, cache, labels, k, filter, locks, targets, gains); break;
        }
        GPART_LAUNCH_CHECK();
        profile_end(ctx, GPART_PROFILE_REFINE_PROPOSALS, 8.0 * n * k + 24.0 * n);
        profile_begin(ctx, GPART_PROFILE_REFINE_FILTER);
        if (c->low_count)
            hipLaunchKernelGGL((priority_filter_kernel<O, W, 8>), dim3(grid_for(c->low_count * 8)), dim3(kBlock), 0,
                               ctx->stream, c->low_count, c->low, off, g->neighbors, w, labels, targets, gains, flags);
        if (c->medium_count)
            hipLaunchKernelGGL((priority_filter_kernel<O, W, 64>), dim3(grid_for(c->medium_count * 64)), dim3(kBlock),
                               0, ctx->stream, c->medium_count, c->medium, off, g->neighbors, w, labels, targets, gains,
                               flags);
        if (c->high_count)
            hipLaunchKernelGGL((priority_filter_kernel<O, W, kBlock>), dim3(static_cast<unsigned>(c->high_count)),
                               dim3(kBlock), 0, ctx->stream, c->high_count, c->high, off, g->neighbors, w, labels,
                               targets, gains, flags);
        GPART_LAUNCH_CHECK();
        const double nd = static_cast<double>(n), m = static_cast<double>(g->m);
        profile_end(ctx, GPART_PROFILE_REFINE_FILTER, (nd + 1) * sizeof(O) + m * (4 + sizeof(W)) + 16 * m + 20 * nd);
        Scratch<std::int64_t> d_count(ctx, 1);
        compact_indices(ctx, flags, n, movers, d_count.get());
        *count = read_scalar(ctx, d_count.get());
    }

    static void apply(gpart_ctx* ctx, const gpart_csr_view* g, const std::int32_t* movers, std::int64_t count,
                      std::int32_t* labels, const std::int32_t* targets, const std::int32_t* flags, int k,
                      unsigned long long* cache, unsigned long long* part_weights, std::int32_t* locks, bool lp,
                      std::int64_t* delta, std::uint64_t* pw_host) {
        const auto* off = static_cast<const O*>(g->offsets);
        const auto* w = static_cast<const W*>(g->edge_weights);
        Scratch<long long> d(ctx, 1);
        fill_bytes(ctx, d.get(), 0, 8);
        if (count > 0) {
            Scratch<std::int32_t> light(ctx, count), heavy(ctx, count);
            Scratch<unsigned long long> counters(ctx, 2);
            fill_bytes(ctx, counters.get(), 0, 16);
            hipLaunchKernelGGL(split_movers_kernel<O>, dim3(grid_for(count)), dim3(kBlock), 0, ctx->stream, count,
                               movers, off, light.get(), heavy.get(), counters.get());
            GPART_LAUNCH_CHECK();
            unsigned long long h[2];
            read_back(ctx, h, counters.get(), sizeof(h));
            const auto nl = static_cast<std::int64_t>(h[0]), nh = static_cast<std::int64_t>(h[1]);
            if (nl)
                hipLaunchKernelGGL((update_cache_kernel<O, W, 8>), dim3(grid_for(nl * 8)), dim3(kBlock), 0, ctx->stream,
                                   nl, light.get(), off, g->neighbors, w, labels, targets, flags, k, cache, d.get());
            if (nh)
                hipLaunchKernelGGL((update_cache_kernel<O, W, kBlock>), dim3(static_cast<unsigned>(nh)), dim3(kBlock), 0,
                                   ctx->stream, nh, heavy.get(), off, g->neighbors, w, labels, targets, flags, k, cache,
                                   d.get());
            hipLaunchKernelGGL(commit_kernel<W>, dim3(grid_for(count)), dim3(kBlock), 0, ctx->stream, count, movers,
                               labels, targets, static_cast<const W*>(g->vertex_weights), part_weights, locks, lp);
            GPART_LAUNCH_CHECK();
        }
        *delta = read_scalar(ctx, d.get());
        read_back(ctx, pw_host, part_weights, sizeof(std::uint64_t) * k);
    }

    static void rebalance_select(gpart_ctx* ctx, const gpart_csr_view* g, const unsigned long long* cache,
                                 const std::int32_t* labels, int k, const unsigned long long* part_weights,
                                 std::uint64_t capacity, std::int32_t* targets, std::int32_t* flags,
                                 std::int32_t* movers, std::int64_t* count) {
        const std::int64_t n = g->n;
        const auto* vw = static_cast<const W*>(g->vertex_weights);
        static const PowerTable powers = make_power_table();
        Scratch<std::int32_t> buckets(ctx, n), list(ctx, n), list_alt(ctx, n);
        Scratch<unsigned long long> ranks(ctx, n), counts(ctx, static_cast<std::int64_t>(k) * kBalanceWidth),
            offsets(ctx, static_cast<std::int64_t>(k) * kBalanceWidth);
        Scratch<std::int64_t> d_count(ctx, 1);
        hipLaunchKernelGGL(overloaded_flags_kernel, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, labels,
                           part_weights, k, capacity, flags, targets, buckets.get());
        GPART_LAUNCH_CHECK();
        compact_indices(ctx, flags, n, list.get(), d_count.get());
        const std::int64_t candidates = read_scalar(ctx, d_count.get());
        const int G = group_for_k(k);
        const unsigned blocks = grid_for(candidates * G);
        const std::int32_t* cl = list.get();
        if (candidates > 0) {
            switch (G) {
                case 2: hipLaunchKernelGGL((rebalance_proposals_kernel<W, 2>), dim3(blocks), dim3(kBlock), 0, ctx->stream, candidates, cl, cache, labels, vw, k, part_weights, capacity, powers, targets, buckets.get()); break;
                case 4: hipLaunchKernelGGL((rebalance_proposals_kernel<W, 4>), dim3(blocks), dim3(kBlock), 0, ctx->stream, candidates, cl, cache, labels, vw, k, part_weights, capacity, powers, targets, buckets.get()); break;
                case 8: hipLaunchKernelGGL((rebalance_proposals_kernel<W, 8>), dim3(blocks), dim3(kBlock), 0, ctx->stream, candidates, cl, cache, labels, vw, k, part_weights, capacity, powers, targets, buckets.get()); break;
                case 16: hipLaunchKernelGGL((rebalance_proposals_kernel<W, 16>), dim3(blocks), dim3(kBlock), 0, ctx->stream, candidates, cl, cache, labels, vw, k, part_weights, capacity, powers, targets, buckets.get()); break;
                default: hipLaunchKernelGGL((rebalance_proposals_kernel<W, 32>), dim3(blocks), dim3(kBlock), 0, ctx->stream, candidates, cl, cache, labels, vw, k, part_weights, capacity, powers, targets, buckets.get()); break;
            }
            GPART_LAUNCH_CHECK();
        }
        const int bucket_bits = std::max(1, bit_width_u64(static_cast<std::uint64_t>(k) * kBalanceWidth - 1));
        // two ranking rounds: by source bucket, then by destination bucket
        for (int round = 0; round < 2; ++round) {
            hipLaunchKernelGGL(bucket_flags_kernel, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, buckets.get(),
                               flags);
            GPART_LAUNCH_CHECK();
            compact_indices(ctx, flags, n, list.get(), d_count.get());
            const std::int64_t p = read_scalar(ctx, d_count.get());
            fill_bytes(ctx, counts.get(), 0, static_cast<std::size_t>(k) * kBalanceWidth * 8);
            std::int32_t* sorted = list.get();
            if (p > 0) {
                if (p > 1) {
                    Scratch<std::uint64_t> keys(ctx, p), keys_alt(ctx, p);
                    hipLaunchKernelGGL(bucket_keys_kernel<W>, dim3(grid_for(p)), dim3(kBlock), 0, ctx->stream, p,
                                       list.get(), buckets.get(), keys.get());
                    GPART_LAUNCH_CHECK();
                    if (radix_sort_pairs(ctx, keys.get(), keys_alt.get(), reinterpret_cast<std::uint32_t*>(list.get()),
                                         reinterpret_cast<std::uint32_t*>(list_alt.get()), p, 0, bucket_bits))
                        sorted = list_alt.get();
                }
                device_scan<SegValue, SegOp>(ctx, RankLoad<W>{sorted, buckets.get(), vw},
                                             RankStore<W>{sorted, buckets.get(), vw, p, ranks.get(), counts.get()}, p,
                                             SegOp{});
            }
            hipLaunchKernelGGL(scan_buckets_kernel, dim3(k), dim3(kBlock), 0, ctx->stream, counts.get(), offsets.get());
            GPART_LAUNCH_CHECK();
            if (round == 0) {
                if (p > 0) {
                    hipLaunchKernelGGL(select_sources_kernel, dim3(grid_for(p)), dim3(kBlock), 0, ctx->stream, p,
                                       sorted, buckets.get(), ranks.get(), offsets.get(), part_weights, capacity,
                                       targets);
                    GPART_LAUNCH_CHECK();
                }
            } else {
                hipLaunchKernelGGL(select_destinations_kernel<W>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n,
                                   buckets.get(), ranks.get(), offsets.get(), part_weights, capacity, vw, flags);
                GPART_LAUNCH_CHECK();
            }
        }
        compact_indices(ctx, flags, n, movers, d_count.get());
        *count = read_scalar(ctx, d_count.get());
    }
};

}  // namespace
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

extern "C" {

int gpart_cut(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* labels, uint64_t* cut) {
    return guarded([&] { GPART_WIDTH_DISPATCH(g->width, OpsT::cut(ctx, g, labels, cut)); });
}

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
                             const int32_t* labels, int k, uint64_t* cache) {
    return guarded([&] {
        GPART_WIDTH_DISPATCH(g->width, OpsT::build_cache(ctx, g, classes, labels, k,
                                                         reinterpret_cast<unsigned long long*>(cache)));
    });
}

int gpart_refine_lp_phase(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* classes,
                          const uint64_t* cache, const int32_t* labels, int k, double filter, int32_t* locks,
                          int32_t* targets, int64_t* gains, int32_t* flags, int32_t* movers, int64_t* count) {
    return guarded([&] {
        GPART_WIDTH_DISPATCH(g->width,
                             OpsT::lp_phase(ctx, g, classes, reinterpret_cast<const unsigned long long*>(cache), labels,
                                            k, filter, locks, targets, reinterpret_cast<long long*>(gains), flags,
                                            movers, count));
    });
}

int gpart_refine_apply(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* movers, int64_t count, int32_t* labels,
                       const int32_t* targets, const int32_t* flags, int k, uint64_t* cache, uint64_t* part_weights,
                       int32_t* locks, int lp, int64_t* delta, uint64_t* part_weights_host) {
    return guarded([&] {
        GPART_WIDTH_DISPATCH(g->width,
                             OpsT::apply(ctx, g, movers, count, labels, targets, flags, k,
                                         reinterpret_cast<unsigned long long*>(cache),
                                         reinterpret_cast<unsigned long long*>(part_weights), locks, lp != 0, delta,
                                         part_weights_host));
    });
}

int gpart_refine_rebalance_select(gpart_ctx* ctx, const gpart_csr_view* g, const uint64_t* cache,
                                  const int32_t* labels, int k, const uint64_t* part_weights, uint64_t capacity,
                                  int32_t* targets, int32_t* flags, int32_t* movers, int64_t* count) {
    return guarded([&] {
        GPART_WIDTH_DISPATCH(g->width,
                             OpsT::rebalance_select(ctx, g, reinterpret_cast<const unsigned long long*>(cache), labels,
                                                    k, reinterpret_cast<const unsigned long long*>(part_weights),
                                                    capacity, targets, flags, movers, count));
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
