// refine_engine.hip -- one level of incremental refinement in a single launch
// (run_incremental_refinement, src/refine_incremental.cuh:235-361 of the reference).
//
// The launch-per-stage path (refine.hip) spends about 17 dispatches and one host
// round trip on every rebalance iteration, and 10 dispatches on every LP round.
// On the small and middle levels of a hierarchy that overhead, not HBM, sets the
// time. Here the whole phase loop -- LP rounds, their rebalance iterations, the
// stale-round rule, checkpoints and the final restore -- runs inside one
// launch of at most kEngineMaxGrid resident workgroups of 512 lanes. Stages
// are separated by a grid barrier, which on MI355X costs 1-3 us for up to 64
// workgroups (measured), below the 2.5 us of a back-to-back dispatch; every
// workgroup evaluates the loop's decisions from the same values read after a
// barrier, so control flow is uniform across the grid.
//
// Work layout: vertex stages give workgroup b the contiguous id slice
// [b * per, (b + 1) * per), so the bucket ranks of the rebalance (weight of the
// lower-id members of a bucket) come from a prefix over workgroups plus an
// in-order pass inside each slice. A step's label commit is deferred to the
// first vertex pass of the next step (flags and targets say what moved); the
// part weights move in the cache-update stage, so they are final after its
// barrier.
//
// Results are identical to the launch-per-stage path: the same proposals,
// filters, bucket quotas, ranks and commits; only the schedule differs.
#include "internal.hpp"
#include "reduce.hpp"
#include "refine_device.hpp"
#include "scan.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>

namespace gpart_hip {
namespace {

constexpr int kEngineThreads = 512;
constexpr int kEngineWaves = kEngineThreads / kWave;
constexpr int kEngineMaxGrid = 128;
constexpr int kEngineHeavyDegree = 64;  // movers above take a wave in the cache update
constexpr unsigned long long kEngineSpinLimit = 1ULL << 22;  // about 4 s of polling per barrier

// Barrier words, each on its own 128-byte line.
struct EngineSync {
    unsigned long long arrive;  // cumulative arrivals
    unsigned long long pad0[15];
    unsigned long long release;  // last released epoch
    unsigned long long pad1[15];
    unsigned int abort;  // set when a barrier wait exceeds kEngineSpinLimit
    unsigned int pad2[31];
};

// Accumulators of one step (an LP round or a rebalance iteration); three rotate so
// that the one of step s + 1 can be cleared while step s - 1's is still read.
struct StepAccum {
    unsigned long long light;  // movers of degree <= kEngineHeavyDegree
    unsigned long long heavy;
    unsigned long long edges;  // CSR entries of all movers
    unsigned long long delta;  // exact cut decrease, two's complement
    unsigned long long touched;  // bit p: a destination bucket of part p is non-empty
    unsigned long long pad[11];
};

}  // namespace
}  // namespace gpart_hip

// The host-visible outcome of a level (include/gpart_hip.h declares the same
// layout as gpart_refine_level_stats).
struct EngineStats {
    int32_t rounds;
    int32_t admission_iterations;
    int32_t admission_fallbacks;
    int32_t error;  // 0 ok, 1 cut underflow, 2 barrier timeout
    uint64_t proposals;
    uint64_t accepted;
    uint64_t mover_entries;
    uint64_t best_cut;
    uint64_t steps;
};
static_assert(sizeof(EngineStats) == sizeof(gpart_refine_level_stats), "stats layout");

namespace gpart_hip {
namespace {

template <typename O, typename W, typename C>
struct EngineArgs {
    std::int64_t n;
    const O* offsets;
    const std::int32_t* neighbors;
    const W* edge_weights;
    const W* vertex_weights;
    const std::int32_t* low;
    const std::int32_t* medium;
    const std::int32_t* high;
    std::int64_t low_count;
    std::int64_t medium_count;
    std::int64_t high_count;
    int k;
    int rounds;
    int stale_rounds;
    int pad;
    std::uint64_t capacity;
    double filter;
    std::uint64_t initial_cut;
    std::uint64_t keep_mantissa;  // 0.999L = keep_mantissa * 2^-64
    C* cache;
    std::int32_t* labels;
    unsigned long long* part_weights;
    std::int32_t* targets;
    long long* gains;
    std::int32_t* flags;
    std::int32_t* locks;
    std::int32_t* best_labels;
    unsigned long long* best_part_weights;
    std::int32_t* light;
    std::int32_t* heavy;
    std::int32_t* cbucket;
    std::int32_t* dbucket;
    unsigned long long* src_counts;   // k * kBalanceWidth
    unsigned long long* dst_counts;
    unsigned long long* src_offsets;
    unsigned long long* dst_offsets;
    unsigned long long* thresholds;   // [0, 32) sources, [32, 64) destinations
    unsigned long long* boundary;     // gridDim.x * 32 per-workgroup boundary weights
    StepAccum* accum;                 // 3
    EngineSync* sync;
    EngineStats* stats;
    PowerTable powers;
};

// LDS of one workgroup.
struct EngineShared {
    BucketTable table;
    unsigned long long wave_total[kEngineWaves][32];
    unsigned long long running[32];
    unsigned long long pw[32];
    unsigned long long gained[32];
    unsigned long long lost[32];
    unsigned long long sums[32];
    unsigned long long chunk_sums[32][16];
    long long reduce[kEngineWaves];
    unsigned long long touched;  // destination parts of this workgroup's taken buckets
    int barrier_ok;
};

// (long double)c < 0.999L * (long double)b on x86-64, exactly: the product of the
// 64-bit significand m (0.999L = m * 2^-64) and b, rounded to 64 significant bits
// to nearest even as the x87 unit does, compared with c. b < 2^63.
__host__ __device__ inline bool below_keep_ratio(std::uint64_t c, std::uint64_t b, std::uint64_t m) {
    const unsigned __int128 x = static_cast<unsigned __int128>(m) * b;
    if (x == 0) return false;
    const std::uint64_t hi = static_cast<std::uint64_t>(x >> 64);
    const int bits = hi ? 128 - __builtin_clzll(hi) : 64 - __builtin_clzll(static_cast<std::uint64_t>(x));
    unsigned __int128 r = x;
    if (bits > 64) {
        const int shift = bits - 64;
        unsigned __int128 q = x >> shift;
        const unsigned __int128 rem = x - (q << shift);
        const unsigned __int128 half = static_cast<unsigned __int128>(1) << (shift - 1);
        if (rem > half || (rem == half && (q & 1))) q += 1;
        r = q << shift;
    }
    return (static_cast<unsigned __int128>(c) << 64) < r;
}

// Grid barrier: every wave drains its memory operations, one lane per workgroup
// writes back the L2 and arrives; the last arrival releases the epoch. A wait
// longer than kEngineSpinLimit polls raises the abort word, which every waiter
// also watches, so a grid can never hang here. Returns false on abort.
__device__ bool engine_barrier(EngineSync* sync, unsigned long long& epoch, EngineShared& sh) {
    ++epoch;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned long long target = epoch * gridDim.x;
        const unsigned long long before =
            __hip_atomic_fetch_add(&sync->arrive, 1ULL, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (before + 1 == target) {
            __hip_atomic_store(&sync->release, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned long long spins = 0;
            while (__hip_atomic_load(&sync->release, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
                if (++spins > kEngineSpinLimit ||
                    __hip_atomic_load(&sync->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    __hip_atomic_store(&sync->abort, 1U, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        sh.barrier_ok = ok;
    }
    __syncthreads();
    return sh.barrier_ok != 0;
}

__device__ __forceinline__ unsigned long long load_u64(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Contiguous slice of count items owned by this workgroup.
__device__ __forceinline__ void engine_slice(std::int64_t count, std::int64_t& lo, std::int64_t& hi) {
    const std::int64_t per = (count + gridDim.x - 1) / gridDim.x;
    lo = min(count, static_cast<std::int64_t>(blockIdx.x) * per);
    hi = min(count, lo + per);
}

// Sum of a per-thread value over the workgroup, returned to every thread.
__device__ long long block_sum_i64(long long v, EngineShared& sh) {
    v = group_sum_i64<kWave>(v);
    if (lane_id() == 0) sh.reduce[threadIdx.x / kWave] = v;
    __syncthreads();
    long long total = 0;
#pragma unroll
    for (int w = 0; w < kEngineWaves; ++w) total += sh.reduce[w];
    __syncthreads();
    return total;
}

// Appends the wave's taking lanes to the light / heavy mover lists and flags them.
// Called by every lane of a wave.
template <typename O>
__device__ void append_movers_wave(bool take, std::int32_t v, const O* offsets, std::int32_t* light,
                                   std::int32_t* heavy, std::int32_t* flags, StepAccum* acc) {
    long long degree = 0;
    if (take) degree = static_cast<long long>(offsets[v + 1] - offsets[v]);
    const bool is_heavy = take && degree > kEngineHeavyDegree;
    const bool is_light = take && !is_heavy;
    const unsigned long long bl = __ballot(is_light), bh = __ballot(is_heavy);
    if (!(bl | bh)) return;
    const long long edges = group_sum_i64<kWave>(degree);
    unsigned long long base_l = 0, base_h = 0;
    if (lane_id() == 0) {
        if (bl) base_l = atomicAdd(&acc->light, static_cast<unsigned long long>(__popcll(bl)));
        if (bh) base_h = atomicAdd(&acc->heavy, static_cast<unsigned long long>(__popcll(bh)));
        atomicAdd(&acc->edges, static_cast<unsigned long long>(edges));
    }
    base_l = shfl_u64(base_l, 0, kWave);
    base_h = shfl_u64(base_h, 0, kWave);
    if (take) {
        const unsigned long long mask = is_heavy ? bh : bl;
        const unsigned long long at = (is_heavy ? base_h : base_l) + __popcll(mask & lanemask_lt());
        (is_heavy ? heavy : light)[at] = v;
        flags[v] = 1;
    }
}

// ---------------------------------------------------------------- stages

// The deferred commit of the previous step and the checkpoint copy, over this
// workgroup's vertex slice; clear_flags also zeroes the slice's flags.
template <typename O, typename W, typename C>
__device__ void pass_commit(const EngineArgs<O, W, C>& a, bool commit, bool lp, bool checkpoint, bool clear_flags,
                            std::int64_t lo, std::int64_t hi) {
    if (!commit && !checkpoint && !clear_flags) return;
    for (std::int64_t v = lo + threadIdx.x; v < hi; v += kEngineThreads) {
        std::int32_t label = 0;
        if (commit || checkpoint) label = a.labels[v];
        if (commit && a.flags[v]) {
            label = a.targets[v];
            a.labels[v] = label;
            if (lp) a.locks[v] = 1;
        }
        if (checkpoint) a.best_labels[v] = label;
        if (clear_flags) a.flags[v] = 0;
    }
    __syncthreads();
}

// incremental_lp_proposals (:27-46) over the slice, G lanes per row.
template <int G, typename O, typename W, typename C>
__device__ void stage_lp_proposals(const EngineArgs<O, W, C>& a, std::int64_t lo, std::int64_t hi) {
    const int j = threadIdx.x & (G - 1);
    for (std::int64_t base = lo; base < hi; base += kEngineThreads / G) {
        const std::int64_t v = base + threadIdx.x / G;
        if (v >= hi) continue;  // uniform per group
        const std::int32_t source = a.labels[v];
        const C* row = a.cache + v * a.k;
        unsigned long long best = 0, current = 0;
        int best_p = 0x7fffffff;
        unsigned long long span[kRowSpan];
        load_span(row, j, a.k, span);
#pragma unroll
        for (int e = 0; e < kRowSpan; ++e) {
            const int p = j * kRowSpan + e;
            const unsigned long long value = span[e];
            if (p == source) current = value;
            else if (value > best) {
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
            current = oc > current ? oc : current;
        }
        if (j != 0) continue;
        if (a.locks[v] != 0) {
            a.locks[v] = 0;
            a.targets[v] = -1;
            a.gains[v] = static_cast<long long>(0x8000000000000000ULL);
            continue;
        }
        std::int32_t target = -1;
        long long gain = static_cast<long long>(0x8000000000000000ULL);
        if (best > 0 &&
            (best >= current ||
             current - best < static_cast<unsigned long long>(a.filter * static_cast<double>(current)))) {
            target = best_p;
            gain = best >= current ? static_cast<long long>(best - current) : -static_cast<long long>(current - best);
        }
        a.targets[v] = target;
        a.gains[v] = gain;
    }
}

// incremental_priority_filter (:48-67) over one class list with Gf lanes per row
// (Gf = 8, 64 or the whole workgroup); flagged vertices become movers.
template <int Gf, typename O, typename W, typename C>
__device__ void stage_filter_list(const EngineArgs<O, W, C>& a, const std::int32_t* list, std::int64_t count,
                                  StepAccum* acc, EngineShared& sh) {
    std::int64_t lo, hi;
    engine_slice(count, lo, hi);
    constexpr int kPer = kEngineThreads / Gf;
    const int sub = threadIdx.x & (Gf - 1);
    for (std::int64_t base = lo; base < hi; base += kPer) {
        const std::int64_t i = base + threadIdx.x / Gf;
        const bool active = i < hi;
        const std::int32_t v = active ? list[i] : 0;
        const std::int32_t target = active ? a.targets[v] : -1;
        long long correction = 0;
        if (target >= 0)
            correction = filter_correction(v, target, a.labels[v], a.gains[v],
                                           static_cast<std::int64_t>(a.offsets[v]) + sub,
                                           static_cast<std::int64_t>(a.offsets[v + 1]), Gf, a.neighbors,
                                           a.edge_weights, a.labels, a.targets, a.gains);
        bool take = false;
        if constexpr (Gf <= kWave) {
            correction = group_sum_i64<Gf>(correction);
            if (active && sub == 0) {
                take = target >= 0 && a.gains[v] + correction >= 0;
                a.flags[v] = take ? 1 : 0;
            }
            append_movers_wave(take, v, a.offsets, a.light, a.heavy, a.flags, acc);
        } else {
            const long long total = block_sum_i64(correction, sh);
            if (threadIdx.x == 0 && active) {
                take = target >= 0 && a.gains[v] + total >= 0;
                a.flags[v] = take ? 1 : 0;
            }
            if (threadIdx.x < kWave) append_movers_wave(take, v, a.offsets, a.light, a.heavy, a.flags, acc);
        }
    }
}

// incremental_update_cache_and_cut (:69-95) for the step's movers, with the part
// weight moves (incremental_commit :97-109, weights only). Light movers take 8
// lanes, heavy movers a wave.
template <typename O, typename W, typename C>
__device__ void stage_update(const EngineArgs<O, W, C>& a, StepAccum* acc, EngineShared& sh) {
    if (threadIdx.x < 32) sh.gained[threadIdx.x] = sh.lost[threadIdx.x] = 0;
    __syncthreads();
    const int k = a.k;
    long long gain = 0;
    auto mover = [&](std::int32_t v, int sub, int step) {
        const std::int32_t source = a.labels[v];
        const std::int32_t target = a.targets[v];
        if (sub == 0) {
            const auto w = static_cast<unsigned long long>(a.vertex_weights[v]);
            atomicAdd(&sh.lost[source], w);
            atomicAdd(&sh.gained[target], w);
        }
        for (std::int64_t e = static_cast<std::int64_t>(a.offsets[v]) + sub; e < a.offsets[v + 1]; e += step) {
            const std::int32_t u = a.neighbors[e];
            if (u == v) continue;
            const auto w = static_cast<unsigned long long>(a.edge_weights[e]);
            atomicAdd(a.cache + static_cast<std::int64_t>(u) * k + source, static_cast<C>(0ULL - w));
            atomicAdd(a.cache + static_cast<std::int64_t>(u) * k + target, static_cast<C>(w));
            const bool fu = a.flags[u] != 0;
            if (!fu || u > v) {
                const std::int32_t lu = a.labels[u];
                const std::int32_t nu = fu ? a.targets[u] : lu;
                gain += static_cast<long long>(w) * ((source != lu) - (target != nu));
            }
        }
    };
    {
        const auto count = static_cast<std::int64_t>(load_u64(&acc->light));
        std::int64_t lo, hi;
        engine_slice(count, lo, hi);
        for (std::int64_t i = lo + threadIdx.x / 8; i < hi; i += kEngineThreads / 8)
            mover(a.light[i], threadIdx.x & 7, 8);
    }
    {
        const auto count = static_cast<std::int64_t>(load_u64(&acc->heavy));
        std::int64_t lo, hi;
        engine_slice(count, lo, hi);
        for (std::int64_t i = lo + threadIdx.x / kWave; i < hi; i += kEngineWaves)
            mover(a.heavy[i], lane_id(), kWave);
    }
    const long long total = block_sum_i64(gain, sh);
    if (threadIdx.x == 0 && total) atomicAdd(&acc->delta, static_cast<unsigned long long>(total));
    if (threadIdx.x < k) {
        const unsigned long long change = sh.gained[threadIdx.x] - sh.lost[threadIdx.x];
        if (change) atomicAdd(a.part_weights + threadIdx.x, change);
    }
}

// incremental_rebalance_proposals (:111-157) + balance_source_histogram (:174-186)
// for the slice's vertices in overloaded parts, G lanes per row.
template <int G, typename O, typename W, typename C>
__device__ void stage_balance_proposals(const EngineArgs<O, W, C>& a, std::int64_t lo, std::int64_t hi,
                                        EngineShared& sh) {
    const int k = a.k;
    const std::uint64_t capacity = a.capacity;
    const int j = threadIdx.x & (G - 1);
    for (std::int64_t base = lo; base < hi; base += kEngineThreads / G) {
        const std::int64_t v = base + threadIdx.x / G;
        if (v >= hi) continue;  // uniform per group
        const std::int32_t source = a.labels[v];
        if (sh.pw[source] <= capacity) continue;  // not a candidate: uniform per group
        const std::uint64_t weight = static_cast<std::uint64_t>(a.vertex_weights[v]);
        const C* row = a.cache + v * k;
        unsigned long long slack[kRowSpan];
        unsigned long long conn = 0, best_slack = 0, current = 0;
        int bp = 0x7fffffff;
        unsigned long long span[kRowSpan];
        load_span(row, j, k, span);
#pragma unroll
        for (int e = 0; e < kRowSpan; ++e) {
            const int p = j * kRowSpan + e;
            const unsigned long long value = span[e];
            if (p == source) current = value;
            const bool ok = p < k && p != source && sh.pw[p] < capacity && weight <= capacity - sh.pw[p];
            slack[e] = ok ? capacity - sh.pw[p] : 0;
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
            current = ocur > current ? ocur : current;
        }
        int best = bp == 0x7fffffff ? -1 : bp;
        if (best >= 0 && conn == 0) {
            // ticket walk over the eligible parts (see balance_proposals_kernel)
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
            const unsigned long long total = shfl_u64(inclusive, G - 1, G);
            const std::uint32_t hash = mix32(static_cast<std::uint32_t>(v));
            const unsigned long long ticket = total > 0xffffffffULL ? hash : hash % static_cast<std::uint32_t>(total);
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
                bucket = source * kBalanceWidth + gain_bucket(per_weight, a.powers) * kBalanceSections +
                         static_cast<int>(mix32(static_cast<std::uint32_t>(v)) % kBalanceSections);
                table_add(sh.table, static_cast<unsigned>(bucket), weight, a.src_counts);
            }
            a.targets[v] = best;
            a.cbucket[v] = bucket;
        }
    }
}

// balance_scan_buckets (:188-202) for the parts this workgroup owns (parts
// blockIdx.x, blockIdx.x + grid, ...), all in one pass: 16 lanes per part, each
// summing a chunk of 50 buckets, then writing its chunk's exclusive prefixes.
// selected(p) names the parts whose buckets can be non-empty.
constexpr int kScanChunks = 16;
constexpr int kScanChunk = kBalanceWidth / kScanChunks;
static_assert(kScanChunk * kScanChunks == kBalanceWidth, "whole chunks");
static_assert(kEngineThreads / kScanChunks >= 32, "every owned part in one pass");
template <typename Selected>
__device__ void stage_bucket_scan(int k, const unsigned long long* counts, unsigned long long* offsets,
                                  unsigned long long* thresholds, bool destinations, std::uint64_t capacity,
                                  Selected selected, EngineShared& sh) {
    const int slot = threadIdx.x / kScanChunks, chunk = threadIdx.x % kScanChunks;
    const int p = static_cast<int>(blockIdx.x) + slot * static_cast<int>(gridDim.x);
    const bool mine = slot < 32 && p < k && selected(p);
    const long long first = static_cast<long long>(p) * kBalanceWidth + chunk * kScanChunk;
    unsigned long long sum = 0;
    if (mine) {
#pragma unroll 10
        for (int b = 0; b < kScanChunk; ++b) sum += counts[first + b];
    }
    if (slot < 32) sh.chunk_sums[slot][chunk] = sum;
    __syncthreads();
    if (mine) {
        unsigned long long running = 0;
        for (int c = 0; c < chunk; ++c) running += sh.chunk_sums[slot][c];
        for (int b = 0; b < kScanChunk; ++b) {
            offsets[first + b] = running;
            running += counts[first + b];
        }
        if (chunk == 0) {
            const unsigned long long w = sh.pw[p];
            thresholds[(destinations ? 32 : 0) + p] =
                destinations ? (w < capacity ? capacity - w : 0) : (w > capacity ? w - capacity : 0);
        }
    }
    __syncthreads();
}

// The exclusive prefix over lower workgroups of the per-part boundary weights,
// into sh.running.
__device__ void boundary_prefix(const unsigned long long* boundary, EngineShared& sh) {
    // thread t: part t % 32, workgroups t / 32, t / 32 + 32, ...
    const int p = threadIdx.x & 31, first = threadIdx.x >> 5;
    unsigned long long s = 0;
    for (unsigned b = first; b < blockIdx.x; b += kEngineThreads / 32) s += load_u64(boundary + b * 32 + p);
    if (threadIdx.x < 32) sh.running[threadIdx.x] = 0;
    __syncthreads();
    if (s) atomicAdd(&sh.running[p], s);
    __syncthreads();
}

// balance_select_sources (:204-217) over the slice: non-boundary buckets decided,
// boundary members summed per part for the ranks.
template <typename O, typename W, typename C>
__device__ void stage_sources(const EngineArgs<O, W, C>& a, std::int64_t lo, std::int64_t hi, EngineShared& sh) {
    for (std::int64_t v = lo + threadIdx.x; v < hi; v += kEngineThreads) {
        int d = kNoBucket;
        const std::int32_t s = a.labels[v];
        if (sh.pw[s] > a.capacity) {
            const int b = a.cbucket[v];
            if (b >= 0) {
                const auto w = static_cast<unsigned long long>(a.vertex_weights[v]);
                const BucketSide side = source_side(a.src_offsets[b], a.src_counts[b], a.thresholds[s]);
                if (side == kTaken) {
                    const int t = a.targets[v];
                    d = t * kBalanceWidth + b % kBalanceWidth;
                    table_add(sh.table, static_cast<unsigned>(d), w, a.dst_counts);
                    atomicOr(&sh.touched, 1ULL << t);
                } else if (side == kBoundary) {
                    d = kSourcePending;
                    atomicAdd(&sh.sums[s], w);
                }
            }
        }
        a.dbucket[v] = d;
    }
}

// Source boundary members in id order: leave while offset + rank < overload.
template <typename O, typename W, typename C>
__device__ void stage_source_ranks(const EngineArgs<O, W, C>& a, std::int64_t lo, std::int64_t hi,
                                   EngineShared& sh) {
    for (std::int64_t base = lo; base < hi; base += kEngineThreads) {
        const std::int64_t v = base + threadIdx.x;
        const bool pending = v < hi && a.dbucket[v] == kSourcePending;
        if (!__syncthreads_or(pending)) continue;
        int b = 0, s = 0;
        unsigned long long w = 0;
        if (pending) {
            b = a.cbucket[v];
            s = b / kBalanceWidth;
            w = static_cast<unsigned long long>(a.vertex_weights[v]);
        }
        const unsigned long long rank = block_key_prefix<kEngineThreads>(pending, s, w, sh.wave_total, sh.running);
        if (pending) {
            if (a.src_offsets[b] + rank < a.thresholds[s]) {
                const int t = a.targets[v];
                const int d = t * kBalanceWidth + b % kBalanceWidth;
                a.dbucket[v] = d;
                table_add(sh.table, static_cast<unsigned>(d), w, a.dst_counts);
                atomicOr(&sh.touched, 1ULL << t);
            } else {
                a.dbucket[v] = kNoBucket;
            }
        }
    }
}

// balance_select_destinations (:219-232) over the slice: taken buckets move,
// boundary members are summed per part.
template <typename O, typename W, typename C>
__device__ void stage_destinations(const EngineArgs<O, W, C>& a, std::int64_t lo, std::int64_t hi, StepAccum* acc,
                                   EngineShared& sh) {
    for (std::int64_t base = lo; base < hi; base += kEngineThreads) {
        const std::int64_t v = base + threadIdx.x;
        bool take = false;
        if (v < hi) {
            const int d = a.dbucket[v];
            if (d >= 0) {
                const int part = d / kBalanceWidth;
                const BucketSide side = destination_side(a.dst_offsets[d], a.dst_counts[d], a.thresholds[32 + part]);
                if (side == kTaken) {
                    take = true;
                } else if (side == kBoundary) {
                    a.dbucket[v] = -3 - d;
                    atomicAdd(&sh.sums[part], static_cast<unsigned long long>(a.vertex_weights[v]));
                }
            }
        }
        append_movers_wave(take, static_cast<std::int32_t>(v), a.offsets, a.light, a.heavy, a.flags, acc);
    }
}

// Destination boundary members in id order: enter while offset + rank + weight <= slack.
template <typename O, typename W, typename C>
__device__ void stage_destination_ranks(const EngineArgs<O, W, C>& a, std::int64_t lo, std::int64_t hi,
                                        StepAccum* acc, EngineShared& sh) {
    for (std::int64_t base = lo; base < hi; base += kEngineThreads) {
        const std::int64_t v = base + threadIdx.x;
        const int code = v < hi ? a.dbucket[v] : kNoBucket;
        const bool pending = code <= -3;
        if (!__syncthreads_or(pending)) continue;
        int d = 0, part = 0;
        unsigned long long w = 0;
        if (pending) {
            d = -3 - code;
            part = d / kBalanceWidth;
            w = static_cast<unsigned long long>(a.vertex_weights[v]);
        }
        const unsigned long long rank = block_key_prefix<kEngineThreads>(pending, part, w, sh.wave_total,
                                                                         sh.running);
        const bool take = pending && a.dst_offsets[d] + rank + w <= a.thresholds[32 + part];
        append_movers_wave(take, static_cast<std::int32_t>(v), a.offsets, a.light, a.heavy, a.flags, acc);
    }
}

// Zeroes a count array of k * kBalanceWidth words, spread over the grid.
__device__ void clear_counts(unsigned long long* counts, int k) {
    const std::int64_t total = static_cast<std::int64_t>(k) * kBalanceWidth;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * kEngineThreads + threadIdx.x; i < total;
         i += static_cast<std::int64_t>(gridDim.x) * kEngineThreads)
        counts[i] = 0;
}

__device__ void clear_accum(StepAccum* acc) {
    if (blockIdx.x == 0 && threadIdx.x < 5) reinterpret_cast<unsigned long long*>(acc)[threadIdx.x] = 0;
}

// ---------------------------------------------------------------- the level

template <int G, typename O, typename W, typename C>
__global__ void __launch_bounds__(kEngineThreads) refine_engine_kernel(EngineArgs<O, W, C> a) {
    __shared__ EngineShared sh;
    const int k = a.k;
    std::int64_t lo, hi;
    engine_slice(a.n, lo, hi);
    unsigned long long epoch = 0;
    std::uint64_t step = 0;
    auto acc_of = [&](std::uint64_t s) { return a.accum + (s % 3); };
    // Part weights as of the last barrier, into LDS; the maximum for the loop tests.
    auto read_weights = [&]() -> unsigned long long {
        if (threadIdx.x < 32) sh.pw[threadIdx.x] = threadIdx.x < k ? load_u64(a.part_weights + threadIdx.x) : 0;
        __syncthreads();
        unsigned long long m = 0;
        for (int p = 0; p < k; ++p) m = sh.pw[p] > m ? sh.pw[p] : m;
        return m;
    };

    std::uint64_t current = a.initial_cut, best = a.initial_cut;
    int stale = 0, empty = 0;
    int rounds = 0, iterations = 0, fallbacks = 0, error = 0;
    std::uint64_t proposals = 0, accepted = 0, entries = 0;
    bool commit = false, commit_lp = false, checkpoint = true;
    bool aborted = false;
    unsigned long long maximum = read_weights();
    if (blockIdx.x == 0 && threadIdx.x < k) a.best_part_weights[threadIdx.x] = sh.pw[threadIdx.x];

    // Absorbs a finished step; false when the loop must stop (underflow).
    auto absorb = [&](StepAccum* acc, unsigned long long movers) -> bool {
        const auto delta = static_cast<long long>(load_u64(&acc->delta));
        if (delta >= 0) {
            if (static_cast<std::uint64_t>(delta) > current) {
                error = 1;
                return false;
            }
            current -= static_cast<std::uint64_t>(delta);
        } else {
            current += static_cast<std::uint64_t>(-delta);
        }
        accepted += movers;
        entries += load_u64(&acc->edges);
        return true;
    };

    for (int phase = 0; phase < a.rounds && stale < a.stale_rounds && !error; ++phase) {
        rounds = phase + 1;
        // ---- LP round (:338-352)
        StepAccum* acc = acc_of(step);
        clear_accum(acc_of(step + 1));
        pass_commit(a, commit, commit_lp, checkpoint, false, lo, hi);
        checkpoint = false;
        stage_lp_proposals<G>(a, lo, hi);
        if ((aborted = !engine_barrier(a.sync, epoch, sh))) break;
        stage_filter_list<8>(a, a.low, a.low_count, acc, sh);
        stage_filter_list<kWave>(a, a.medium, a.medium_count, acc, sh);
        stage_filter_list<kEngineThreads>(a, a.high, a.high_count, acc, sh);
        if ((aborted = !engine_barrier(a.sync, epoch, sh))) break;
        stage_update(a, acc, sh);
        if ((aborted = !engine_barrier(a.sync, epoch, sh))) break;
        ++step;
        maximum = read_weights();
        const unsigned long long movers = load_u64(&acc->light) + load_u64(&acc->heavy);
        commit = movers > 0;
        commit_lp = true;
        proposals += movers;
        if (!movers) {
            if (++empty >= 2) break;
            ++stale;
            continue;
        }
        empty = 0;
        if (!absorb(acc, movers)) break;
        // ---- rebalance (:303-331)
        bool balanced = true;
        for (int it = 0; it < 2 * k + 4 && maximum > a.capacity; ++it) {
            ++iterations;
            acc = acc_of(step);
            clear_accum(acc_of(step + 1));
            pass_commit(a, commit, commit_lp, false, true, lo, hi);
            table_clear(sh.table);
            if (threadIdx.x < 32) sh.sums[threadIdx.x] = 0;
            __syncthreads();
            stage_balance_proposals<G>(a, lo, hi, sh);
            __syncthreads();
            table_flush(sh.table, a.src_counts);
            if ((aborted = !engine_barrier(a.sync, epoch, sh))) break;
            stage_bucket_scan(k, a.src_counts, a.src_offsets, a.thresholds, false, a.capacity,
                              [&](int p) { return sh.pw[p] > a.capacity; }, sh);
            if ((aborted = !engine_barrier(a.sync, epoch, sh))) break;
            table_clear(sh.table);
            if (threadIdx.x == 0) sh.touched = 0;
            __syncthreads();
            stage_sources(a, lo, hi, sh);
            __syncthreads();
            table_flush(sh.table, a.dst_counts);
            if (threadIdx.x < 32) a.boundary[static_cast<std::int64_t>(blockIdx.x) * 32 + threadIdx.x] = sh.sums[threadIdx.x];
            if (threadIdx.x == 0 && sh.touched) atomicOr(&acc->touched, sh.touched);
            if ((aborted = !engine_barrier(a.sync, epoch, sh))) break;
            boundary_prefix(a.boundary, sh);
            table_clear(sh.table);
            if (threadIdx.x == 0) sh.touched = 0;
            __syncthreads();
            stage_source_ranks(a, lo, hi, sh);
            __syncthreads();
            table_flush(sh.table, a.dst_counts);
            if (threadIdx.x == 0 && sh.touched) atomicOr(&acc->touched, sh.touched);
            clear_counts(a.src_counts, k);  // read for the last time by the sources stage
            if ((aborted = !engine_barrier(a.sync, epoch, sh))) break;
            const unsigned long long touched = load_u64(&acc->touched);
            stage_bucket_scan(k, a.dst_counts, a.dst_offsets, a.thresholds, true, a.capacity,
                              [&](int p) { return ((touched >> p) & 1ULL) != 0; }, sh);
            if ((aborted = !engine_barrier(a.sync, epoch, sh))) break;
            if (threadIdx.x < 32) sh.sums[threadIdx.x] = 0;
            __syncthreads();
            stage_destinations(a, lo, hi, acc, sh);
            __syncthreads();
            if (threadIdx.x < 32) a.boundary[static_cast<std::int64_t>(blockIdx.x) * 32 + threadIdx.x] = sh.sums[threadIdx.x];
            if ((aborted = !engine_barrier(a.sync, epoch, sh))) break;
            boundary_prefix(a.boundary, sh);
            stage_destination_ranks(a, lo, hi, acc, sh);
            clear_counts(a.dst_counts, k);  // read for the last time by the destinations stage
            if ((aborted = !engine_barrier(a.sync, epoch, sh))) break;
            stage_update(a, acc, sh);
            if ((aborted = !engine_barrier(a.sync, epoch, sh))) break;
            ++step;
            maximum = read_weights();
            const unsigned long long moved = load_u64(&acc->light) + load_u64(&acc->heavy);
            commit = moved > 0;
            commit_lp = false;
            if (!moved) {
                balanced = false;
                break;
            }
            if (!absorb(acc, moved)) break;
        }
        if (aborted || error) break;
        if (balanced) balanced = maximum <= a.capacity;
        if (!balanced) {
            ++fallbacks;
            break;
        }
        ++stale;
        if (current < best) {
            if (below_keep_ratio(current, best, a.keep_mantissa)) stale = 0;
            best = current;
            checkpoint = true;
            if (blockIdx.x == 0 && threadIdx.x < k) a.best_part_weights[threadIdx.x] = sh.pw[threadIdx.x];
        }
    }
    if (aborted) {
        if (blockIdx.x == 0 && threadIdx.x == 0) a.stats->error = 2;
        return;
    }
    // Final labels: the pending commit, then either this state is the best
    // (checkpoint pending) or the best copy comes back.
    for (std::int64_t v = lo + threadIdx.x; v < hi; v += kEngineThreads) {
        if (error) break;
        if (checkpoint) {
            if (commit && a.flags[v]) a.labels[v] = a.targets[v];
        } else {
            a.labels[v] = a.best_labels[v];
        }
    }
    if (blockIdx.x == 0) {
        if (!checkpoint && threadIdx.x < k) a.part_weights[threadIdx.x] = a.best_part_weights[threadIdx.x];
        if (threadIdx.x == 0) {
            EngineStats* s = a.stats;
            s->rounds = rounds;
            s->admission_iterations = iterations;
            s->admission_fallbacks = fallbacks;
            s->error = error;
            s->proposals = proposals;
            s->accepted = accepted;
            s->mover_entries = entries;
            s->best_cut = best;
            s->steps = step;
        }
    }
}

// Engine tuning of a context (levels above max_vertices or with rows above
// max_degree take the launch-per-stage path). Off by default: on the 128^3 mesh
// at k = 32 every level measured slower than the launch-per-stage path (at 512
// vertices per workgroup: 3,231 vertices 4.6 vs 4.1 ms, 137,845 vertices 18.5 vs
// 10.3 ms, 538,010 vertices 15.8 vs 5.6 ms). A step costs its chain of dependent
// stages either way -- each stage one or two memory latencies after a barrier
// that writes back and invalidates the L2 -- and the engine's few resident
// workgroups hide less latency than full-device launches.
struct EngineTuning {
    std::int64_t max_vertices = 0;
    std::int64_t vertices_per_group = 512;
    int max_grid = 128;
    std::int64_t max_degree = 4096;
};
std::mutex& tuning_mutex() {
    static std::mutex m;
    return m;
}
std::map<const gpart_ctx*, EngineTuning>& tunings() {
    static std::map<const gpart_ctx*, EngineTuning> t;
    return t;
}
EngineTuning tuning_of(const gpart_ctx* ctx) {
    std::lock_guard<std::mutex> lock(tuning_mutex());
    auto it = tunings().find(ctx);
    return it == tunings().end() ? EngineTuning{} : it->second;
}

std::uint64_t keep_mantissa() {
    int e = 0;
    const long double f = std::frexp(0.999L, &e);  // f in [0.5, 1), 0.999L = f * 2^e
    if (e != 0) throw Error(GPART_ERR_RUNTIME, "unexpected exponent of 0.999L");
    const long double scaled = std::ldexp(f, 64);
    const auto m = static_cast<std::uint64_t>(scaled);
    if (static_cast<long double>(m) != scaled || std::ldexp(static_cast<long double>(m), -64) != 0.999L)
        throw Error(GPART_ERR_RUNTIME, "0.999L does not have a 64-bit significand");
    return m;
}

template <typename O, typename W, typename C>
void run_engine(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* c, C* cache,
                std::int32_t* labels, int k, unsigned long long* part_weights, std::uint64_t capacity, double filter,
                int rounds, int stale_rounds, std::uint64_t initial_cut, std::int32_t* targets, long long* gains,
                std::int32_t* flags, std::int32_t* locks, int grid, gpart_refine_level_stats* out) {
    const std::int64_t n = g->n;
    const std::int64_t width = static_cast<std::int64_t>(k) * kBalanceWidth;
    Scratch<std::int32_t> best_labels(ctx, n), light(ctx, n), heavy(ctx, n), cbucket(ctx, n), dbucket(ctx, n);
    // zeroed together: both count arrays, accumulators, barrier words, stats
    const std::size_t zero_bytes = static_cast<std::size_t>(2 * width) * 8 + 3 * sizeof(StepAccum) +
                                   sizeof(EngineSync) + sizeof(EngineStats);
    Scratch<unsigned char> zeroed(ctx, static_cast<std::int64_t>(zero_bytes));
    Scratch<unsigned long long> offsets(ctx, 2 * width), thresholds(ctx, 64), best_pw(ctx, 32),
        boundary(ctx, static_cast<std::int64_t>(grid) * 32);
    fill_bytes(ctx, zeroed.get(), 0, zero_bytes);
    auto* base = zeroed.get();
    auto* src_counts = reinterpret_cast<unsigned long long*>(base);
    auto* dst_counts = src_counts + width;
    auto* accum = reinterpret_cast<StepAccum*>(dst_counts + width);
    auto* sync = reinterpret_cast<EngineSync*>(accum + 3);
    auto* stats = reinterpret_cast<EngineStats*>(sync + 1);
    static const PowerTable powers = make_power_table();
    static const std::uint64_t mantissa = keep_mantissa();
    EngineArgs<O, W, C> a{};
    a.n = n;
    a.offsets = static_cast<const O*>(g->offsets);
    a.neighbors = g->neighbors;
    a.edge_weights = static_cast<const W*>(g->edge_weights);
    a.vertex_weights = static_cast<const W*>(g->vertex_weights);
    a.low = c->low;
    a.medium = c->medium;
    a.high = c->high;
    a.low_count = c->low_count;
    a.medium_count = c->medium_count;
    a.high_count = c->high_count;
    a.k = k;
    a.rounds = rounds;
    a.stale_rounds = stale_rounds;
    a.capacity = capacity;
    a.filter = filter;
    a.initial_cut = initial_cut;
    a.keep_mantissa = mantissa;
    a.cache = cache;
    a.labels = labels;
    a.part_weights = part_weights;
    a.targets = targets;
    a.gains = gains;
    a.flags = flags;
    a.locks = locks;
    a.best_labels = best_labels.get();
    a.best_part_weights = best_pw.get();
    a.light = light.get();
    a.heavy = heavy.get();
    a.cbucket = cbucket.get();
    a.dbucket = dbucket.get();
    a.src_counts = src_counts;
    a.dst_counts = dst_counts;
    a.src_offsets = offsets.get();
    a.dst_offsets = offsets.get() + width;
    a.thresholds = thresholds.get();
    a.boundary = boundary.get();
    a.accum = accum;
    a.sync = sync;
    a.stats = stats;
    a.powers = powers;
    // An ordinary launch: the grid is at most the device's resident capacity for this
    // kernel (one workgroup per CU, gpart_refine_engine_grid), so every workgroup
    // is dispatched once earlier work on the device drains. (The runtime's
    // cooperative launch refuses a second process on the device.) Should a
    // workgroup still not be scheduled, the barriers' spin limit ends the launch
    // with an error instead of a hang.
    const int G = group_for_k(k);
    const dim3 blocks(static_cast<unsigned>(grid)), threads(kEngineThreads);
    if (G == 1) hipLaunchKernelGGL((refine_engine_kernel<1, O, W, C>), blocks, threads, 0, ctx->stream, a);
    else if (G == 2) hipLaunchKernelGGL((refine_engine_kernel<2, O, W, C>), blocks, threads, 0, ctx->stream, a);
    else hipLaunchKernelGGL((refine_engine_kernel<4, O, W, C>), blocks, threads, 0, ctx->stream, a);
    GPART_LAUNCH_CHECK();
    EngineStats h{};
    read_back(ctx, &h, stats, sizeof(h));
    if (h.error == 2) throw Error(GPART_ERR_RUNTIME, "refinement engine: a grid barrier timed out");
    if (h.error == 1) throw Error(GPART_ERR_RUNTIME, "incremental cut underflow");
    std::memcpy(out, &h, sizeof(h));
}

// Largest grid the device keeps resident for the engine kernel.
int resident_limit(gpart_ctx* ctx, const void* fn) {
    static std::mutex mutex;
    static std::map<int, int> limits;
    std::lock_guard<std::mutex> lock(mutex);
    auto it = limits.find(ctx->device);
    if (it != limits.end()) return it->second;
    int cus = 0, per_cu = 0;
    GPART_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    GPART_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kEngineThreads, 0));
    // at most half the CUs, one workgroup each: room for a second process's engine
    const int limit = std::min(cus / 2, cus * per_cu);
    limits[ctx->device] = limit;
    return limit;
}

}  // namespace
}  // namespace gpart_hip

using namespace gpart_hip;

extern "C" {

int gpart_refine_set_engine(gpart_ctx* ctx, int64_t max_vertices, int64_t vertices_per_group, int max_grid,
                            int64_t max_degree) {
    return guarded([&] {
        if (!ctx) throw Error(GPART_ERR_INVALID_ARGUMENT, "null context");
        if (vertices_per_group < 1 || max_grid < 1 || max_grid > kEngineMaxGrid)
            throw Error(GPART_ERR_INVALID_ARGUMENT, "invalid engine tuning");
        std::lock_guard<std::mutex> lock(tuning_mutex());
        EngineTuning& t = tunings()[ctx];
        t.max_vertices = max_vertices;
        t.vertices_per_group = vertices_per_group;
        t.max_grid = max_grid;
        t.max_degree = max_degree;
    });
}

int gpart_refine_engine_grid(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* classes,
                             int cache_bytes, int* grid) {
    return guarded([&] {
        if (!ctx || !g || !classes || !grid) throw Error(GPART_ERR_INVALID_ARGUMENT, "null argument");
        *grid = 0;
        const EngineTuning t = tuning_of(ctx);
        if (g->n < 1 || g->n > t.max_vertices || classes->max_degree > t.max_degree) return;
        const void* fn = g->width == GPART_WIDTH_STANDARD
                             ? (cache_bytes == 4
                                    ? reinterpret_cast<const void*>(refine_engine_kernel<4, std::int32_t, std::uint32_t, unsigned>)
                                    : reinterpret_cast<const void*>(
                                          refine_engine_kernel<4, std::int32_t, std::uint32_t, unsigned long long>))
                             : (cache_bytes == 4
                                    ? reinterpret_cast<const void*>(refine_engine_kernel<4, std::int64_t, std::uint64_t, unsigned>)
                                    : reinterpret_cast<const void*>(
                                          refine_engine_kernel<4, std::int64_t, std::uint64_t, unsigned long long>));
        const int limit = resident_limit(ctx, fn);
        if (limit < 1) return;
        const std::int64_t want = (g->n + t.vertices_per_group - 1) / t.vertices_per_group;
        *grid = static_cast<int>(std::max<std::int64_t>(1, std::min<std::int64_t>({want, t.max_grid, limit})));
    });
}

int gpart_refine_level(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* classes, void* cache,
                       int cache_bytes, int32_t* labels, int k, uint64_t* part_weights, uint64_t capacity,
                       double filter, int rounds, int stale_rounds, uint64_t initial_cut, int32_t* targets,
                       int64_t* gains, int32_t* flags, int32_t* locks, int grid, gpart_refine_level_stats* out) {
    return guarded([&] {
        if (!ctx || !g || !classes || !out) throw Error(GPART_ERR_INVALID_ARGUMENT, "null argument");
        if (k < 2 || k > 32) throw Error(GPART_ERR_INVALID_ARGUMENT, "part count must be in [2, 32]");
        if (grid < 1 || grid > kEngineMaxGrid) throw Error(GPART_ERR_INVALID_ARGUMENT, "invalid engine grid");
        if (cache_bytes != 4 && cache_bytes != 8) throw Error(GPART_ERR_INVALID_ARGUMENT, "cache entries of 4 or 8 bytes");
        if (initial_cut >= (std::uint64_t{1} << 62)) throw Error(GPART_ERR_INVALID_ARGUMENT, "cut beyond the engine's range");
        auto* pw = reinterpret_cast<unsigned long long*>(part_weights);
        auto* gn = reinterpret_cast<long long*>(gains);
#define GPART_ENGINE(O, W)                                                                                           \
    do {                                                                                                             \
        if (cache_bytes == 4)                                                                                        \
            run_engine<O, W, unsigned>(ctx, g, classes, static_cast<unsigned*>(cache), labels, k, pw, capacity,    \
                                       filter, rounds, stale_rounds, initial_cut, targets, gn, flags, locks, grid,  \
                                       out);                                                                         \
        else                                                                                                         \
            run_engine<O, W, unsigned long long>(ctx, g, classes, static_cast<unsigned long long*>(cache), labels,  \
                                                 k, pw, capacity, filter, rounds, stale_rounds, initial_cut,       \
                                                 targets, gn, flags, locks, grid, out);                             \
    } while (0)
        if (g->width == GPART_WIDTH_STANDARD) GPART_ENGINE(std::int32_t, std::uint32_t);
        else GPART_ENGINE(std::int64_t, std::uint64_t);
#undef GPART_ENGINE
    });
}

int gpart_refine_keep_ratio(uint64_t current, uint64_t best) {
    static const std::uint64_t mantissa = keep_mantissa();
    return below_keep_ratio(current, best, mantissa) ? 1 : 0;
}

}  // extern "C"
