// contract.hip -- CSR contraction for gfx950 (contract_typed, src/coarsen.cu:1674-1814).
//
// 1. encode: every cross entry (map[v] != map[u]) becomes a record in its row's
//    range (record order inside the range is irrelevant: the sort defines it and
//    duplicate keys are summed exactly);
// 2. LSD radix sort of the records by (source, target): 8-byte keys only in the
//    packed layout (weight in the low bits: the row bound fits them, or else the
//    largest entry weight does), key + value otherwise (4-byte values whenever
//    the row bound fits them);
// 3. one scan fuses segment counting and segmented weight sums and emits each
//    (source, target, sum) at its segment end, staged per tile for coalesced
//    stores;
// 4. row offsets are written once per row from row boundaries.
#include "internal.hpp"
#include "reduce.hpp"
#include "scan.hpp"

#include <algorithm>

namespace gpart_hip {
namespace {

constexpr int kBlock = 256;
// Row entries a lane loads per step in the count and encode passes: groups of 8
// lanes serve rows of at most 8 entries, wider groups take 4 per lane.
template <int G>
constexpr int entries_per_lane() { return G >= kWave ? 4 : 1; }

struct Layout {
    int vertex_bits;
    int weight_bits;  // 0 = unpacked
    std::uint64_t weight_mask;
    std::uint64_t vertex_mask;
};

// SclpRecordLayout::checked(n, bound, 1) (include/detail/sclp_record.hpp:20-35)
Layout choose_layout(std::int64_t vertices, std::uint64_t bound) {
    int bits = 0;
    for (auto v = vertices; v; v >>= 1) ++bits;
    Layout l{bits, 0, 0, (bits >= 64 ? ~0ULL : (1ULL << bits) - 1)};
    if (vertices <= 0 || vertices > 0x7fffffff || bound == ~0ULL) return l;
    const int wb = 64 - 2 * bits;
    if (wb <= 0) return l;
    const std::uint64_t mask = (std::uint64_t{1} << wb) - 1;
    if (bound > mask) return l;
    l.weight_bits = wb;
    l.weight_mask = mask;
    return l;
}

// Pass 1: number of cross entries of each vertex (G lanes per vertex, G == kBlock
// for one workgroup per vertex).
template <typename O, int G>
__global__ void __launch_bounds__(kBlock) cross_count_kernel(std::int64_t count, const std::int32_t* list,
                                                             const O* offsets, const std::int32_t* neighbors,
                                                             const std::int32_t* map, std::int64_t* cross,
                                                             std::int32_t* targets) {
    __shared__ unsigned long long waves[kBlock / kWave];
    const std::int64_t group = (static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
    const int sub = threadIdx.x & (G - 1);
    const bool active = group < count;
    const std::int32_t v = active ? list[group] : 0;
    unsigned long long c = 0;
    if (active) {
        const std::int32_t source = map[v];
        const std::int64_t end = static_cast<std::int64_t>(offsets[v + 1]);
        // independent neighbour loads, then their map gathers
        constexpr int kEntriesPerLane = entries_per_lane<G>();
        for (std::int64_t e0 = static_cast<std::int64_t>(offsets[v]) + sub; e0 < end; e0 += G * kEntriesPerLane) {
            std::int32_t nb[kEntriesPerLane];
#pragma unroll
            for (int u = 0; u < kEntriesPerLane; ++u) nb[u] = e0 + u * G < end ? neighbors[e0 + u * G] : 0;
#pragma unroll
            for (int u = 0; u < kEntriesPerLane; ++u) {
                if (e0 + u * G < end) {
                    const std::int32_t target = map[nb[u]];
                    if (targets) targets[e0 + u * G] = target;
                    c += target != source;
                }
            }
        }
    }
    if constexpr (G <= kWave) {
        c = group_sum_u64<G>(c);
        if (active && sub == 0) cross[v] = static_cast<std::int64_t>(c);
    } else {
        c = group_sum_u64<kWave>(c);
        if (lane_id() == 0) waves[threadIdx.x / kWave] = c;
        __syncthreads();
        if (threadIdx.x == 0 && active) {
            unsigned long long t = 0;
            for (int w = 0; w < kBlock / kWave; ++w) t += waves[w];
            cross[v] = static_cast<std::int64_t>(t);
        }
    }
}

// Pass 2: write each vertex's cross records into [base[v], base[v] + cross[v]).
// Their order inside a row is free (the sort that follows orders them, and
// records with equal keys are summed exactly): groups of at most a wave place
// them by ballot rank; a workgroup's waves (G == kBlock) reserve their slots
// with one LDS atomic per wave and step, without workgroup barriers.
template <typename O, typename W, typename V, int G>
__global__ void __launch_bounds__(kBlock) encode_kernel(std::int64_t count, const std::int32_t* list, const O* offsets,
                                                        const std::int32_t* neighbors, const W* weights,
                                                        const std::int32_t* map, const std::int32_t* targets,
                                                        const std::int64_t* base, Layout layout, std::uint64_t* keys,
                                                        V* values) {
    __shared__ unsigned long long s_fill;
    const std::int64_t group = (static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
    const int sub = threadIdx.x & (G - 1);
    const bool active = group < count;
    const std::int32_t v = active ? list[group] : 0;
    const std::int64_t begin = active ? static_cast<std::int64_t>(offsets[v]) : 0;
    const std::int64_t end = active ? static_cast<std::int64_t>(offsets[v + 1]) : 0;
    const std::uint64_t source = active ? static_cast<std::uint32_t>(map[v]) : 0;
    std::int64_t next = active ? base[v] : 0;
    const int lane = lane_id();
    if constexpr (G > kWave) {
        if (threadIdx.x == 0) s_fill = 0;
        __syncthreads();
    }
    constexpr int kEntriesPerLane = entries_per_lane<G>();
    constexpr int kStep = G * kEntriesPerLane;
    // group-uniform trip count
    const std::int64_t trips = end - begin > 0 ? (end - begin + kStep - 1) / kStep : 0;
    for (std::int64_t it = 0; it < trips; ++it) {
        const std::int64_t e0 = begin + sub + it * kStep;
        std::int32_t nb[kEntriesPerLane];
        W w[kEntriesPerLane];
#pragma unroll
        for (int u = 0; u < kEntriesPerLane; ++u) {
            const std::int64_t e = e0 + u * G;
            nb[u] = e < end ? (targets ? targets[e] : neighbors[e]) : 0;
            w[u] = e < end ? weights[e] : W(0);
        }
        bool emit[kEntriesPerLane];
        std::uint64_t key[kEntriesPerLane];
        std::uint64_t mask[kEntriesPerLane];
#pragma unroll
        for (int u = 0; u < kEntriesPerLane; ++u) {
            emit[u] = false;
            key[u] = 0;
            if (e0 + u * G < end) {
                const std::uint64_t target = static_cast<std::uint32_t>(targets ? nb[u] : map[nb[u]]);
                if (target != source) {
                    emit[u] = true;
                    key[u] = layout.weight_bits ? (((source << layout.vertex_bits) | target) << layout.weight_bits) |
                                                      static_cast<std::uint64_t>(w[u])
                                                : (source << layout.vertex_bits) | target;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kEntriesPerLane; ++u) mask[u] = __ballot(emit[u]);
        if constexpr (G <= kWave) {
            const int first = lane & ~(G - 1);
            const std::uint64_t group_bits = G == kWave ? ~0ULL : (((1ULL << G) - 1) << first);
#pragma unroll
            for (int u = 0; u < kEntriesPerLane; ++u) {
                const std::uint64_t gmask = mask[u] & group_bits;
                if (emit[u]) {
                    const std::int64_t at = next + __popcll(gmask & lanemask_lt());
                    keys[at] = key[u];
                    if (layout.weight_bits == 0) values[at] = static_cast<V>(w[u]);
                }
                next += __popcll(gmask);
            }
        } else {
            unsigned long long wave_total = 0;
#pragma unroll
            for (int u = 0; u < kEntriesPerLane; ++u) wave_total += __popcll(mask[u]);
            unsigned long long start = 0;
            if (lane == 0 && wave_total) start = atomicAdd(&s_fill, wave_total);
            const unsigned lo = __shfl(static_cast<unsigned>(start), 0, kWave);
            const unsigned hi = __shfl(static_cast<unsigned>(start >> 32), 0, kWave);
            std::int64_t at_wave = next + static_cast<std::int64_t>((static_cast<unsigned long long>(hi) << 32) | lo);
#pragma unroll
            for (int u = 0; u < kEntriesPerLane; ++u) {
                if (emit[u]) {
                    const std::int64_t at = at_wave + __popcll(mask[u] & lanemask_lt());
                    keys[at] = key[u];
                    if (layout.weight_bits == 0) values[at] = static_cast<V>(w[u]);
                }
                at_wave += __popcll(mask[u]);
            }
        }
    }
}

struct RunValue {
    unsigned long long heads;  // segment count
    unsigned long long value;  // segmented weight sum
    unsigned int head;
    unsigned int pad;
};
struct RunOp {
    __device__ __forceinline__ RunValue operator()(const RunValue& a, const RunValue& b) const {
        return RunValue{a.heads + b.heads, b.head ? b.value : a.value + b.value, a.head | b.head, 0};
    }
    __device__ __forceinline__ RunValue identity() const { return RunValue{0, 0, 0, 0}; }
};

// The reduce pass runs over the sorted keys in tiles of kScanTile entries. A tile
// is loaded lane after lane (coalesced) into LDS, one padding slot per
// kScanItems entries so that the blocked reads below spread over the banks, and
// each lane then takes kScanItems consecutive entries.
constexpr int kRunStage = kScanTile + kScanTile / kScanItems;
__device__ __forceinline__ int run_slot(int local) { return local + local / kScanItems; }

template <typename V, bool Packed>
__device__ __forceinline__ void stage_run_tile(const std::uint64_t* keys, const V* values, std::int64_t tile_start,
                                               int count, std::uint64_t* s_keys, V* s_values) {
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int local = i * kScanThreads + static_cast<int>(threadIdx.x);
        if (local < count) {
            s_keys[run_slot(local)] = keys[tile_start + local];
            if constexpr (!Packed) s_values[run_slot(local)] = values[tile_start + local];
        }
    }
    __syncthreads();
}

// This lane's kScanItems entries as (head, weight) items: an entry heads a
// segment when its (source, target) pair differs from the entry before it.
// Returns the lanes' combined item; pairs[] receives the entries' pairs.
template <typename V, bool Packed>
__device__ __forceinline__ RunValue run_items(const std::uint64_t* keys, std::int64_t tile_start, int count,
                                              const std::uint64_t* s_keys, const V* s_values, int shift,
                                              std::uint64_t weight_mask, RunValue (&items)[kScanItems],
                                              std::uint64_t (&pairs)[kScanItems]) {
    const int first = static_cast<int>(threadIdx.x) * kScanItems;
    bool has_prev = first > 0 || tile_start > 0;
    std::uint64_t prev = 0;
    if (first > 0) {
        if (first - 1 < count) prev = s_keys[run_slot(first - 1)] >> shift;
    } else if (tile_start > 0) {
        prev = keys[tile_start - 1] >> shift;
    }
    RunValue acc = RunOp{}.identity();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int local = first + i;
        if (local < count) {
            const std::uint64_t k = s_keys[run_slot(local)];
            const std::uint64_t pair = k >> shift;
            const bool head = !has_prev || pair != prev;
            const std::uint64_t w = Packed ? (k & weight_mask) : static_cast<std::uint64_t>(s_values[run_slot(local)]);
            items[i] = RunValue{head ? 1ULL : 0ULL, w, head ? 1U : 0U, 0};
            pairs[i] = pair;
            prev = pair;
            has_prev = true;
        } else {
            items[i] = RunOp{}.identity();
            pairs[i] = 0;
        }
        acc = RunOp{}(acc, items[i]);
    }
    return acc;
}

// Pass 1: one (segment count, open-segment sum, any head) total per tile.
template <typename V, bool Packed>
__global__ void __launch_bounds__(kScanThreads)
    run_reduce_kernel(const std::uint64_t* keys, const V* values, std::int64_t n, int shift,
                      std::uint64_t weight_mask, RunValue* partials) {
    __shared__ std::uint64_t s_keys[kRunStage];
    __shared__ V s_values[Packed ? 1 : kRunStage];
    __shared__ RunValue shared_waves[kScanThreads / kWave + 1];
    const std::int64_t tile_start = static_cast<std::int64_t>(blockIdx.x) * kScanTile;
    const int count = static_cast<int>(std::min<std::int64_t>(kScanTile, n - tile_start));
    stage_run_tile<V, Packed>(keys, values, tile_start, count, s_keys, s_values);
    RunValue items[kScanItems];
    std::uint64_t pairs[kScanItems];
    const RunValue acc =
        run_items<V, Packed>(keys, tile_start, count, s_keys, s_values, shift, weight_mask, items, pairs);
    RunValue total;
    (void)block_exclusive_scan(acc, RunOp{}, &total, shared_waves);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// Pass 3: the segmented scan of each tile with its carried-in prefix. Every
// segment end j = (segments up to it) - 1 gets (source, target, weight sum); the
// tile's ends form one contiguous range of j, staged in LDS and written out lane
// after lane. partial_prefix holds either the exclusive prefixes of the tile
// totals or (tile_totals) the totals themselves, reduced below the tile here.
template <typename V, typename W, bool Packed>
__global__ void __launch_bounds__(kScanThreads)
    run_emit_kernel(const std::uint64_t* keys, const V* values, std::int64_t n, int shift, std::uint64_t weight_mask,
                    std::uint64_t vertex_mask, int vertex_bits, const RunValue* partial_prefix, int tile_totals,
                    std::uint32_t* sources, std::int32_t* neighbors, W* weights, unsigned long long* segments) {
    __shared__ std::uint64_t s_keys[kRunStage];
    __shared__ V s_values[Packed ? 1 : kRunStage];
    __shared__ std::uint64_t s_ends[kScanTile + 1];  // source << 32 | target
    __shared__ W s_sums[kScanTile + 1];
    __shared__ RunValue shared_waves[kScanThreads / kWave + 1];
    __shared__ int s_first_head, s_last_end;
    const std::int64_t tile_start = static_cast<std::int64_t>(blockIdx.x) * kScanTile;
    const int count = static_cast<int>(std::min<std::int64_t>(kScanTile, n - tile_start));
    stage_run_tile<V, Packed>(keys, values, tile_start, count, s_keys, s_values);
    RunValue items[kScanItems];
    std::uint64_t pairs[kScanItems];
    const RunValue acc =
        run_items<V, Packed>(keys, tile_start, count, s_keys, s_values, shift, weight_mask, items, pairs);
    const int first = static_cast<int>(threadIdx.x) * kScanItems;
    unsigned last_mask = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int local = first + i;
        if (local >= count) continue;
        bool last = tile_start + local == n - 1;
        if (!last) {
            const std::uint64_t next = local + 1 < count ? s_keys[run_slot(local + 1)] >> shift
                                                         : keys[tile_start + local + 1] >> shift;
            last = next != pairs[i];
        }
        if (last) last_mask |= 1U << i;
        if (local == count - 1) s_last_end = last ? 1 : 0;
    }
    if (threadIdx.x == 0) s_first_head = static_cast<int>(items[0].head);
    RunValue carry = RunOp{}.identity();
    if (partial_prefix && tile_totals) {
        const int below = static_cast<int>(blockIdx.x);
        const int per = (below + kScanThreads - 1) / kScanThreads;
        RunValue part = RunOp{}.identity();
        for (int i = 0; i < per; ++i) {
            const int idx = static_cast<int>(threadIdx.x) * per + i;
            if (idx < below) part = RunOp{}(part, partial_prefix[idx]);
        }
        (void)block_exclusive_scan(part, RunOp{}, &carry, shared_waves);
    } else if (partial_prefix) {
        carry = partial_prefix[blockIdx.x];
    }
    RunValue tile_total;
    RunValue prefix = RunOp{}(carry, block_exclusive_scan(acc, RunOp{}, &tile_total, shared_waves));
    // local slot 0 is segment end j = carry.heads - 1 (the segment open at the tile start)
    const std::int64_t base_j = static_cast<std::int64_t>(carry.heads) - 1;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int local = first + i;
        if (local >= count) continue;
        const RunValue incl = RunOp{}(prefix, items[i]);
        if (tile_start + local == n - 1) *segments = incl.heads;
        if (last_mask & (1U << i)) {
            const int slot = static_cast<int>(static_cast<std::int64_t>(incl.heads) - 1 - base_j);
            s_ends[slot] = ((pairs[i] >> vertex_bits) << 32) | (pairs[i] & vertex_mask);
            s_sums[slot] = static_cast<W>(incl.value);
        }
        prefix = incl;
    }
    __syncthreads();
    const int lo = s_first_head ? 1 : 0;
    const int ends = static_cast<int>(tile_total.heads) + (1 - lo) - (s_last_end ? 0 : 1);
    for (int x = static_cast<int>(threadIdx.x); x < ends; x += kScanThreads) {
        const std::int64_t j = base_j + lo + x;
        const std::uint64_t e = s_ends[lo + x];
        sources[j] = static_cast<std::uint32_t>(e >> 32);
        neighbors[j] = static_cast<std::int32_t>(static_cast<std::uint32_t>(e));
        weights[j] = s_sums[lo + x];
    }
}

// Segment count and (source, target, summed weight) of every segment of the n
// sorted keys: reduce, scan of the tile totals (inline up to kScanInlineTiles),
// emit.
template <typename V, typename W, bool Packed>
void run_segments(gpart_ctx* ctx, const Layout& layout, const std::uint64_t* keys, const V* values, std::int64_t n,
                  std::uint32_t* sources, std::int32_t* neighbors, W* weights, unsigned long long* segments) {
    const std::int64_t tiles = (n + kScanTile - 1) / kScanTile;
    const int shift = layout.weight_bits;
    Scratch<RunValue> partials;
    int inline_totals = 0;
    if (tiles > 1) {
        partials.reset(ctx, tiles);
        hipLaunchKernelGGL((run_reduce_kernel<V, Packed>), dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0,
                           ctx->stream, keys, values, n, shift, layout.weight_mask, partials.get());
        GPART_LAUNCH_CHECK();
        inline_totals = tiles <= kScanInlineTiles ? 1 : 0;
        if (!inline_totals)
            device_scan<RunValue, RunOp>(ctx, ArrayLoad<RunValue>{partials.get()},
                                         ExclusiveStore<RunValue>{partials.get()}, tiles, RunOp{});
    }
    hipLaunchKernelGGL((run_emit_kernel<V, W, Packed>), dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0,
                       ctx->stream, keys, values, n, shift, layout.weight_mask, layout.vertex_mask,
                       layout.vertex_bits, static_cast<const RunValue*>(partials.get()), inline_totals, sources,
                       neighbors, weights, segments);
    GPART_LAUNCH_CHECK();
}

template <typename O>
__global__ void row_offsets_kernel(std::int64_t m, const std::uint32_t* sources, std::int64_t n_coarse, O* offsets) {
    const std::int64_t j = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j > m) return;
    // rows (previous source, this source] start at j; the sentinel j == m closes the tail
    const std::int64_t here = j < m ? static_cast<std::int64_t>(sources[j]) : n_coarse;
    const std::int64_t prev = j == 0 ? -1 : static_cast<std::int64_t>(sources[j - 1]);
    for (std::int64_t r = prev + 1; r <= here; ++r) offsets[r] = static_cast<O>(j);
}

inline unsigned grid_for(std::int64_t threads) { return static_cast<unsigned>((threads + kBlock - 1) / kBlock); }

// V: the record value type of the unpacked layout, 4 bytes whenever the row bound
// (which bounds every entry's weight and every summed coarse weight) fits them.
template <typename O, typename W, typename V>
void contract_impl(gpart_ctx* ctx, const gpart_csr_view* fine, const gpart_degree_classes* classes,
                   const std::int32_t* map, std::int64_t n_coarse, std::uint64_t bound, void** out_offsets,
                   std::int32_t** out_neighbors, void** out_weights, std::int64_t* out_m) {
    Layout layout = choose_layout(n_coarse, bound);
    const std::int64_t m = fine->m;
    const auto* off = static_cast<const O*>(fine->offsets);
    const auto* ew = static_cast<const W*>(fine->edge_weights);
    profile_begin(ctx, GPART_PROFILE_CONTRACT_SORT);
    if (!layout.weight_bits && m > 0) {
        // A packed key holds one fine entry's weight, never a sum (sums are taken
        // in the run scan): when the row bound rules the packed layout out, it
        // still fits if the largest entry weight does (8- instead of 12- or
        // 16-byte records through the sort; the result is the same).
        const Layout packed = choose_layout(n_coarse, 0);
        if (packed.weight_bits) {
            Scratch<unsigned long long> largest(ctx, 1);
            device_reduce<unsigned long long>(ctx, WidenLoad<W>{ew}, m, MaxU64{}, largest.get());
            if (read_scalar(ctx, largest.get()) <= packed.weight_mask) layout = packed;
        }
    }
    Scratch<std::uint64_t> keys(ctx, std::max<std::int64_t>(m, 1)), keys_alt(ctx, std::max<std::int64_t>(m, 1));
    Scratch<V> values, values_alt;
    if (!layout.weight_bits) {
        values.reset(ctx, std::max<std::int64_t>(m, 1));
        values_alt.reset(ctx, std::max<std::int64_t>(m, 1));
    }
    const std::int64_t n_fine = fine->n;
    Scratch<std::int64_t> base(ctx, n_fine + 1);
    // From 2^24 fine vertices on (the default of ctx->contract_stash_vertices), map
    // (64 MB and up) no longer stays in the caches and its random gathers dominate
    // both passes: the count pass then stores every entry's coarse neighbour and
    // the encode pass reads it sequentially.
    Scratch<std::int32_t> targets;
    if (m > 0 && n_fine >= ctx->contract_stash_vertices) targets.reset(ctx, m);
    fill_bytes(ctx, base.get() + n_fine, 0, 8);
    if (classes->low_count)
        hipLaunchKernelGGL((cross_count_kernel<O, 8>), dim3(grid_for(classes->low_count * 8)), dim3(kBlock), 0,
                           ctx->stream, classes->low_count, classes->low, off, fine->neighbors, map, base.get(),
                           targets.get());
    if (classes->medium_count)
        hipLaunchKernelGGL((cross_count_kernel<O, 64>), dim3(grid_for(classes->medium_count * 64)), dim3(kBlock), 0,
                           ctx->stream, classes->medium_count, classes->medium, off, fine->neighbors, map, base.get(),
                           targets.get());
    if (classes->high_count)
        hipLaunchKernelGGL((cross_count_kernel<O, kBlock>), dim3(static_cast<unsigned>(classes->high_count)),
                           dim3(kBlock), 0, ctx->stream, classes->high_count, classes->high, off, fine->neighbors, map,
                           base.get(), targets.get());
    GPART_LAUNCH_CHECK();
    // exclusive scan over n + 1 entries: base[n] becomes the cross-entry total
    exclusive_scan<std::int64_t>(ctx, base.get(), base.get(), n_fine + 1);
    if (classes->low_count)
        hipLaunchKernelGGL((encode_kernel<O, W, V, 8>), dim3(grid_for(classes->low_count * 8)), dim3(kBlock), 0,
                           ctx->stream, classes->low_count, classes->low, off, fine->neighbors, ew, map, targets.get(),
                           base.get(), layout, keys.get(), values.get());
    if (classes->medium_count)
        hipLaunchKernelGGL((encode_kernel<O, W, V, 64>), dim3(grid_for(classes->medium_count * 64)), dim3(kBlock), 0,
                           ctx->stream, classes->medium_count, classes->medium, off, fine->neighbors, ew, map,
                           targets.get(), base.get(), layout, keys.get(), values.get());
    if (classes->high_count)
        hipLaunchKernelGGL((encode_kernel<O, W, V, kBlock>), dim3(static_cast<unsigned>(classes->high_count)),
                           dim3(kBlock), 0, ctx->stream, classes->high_count, classes->high, off, fine->neighbors, ew,
                           map, targets.get(), base.get(), layout, keys.get(), values.get());
    GPART_LAUNCH_CHECK();
    const auto cross = read_scalar(ctx, base.get() + n_fine);
    const int shift = layout.weight_bits;
    const int key_bits = 2 * layout.vertex_bits;
    std::uint64_t* sorted_keys = keys.get();
    V* sorted_values = values.get();
    if (cross > 1) {
        bool alt;
        if (layout.weight_bits)
            alt = radix_sort_keys(ctx, keys.get(), keys_alt.get(), cross, shift, shift + key_bits);
        else
            alt = radix_sort_pairs(ctx, keys.get(), keys_alt.get(), values.get(), values_alt.get(), cross, 0, key_bits);
        if (alt) {
            sorted_keys = keys_alt.get();
            sorted_values = values_alt.get();
        }
    }
    // reduce: one scan counts (source, target) segments and sums their weights
    // The coarse entries are emitted straight into the output arrays, sized to the
    // cross-entry count (their upper bound), so no copy to exact sizes follows.
    Scratch<std::uint32_t> sources(ctx, std::max<std::int64_t>(cross, 1));
    Scratch<std::int32_t> out_nbr(ctx, std::max<std::int64_t>(cross, 1));
    Scratch<W> out_w(ctx, std::max<std::int64_t>(cross, 1));
    Scratch<unsigned long long> segments(ctx, 1);
    fill_bytes(ctx, segments.get(), 0, 8);
    if (cross > 0) {
        if (layout.weight_bits)
            run_segments<V, W, true>(ctx, layout, sorted_keys, sorted_values, cross, sources.get(), out_nbr.get(),
                                     out_w.get(), segments.get());
        else
            run_segments<V, W, false>(ctx, layout, sorted_keys, sorted_values, cross, sources.get(), out_nbr.get(),
                                      out_w.get(), segments.get());
    }
    const auto coarse_m = static_cast<std::int64_t>(read_scalar(ctx, segments.get()));
    Scratch<O> offsets(ctx, n_coarse + 1);
    hipLaunchKernelGGL(row_offsets_kernel<O>, dim3(grid_for(coarse_m + 1)), dim3(kBlock), 0, ctx->stream, coarse_m,
                       sources.get(), n_coarse, offsets.get());
    GPART_LAUNCH_CHECK();
    const double n = static_cast<double>(fine->n), md = static_cast<double>(m), mc = static_cast<double>(coarse_m);
    const double bytes = (n + 1) * sizeof(O) + md * (4 + sizeof(W)) + 4 * md + 4 * n + 32 * md + 16 * mc +
                         mc * (4 + sizeof(W)) + (n_coarse + 1.0) * sizeof(O) +
                         (layout.weight_bits ? 0.0 : 2.0 * sizeof(V) * md);
    profile_end(ctx, GPART_PROFILE_CONTRACT_SORT, bytes);
    *out_offsets = offsets.take();
    *out_neighbors = out_nbr.take();
    *out_weights = out_w.take();
    *out_m = coarse_m;
}

}  // namespace
}  // namespace gpart_hip

using namespace gpart_hip;

extern "C" void gpart_record_layout(int64_t vertices, uint64_t bound, int* vertex_bits, int* weight_bits,
                                    uint64_t* weight_mask) {
    const Layout l = choose_layout(vertices, bound);
    // the reference reports 32 vertex bits for the unpacked layout
    *vertex_bits = l.weight_bits ? l.vertex_bits : 32;
    *weight_bits = l.weight_bits;
    *weight_mask = l.weight_mask;
}

extern "C" int gpart_contract_set_stash(gpart_ctx* ctx, int64_t min_vertices) {
    return guarded([&] {
        if (min_vertices < 0) throw Error(GPART_ERR_INVALID_ARGUMENT, "negative vertex count");
        ctx->contract_stash_vertices = min_vertices;
    });
}

extern "C" int gpart_contract(gpart_ctx* ctx, const gpart_csr_view* fine, const gpart_degree_classes* classes,
                              const int32_t* map, int64_t n_coarse, uint64_t coarse_row_bound, void** coarse_offsets,
                              int32_t** coarse_neighbors, void** coarse_edge_weights, int64_t* coarse_m) {
    return guarded([&] {
        if (n_coarse <= 0 || n_coarse > 0x7fffffff)
            throw Error(GPART_ERR_INVALID_ARGUMENT, "invalid coarse vertex count");
        if (fine->width == GPART_WIDTH_STANDARD)
            contract_impl<std::int32_t, std::uint32_t, std::uint32_t>(ctx, fine, classes, map, n_coarse,
                                                                      coarse_row_bound, coarse_offsets,
                                                                      coarse_neighbors, coarse_edge_weights, coarse_m);
        else if (coarse_row_bound <= 0xffffffffULL)
            contract_impl<std::int64_t, std::uint64_t, std::uint32_t>(ctx, fine, classes, map, n_coarse,
                                                                      coarse_row_bound, coarse_offsets,
                                                                      coarse_neighbors, coarse_edge_weights, coarse_m);
        else
            contract_impl<std::int64_t, std::uint64_t, std::uint64_t>(ctx, fine, classes, map, n_coarse,
                                                                      coarse_row_bound, coarse_offsets,
                                                                      coarse_neighbors, coarse_edge_weights, coarse_m);
    });
}
