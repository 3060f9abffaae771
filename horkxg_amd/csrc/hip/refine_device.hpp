// refine_device.hpp -- device pieces of incremental refinement shared by the
// launch-per-stage kernels (refine.hip) and the single-launch level engine
// (refine_engine.hip): cache-row spans, the afterburner correction, the gain
// buckets and bucket tables of the rebalance, and the bucket-order ranks.
#pragma once

#include "common.hpp"

#include <cmath>
#include <cstdint>

namespace gpart_hip {
namespace {

constexpr int kBalanceBuckets = 50;
constexpr int kBalanceSections = 16;
constexpr int kBalanceWidth = kBalanceBuckets * kBalanceSections;

// Rows of the n x k cache are read by groups of G lanes, lane j owning the
// kRowSpan consecutive parts [kRowSpan*j, kRowSpan*j + kRowSpan): G = 1, 2 or 4
// for k <= 8, 16, 32. A wave covers 64/G rows per pass.
constexpr int kRowSpan = 8;

// The lane's kRowSpan entries of a cache row, entries at or past k as 0: 16-byte
// loads when the span lies inside the row and is 16-byte aligned (k a multiple of
// 4 or 2 entries), else one load per entry.
template <typename C>
__device__ __forceinline__ void load_span(const C* row, int j, int k, unsigned long long (&out)[kRowSpan]) {
    static_assert(sizeof(C) == 4 || sizeof(C) == 8, "4- or 8-byte cache entries");
    const int first = j * kRowSpan;
    const C* span = row + first;
    if (first + kRowSpan <= k && (reinterpret_cast<std::uintptr_t>(span) & 15) == 0) {
        constexpr int kPer = 16 / static_cast<int>(sizeof(C));
#pragma unroll
        for (int q = 0; q < kRowSpan / kPer; ++q) {
            const uint4 x = reinterpret_cast<const uint4*>(span)[q];
            if constexpr (sizeof(C) == 4) {
                out[4 * q] = x.x;
                out[4 * q + 1] = x.y;
                out[4 * q + 2] = x.z;
                out[4 * q + 3] = x.w;
            } else {
                out[2 * q] = static_cast<unsigned long long>(x.x) | static_cast<unsigned long long>(x.y) << 32;
                out[2 * q + 1] = static_cast<unsigned long long>(x.z) | static_cast<unsigned long long>(x.w) << 32;
            }
        }
    } else {
#pragma unroll
        for (int e = 0; e < kRowSpan; ++e) out[e] = first + e < k ? static_cast<unsigned long long>(span[e]) : 0;
    }
}

// incremental_priority_filter (:48-67). The correction of a proposing vertex v
// summed over its entries [e, end) with stride `step`: neighbours that propose
// too and rank before v (higher gain, ties to the lower id).
template <typename W>
__device__ __forceinline__ long long filter_correction(std::int32_t v, std::int32_t target, std::int32_t source,
                                                       long long gain, std::int64_t e, std::int64_t end, int step,
                                                       const std::int32_t* neighbors, const W* weights,
                                                       const std::int32_t* labels, const std::int32_t* targets,
                                                       const long long* gains) {
    long long correction = 0;
    for (; e < end; e += step) {
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
    return correction;
}

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

// Per-workgroup histogram of bucket weights in LDS (open addressing); a full
// table spills to the global counters. Workgroups flush into one of kReplicas
// copies of the counter array (blockIdx % kReplicas): a few hundred hot buckets
// would otherwise take every workgroup's atomics on the same addresses. The
// bucket scan sums the copies.
constexpr int kReplicas = 16;
constexpr int kTableBits = 9;
constexpr int kTableSlots = 1 << kTableBits;
constexpr unsigned kTableEmpty = 0xffffffffU;
struct BucketTable {
    unsigned key[kTableSlots];
    unsigned long long value[kTableSlots];
};

__device__ __forceinline__ void table_clear(BucketTable& t) {
    for (int i = threadIdx.x; i < kTableSlots; i += blockDim.x) {
        t.key[i] = kTableEmpty;
        t.value[i] = 0;
    }
}

__device__ __forceinline__ void table_add(BucketTable& t, unsigned key, unsigned long long w,
                                          unsigned long long* global) {
    const unsigned h = (key * 2654435761U) >> (32 - kTableBits);  // multiplicative hash, top bits
#pragma unroll 1
    for (int probe = 0; probe < 64; ++probe) {
        const unsigned slot = (h + probe) & (kTableSlots - 1);
        const unsigned prev = atomicCAS(&t.key[slot], kTableEmpty, key);
        if (prev == kTableEmpty || prev == key) {
            atomicAdd(&t.value[slot], w);
            return;
        }
    }
    atomicAdd(global + key, w);
}

__device__ __forceinline__ void table_flush(const BucketTable& t, unsigned long long* global) {
    for (int i = threadIdx.x; i < kTableSlots; i += blockDim.x)
        if (t.key[i] != kTableEmpty && t.value[i]) atomicAdd(global + t.key[i], t.value[i]);
}

// Contiguous slice [begin, end) of count items owned by this workgroup.
__device__ __forceinline__ void block_slice(long long count, long long& begin, long long& end) {
    const long long per = (count + gridDim.x - 1) / gridDim.x;
    begin = min(count, static_cast<long long>(blockIdx.x) * per);
    end = min(count, begin + per);
}

// For the workgroup's threads in index order: a pending thread with key (< 32)
// and weight w receives running[key] plus the weight of the pending threads
// with the same key and a lower index; running[] then advances by the pass.
// Called by every thread of the workgroup.
template <int Threads>
__device__ unsigned long long block_key_prefix(bool pending, int key, unsigned long long w,
                                               unsigned long long (*wave_total)[32], unsigned long long* running) {
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    if (lane < 32) wave_total[wave][lane] = 0;
    unsigned long long mine = 0;
    unsigned long long open = __ballot(pending);
    while (open) {
        const int leader = __ffsll(static_cast<unsigned long long>(open)) - 1;
        const int current = __shfl(key, leader, kWave);
        const bool member = pending && key == current;
        const unsigned long long members = __ballot(member);
        unsigned long long x = member ? w : 0;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const auto other = shfl_up_u64(x, d, kWave);
            if (lane >= d) x += other;
        }
        if (member) mine = x - w;
        const auto total = shfl_u64(x, kWave - 1, kWave);
        if (lane == 0) wave_total[wave][current] = total;
        open &= ~members;
    }
    __syncthreads();
    if (pending) {
        for (int q = 0; q < wave; ++q) mine += wave_total[q][key];
        mine += running[key];
    }
    __syncthreads();
    if (threadIdx.x < 32) {
        unsigned long long pass = 0;
        for (int q = 0; q < Threads / kWave; ++q) pass += wave_total[q][threadIdx.x];
        running[threadIdx.x] += pass;
    }
    __syncthreads();
    return mine;
}

// Where a bucket stands against its part's threshold T, exactly for any
// non-negative weights (ranks of members lie in [0, count]):
//   sources      member leaves iff offset + rank < T
//                taken iff offset + count < T, refused iff offset >= T
//   destinations member enters iff offset + rank + weight <= T
//                taken iff offset + count <= T, refused iff offset > T
// The in-between intervals of a part's non-empty buckets are disjoint, so at
// most one bucket per part is a boundary bucket.
enum BucketSide { kTaken, kRefused, kBoundary };
__device__ __forceinline__ BucketSide source_side(unsigned long long offset, unsigned long long count,
                                                  unsigned long long threshold) {
    if (offset + count < threshold) return kTaken;
    if (offset >= threshold) return kRefused;
    return kBoundary;
}
__device__ __forceinline__ BucketSide destination_side(unsigned long long offset, unsigned long long count,
                                                       unsigned long long threshold) {
    if (offset + count <= threshold) return kTaken;
    if (offset > threshold) return kRefused;
    return kBoundary;
}

// Destination bucket codes in dbucket[i]: d >= 0 selected; -1 none; -2 source
// boundary (rank pending); -3 - d destination boundary (rank pending).
constexpr int kNoBucket = -1;
constexpr int kSourcePending = -2;

// Lanes per cache row: G * kRowSpan >= k.
inline int group_for_k(int k) {
    int g = 1;
    while (g * kRowSpan < k) g <<= 1;
    return g;
}

inline PowerTable make_power_table() {
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

}  // namespace
}  // namespace gpart_hip
