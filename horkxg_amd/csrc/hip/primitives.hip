// primitives.hip -- project-written sort, scan and compaction for gfx950.
//
// Replaces the reference's CUB/Thrust calls (cub::DeviceRadixSort SortKeys /
// SortPairs, thrust scans, copy_if; src/coarsen.cu:898-934,1017,1069-1111,
// 1336-1339,1550; src/refine_incremental.cuh:277).
//
// Radix sort: stable LSD over 8-bit digits. Per pass:
//   1. digit histogram of each 2048-key tile (LDS counters);
//   2. exclusive scan of the digit-major [256 x tiles] histogram (global bases);
//   3. scatter: each wave ranks its keys with a 64-lane multisplit (eight
//      ballots give the lanes sharing a digit), waves are combined in LDS, the
//      tile is staged in LDS in digit order and written out so that consecutive
//      threads store consecutive addresses of one digit run.
// Inputs of at most 8192 keys (four tiles) are sorted by one workgroup instead:
// three or more launches per 8-bit pass cost more than the whole sort in LDS.
#include "internal.hpp"
#include "scan.hpp"

namespace gpart_hip {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kSortItems = 8;  // per lane
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kDigits = 256;
// Sorts of at most this many tiles scan the digit histograms inside each scatter
// workgroup (at most 256 x 64 loads) instead of two scan launches per pass.
constexpr std::int64_t kInlineHistogramTiles = 64;

__device__ __forceinline__ std::int64_t tile_item(int wave, int it, int lane) {
    return static_cast<std::int64_t>(wave) * (kWave * kSortItems) + it * kWave + lane;
}

__global__ void __launch_bounds__(kSortThreads) radix_histogram_kernel(const std::uint64_t* keys, std::int64_t n, int shift,
                                                                       std::uint64_t mask, std::uint32_t* hist,
                                                                       std::int64_t tiles) {
    __shared__ std::uint32_t counts[kSortWaves][kDigits];
    for (int i = threadIdx.x; i < kSortWaves * kDigits; i += kSortThreads)
        (&counts[0][0])[i] = 0;
    __syncthreads();
    const int wave = threadIdx.x / kWave, lane = lane_id();
    const std::int64_t tile_start = static_cast<std::int64_t>(blockIdx.x) * kSortTile;
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const std::int64_t idx = tile_start + tile_item(wave, it, lane);
        if (idx < n) {
            const int d = static_cast<int>((keys[idx] >> shift) & mask);
            atomicAdd(&counts[wave][d], 1U);
        }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kDigits; d += kSortThreads) {
        std::uint32_t total = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) total += counts[w][d];
        hist[static_cast<std::int64_t>(d) * tiles + blockIdx.x] = total;
    }
}

// offsets: the digit-major exclusive scan of the tiles' histograms, or, with
// raw_histogram, the histograms themselves, which each workgroup then scans
// for its own tile (grids of at most kInlineHistogramTiles tiles).
template <typename V, bool HasValues>
__global__ void __launch_bounds__(kSortThreads) radix_scatter_kernel(const std::uint64_t* keys_in, std::uint64_t* keys_out,
                                                                     const V* values_in, V* values_out, std::int64_t n,
                                                                     int shift, std::uint64_t mask,
                                                                     const std::uint32_t* offsets, std::int64_t tiles,
                                                                     int raw_histogram) {
    __shared__ std::uint32_t wave_counts[kSortWaves][kDigits];
    __shared__ std::uint32_t digit_start[kDigits + 1];
    __shared__ std::uint32_t global_base[kDigits];
    __shared__ std::uint64_t stage_keys[kSortTile];
    __shared__ V stage_values[HasValues ? kSortTile : 1];
    __shared__ std::uint32_t scan_waves[kSortWaves + 1];

    const int wave = threadIdx.x / kWave, lane = lane_id();
    for (int i = threadIdx.x; i < kSortWaves * kDigits; i += kSortThreads)
        (&wave_counts[0][0])[i] = 0;
    __syncthreads();
    const std::int64_t tile_start = static_cast<std::int64_t>(blockIdx.x) * kSortTile;
    const int tile_valid = static_cast<int>(n - tile_start < kSortTile ? n - tile_start : kSortTile);

    std::uint64_t key[kSortItems];
    V value[kSortItems];
    int digit[kSortItems];
    std::uint32_t rank[kSortItems];
    const std::uint64_t below = lanemask_lt();
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const std::int64_t idx = tile_start + tile_item(wave, it, lane);
        const bool valid = idx < n;
        key[it] = valid ? keys_in[idx] : 0;
        if constexpr (HasValues) value[it] = valid ? values_in[idx] : V{};
        const int d = valid ? static_cast<int>((key[it] >> shift) & mask) : 0;
        digit[it] = valid ? d : -1;
        std::uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const std::uint64_t bits = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? bits : ~bits;
        }
        std::uint32_t before = 0;
        if (valid) before = wave_counts[wave][d];
        const std::uint32_t mine = static_cast<std::uint32_t>(__popcll(peers & below));
        rank[it] = before + mine;
        // the lowest peer publishes the new running count
        if (valid && mine == 0) wave_counts[wave][d] = before + static_cast<std::uint32_t>(__popcll(peers));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: wave offsets within the tile, digit totals
    std::uint32_t digit_total = 0;
    std::uint32_t all_tiles = 0, tiles_before = 0;  // raw_histogram: the digit over all tiles, and before this one
    if (threadIdx.x < kDigits) {
        const int d = threadIdx.x;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            const std::uint32_t c = wave_counts[w][d];
            wave_counts[w][d] = digit_total;
            digit_total += c;
        }
        if (raw_histogram) {
            const std::uint32_t* row = offsets + static_cast<std::int64_t>(d) * tiles;
            for (std::int64_t b = 0; b < tiles; ++b) {
                const std::uint32_t c = row[b];
                all_tiles += c;
                tiles_before += b < static_cast<std::int64_t>(blockIdx.x) ? c : 0;
            }
        } else {
            global_base[d] = offsets[static_cast<std::int64_t>(d) * tiles + blockIdx.x];
        }
    }
    if (raw_histogram) {  // the same branch in every thread
        std::uint32_t everything;
        const std::uint32_t digit_first = block_exclusive_scan(all_tiles, SumOp<std::uint32_t>{}, &everything,
                                                               scan_waves);
        if (threadIdx.x < kDigits) global_base[threadIdx.x] = digit_first + tiles_before;
    }
    std::uint32_t total;
    const std::uint32_t start = block_exclusive_scan(digit_total, SumOp<std::uint32_t>{}, &total, scan_waves);
    if (threadIdx.x < kDigits) digit_start[threadIdx.x] = start;
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        if (digit[it] < 0) continue;
        const int d = digit[it];
        const std::uint32_t pos = digit_start[d] + wave_counts[wave][d] + rank[it];
        stage_keys[pos] = key[it];
        if constexpr (HasValues) stage_values[pos] = value[it];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < tile_valid; j += kSortThreads) {
        const std::uint64_t k = stage_keys[j];
        const int d = static_cast<int>((k >> shift) & mask);
        const std::int64_t out = static_cast<std::int64_t>(global_base[d]) + (j - static_cast<int>(digit_start[d]));
        keys_out[out] = k;
        if constexpr (HasValues) values_out[out] = stage_values[j];
    }
}

// Stable sort of at most kLdsSortMax keys by their bits [shift, shift + width) in
// one workgroup: a bitonic network in LDS over (digits, input position) pairs.
// Positions are distinct, so the order is total and equal digits keep their
// input order, as in the multi-pass sort. The m - n padding entries carry the
// largest digits and positions and end up past the n real ones.
//
// For m >= 2 * threads each thread owns R = m / threads consecutive entries: the
// network's stages with strides below R compare entries of one thread and run in
// registers (one LDS round trip per merge size instead of one per stage); the
// longer strides go through LDS by pairs.
constexpr int kLdsSortMax = 4 * kSortTile;
constexpr int kLdsSortThreads = 1024;
static_assert(kLdsSortMax <= 65536, "positions are 16-bit");

// Compare-exchange of entries i < j of the network's stage for merge size `size`:
// ascending where bit `size` of i is clear.
__device__ __forceinline__ void bitonic_exchange(std::uint64_t& di, std::uint16_t& pi, std::uint64_t& dj,
                                                 std::uint16_t& pj, bool ascending) {
    const bool greater = di > dj || (di == dj && pi > pj);
    if (greater == ascending) {
        const std::uint64_t d = di;
        di = dj;
        dj = d;
        const std::uint16_t q = pi;
        pi = pj;
        pj = q;
    }
}

// The stages of merge size `size` with strides below R, on one thread's R entries
// starting at index base (sizes below R run every stride, from size / 2 down).
template <int R>
__device__ __forceinline__ void bitonic_registers(std::uint64_t (&d)[R], std::uint16_t (&q)[R], int base,
                                                  int size) {
#pragma unroll
    for (int stride = R / 2; stride > 0; stride >>= 1) {
        if (stride >= size) continue;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int j = i ^ stride;
            if (j > i) bitonic_exchange(d[i], q[i], d[j], q[j], ((base + i) & size) == 0);
        }
    }
}

template <typename V, bool HasValues, int R>
__device__ void bitonic_blocked(std::uint64_t* digit, std::uint16_t* position, const std::uint64_t* keys_in, int n,
                                int m, int shift, std::uint64_t mask) {
    const int base = threadIdx.x * R;
    std::uint64_t d[R];
    std::uint16_t q[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int at = base + i;
        d[i] = at < n ? (keys_in[at] >> shift) & mask : ~0ULL;
        q[i] = static_cast<std::uint16_t>(at);
    }
    // merge sizes up to R inside the thread
    for (int size = 2; size <= R; size <<= 1) bitonic_registers<R>(d, q, base, size);
#pragma unroll
    for (int i = 0; i < R; ++i) {
        digit[base + i] = d[i];
        position[base + i] = q[i];
    }
    __syncthreads();
    for (int size = 2 * R; size <= m; size <<= 1) {
        for (int stride = size >> 1; stride >= R; stride >>= 1) {
            for (int p = threadIdx.x; p < (m >> 1); p += kLdsSortThreads) {
                const int i = ((p & ~(stride - 1)) << 1) | (p & (stride - 1));
                const int j = i + stride;
                std::uint64_t di = digit[i], dj = digit[j];
                std::uint16_t pi = position[i], pj = position[j];
                bitonic_exchange(di, pi, dj, pj, (i & size) == 0);
                digit[i] = di;
                digit[j] = dj;
                position[i] = pi;
                position[j] = pj;
            }
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < R; ++i) {
            d[i] = digit[base + i];
            q[i] = position[base + i];
        }
        bitonic_registers<R>(d, q, base, size);
#pragma unroll
        for (int i = 0; i < R; ++i) {
            digit[base + i] = d[i];
            position[base + i] = q[i];
        }
        __syncthreads();
    }
}

template <typename V, bool HasValues>
__global__ void __launch_bounds__(kLdsSortThreads) lds_sort_kernel(const std::uint64_t* keys_in,
                                                                   std::uint64_t* keys_out, const V* values_in,
                                                                   V* values_out, int n, int m, int shift,
                                                                   std::uint64_t mask) {
    __shared__ std::uint64_t digit[kLdsSortMax];
    __shared__ std::uint16_t position[kLdsSortMax];
    if (m >= 8 * kLdsSortThreads) {
        bitonic_blocked<V, HasValues, 8>(digit, position, keys_in, n, m, shift, mask);
    } else if (m >= 4 * kLdsSortThreads) {
        bitonic_blocked<V, HasValues, 4>(digit, position, keys_in, n, m, shift, mask);
    } else if (m >= 2 * kLdsSortThreads) {
        bitonic_blocked<V, HasValues, 2>(digit, position, keys_in, n, m, shift, mask);
    } else {
        for (int i = threadIdx.x; i < m; i += kLdsSortThreads) {
            digit[i] = i < n ? (keys_in[i] >> shift) & mask : ~0ULL;
            position[i] = static_cast<std::uint16_t>(i);
        }
        __syncthreads();
        for (int size = 2; size <= m; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int p = threadIdx.x; p < (m >> 1); p += kLdsSortThreads) {
                    const int i = ((p & ~(stride - 1)) << 1) | (p & (stride - 1));
                    const int j = i + stride;
                    std::uint64_t di = digit[i], dj = digit[j];
                    std::uint16_t pi = position[i], pj = position[j];
                    bitonic_exchange(di, pi, dj, pj, (i & size) == 0);
                    digit[i] = di;
                    digit[j] = dj;
                    position[i] = pi;
                    position[j] = pj;
                }
                __syncthreads();
            }
        }
    }
    for (int r = threadIdx.x; r < n; r += kLdsSortThreads) {
        const int from = position[r];
        keys_out[r] = keys_in[from];
        if constexpr (HasValues) values_out[r] = values_in[from];
    }
}

template <typename V, bool HasValues>
bool radix_sort_impl(gpart_ctx* ctx, std::uint64_t* keys, std::uint64_t* keys_alt, V* values, V* values_alt,
                     std::int64_t n, int begin_bit, int end_bit) {
    if (n <= 1 || end_bit <= begin_bit) return false;
    if (n > (std::int64_t{1} << 32) - kSortTile)
        throw Error(GPART_ERR_INVALID_ARGUMENT, "radix sort supports fewer than 2^32 keys per call");
    if (n <= kLdsSortMax) {
        int m = 2;
        while (m < n) m <<= 1;
        const int width = end_bit - begin_bit;
        const std::uint64_t mask = width >= 64 ? ~0ULL : ((1ULL << width) - 1);
        hipLaunchKernelGGL((lds_sort_kernel<V, HasValues>), dim3(1), dim3(kLdsSortThreads), 0, ctx->stream, keys,
                           keys_alt, values, values_alt, static_cast<int>(n), m, begin_bit, mask);
        GPART_LAUNCH_CHECK();
        return true;
    }
    const std::int64_t tiles = (n + kSortTile - 1) / kSortTile;
    Scratch<std::uint32_t> hist(ctx, tiles * kDigits);
    std::uint64_t* src_k = keys;
    std::uint64_t* dst_k = keys_alt;
    V* src_v = values;
    V* dst_v = values_alt;
    bool in_alt = false;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        const int bits = end_bit - shift < 8 ? end_bit - shift : 8;
        const std::uint64_t mask = (1ULL << bits) - 1;
        hipLaunchKernelGGL(radix_histogram_kernel, dim3(static_cast<unsigned>(tiles)), dim3(kSortThreads), 0,
                           ctx->stream, src_k, n, shift, mask, hist.get(), tiles);
        GPART_LAUNCH_CHECK();
        const bool raw = tiles <= kInlineHistogramTiles;
        if (!raw)
            device_scan<std::uint32_t, SumOp<std::uint32_t>>(ctx, ArrayLoad<std::uint32_t>{hist.get()},
                                                             ExclusiveStore<std::uint32_t>{hist.get()},
                                                             tiles * kDigits, SumOp<std::uint32_t>{});
        hipLaunchKernelGGL((radix_scatter_kernel<V, HasValues>), dim3(static_cast<unsigned>(tiles)), dim3(kSortThreads),
                           0, ctx->stream, src_k, dst_k, src_v, dst_v, n, shift, mask, hist.get(), tiles, raw ? 1 : 0);
        GPART_LAUNCH_CHECK();
        std::swap(src_k, dst_k);
        std::swap(src_v, dst_v);
        in_alt = !in_alt;
    }
    return in_alt;
}

struct FlagLoad {
    const std::int32_t* flags;
    __device__ __forceinline__ std::int64_t operator()(std::int64_t i) const { return flags[i] != 0 ? 1 : 0; }
};

struct CompactStore {
    std::int32_t* out;
    std::int64_t n;
    std::int64_t* count;
    __device__ __forceinline__ void operator()(std::int64_t i, std::int64_t prefix, std::int64_t item) const {
        if (item) out[prefix] = static_cast<std::int32_t>(i);
        if (i == n - 1) *count = prefix + item;
    }
};

}  // namespace

template <typename V>
bool radix_sort_pairs(gpart_ctx* ctx, std::uint64_t* keys, std::uint64_t* keys_alt, V* values, V* values_alt,
                      std::int64_t n, int begin_bit, int end_bit) {
    return radix_sort_impl<V, true>(ctx, keys, keys_alt, values, values_alt, n, begin_bit, end_bit);
}
template bool radix_sort_pairs<std::uint32_t>(gpart_ctx*, std::uint64_t*, std::uint64_t*, std::uint32_t*,
                                              std::uint32_t*, std::int64_t, int, int);
template bool radix_sort_pairs<std::uint64_t>(gpart_ctx*, std::uint64_t*, std::uint64_t*, std::uint64_t*,
                                              std::uint64_t*, std::int64_t, int, int);

bool radix_sort_keys(gpart_ctx* ctx, std::uint64_t* keys, std::uint64_t* keys_alt, std::int64_t n, int begin_bit,
                     int end_bit) {
    return radix_sort_impl<std::uint32_t, false>(ctx, keys, keys_alt, nullptr, nullptr, n, begin_bit, end_bit);
}

template <typename T>
void exclusive_scan(gpart_ctx* ctx, const T* in, T* out, std::int64_t n, T* total_device) {
    device_scan<T, SumOp<T>>(ctx, ArrayLoad<T>{in}, ExclusiveStore<T>{out}, n, SumOp<T>{});
    (void)total_device;
}
template <typename T>
void inclusive_scan(gpart_ctx* ctx, const T* in, T* out, std::int64_t n) {
    device_scan<T, SumOp<T>>(ctx, ArrayLoad<T>{in}, InclusiveStore<T>{out}, n, SumOp<T>{});
}
template void exclusive_scan<std::uint64_t>(gpart_ctx*, const std::uint64_t*, std::uint64_t*, std::int64_t,
                                            std::uint64_t*);
template void exclusive_scan<std::int64_t>(gpart_ctx*, const std::int64_t*, std::int64_t*, std::int64_t,
                                           std::int64_t*);
template void exclusive_scan<std::uint32_t>(gpart_ctx*, const std::uint32_t*, std::uint32_t*, std::int64_t,
                                            std::uint32_t*);
template void inclusive_scan<std::uint64_t>(gpart_ctx*, const std::uint64_t*, std::uint64_t*, std::int64_t);

void compact_indices(gpart_ctx* ctx, const std::int32_t* flags, std::int64_t n, std::int32_t* out,
                     std::int64_t* d_count) {
    if (n <= 0) {
        fill_bytes(ctx, d_count, 0, sizeof(std::int64_t));
        return;
    }
    device_scan<std::int64_t, SumOp<std::int64_t>>(ctx, FlagLoad{flags}, CompactStore{out, n, d_count}, n,
                                                   SumOp<std::int64_t>{});
}

}  // namespace gpart_hip

using namespace gpart_hip;

extern "C" {

int gpart_sort_pairs_u64(gpart_ctx* ctx, uint64_t* keys, void* values, int value_bytes, int64_t n,
                         int begin_bit, int end_bit) {
    return guarded([&] {
        if (begin_bit < 0 || end_bit > 64 || begin_bit > end_bit)
            throw Error(GPART_ERR_INVALID_ARGUMENT, "invalid radix sort bit range");
        if (n <= 1) return;
        Scratch<std::uint64_t> keys_alt(ctx, n);
        bool alt = false;
        if (!values) {
            alt = radix_sort_keys(ctx, keys, keys_alt.get(), n, begin_bit, end_bit);
            if (alt) copy_device(ctx, keys, keys_alt.get(), n * 8);
        } else if (value_bytes == 4) {
            Scratch<std::uint32_t> values_alt(ctx, n);
            alt = radix_sort_pairs(ctx, keys, keys_alt.get(), static_cast<std::uint32_t*>(values), values_alt.get(), n,
                                   begin_bit, end_bit);
            if (alt) {
                copy_device(ctx, keys, keys_alt.get(), n * 8);
                copy_device(ctx, values, values_alt.get(), n * 4);
            }
        } else if (value_bytes == 8) {
            Scratch<std::uint64_t> values_alt(ctx, n);
            alt = radix_sort_pairs(ctx, keys, keys_alt.get(), static_cast<std::uint64_t*>(values), values_alt.get(), n,
                                   begin_bit, end_bit);
            if (alt) {
                copy_device(ctx, keys, keys_alt.get(), n * 8);
                copy_device(ctx, values, values_alt.get(), n * 8);
            }
        } else {
            throw Error(GPART_ERR_INVALID_ARGUMENT, "value_bytes must be 4 or 8");
        }
    });
}

int gpart_exclusive_scan_u64(gpart_ctx* ctx, const uint64_t* in, uint64_t* out, int64_t n) {
    return guarded([&] { exclusive_scan<std::uint64_t>(ctx, in, out, n); });
}

int gpart_compact_flags(gpart_ctx* ctx, const int32_t* flags, int64_t n, int32_t* out, int64_t* count) {
    return guarded([&] {
        Scratch<std::int64_t> d_count(ctx, 1);
        compact_indices(ctx, flags, n, out, d_count.get());
        *count = read_scalar(ctx, d_count.get());
    });
}

}  // extern "C"
