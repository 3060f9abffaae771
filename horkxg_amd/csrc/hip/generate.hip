// generate.hip -- synthetic input graphs built on the device.
//
// The same graphs as the host generators of host/generate.cpp, entry for entry:
// 2D 4-neighbour grids, 3D 6-neighbour meshes and symmetrised R-MAT graphs (self
// loops dropped, {min, max} pairs deduplicated, both directions emitted, rows
// sorted; graph_untils/csr_symmetrize.cpp:277-336 of the reference). R-MAT
// sample i draws its quadrant bits from splitmix64 states seeded with
// splitmix64(seed ^ splitmix64(i)), as the host generator does, so the two
// agree bit for bit. On the device the pairs go through two radix sorts (pairs,
// then directed entries) instead of the host's comparison sorts: R-MAT scale 26
// (2^30 samples) takes seconds instead of minutes.
#include "internal.hpp"
#include "scan.hpp"

#include <algorithm>

namespace gpart_hip {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(std::int64_t threads) { return static_cast<unsigned>((threads + kBlock - 1) / kBlock); }
inline unsigned capped_grid(std::int64_t threads) {
    return static_cast<unsigned>(std::max<std::int64_t>(1, std::min<std::int64_t>(grid_for(threads), 65536)));
}

__device__ __forceinline__ std::uint64_t splitmix64(std::uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

// A 53-bit integer times 2^-53: exact in FP64, the same value as on the host.
__device__ __forceinline__ double unit_interval(std::uint64_t x) {
    return static_cast<double>(x >> 11) * (1.0 / 9007199254740992.0);
}

// Sample i -> key (min << scale | max) + 1, or 0 for a self loop (sorted first,
// dropped with the duplicates).
__global__ void rmat_samples_kernel(std::int64_t samples, int scale, double a, double ab, double abc,
                                    std::uint64_t seed, std::uint64_t* keys) {
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < samples;
         i += static_cast<std::int64_t>(gridDim.x) * blockDim.x) {
        std::uint64_t state = splitmix64(seed ^ splitmix64(static_cast<std::uint64_t>(i)));
        std::uint64_t u = 0, v = 0;
        for (int level = 0; level < scale; ++level) {
            state = splitmix64(state);
            const double r = unit_interval(state);
            const std::uint64_t bit = std::uint64_t{1} << (scale - 1 - level);
            if (r < a) {
            } else if (r < ab) {
                v |= bit;
            } else if (r < abc) {
                u |= bit;
            } else {
                u |= bit;
                v |= bit;
            }
        }
        keys[i] = u == v ? 0 : (((u < v ? u : v) << scale) | (u < v ? v : u)) + 1;
    }
}

struct FirstOfRunLoad {
    const std::uint64_t* keys;
    __device__ __forceinline__ long long operator()(std::int64_t i) const {
        return keys[i] != 0 && (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
    }
};

// Unique pair j -> directed entries 2j = (min -> max) and 2j + 1 = (max -> min),
// keys (source << scale | target).
struct DirectedStore {
    const std::uint64_t* keys;
    std::uint64_t* directed;
    int scale;
    long long* count;
    std::int64_t n;
    __device__ __forceinline__ void operator()(std::int64_t i, long long prefix, long long item) const {
        if (item) {
            const std::uint64_t pair = keys[i] - 1;
            const std::uint64_t lo = pair >> scale, hi = pair & ((std::uint64_t{1} << scale) - 1);
            directed[2 * prefix] = (lo << scale) | hi;
            directed[2 * prefix + 1] = (hi << scale) | lo;
        }
        if (i == n - 1) *count = prefix + item;
    }
};

template <typename O>
__global__ void csr_from_keys_kernel(std::int64_t m, const std::uint64_t* keys, int scale, std::int64_t n,
                                     O* offsets, std::int32_t* neighbors) {
    const std::uint64_t mask = (std::uint64_t{1} << scale) - 1;
    for (std::int64_t j = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j <= m;
         j += static_cast<std::int64_t>(gridDim.x) * blockDim.x) {
        if (j < m) neighbors[j] = static_cast<std::int32_t>(keys[j] & mask);
        // rows (previous source, this source] start at j; j == m closes the tail
        const std::int64_t here = j < m ? static_cast<std::int64_t>(keys[j] >> scale) : n;
        const std::int64_t prev = j == 0 ? -1 : static_cast<std::int64_t>(keys[j - 1] >> scale);
        for (std::int64_t r = prev + 1; r <= here; ++r) offsets[r] = static_cast<O>(j);
    }
}

// Grid and mesh: vertex v's degree, then its neighbours in ascending order.
struct Lattice {
    std::int64_t nx, ny, nz;  // nz == 1: 2D grid (rows = ny, cols = nx)
    bool mesh;
    __device__ __forceinline__ int degree(std::int64_t v) const {
        const std::int64_t x = v % nx, y = (v / nx) % ny, z = v / (nx * ny);
        int d = (x > 0) + (x + 1 < nx) + (y > 0) + (y + 1 < ny);
        if (mesh) d += (z > 0) + (z + 1 < nz);
        return d;
    }
};

struct LatticeDegreeLoad {
    Lattice l;
    std::int64_t n;
    __device__ __forceinline__ long long operator()(std::int64_t v) const { return v < n ? l.degree(v) : 0; }
};

template <typename O>
struct OffsetStore {
    O* offsets;
    __device__ __forceinline__ void operator()(std::int64_t v, long long prefix, long long) const {
        offsets[v] = static_cast<O>(prefix);
    }
};

template <typename O>
__global__ void lattice_neighbors_kernel(Lattice l, std::int64_t n, const O* offsets, std::int32_t* neighbors) {
    for (std::int64_t v = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<std::int64_t>(gridDim.x) * blockDim.x) {
        const std::int64_t x = v % l.nx, y = (v / l.nx) % l.ny, z = v / (l.nx * l.ny);
        const std::int64_t plane = l.nx * l.ny;
        std::int64_t e = static_cast<std::int64_t>(offsets[v]);
        if (l.mesh && z > 0) neighbors[e++] = static_cast<std::int32_t>(v - plane);
        if (y > 0) neighbors[e++] = static_cast<std::int32_t>(v - l.nx);
        if (x > 0) neighbors[e++] = static_cast<std::int32_t>(v - 1);
        if (x + 1 < l.nx) neighbors[e++] = static_cast<std::int32_t>(v + 1);
        if (y + 1 < l.ny) neighbors[e++] = static_cast<std::int32_t>(v + l.nx);
        if (l.mesh && z + 1 < l.nz) neighbors[e++] = static_cast<std::int32_t>(v + plane);
    }
}

template <typename O>
void rmat_impl(gpart_ctx* ctx, int scale, std::int64_t edge_factor, double a, double b, double c, std::uint64_t seed,
               void** offsets_out, std::int32_t** neighbors_out, std::int64_t* n_out, std::int64_t* m_out) {
    const std::int64_t n = std::int64_t{1} << scale;
    const std::int64_t samples = edge_factor * n;
    if (samples >= (std::int64_t{1} << 32) - 4096)
        throw Error(GPART_ERR_INVALID_ARGUMENT, "R-MAT sample count exceeds the device sort");
    const int key_bits = 2 * scale + 1;  // pairs + 1, zero for self loops
    Scratch<std::uint64_t> keys(ctx, samples), keys_alt(ctx, samples);
    hipLaunchKernelGGL(rmat_samples_kernel, dim3(capped_grid(samples)), dim3(kBlock), 0, ctx->stream, samples, scale,
                       a, a + b, a + b + c, seed, keys.get());
    GPART_LAUNCH_CHECK();
    const std::uint64_t* sorted =
        radix_sort_keys(ctx, keys.get(), keys_alt.get(), samples, 0, key_bits) ? keys_alt.get() : keys.get();
    // unique pairs -> both directions; at most 2 * samples entries
    Scratch<std::uint64_t> directed(ctx, 2 * samples), directed_alt;
    Scratch<long long> count(ctx, 1);
    device_scan<long long, SumOp<long long>>(ctx, FirstOfRunLoad{sorted},
                                             DirectedStore{sorted, directed.get(), scale, count.get(), samples},
                                             samples, SumOp<long long>{});
    const std::int64_t pairs = read_scalar(ctx, count.get());
    keys.release();
    keys_alt.release();
    const std::int64_t m = 2 * pairs;
    if (m > std::int64_t{0x7fffffff} && sizeof(O) == 4)
        throw Error(GPART_ERR_OVERFLOW, "directed entries exceed the standard width; use the big width");
    directed_alt.reset(ctx, std::max<std::int64_t>(m, 1));
    const std::uint64_t* rows = directed.get();
    if (m > 1 && radix_sort_keys(ctx, directed.get(), directed_alt.get(), m, 0, 2 * scale)) rows = directed_alt.get();
    auto* offsets = static_cast<O*>(ctx_alloc(ctx, sizeof(O) * static_cast<std::size_t>(n + 1)));
    auto* neighbors = static_cast<std::int32_t*>(ctx_alloc(ctx, sizeof(std::int32_t) * std::max<std::int64_t>(m, 1)));
    hipLaunchKernelGGL(csr_from_keys_kernel<O>, dim3(capped_grid(m + 1)), dim3(kBlock), 0, ctx->stream, m, rows,
                       scale, n, offsets, neighbors);
    GPART_LAUNCH_CHECK();
    *offsets_out = offsets;
    *neighbors_out = neighbors;
    *n_out = n;
    *m_out = m;
}

template <typename O>
void lattice_impl(gpart_ctx* ctx, Lattice l, void** offsets_out, std::int32_t** neighbors_out, std::int64_t* n_out,
                  std::int64_t* m_out) {
    const std::int64_t n = l.nx * l.ny * l.nz;
    auto* offsets = static_cast<O*>(ctx_alloc(ctx, sizeof(O) * static_cast<std::size_t>(n + 1)));
    device_scan<long long, SumOp<long long>>(ctx, LatticeDegreeLoad{l, n}, OffsetStore<O>{offsets}, n + 1,
                                             SumOp<long long>{});
    O last{};
    read_back(ctx, &last, offsets + n, sizeof(O));
    const std::int64_t m = static_cast<std::int64_t>(last);
    auto* neighbors = static_cast<std::int32_t*>(ctx_alloc(ctx, sizeof(std::int32_t) * std::max<std::int64_t>(m, 1)));
    hipLaunchKernelGGL(lattice_neighbors_kernel<O>, dim3(capped_grid(n)), dim3(kBlock), 0, ctx->stream, l, n, offsets,
                       neighbors);
    GPART_LAUNCH_CHECK();
    *offsets_out = offsets;
    *neighbors_out = neighbors;
    *n_out = n;
    *m_out = m;
}

void check_width(int width) {
    if (width != GPART_WIDTH_STANDARD && width != GPART_WIDTH_BIG)
        throw Error(GPART_ERR_INVALID_ARGUMENT, "unknown graph width");
}

}  // namespace
}  // namespace gpart_hip

using namespace gpart_hip;

extern "C" {

int gpart_generate_rmat_device(gpart_ctx* ctx, int width, int scale, int64_t edge_factor, double a, double b,
                               double c, uint64_t seed, void** offsets, int32_t** neighbors, int64_t* n, int64_t* m) {
    return guarded([&] {
        check_width(width);
        if (scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
            throw Error(GPART_ERR_INVALID_ARGUMENT, "invalid R-MAT parameters");
        if (width == GPART_WIDTH_STANDARD)
            rmat_impl<std::int32_t>(ctx, scale, edge_factor, a, b, c, seed, offsets, neighbors, n, m);
        else
            rmat_impl<std::int64_t>(ctx, scale, edge_factor, a, b, c, seed, offsets, neighbors, n, m);
    });
}

int gpart_generate_lattice_device(gpart_ctx* ctx, int width, int64_t nx, int64_t ny, int64_t nz, void** offsets,
                                  int32_t** neighbors, int64_t* n, int64_t* m) {
    return guarded([&] {
        check_width(width);
        if (nx <= 0 || ny <= 0 || nz <= 0 || nx * ny * nz > 0x7fffffff)
            throw Error(GPART_ERR_INVALID_ARGUMENT, "invalid lattice size");
        const Lattice l{nx, ny, nz, nz > 1};
        if (width == GPART_WIDTH_STANDARD) lattice_impl<std::int32_t>(ctx, l, offsets, neighbors, n, m);
        else lattice_impl<std::int64_t>(ctx, l, offsets, neighbors, n, m);
    });
}

}  // extern "C"
