// generate.cpp -- synthetic input graphs as int64 CSR (the reference's input format,
// README.md:52-55): 2D 4-neighbour grids, 3D 6-neighbour meshes and symmetrized
// R-MAT graphs. R-MAT preprocessing follows the reference's csr_symmetrize
// semantics (graph_untils/csr_symmetrize.cpp:277-336,461-538): self loops
// dropped, {min,max} pairs deduplicated, both directions emitted, rows sorted.
#include "gpart.h"

#include <parallel/algorithm>

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <new>
#include <vector>

struct gpart_host_csr {
    std::vector<std::int64_t> indptr;
    std::vector<std::int64_t> indices;
};

namespace {

std::uint64_t splitmix64(std::uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

double unit(std::uint64_t x) { return static_cast<double>(x >> 11) * (1.0 / 9007199254740992.0); }

// CSR from sorted unique directed keys (source << 32 | target).
gpart_host_csr* from_directed_keys(std::int64_t n, const std::vector<std::uint64_t>& keys) {
    auto* out = new gpart_host_csr;
    out->indptr.assign(static_cast<std::size_t>(n) + 1, 0);
    out->indices.resize(keys.size());
#pragma omp parallel for schedule(static)
    for (std::int64_t i = 0; i < static_cast<std::int64_t>(keys.size()); ++i)
        out->indices[i] = static_cast<std::int64_t>(keys[i] & 0xffffffffULL);
    for (const auto k : keys) ++out->indptr[(k >> 32) + 1];
    for (std::int64_t v = 0; v < n; ++v) out->indptr[v + 1] += out->indptr[v];
    return out;
}

}  // namespace

extern "C" {

gpart_host_csr* gpart_generate_grid2d(int64_t rows, int64_t cols) {
    if (rows <= 0 || cols <= 0) return nullptr;
    auto* out = new (std::nothrow) gpart_host_csr;
    if (!out) return nullptr;
    const std::int64_t n = rows * cols;
    out->indptr.resize(static_cast<std::size_t>(n) + 1);
    out->indptr[0] = 0;
    for (std::int64_t r = 0; r < rows; ++r)
        for (std::int64_t c = 0; c < cols; ++c) {
            const std::int64_t v = r * cols + c;
            const std::int64_t deg = (r > 0) + (c > 0) + (c + 1 < cols) + (r + 1 < rows);
            out->indptr[v + 1] = deg;
        }
    for (std::int64_t v = 0; v < n; ++v) out->indptr[v + 1] += out->indptr[v];
    out->indices.resize(static_cast<std::size_t>(out->indptr[n]));
#pragma omp parallel for schedule(static)
    for (std::int64_t r = 0; r < rows; ++r)
        for (std::int64_t c = 0; c < cols; ++c) {
            const std::int64_t v = r * cols + c;
            auto e = out->indptr[v];
            if (r > 0) out->indices[e++] = v - cols;
            if (c > 0) out->indices[e++] = v - 1;
            if (c + 1 < cols) out->indices[e++] = v + 1;
            if (r + 1 < rows) out->indices[e++] = v + cols;
        }
    return out;
}

gpart_host_csr* gpart_generate_mesh3d(int64_t nx, int64_t ny, int64_t nz) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return nullptr;
    auto* out = new (std::nothrow) gpart_host_csr;
    if (!out) return nullptr;
    const std::int64_t n = nx * ny * nz, plane = nx * ny;
    out->indptr.resize(static_cast<std::size_t>(n) + 1);
    out->indptr[0] = 0;
    for (std::int64_t z = 0; z < nz; ++z)
        for (std::int64_t y = 0; y < ny; ++y)
            for (std::int64_t x = 0; x < nx; ++x) {
                const std::int64_t v = (z * ny + y) * nx + x;
                out->indptr[v + 1] = (z > 0) + (y > 0) + (x > 0) + (x + 1 < nx) + (y + 1 < ny) + (z + 1 < nz);
            }
    for (std::int64_t v = 0; v < n; ++v) out->indptr[v + 1] += out->indptr[v];
    out->indices.resize(static_cast<std::size_t>(out->indptr[n]));
#pragma omp parallel for schedule(static)
    for (std::int64_t z = 0; z < nz; ++z)
        for (std::int64_t y = 0; y < ny; ++y)
            for (std::int64_t x = 0; x < nx; ++x) {
                const std::int64_t v = (z * ny + y) * nx + x;
                auto e = out->indptr[v];
                if (z > 0) out->indices[e++] = v - plane;
                if (y > 0) out->indices[e++] = v - nx;
                if (x > 0) out->indices[e++] = v - 1;
                if (x + 1 < nx) out->indices[e++] = v + 1;
                if (y + 1 < ny) out->indices[e++] = v + nx;
                if (z + 1 < nz) out->indices[e++] = v + plane;
            }
    return out;
}

gpart_host_csr* gpart_generate_rmat(int scale, int64_t edge_factor, double a, double b, double c, uint64_t seed) {
    if (scale < 1 || scale > 31 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return nullptr;
    const std::int64_t n = std::int64_t{1} << scale;
    const std::int64_t samples = edge_factor * n;
    const double ab = a + b, abc = a + b + c;
    std::vector<std::uint64_t> pairs(static_cast<std::size_t>(samples));
#pragma omp parallel for schedule(static)
    for (std::int64_t i = 0; i < samples; ++i) {
        std::uint64_t state = splitmix64(seed ^ splitmix64(static_cast<std::uint64_t>(i)));
        std::uint64_t u = 0, v = 0;
        for (int level = 0; level < scale; ++level) {
            state = splitmix64(state);
            const double r = unit(state);
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
        // undirected key {min, max}; self loops marked for removal
        pairs[i] = u == v ? ~0ULL : ((u < v ? u : v) << 32) | (u < v ? v : u);
    }
    __gnu_parallel::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    if (!pairs.empty() && pairs.back() == ~0ULL) pairs.pop_back();
    std::vector<std::uint64_t> directed(pairs.size() * 2);
#pragma omp parallel for schedule(static)
    for (std::int64_t i = 0; i < static_cast<std::int64_t>(pairs.size()); ++i) {
        const auto k = pairs[i];
        directed[2 * i] = k;
        directed[2 * i + 1] = ((k & 0xffffffffULL) << 32) | (k >> 32);
    }
    std::vector<std::uint64_t>().swap(pairs);
    __gnu_parallel::sort(directed.begin(), directed.end());
    return from_directed_keys(n, directed);
}

int64_t gpart_host_csr_vertices(const gpart_host_csr* h) { return static_cast<int64_t>(h->indptr.size()) - 1; }
int64_t gpart_host_csr_entries(const gpart_host_csr* h) { return static_cast<int64_t>(h->indices.size()); }
const int64_t* gpart_host_csr_indptr(const gpart_host_csr* h) { return h->indptr.data(); }
const int64_t* gpart_host_csr_indices(const gpart_host_csr* h) { return h->indices.data(); }

int gpart_host_csr_write(const gpart_host_csr* h, const char* indptr_path, const char* indices_path) {
    std::ofstream a(indptr_path, std::ios::binary | std::ios::trunc);
    std::ofstream b(indices_path, std::ios::binary | std::ios::trunc);
    if (!a || !b) return GPART_ERR_RUNTIME;
    a.write(reinterpret_cast<const char*>(h->indptr.data()), static_cast<std::streamsize>(h->indptr.size() * 8));
    b.write(reinterpret_cast<const char*>(h->indices.data()), static_cast<std::streamsize>(h->indices.size() * 8));
    return (a && b) ? GPART_OK : GPART_ERR_RUNTIME;
}

void gpart_host_csr_free(gpart_host_csr* h) { delete h; }

}  // extern "C"
