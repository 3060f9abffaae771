// graph.cpp -- input boundary: int64 CSR files, validation, host cut, transfers
// (reference: src/graph.cpp:47-241; same checks, messages and unit weights).
#include "log.hpp"
#include "check.hpp"
#include "graph.hpp"

#include <algorithm>
#include <fstream>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <type_traits>

namespace gpart {
namespace {

std::int64_t element_count(const std::string& path, std::size_t element_size) {
    std::ifstream input(path, std::ios::binary | std::ios::ate);
    if (!input) throw std::runtime_error("cannot open " + path);
    const auto bytes = input.tellg();
    if (bytes < 0 || bytes % static_cast<std::streamoff>(element_size) != 0)
        throw std::runtime_error("invalid binary size " + path);
    return static_cast<std::int64_t>(bytes / static_cast<std::streamoff>(element_size));
}

[[noreturn]] void standard_limit_error() {
    throw std::runtime_error("graph exceeds standard CSR limits; use gpart_coarsen_big");
}

template <typename Types>
void check_dimensions(std::int64_t vertex_count, std::int64_t edge_count) {
    using VertexT = typename Types::VertexT;
    using OffsetT = typename Types::OffsetT;
    using WeightT = typename Types::WeightT;
    if constexpr (std::is_same_v<Types, StandardTypes>) {
        if (vertex_count > std::numeric_limits<VertexT>::max() || edge_count > std::numeric_limits<OffsetT>::max() ||
            static_cast<std::uint64_t>(vertex_count) > std::numeric_limits<WeightT>::max() ||
            static_cast<std::uint64_t>(edge_count) > std::numeric_limits<WeightT>::max())
            standard_limit_error();
    } else if (vertex_count > std::numeric_limits<VertexT>::max()) {
        throw std::runtime_error("graph vertex ids exceed configured graph type");
    }
}

template <typename Types>
void narrow_offsets(const std::int64_t* disk, std::int64_t count, WeightedGraph<Types>& graph) {
    using OffsetT = typename Types::OffsetT;
    graph.offsets.resize(static_cast<std::size_t>(count));
    for (std::int64_t i = 0; i < count; ++i) {
        const auto offset = disk[i];
        if (offset < 0 || (i > 0 && offset < disk[i - 1])) throw std::runtime_error("CSR offsets are not monotone");
        if (offset > static_cast<std::int64_t>(std::numeric_limits<OffsetT>::max())) {
            if constexpr (std::is_same_v<Types, StandardTypes>) standard_limit_error();
            throw std::runtime_error("CSR offset exceeds configured graph type");
        }
        graph.offsets[static_cast<std::size_t>(i)] = static_cast<OffsetT>(offset);
    }
}

template <typename Types>
void narrow_neighbors(const std::int64_t* raw, std::int64_t take, std::int64_t vertex_count, std::int64_t done,
                      WeightedGraph<Types>& graph) {
    using VertexT = typename Types::VertexT;
    std::int64_t bad = 0;
#pragma omp parallel for reduction(| : bad) schedule(static)
    for (std::int64_t i = 0; i < take; ++i) {
        const auto u = raw[i];
        bad |= static_cast<std::int64_t>(u < 0 || u >= vertex_count);
        graph.neighbors[static_cast<std::size_t>(done + i)] = static_cast<VertexT>(u);
    }
    if (bad) throw std::runtime_error("CSR contains an invalid neighbor");
}

}  // namespace

template <typename Types>
WeightedGraph<Types> load_weighted_graph(const std::string& indptr_path, const std::string& indices_path) {
    using WeightT = typename Types::WeightT;
    const auto offset_count = element_count(indptr_path, sizeof(std::int64_t));
    const auto edge_count = element_count(indices_path, sizeof(std::int64_t));
    if (offset_count < 1) throw std::runtime_error("empty CSR indptr");
    const auto vertex_count = offset_count - 1;
    check_dimensions<Types>(vertex_count, edge_count);
    std::vector<std::int64_t> disk(static_cast<std::size_t>(offset_count));
    {
        std::ifstream input(indptr_path, std::ios::binary);
        if (!input) throw std::runtime_error("cannot open " + indptr_path);
        input.read(reinterpret_cast<char*>(disk.data()), static_cast<std::streamsize>(offset_count * 8));
        if (!input) throw std::runtime_error("cannot read " + indptr_path);
    }
    if (disk.front() != 0 || disk.back() != edge_count) throw std::runtime_error("CSR offsets do not match indices");
    WeightedGraph<Types> graph;
    narrow_offsets(disk.data(), offset_count, graph);
    std::ifstream input(indices_path, std::ios::binary);
    if (!input) throw std::runtime_error("cannot open " + indices_path);
    graph.neighbors.resize(static_cast<std::size_t>(edge_count));
    constexpr std::int64_t chunk = std::int64_t{1} << 22;
    std::vector<std::int64_t> raw(static_cast<std::size_t>(std::min(chunk, std::max<std::int64_t>(edge_count, 1))));
    std::int64_t done = 0;
    while (done < edge_count) {
        const auto take = std::min(chunk, edge_count - done);
        input.read(reinterpret_cast<char*>(raw.data()), static_cast<std::streamsize>(take * 8));
        if (!input) throw std::runtime_error("cannot read " + indices_path);
        narrow_neighbors(raw.data(), take, vertex_count, done, graph);
        done += take;
    }
    graph.edge_weights.assign(static_cast<std::size_t>(edge_count), WeightT{1});
    graph.vertex_weights.assign(static_cast<std::size_t>(vertex_count), WeightT{1});
    gpart::detail::log() << "vertices=" << vertex_count << " edge_entries=" << edge_count << " device_count=1\n";
    return graph;
}

template <typename Types>
WeightedGraph<Types> graph_from_int64(std::int64_t vertex_count, const std::int64_t* indptr, std::int64_t entries,
                                      const std::int64_t* indices) {
    using WeightT = typename Types::WeightT;
    if (vertex_count < 0) throw std::runtime_error("empty CSR indptr");
    check_dimensions<Types>(vertex_count, entries);
    if (indptr[0] != 0 || indptr[vertex_count] != entries) throw std::runtime_error("CSR offsets do not match indices");
    WeightedGraph<Types> graph;
    narrow_offsets(indptr, vertex_count + 1, graph);
    graph.neighbors.resize(static_cast<std::size_t>(entries));
    narrow_neighbors(indices, entries, vertex_count, 0, graph);
    graph.edge_weights.assign(static_cast<std::size_t>(entries), WeightT{1});
    graph.vertex_weights.assign(static_cast<std::size_t>(vertex_count), WeightT{1});
    return graph;
}

template <typename Types>
void validate_weighted_shape(const WeightedGraph<Types>& graph, bool allow_self_loops) {
    using OffsetT = typename Types::OffsetT;
    if (graph.offsets.size() != graph.vertex_weights.size() + 1 || graph.offsets.empty() ||
        graph.offsets.front() != OffsetT{0} || static_cast<std::int64_t>(graph.offsets.back()) != graph.edges() ||
        graph.neighbors.size() != graph.edge_weights.size())
        throw std::runtime_error("invalid weighted CSR shape");
    const auto n = graph.vertices();
    for (std::int64_t v = 0; v < n; ++v)
        if (graph.offsets[v] > graph.offsets[v + 1]) throw std::runtime_error("weighted CSR offsets are not monotone");
    const auto m = graph.edges();
    unsigned invalid = 0;
#pragma omp parallel for reduction(| : invalid) schedule(static)
    for (std::int64_t e = 0; e < m; ++e) invalid |= static_cast<unsigned>(graph.neighbors[e] < 0 || graph.neighbors[e] >= n);
    if (invalid) throw std::runtime_error("weighted CSR contains an invalid neighbor");
    if (allow_self_loops) return;
    unsigned loops = 0;
#pragma omp parallel for reduction(| : loops) schedule(dynamic, 4096)
    for (std::int64_t v = 0; v < n; ++v)
        for (auto e = graph.offsets[v]; e < graph.offsets[v + 1]; ++e) loops |= static_cast<unsigned>(graph.neighbors[e] == v);
    if (loops) throw std::runtime_error("contracted graph contains a self loop");
}

template <typename Types>
void validate_weighted_csr(const WeightedGraph<Types>& graph, bool allow_self_loops) {
    validate_weighted_shape(graph, allow_self_loops);
    const auto n = graph.vertices();
    // 0 when every entry of row v has its reverse entry with the same weight, else 1
    // (no reverse entry) or 2 (weights differ) for the row's first failing entry
    auto row_error = [&](std::int64_t v) -> int {
        for (auto e = graph.offsets[v]; e < graph.offsets[v + 1]; ++e) {
            const auto u = graph.neighbors[e];
            const auto rb = graph.neighbors.begin() + graph.offsets[u];
            const auto re = graph.neighbors.begin() + graph.offsets[u + 1];
            const auto r = std::lower_bound(rb, re, static_cast<typename Types::VertexT>(v));
            if (r == re || *r != v) return 1;
            if (graph.edge_weights[e] != graph.edge_weights[static_cast<std::size_t>(r - graph.neighbors.begin())])
                return 2;
        }
        return 0;
    };
    // the first failing row in vertex order decides the message, as a serial scan would
    std::int64_t first = n;
#pragma omp parallel for reduction(min : first) schedule(dynamic, 4096)
    for (std::int64_t v = 0; v < n; ++v)
        if (v < first && row_error(v)) first = v;
    if (first < n) {
        if (row_error(first) == 1) throw std::runtime_error("weighted CSR is not bidirectional");
        throw std::runtime_error("weighted CSR reverse edge weights differ");
    }
}

template <typename Types>
std::uint64_t host_cut(const WeightedGraph<Types>& graph, const std::vector<typename Types::VertexT>& labels) {
    std::uint64_t directed = 0;
    const auto n = graph.vertices();
#pragma omp parallel for reduction(+ : directed) schedule(static)
    for (std::int64_t v = 0; v < n; ++v)
        for (auto e = graph.offsets[v]; e < graph.offsets[v + 1]; ++e)
            if (labels[v] != labels[graph.neighbors[e]]) directed += static_cast<std::uint64_t>(graph.edge_weights[e]);
    if (directed & 1ULL) throw std::runtime_error("weighted CSR cut is not symmetric");
    return directed / 2;
}

namespace {
// gpart_csr_from_int64 status bits, mapped to the host loader's messages in the
// host loader's order of checks.
template <typename Types>
[[noreturn]] void throw_int64_status(int status) {
    if (status & 8) throw std::runtime_error("CSR offsets do not match indices");
    if (status & 1) throw std::runtime_error("CSR offsets are not monotone");
    if (status & 2) {
        if constexpr (std::is_same_v<Types, StandardTypes>) standard_limit_error();
        throw std::runtime_error("CSR offset exceeds configured graph type");
    }
    throw std::runtime_error("CSR contains an invalid neighbor");
}

template <typename Types>
DeviceWeightedGraph<Types> adopt_csr(const gpart_device_csr& csr) {
    using O = typename Types::OffsetT;
    using W = typename Types::WeightT;
    DeviceWeightedGraph<Types> out;
    const auto n = static_cast<std::size_t>(csr.n), m = static_cast<std::size_t>(csr.m);
    out.offsets = DeviceVector<O>::adopt(static_cast<O*>(csr.offsets), n + 1);
    out.neighbors = DeviceVector<std::int32_t>::adopt(csr.neighbors, m);
    out.edge_weights = DeviceVector<W>::adopt(static_cast<W*>(csr.edge_weights), m);
    out.vertex_weights = DeviceVector<W>::adopt(static_cast<W*>(csr.vertex_weights), n);
    return out;
}
}  // namespace

template <typename Types>
DeviceWeightedGraph<Types> device_graph_from_int64(std::int64_t vertex_count, const std::int64_t* indptr,
                                                   std::int64_t entries, const std::int64_t* indices) {
    if (vertex_count < 0) throw std::runtime_error("empty CSR indptr");
    check_dimensions<Types>(vertex_count, entries);
    gpart_device_csr csr{};
    std::int32_t status = 0;
    GPART_CHECK(gpart_csr_from_int64(detail::context(), Types::kWidth, vertex_count, indptr, entries, indices, &csr,
                                     &status));
    if (status) throw_int64_status<Types>(status);
    return adopt_csr<Types>(csr);
}

template <typename Types>
DeviceWeightedGraph<Types> load_device_graph(const std::string& indptr_path, const std::string& indices_path) {
    const auto offset_count = element_count(indptr_path, sizeof(std::int64_t));
    const auto edge_count = element_count(indices_path, sizeof(std::int64_t));
    if (offset_count < 1) throw std::runtime_error("empty CSR indptr");
    const auto vertex_count = offset_count - 1;
    check_dimensions<Types>(vertex_count, edge_count);
    gpart_device_csr csr{};
    std::int32_t status = 0;
    GPART_CHECK(gpart_csr_from_int64_files(detail::context(), Types::kWidth, indptr_path.c_str(),
                                           indices_path.c_str(), vertex_count, edge_count, &csr, &status));
    if (status) throw_int64_status<Types>(status);
    detail::log() << "vertices=" << vertex_count << " edge_entries=" << edge_count << " device_count=1\n";
    return adopt_csr<Types>(csr);
}

template <typename Types>
DeviceWeightedGraph<Types> make_device_weighted(const WeightedGraph<Types>& graph) {
    DeviceWeightedGraph<Types> out;
    out.offsets.assign(graph.offsets);
    out.neighbors.assign(graph.neighbors);
    out.edge_weights.assign(graph.edge_weights);
    out.vertex_weights.assign(graph.vertex_weights);
    return out;
}

template <typename Types>
WeightedGraph<Types> copy_device_weighted(const DeviceWeightedGraph<Types>& graph) {
    WeightedGraph<Types> out;
    out.offsets = graph.offsets.to_host();
    out.neighbors = graph.neighbors.to_host();
    out.edge_weights = graph.edge_weights.to_host();
    out.vertex_weights = graph.vertex_weights.to_host();
    return out;
}

#define GPART_INSTANTIATE_GRAPH(T)                                                                              \
    template WeightedGraph<T> load_weighted_graph<T>(const std::string&, const std::string&);                  \
    template WeightedGraph<T> graph_from_int64<T>(std::int64_t, const std::int64_t*, std::int64_t,              \
                                                  const std::int64_t*);                                         \
    template void validate_weighted_shape<T>(const WeightedGraph<T>&, bool);                                   \
    template void validate_weighted_csr<T>(const WeightedGraph<T>&, bool);                                     \
    template std::uint64_t host_cut<T>(const WeightedGraph<T>&, const std::vector<typename T::VertexT>&);     \
    template DeviceWeightedGraph<T> make_device_weighted<T>(const WeightedGraph<T>&);                          \
    template DeviceWeightedGraph<T> device_graph_from_int64<T>(std::int64_t, const std::int64_t*, std::int64_t, \
                                                               const std::int64_t*);                            \
    template DeviceWeightedGraph<T> load_device_graph<T>(const std::string&, const std::string&);              \
    template WeightedGraph<T> copy_device_weighted<T>(const DeviceWeightedGraph<T>&);
GPART_INSTANTIATE_GRAPH(StandardTypes)
GPART_INSTANTIATE_GRAPH(BigTypes)

}  // namespace gpart
