// graph.hpp -- weighted CSR containers and the input boundary
// (reference: include/graph.hpp:13-74, src/graph.cpp:47-241).
#pragma once

#include "device_vector.hpp"
#include "graph_types.hpp"

#include <cstdint>
#include <string>
#include <vector>

namespace gpart {

template <typename Types>
struct WeightedGraph {
    using VertexT = typename Types::VertexT;
    using OffsetT = typename Types::OffsetT;
    using WeightT = typename Types::WeightT;

    std::vector<OffsetT> offsets;
    std::vector<VertexT> neighbors;
    std::vector<WeightT> edge_weights;
    std::vector<WeightT> vertex_weights;

    std::int64_t vertices() const { return static_cast<std::int64_t>(vertex_weights.size()); }
    std::int64_t edges() const { return static_cast<std::int64_t>(neighbors.size()); }
};

template <typename Types>
struct DeviceWeightedGraph {
    using VertexT = typename Types::VertexT;
    using OffsetT = typename Types::OffsetT;
    using WeightT = typename Types::WeightT;

    DeviceVector<OffsetT> offsets;
    DeviceVector<VertexT> neighbors;
    DeviceVector<WeightT> edge_weights;
    DeviceVector<WeightT> vertex_weights;

    std::int64_t vertices() const { return static_cast<std::int64_t>(vertex_weights.size()); }
    std::int64_t edges() const { return static_cast<std::int64_t>(neighbors.size()); }

    // Plain view for the kernel library's C ABI.
    gpart_csr_view view() const {
        gpart_csr_view v;
        v.n = vertices();
        v.m = edges();
        v.width = Types::kWidth;
        v.offsets = offsets.data();
        v.neighbors = neighbors.data();
        v.edge_weights = edge_weights.data();
        v.vertex_weights = vertex_weights.data();
        return v;
    }
};

// int64 indptr/indices files -> typed CSR with unit weights (src/graph.cpp:47-122).
template <typename Types>
WeightedGraph<Types> load_weighted_graph(const std::string& indptr_path, const std::string& indices_path);

// Same checks on in-memory int64 arrays (the C ABI's entry for FFI callers).
template <typename Types>
WeightedGraph<Types> graph_from_int64(std::int64_t vertices, const std::int64_t* indptr,
                                      std::int64_t entries, const std::int64_t* indices);

// The same inputs and checks (and messages), narrowed on the device: only the int64
// arrays cross PCIe, the weights are filled on the device, and no host copy is
// kept. Shape validation (monotone offsets, neighbours in range) is included;
// symmetry is not (validate_weighted_csr needs a host copy).
template <typename Types>
DeviceWeightedGraph<Types> device_graph_from_int64(std::int64_t vertices, const std::int64_t* indptr,
                                                   std::int64_t entries, const std::int64_t* indices);
template <typename Types>
DeviceWeightedGraph<Types> load_device_graph(const std::string& indptr_path, const std::string& indices_path);

template <typename Types>
void validate_weighted_shape(const WeightedGraph<Types>& graph, bool allow_self_loops);

template <typename Types>
void validate_weighted_csr(const WeightedGraph<Types>& graph, bool allow_self_loops);

template <typename Types>
std::uint64_t host_cut(const WeightedGraph<Types>& graph, const std::vector<typename Types::VertexT>& labels);

template <typename Types>
DeviceWeightedGraph<Types> make_device_weighted(const WeightedGraph<Types>& graph);

template <typename Types>
WeightedGraph<Types> copy_device_weighted(const DeviceWeightedGraph<Types>& graph);

}  // namespace gpart
