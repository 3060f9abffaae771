// hierarchy.hpp -- coarsening hierarchy and its int32 export
// (reference: include/hierarchy.hpp:15-126; same on-disk format).
#pragma once

#include "graph.hpp"

#include <string>
#include <vector>

namespace gpart {

template <typename Types>
struct Hierarchy {
    using VertexT = typename Types::VertexT;
    std::vector<WeightedGraph<Types>> levels;
    std::vector<std::vector<VertexT>> fine_to_coarse;
    std::string stop_reason = "capacity_floor";
    double hierarchy_loop_seconds = 0.0;
    double snapshot_seconds = 0.0;
    std::size_t layer_count() const { return levels.size(); }
};

// Levels and projection maps stay on the device for the production pipeline.
template <typename Types>
struct DeviceHierarchy {
    std::vector<DeviceWeightedGraph<Types>> levels;
    std::vector<DeviceVector<typename Types::VertexT>> fine_to_coarse;
    std::string stop_reason = "capacity_floor";
    double hierarchy_loop_seconds = 0.0;
    double snapshot_seconds = 0.0;
    std::size_t layer_count() const { return levels.size(); }
};

// int32 file: level count, then per level n, m, offsets[n+1], neighbors[m],
// edge weights[m], vertex weights[n] and, for l > 0, the map of level l-1.
// Written to path.tmp and renamed.
template <typename Types>
void write_hierarchy(const Hierarchy<Types>& hierarchy, const std::string& path);

template <typename Types>
Hierarchy<Types> read_hierarchy(const std::string& path);

}  // namespace gpart
