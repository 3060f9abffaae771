// uncoarsen.hpp -- projection + refinement over a hierarchy
// (reference: include/uncoarsen.hpp:10-22).
#pragma once

#include "hierarchy.hpp"
#include "refine.hpp"

#include <vector>

namespace gpart {

template <typename Types>
std::vector<typename Types::VertexT> uncoarsen(const Hierarchy<Types>& hierarchy,
                                               const std::vector<typename Types::VertexT>& coarsest_partition,
                                               const RefineOptions& options);

// Consumes device levels/maps and labels; the final partition stays on device.
template <typename Types>
DeviceUncoarsenResult<Types> uncoarsen(DeviceHierarchy<Types>& hierarchy,
                                       DeviceVector<typename Types::VertexT>&& coarsest_partition,
                                       const RefineOptions& options);

}  // namespace gpart
