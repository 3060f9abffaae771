// graph_types.hpp -- compile-time graph widths (reference: include/graph_types.hpp:7-27).
// Both widths are instantiated in libgpart.so; the CLIs pick one with
// GPART_STANDARD_GRAPH / GPART_BIG_GRAPH like the reference's per-width builds.
#pragma once

#include "gpart_hip.h"

#include <cstdint>

namespace gpart {

struct StandardTypes {
    using VertexT = std::int32_t;
    using OffsetT = std::int32_t;
    using WeightT = std::uint32_t;
    static constexpr int kWidth = GPART_WIDTH_STANDARD;
};

struct BigTypes {
    using VertexT = std::int32_t;
    using OffsetT = std::int64_t;
    using WeightT = std::uint64_t;
    static constexpr int kWidth = GPART_WIDTH_BIG;
};

#if defined(GPART_STANDARD_GRAPH) && defined(GPART_BIG_GRAPH)
#error "Define exactly one of GPART_STANDARD_GRAPH or GPART_BIG_GRAPH"
#elif defined(GPART_STANDARD_GRAPH)
using ActiveTypes = StandardTypes;
inline constexpr const char* kActiveGraphName = "standard";
#elif defined(GPART_BIG_GRAPH)
using ActiveTypes = BigTypes;
inline constexpr const char* kActiveGraphName = "big";
#endif

}  // namespace gpart
