// log.hpp -- the report lines of the host library.
//
// The reference's drivers print their progress lines to stdout unconditionally.
// The CLIs here do the same; library callers (bindings, benchmarks) switch them
// off with gpart_set_log(0) so their own stdout stays parseable.
#pragma once

#include <ostream>

namespace gpart::detail {
// stdout while logging is on, a discarding stream otherwise.
std::ostream& log();
}  // namespace gpart::detail
