// check.hpp -- status-to-exception translation at the C ABI boundary
// (reference: include/check.hpp:8-18, CUDA_CHECK).
#pragma once

#include "gpart_hip.h"

#include <sstream>
#include <stdexcept>
#include <string>

namespace gpart {

[[noreturn]] inline void throw_status(int status, const char* call) {
    const std::string message = gpart_last_error();
    switch (status) {
        case GPART_ERR_INVALID_ARGUMENT: throw std::invalid_argument(message);
        case GPART_ERR_OVERFLOW: throw std::overflow_error(message);
        default: {
            std::ostringstream s;
            s << message << " (in " << call << ")";
            throw std::runtime_error(s.str());
        }
    }
}

}  // namespace gpart

#define GPART_CHECK(call)                                              \
    do {                                                               \
        const int status__ = (call);                                   \
        if (status__ != GPART_OK) ::gpart::throw_status(status__, #call); \
    } while (0)
