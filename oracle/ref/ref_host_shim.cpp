// ref_host_shim.cpp -- C entry points over the reference's own host-only headers
// (include/partition_capacity.hpp, include/detail/sclp_record.hpp), compiled
// from the reference tree into oracle/_ref/. TEST INFRASTRUCTURE ONLY: used to
// check the oracle's and the product's restatements against the reference code.
#include "partition_capacity.hpp"
#include "detail/sclp_record.hpp"

#include <cstdint>
#include <stdexcept>

extern "C" {

// 0 ok, 1 invalid_argument, 2 overflow_error
int ref_partition_capacity(std::uint64_t total, int parts, double ratio, std::uint64_t* out) {
    try {
        *out = gpart::partition_capacity(total, parts, ratio);
        return 0;
    } catch (const std::invalid_argument&) {
        return 1;
    } catch (const std::overflow_error&) {
        return 2;
    }
}

void ref_record_layout(std::int64_t vertices, std::uint64_t degree_per_weight,
                       std::uint64_t maximum_vertex_weight, int* vertex_bits,
                       int* weight_bits, std::uint64_t* weight_mask) {
    const auto layout = gpart::detail::SclpRecordLayout::checked(
        vertices, degree_per_weight, maximum_vertex_weight);
    *vertex_bits = layout.vertex_bits;
    *weight_bits = layout.weight_bits;
    *weight_mask = layout.weight_mask;
}

std::uint64_t ref_record_encode(int vertex_bits, int weight_bits, std::uint64_t weight_mask,
                                std::uint32_t source, std::uint32_t target,
                                std::uint64_t weight) {
    gpart::detail::SclpRecordLayout layout{vertex_bits, weight_bits, weight_mask};
    return layout.encode(source, target, weight);
}

}  // extern "C"
