// initial_partition.hpp -- GPU recursive spectral bisection
// (reference: include/initial_partition.hpp:9-12).
#pragma once

#include "graph.hpp"

#include <cstdint>
#include <vector>

namespace gpart {

// How the bisection tasks run: the reference's per-task launch sequence in FP64,
// every task of one recursion depth in one launch (FP64, identical results), or
// the batched form with FP32 diffusion vectors and FP64 moments / prefix.
enum class SpectralMode { PlainFp64 = 0, BatchedFp64 = 1, BatchedFp32 = 2 };

// coords, when given, receives each vertex's final diffusion coordinate (the
// normalised vector of the deepest task that split it).
template <typename Types>
DeviceVector<typename Types::VertexT> initial_partition_gpu(const DeviceWeightedGraph<Types>& graph, int parts,
                                                            double imbalance_ratio, std::uint32_t seed = 0,
                                                            SpectralMode mode = SpectralMode::BatchedFp64,
                                                            DeviceVector<double>* coords = nullptr);

// The batched bisection of several graphs (one width, one mode): every task of
// at most gpart_spectral_depth_max_task() vertices of every graph runs in one
// persistent launch (gpart_spectral_queue); each graph's labels are those
// initial_partition_gpu gives it alone. PlainFp64 runs the graphs one by one.
template <typename Types>
std::vector<DeviceVector<typename Types::VertexT>> initial_partition_group(
    const std::vector<const DeviceWeightedGraph<Types>*>& graphs, const std::vector<int>& parts,
    const std::vector<double>& imbalance_ratios, const std::vector<std::uint32_t>& seeds,
    SpectralMode mode = SpectralMode::BatchedFp64);

// FP32 against FP64 diffusion on one graph (no reference counterpart: the
// reference has only FP64). Each recursion depth runs both precisions on the same
// tasks, from the FP64 state of the depth before; the FP64 result carries on.
struct SpectralDepthError {
    int depth = 0;
    int tasks = 0;                   // batched tasks compared at this depth
    std::int64_t vertices = 0;       // their vertices
    double max_abs_error = 0.0;      // max |x32 - x64| of the normalised vectors
    double rms_error = 0.0;
    int splits_differ = 0;           // tasks whose split boundary differs
    std::int64_t labels_differ = 0;  // vertices on a different side after the depth
};
struct SpectralComparison {
    std::vector<SpectralDepthError> depths;
    std::uint64_t cut_fp64 = 0, cut_fp32 = 0;  // full batched recursions in each precision
    std::uint64_t max_load_fp64 = 0, max_load_fp32 = 0;
    std::int64_t labels_differ = 0;
};

template <typename Types>
SpectralComparison compare_spectral_precision(const DeviceWeightedGraph<Types>& graph, int parts,
                                              double imbalance_ratio, std::uint32_t seed = 0);

}  // namespace gpart
