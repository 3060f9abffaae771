// initial_partition.hpp -- GPU recursive spectral bisection
// (reference: include/initial_partition.hpp:9-12).
#pragma once

#include "graph.hpp"

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

}  // namespace gpart
