// partition_main.cpp -- gpart_partition_{standard,big}: the drop-in CLI
// (reference: app/partition_main.cu:20-142; same arguments, exit codes, timed
// scope and key log lines).
#include "coarsen.hpp"
#include "graph.hpp"
#include "initial_partition.hpp"
#include "partition_capacity.hpp"
#include "uncoarsen.hpp"

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>

int main(int argc, char** argv) {
    try {
        const auto wall_start = std::chrono::steady_clock::now();
        bool verify = false;
        int positional_argc = argc;
        while (positional_argc > 1) {
            const std::string flag(argv[positional_argc - 1]);
            if (flag == "--verify") verify = true;
            else break;
            --positional_argc;
        }
        if (positional_argc < 5 || positional_argc > 7) {
            std::cerr << "Usage: " << argv[0] << " <indptr.bin> <indices.bin> <parts> <partition.out>"
                      << " [imbalance_ratio=1.03] [seed] [--verify]\n";
            return 2;
        }
        using Types = gpart::ActiveTypes;
        const int parts = std::stoi(argv[3]);
        const double imbalance = positional_argc >= 6 ? std::stod(argv[5]) : 1.03;
        const auto seed = positional_argc >= 7 ? static_cast<std::uint32_t>(std::stoul(argv[6])) : 0U;
        // GPART_SPECTRAL_MODE selects the bisection form: 0 plain FP64 (the
        // reference's launch sequence), 1 batched FP64 (default), 2 batched FP32.
        const char* mode_env = std::getenv("GPART_SPECTRAL_MODE");
        const int mode = mode_env ? std::atoi(mode_env) : 1;
        if (mode < 0 || mode > 2) throw std::invalid_argument("GPART_SPECTRAL_MODE must be 0, 1 or 2");
        const auto load_start = std::chrono::steady_clock::now();
        // Without --verify the input is narrowed and shape-checked on the device (only
        // the int64 files cross PCIe); --verify keeps the host copy for the symmetry check.
        gpart::WeightedGraph<Types> graph;
        gpart::DeviceWeightedGraph<Types> device_graph;
        if (verify) graph = gpart::load_weighted_graph<Types>(argv[1], argv[2]);
        else device_graph = gpart::load_device_graph<Types>(argv[1], argv[2]);
        const std::int64_t vertices = verify ? graph.vertices() : device_graph.vertices();
        gpart::partition_capacity(static_cast<std::uint64_t>(vertices), parts, imbalance);
        const auto load_end = std::chrono::steady_clock::now();
        std::cout << "verify=" << (verify ? 1 : 0) << '\n';
        gpart::CoarsenOptions coarsen_options;
        coarsen_options.parts = parts;
        coarsen_options.seed = seed;
        coarsen_options.strict_verify = verify;
        // at least 64 coarse clusters per part (app/partition_main.cu:50-56)
        coarsen_options.cluster_size_factor = std::max(64, (512 + parts - 1) / parts);
        coarsen_options.pack_isolated = true;
        std::cout << "pipeline_mode=incremental_bucket cluster_size_factor=" << coarsen_options.cluster_size_factor
                  << " pack_isolated=" << coarsen_options.pack_isolated << " spectral_mode=" << mode << '\n';
        const auto prepare_start = std::chrono::steady_clock::now();
        if (verify) {
            gpart::validate_weighted_csr(graph, true);
            device_graph = gpart::make_device_weighted(graph);
        }
        GPART_CHECK(gpart_ctx_synchronize(gpart::detail::context()));
        const auto coarsen_start = std::chrono::steady_clock::now();
        auto hierarchy = gpart::coarsen<Types>(std::move(device_graph), coarsen_options);
        GPART_CHECK(gpart_ctx_synchronize(gpart::detail::context()));
        const auto coarsen_end = std::chrono::steady_clock::now();
        auto coarse_partition = gpart::initial_partition_gpu<Types>(hierarchy.levels.back(), parts, imbalance, seed,
                                                                    static_cast<gpart::SpectralMode>(mode));
        GPART_CHECK(gpart_ctx_synchronize(gpart::detail::context()));
        const auto initial_end = std::chrono::steady_clock::now();
        gpart::RefineOptions refine_options;
        refine_options.parts = parts;
        refine_options.imbalance_ratio = imbalance;
        refine_options.max_rounds = 32;
        refine_options.strict_verify = verify;
        refine_options.refinement_growth_ratio = 2.0;
        auto result = gpart::uncoarsen<Types>(hierarchy, std::move(coarse_partition), refine_options);
        GPART_CHECK(gpart_ctx_synchronize(gpart::detail::context()));
        const auto uncoarsen_end = std::chrono::steady_clock::now();
        hierarchy.levels.clear();
        hierarchy.fine_to_coarse.clear();
        auto device_partition = std::move(result.device_partition);
        gpart::detail::release_context_cache();
        const auto free_end = std::chrono::steady_clock::now();
        const auto partition = device_partition.to_host();
        const auto final_d2h_end = std::chrono::steady_clock::now();
        const auto output_start = std::chrono::steady_clock::now();
        std::ofstream output(argv[4], std::ios::binary | std::ios::trunc);
        if (!output) throw std::runtime_error("cannot open partition output");
        output.write(reinterpret_cast<const char*>(partition.data()),
                     static_cast<std::streamsize>(partition.size() * sizeof(typename Types::VertexT)));
        output.close();
        if (!output) throw std::runtime_error("cannot write partition output");
        const auto output_end = std::chrono::steady_clock::now();
        auto secs = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
        const double algorithm_seconds = secs(coarsen_start, free_end);
        std::cout << std::setprecision(10) << "partition_timing input_load_seconds=" << secs(load_start, load_end)
                  << " coarsen_seconds=" << secs(coarsen_start, coarsen_end)
                  << " initial_partition_seconds=" << secs(coarsen_end, initial_end)
                  << " uncoarsen_seconds=" << secs(initial_end, uncoarsen_end)
                  << " input_prepare_seconds=" << secs(prepare_start, coarsen_start)
                  << " hierarchy_free_seconds=" << secs(uncoarsen_end, free_end)
                  << " final_partition_d2h_seconds=" << secs(free_end, final_d2h_end)
                  << " timing_scope=jet_device_input_to_device_partition"
                  << " partition_output_seconds=" << secs(output_start, output_end)
                  << " partition_total_seconds=" << algorithm_seconds
                  << " algorithm_total_seconds=" << algorithm_seconds
                  << " wall_total_seconds=" << secs(wall_start, output_end) << '\n';
        return 0;
    } catch (const std::exception& error) {
        std::cerr << "gpart_partition: " << error.what() << '\n';
        return 1;
    }
}
