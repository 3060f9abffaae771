// partition_main.cpp -- gpart_partition_{standard,big}: the drop-in partitioning tool.
//
//   gpart_partition_<width> <indptr.bin> <indices.bin> <parts> <partition.out>
//                           [imbalance_ratio=1.03] [seed] [--verify]
//
// Same arguments, defaults, exit codes, timed scope and report keys as
// app/partition_main.cu:20-142. algorithm_total_seconds runs from the start of
// coarsening (input resident on the device, device idle) to the hierarchy's
// release; it leaves out loading, validation, the upload, the final label
// download and the file write.
//
// GPART_SPECTRAL_MODE selects the bisection form: 0 plain FP64 (one task at a
// time), 1 batched FP64 (default, bit-identical labels), 2 batched FP32.
#include "cli_support.hpp"
#include "coarsen.hpp"
#include "graph.hpp"
#include "initial_partition.hpp"
#include "partition_capacity.hpp"
#include "uncoarsen.hpp"

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <stdexcept>

namespace {

using gpart::cli::Arguments;
using gpart::detail::ReportLine;
using Types = gpart::ActiveTypes;
using Vertex = Types::VertexT;

const gpart::cli::Grammar kGrammar{
    {"indptr.bin", "indices.bin", "parts", "partition.out"},
    {"imbalance_ratio=1.03", "seed"},
    "--verify",
    true,
};

enum Position : std::size_t { kIndptr, kIndices, kParts, kOutput, kRatio, kSeed };

struct Settings {
    int parts = 0;
    double ratio = 1.03;
    std::uint32_t seed = 0;
    bool verify = false;
    int spectral_mode = 1;
};

Settings settings_from(const Arguments& args) {
    Settings s;
    s.parts = args.value_or(kParts, 0, gpart::cli::to_int);
    s.ratio = args.value_or(kRatio, 1.03, gpart::cli::to_double);
    s.seed = args.value_or<std::uint32_t>(kSeed, 0u, gpart::cli::to_ulong);
    s.verify = args.flag();
    if (const char* mode = std::getenv("GPART_SPECTRAL_MODE")) s.spectral_mode = std::atoi(mode);
    if (s.spectral_mode < 0 || s.spectral_mode > 2)
        throw std::invalid_argument("GPART_SPECTRAL_MODE must be 0, 1 or 2");
    return s;
}

// The production coarsening options: at least 64 coarse clusters per part and
// packed isolated vertices (app/partition_main.cu:47-56).
gpart::CoarsenOptions coarsen_options(const Settings& s) {
    gpart::CoarsenOptions o;
    o.parts = s.parts;
    o.seed = s.seed;
    o.strict_verify = s.verify;
    o.cluster_size_factor = std::max(64, (512 + s.parts - 1) / s.parts);
    o.pack_isolated = true;
    return o;
}

gpart::RefineOptions refine_options(const Settings& s) {
    gpart::RefineOptions o;
    o.parts = s.parts;
    o.imbalance_ratio = s.ratio;
    o.max_rounds = 32;
    o.strict_verify = s.verify;
    o.refinement_growth_ratio = 2.0;
    return o;
}

void device_idle() { GPART_CHECK(gpart_ctx_synchronize(gpart::detail::context())); }

void write_labels(const std::string& path, const std::vector<Vertex>& labels) {
    std::ofstream file(path, std::ios::binary | std::ios::trunc);
    if (!file) throw std::runtime_error("cannot open partition output");
    file.write(reinterpret_cast<const char*>(labels.data()),
               static_cast<std::streamsize>(labels.size() * sizeof(Vertex)));
    file.close();
    if (!file) throw std::runtime_error("cannot write partition output");
}

int partition_tool(int argc, char** argv) {
    gpart::cli::Marks t;
    t.mark("wall");
    const auto args = Arguments::parse(argc, argv, kGrammar);
    if (!args) {
        std::cerr << gpart::cli::usage_text(argv[0], kGrammar);
        return gpart::cli::kUsage;
    }
    const Settings s = settings_from(*args);

    // Input. Without --verify the int64 files stream to the device, where they are
    // narrowed and shape-checked (only the files cross PCIe). --verify keeps a host
    // copy for the full CSR validation (sorted rows, symmetry).
    t.mark("load");
    gpart::WeightedGraph<Types> host_graph;
    gpart::DeviceWeightedGraph<Types> graph;
    if (s.verify) {
        host_graph = gpart::load_weighted_graph<Types>((*args)[kIndptr], (*args)[kIndices]);
    } else {
        graph = gpart::load_device_graph<Types>((*args)[kIndptr], (*args)[kIndices]);
    }
    const std::int64_t n = s.verify ? host_graph.vertices() : graph.vertices();
    gpart::partition_capacity(static_cast<std::uint64_t>(n), s.parts, s.ratio);
    t.mark("loaded");
    ReportLine().field("verify", s.verify ? 1 : 0).write(std::cout);

    const gpart::CoarsenOptions co = coarsen_options(s);
    ReportLine()
        .field("pipeline_mode", "incremental_bucket")
        .field("cluster_size_factor", co.cluster_size_factor)
        .field("pack_isolated", co.pack_isolated ? 1 : 0)
        .field("spectral_mode", s.spectral_mode)
        .write(std::cout);

    t.mark("prepare");
    if (s.verify) {
        gpart::validate_weighted_csr(host_graph, true);
        graph = gpart::make_device_weighted(host_graph);
    }
    device_idle();

    // ---- timed scope ----
    t.mark("coarsen");
    auto hierarchy = gpart::coarsen<Types>(std::move(graph), co);
    device_idle();
    t.mark("initial");
    auto coarsest = gpart::initial_partition_gpu<Types>(hierarchy.levels.back(), s.parts, s.ratio, s.seed,
                                                        static_cast<gpart::SpectralMode>(s.spectral_mode));
    device_idle();
    t.mark("uncoarsen");
    auto result = gpart::uncoarsen<Types>(hierarchy, std::move(coarsest), refine_options(s));
    device_idle();
    t.mark("free");
    hierarchy.levels.clear();
    hierarchy.fine_to_coarse.clear();
    auto labels_on_device = std::move(result.device_partition);
    gpart::detail::release_context_cache();
    t.mark("freed");
    // ---- end of the timed scope ----

    const std::vector<Vertex> labels = labels_on_device.to_host();
    t.mark("downloaded");
    t.mark("output");
    write_labels((*args)[kOutput], labels);
    t.mark("written");

    const double algorithm = t.span("coarsen", "freed");
    ReportLine("partition_timing")
        .field("input_load_seconds", t.span("load", "loaded"))
        .field("coarsen_seconds", t.span("coarsen", "initial"))
        .field("initial_partition_seconds", t.span("initial", "uncoarsen"))
        .field("uncoarsen_seconds", t.span("uncoarsen", "free"))
        .field("input_prepare_seconds", t.span("prepare", "coarsen"))
        .field("hierarchy_free_seconds", t.span("free", "freed"))
        .field("final_partition_d2h_seconds", t.span("freed", "downloaded"))
        .field("timing_scope", "jet_device_input_to_device_partition")
        .field("partition_output_seconds", t.span("output", "written"))
        .field("partition_total_seconds", algorithm)
        .field("algorithm_total_seconds", algorithm)
        .field("wall_total_seconds", t.span("wall", "written"))
        .write(std::cout);
    return gpart::cli::kSuccess;
}

}  // namespace

int main(int argc, char** argv) {
    return gpart::cli::run_tool("gpart_partition", [&] { return partition_tool(argc, argv); });
}
