// coarsen_main.cpp -- gpart_coarsen_{standard,big}: coarsen an int64 CSR graph with
// the library's default options and export the int32 hierarchy file.
//
//   gpart_coarsen_<width> <indptr.bin> <indices.bin> <parts> <hierarchy.out>
//                         [seed=0] [stop_contraction_ratio=0.90] [max_levels=24] [--verify]
//
// Drop-in for app/coarsen_main.cu:13-79: same arguments, defaults, checks, exit
// codes and ml_* report keys; ML_SKIP_HIERARCHY_EXPORT skips the file write.
#include "cli_support.hpp"
#include "coarsen.hpp"
#include "graph.hpp"
#include "hierarchy.hpp"

#include <cstdlib>
#include <stdexcept>

namespace {

using gpart::cli::Arguments;
using gpart::detail::ReportLine;

const gpart::cli::Grammar kGrammar{
    {"indptr.bin", "indices.bin", "parts", "hierarchy.out"},
    {"seed=0", "stop_contraction_ratio=0.90", "max_levels=24"},
    "--verify",
    false,
};

// Positions of the arguments after the program name.
enum Position : std::size_t { kIndptr, kIndices, kParts, kOutput, kSeed, kStopRatio, kMaxLevels };

gpart::CoarsenOptions options_from(const Arguments& args) {
    gpart::CoarsenOptions o;
    o.parts = args.value_or(kParts, 0, gpart::cli::to_int);
    o.seed = args.value_or<std::uint32_t>(kSeed, 0u, gpart::cli::to_ulong);
    o.stop_contraction_ratio = args.value_or(kStopRatio, 0.90, gpart::cli::to_double);
    o.max_levels = args.value_or(kMaxLevels, 24, gpart::cli::to_int);
    o.strict_verify = args.flag();
    // the tool's own limits, checked before the input is read
    if (o.parts < 2 || o.parts > 32) throw std::runtime_error("parts must be between 2 and 32");
    if (!(o.stop_contraction_ratio > 0.0 && o.stop_contraction_ratio <= 1.0))
        throw std::runtime_error("invalid coarsening stop ratio");
    if (o.max_levels < 1 || o.max_levels > 200) throw std::runtime_error("invalid coarsening max levels");
    return o;
}

int coarsen_tool(int argc, char** argv) {
    const auto args = Arguments::parse(argc, argv, kGrammar);
    if (!args) {
        std::cerr << gpart::cli::usage_text(argv[0], kGrammar);
        return gpart::cli::kUsage;
    }
    const gpart::CoarsenOptions options = options_from(*args);
    using Types = gpart::ActiveTypes;
    const auto graph = gpart::load_weighted_graph<Types>((*args)[kIndptr], (*args)[kIndices]);

    gpart::cli::Marks t;
    t.mark("start");
    const auto hierarchy = gpart::coarsen<Types>(graph, options);
    const bool export_file = std::getenv("ML_SKIP_HIERARCHY_EXPORT") == nullptr;
    const std::string& path = (*args)[kOutput];
    t.mark("export");
    if (export_file) gpart::write_hierarchy(hierarchy, path);
    t.mark("exported");

    auto& out = std::cout;
    ReportLine()
        .field("ml_hierarchy_export_seconds", t.span("export", "exported"))
        .field("skipped", export_file ? 0 : 1)
        .write(out);
    ReportLine()
        .field("ml_hierarchy_levels", hierarchy.layer_count())
        .field("coarsest_vertices", hierarchy.levels.back().vertices())
        .field("stop_reason", hierarchy.stop_reason)
        .field("stop_contraction_ratio", options.stop_contraction_ratio)
        .field("method", "sclp")
        .write(out);
    ReportLine()
        .field("ml_gpu_device_snapshot_total_seconds", hierarchy.snapshot_seconds)
        .field("method", "sclp")
        .write(out);
    t.mark("end");
    ReportLine()
        .field("ml_total_seconds", t.span("start", "end"))
        .field("hierarchy", export_file ? path : std::string("skipped"))
        .field("coarsen_only", 1)
        .field("method", "sclp")
        .field("graph_type", gpart::kActiveGraphName)
        .write(out);
    return gpart::cli::kSuccess;
}

}  // namespace

int main(int argc, char** argv) {
    return gpart::cli::run_tool("gpart_coarsen", [&] { return coarsen_tool(argc, argv); });
}
