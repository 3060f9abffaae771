// hierarchy.cpp -- int32 hierarchy file writer and reader
// (reference: include/hierarchy.hpp:42-126; reader format as parsed by
// tests/test_sclp_small.py:72-87).
#include "hierarchy.hpp"

#include <cstdio>
#include <fstream>
#include <limits>
#include <stdexcept>

namespace gpart {

template <typename Types>
void write_hierarchy(const Hierarchy<Types>& hierarchy, const std::string& path) {
    const auto& levels = hierarchy.levels;
    const auto& maps = hierarchy.fine_to_coarse;
    if (levels.empty() || maps.size() + 1 != levels.size())
        throw std::runtime_error("cannot write an incomplete hierarchy");
    const auto temporary = path + ".tmp";
    try {
        std::ofstream out(temporary, std::ios::binary | std::ios::trunc);
        if (!out) throw std::runtime_error("cannot open hierarchy temporary file");
        std::vector<std::int32_t> buffer;
        auto put = [&](std::int32_t v) { buffer.push_back(v); };
        put(static_cast<std::int32_t>(levels.size()));
        for (std::size_t i = 0; i < levels.size(); ++i) {
            const auto& g = levels[i];
            if (g.vertices() > std::numeric_limits<std::int32_t>::max() ||
                g.edges() > std::numeric_limits<std::int32_t>::max())
                throw std::runtime_error("hierarchy export requires 32-bit dimensions");
            const auto n = static_cast<std::int32_t>(g.vertices());
            put(n);
            put(static_cast<std::int32_t>(g.edges()));
            for (const auto o : g.offsets) {
                const auto wide = static_cast<std::int64_t>(o);
                if (wide < 0 || wide > std::numeric_limits<std::int32_t>::max())
                    throw std::runtime_error("hierarchy row offset exceeds int32");
                put(static_cast<std::int32_t>(wide));
            }
            for (const auto u : g.neighbors) put(u);
            for (const auto w : g.edge_weights) {
                if (static_cast<std::uint64_t>(w) > static_cast<std::uint64_t>(std::numeric_limits<std::int32_t>::max()))
                    throw std::runtime_error("hierarchy edge weight exceeds int32");
                put(static_cast<std::int32_t>(w));
            }
            for (const auto w : g.vertex_weights) {
                if (static_cast<std::uint64_t>(w) > static_cast<std::uint64_t>(std::numeric_limits<std::int32_t>::max()))
                    throw std::runtime_error("hierarchy vertex weight exceeds int32");
                put(static_cast<std::int32_t>(w));
            }
            if (i > 0) {
                const auto& map = maps[i - 1];
                if (map.size() != static_cast<std::size_t>(levels[i - 1].vertices()))
                    throw std::runtime_error("hierarchy map has the wrong fine-level length");
                for (const auto c : map) {
                    if (c < 0 || c >= n) throw std::runtime_error("hierarchy map contains an invalid coarse id");
                    put(c);
                }
            }
            out.write(reinterpret_cast<const char*>(buffer.data()),
                      static_cast<std::streamsize>(buffer.size() * sizeof(std::int32_t)));
            buffer.clear();
        }
        out.flush();
        if (!out) throw std::runtime_error("failed while writing hierarchy");
        out.close();
        if (std::rename(temporary.c_str(), path.c_str()) != 0)
            throw std::runtime_error("cannot commit hierarchy output");
    } catch (...) {
        std::remove(temporary.c_str());
        throw;
    }
}

template <typename Types>
Hierarchy<Types> read_hierarchy(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open hierarchy " + path);
    auto get = [&](std::size_t count) {
        std::vector<std::int32_t> v(count);
        if (count) in.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(count * 4));
        if (!in) throw std::runtime_error("truncated hierarchy");
        return v;
    };
    const auto count = get(1)[0];
    if (count < 1) throw std::runtime_error("hierarchy has no levels");
    Hierarchy<Types> h;
    for (std::int32_t l = 0; l < count; ++l) {
        const auto nm = get(2);
        const auto n = nm[0], m = nm[1];
        if (n < 0 || m < 0) throw std::runtime_error("invalid hierarchy dimensions");
        WeightedGraph<Types> g;
        for (auto o : get(static_cast<std::size_t>(n) + 1)) g.offsets.push_back(static_cast<typename Types::OffsetT>(o));
        g.neighbors = get(static_cast<std::size_t>(m));
        for (auto w : get(static_cast<std::size_t>(m))) g.edge_weights.push_back(static_cast<typename Types::WeightT>(w));
        for (auto w : get(static_cast<std::size_t>(n))) g.vertex_weights.push_back(static_cast<typename Types::WeightT>(w));
        if (l > 0) h.fine_to_coarse.push_back(get(static_cast<std::size_t>(h.levels.back().vertices())));
        h.levels.push_back(std::move(g));
    }
    return h;
}

template void write_hierarchy<StandardTypes>(const Hierarchy<StandardTypes>&, const std::string&);
template void write_hierarchy<BigTypes>(const Hierarchy<BigTypes>&, const std::string&);
template Hierarchy<StandardTypes> read_hierarchy<StandardTypes>(const std::string&);
template Hierarchy<BigTypes> read_hierarchy<BigTypes>(const std::string&);

}  // namespace gpart
