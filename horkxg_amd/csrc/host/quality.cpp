// quality.cpp -- the verifier (tools/verify_partition.py:22-65) and the quality
// metrics (graph_untils/analyze_partition.cpp:176-294) over the device counters.
#include "quality.hpp"

#include "check.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace gpart {

namespace {
using u128 = unsigned __int128;

// A decimal literal as Python's Decimal accepts it in plain form:
// [ws][+|-]digits[.digits][(e|E)[+|-]digits][ws], at least one digit in the
// mantissa. Value = sign * mantissa * 10^exponent.
bool parse_decimal(const std::string& s, bool* negative, u128* mantissa, long long* exponent) {
    std::size_t i = 0, end = s.size();
    while (i < end && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
    while (end > i && std::isspace(static_cast<unsigned char>(s[end - 1]))) --end;
    *negative = false;
    if (i < end && (s[i] == '+' || s[i] == '-')) *negative = s[i++] == '-';
    u128 m = 0;
    long long e = 0;
    int digits = 0, significant = 0;
    bool point = false;
    for (; i < end; ++i) {
        const char c = s[i];
        if (c == '.' && !point) {
            point = true;
            continue;
        }
        if (!std::isdigit(static_cast<unsigned char>(c))) break;
        ++digits;
        if (m != 0 || c != '0') ++significant;
        if (significant > 36) return false;  // beyond what the exact arithmetic below holds
        m = m * 10 + static_cast<unsigned>(c - '0');
        if (point) --e;
    }
    if (digits == 0) return false;
    if (i < end && (s[i] == 'e' || s[i] == 'E')) {
        ++i;
        bool neg = false;
        if (i < end && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
        long long x = 0;
        int xd = 0;
        for (; i < end && std::isdigit(static_cast<unsigned char>(s[i])); ++i) {
            if (++xd > 6) return false;
            x = x * 10 + (s[i] - '0');
        }
        if (xd == 0) return false;
        e += neg ? -x : x;
    }
    if (i != end) return false;
    *mantissa = m;
    *exponent = e;
    return true;
}

bool mul_ok(u128 a, u128 b, u128* out) {
    if (a != 0 && b > (~u128{0}) / a) return false;
    *out = a * b;
    return true;
}

double deviation_ratio(double value, double average) {
    if (average <= 0.0) return value == 0.0 ? 0.0 : std::numeric_limits<double>::infinity();
    return std::abs(value - average) / average;
}

double imbalance_ratio(double max_value, double average) { return average > 0.0 ? max_value / average : 0.0; }

}  // namespace

bool decimal_capacity(const std::string& ratio, std::int64_t n, int parts, std::int64_t* capacity) {
    bool negative = false;
    u128 m = 0;
    long long e = 0;
    if (!parse_decimal(ratio, &negative, &m, &e) || negative || parts <= 0 || n < 0) return false;
    // ratio >= 1 and floor(m * 10^e * n / parts), exactly
    u128 num = m, den = static_cast<u128>(parts);
    u128 one = 1;
    if (e >= 0) {
        for (long long i = 0; i < e; ++i)
            if (!mul_ok(num, 10, &num)) return false;
    } else {
        for (long long i = 0; i < -e; ++i)
            if (!mul_ok(one, 10, &one)) return false;
        if (!mul_ok(den, one, &den)) return false;
    }
    // num / one is the ratio (one = 10^-e when e < 0, else 1)
    if (num < one) return false;
    if (!mul_ok(num, static_cast<u128>(n), &num)) return false;
    const u128 q = num / den;
    if (q > static_cast<u128>(std::numeric_limits<std::int64_t>::max())) return false;
    *capacity = static_cast<std::int64_t>(q);
    return true;
}

VerifyResult verify_partition(gpart_ctx* ctx, const gpart_csr_view& g, const std::int32_t* labels,
                              const std::int32_t* host_labels, int parts, const std::string& ratio) {
    std::int64_t capacity = 0;
    if (parts < 2 || parts > 32 || !decimal_capacity(ratio, g.n, parts, &capacity))
        throw std::invalid_argument("invalid parts or imbalance ratio");
    gpart_part_counters_t c{};
    std::vector<std::uint64_t> per_part(static_cast<std::size_t>(7) * parts);
    GPART_CHECK(gpart_part_counters(ctx, &g, labels, parts, &c, per_part.data()));
    if (c.first_invalid_label >= 0) {
        std::int32_t value = 0;
        if (host_labels) value = host_labels[c.first_invalid_label];
        else GPART_CHECK(gpart_memcpy_dtoh(ctx, &value, labels + c.first_invalid_label, sizeof value));
        throw std::invalid_argument("invalid label: " + std::to_string(value));
    }
    if (c.first_invalid_neighbor >= 0) throw std::invalid_argument("invalid neighbor id");
    if (c.directed_cut % 2 || c.nonloop_entries % 2)
        throw std::invalid_argument("odd undirected edge/cut count; input must be symmetric");
    VerifyResult r;
    r.vertices = g.n;
    r.cut = static_cast<std::int64_t>(c.directed_cut / 2);
    r.undirected_edges = static_cast<std::int64_t>(c.nonloop_entries / 2);
    r.cut_ratio = r.undirected_edges ? static_cast<double>(r.cut) / static_cast<double>(r.undirected_edges) : 0.0;
    r.loads.assign(per_part.begin(), per_part.begin() + parts);
    r.max_load = *std::max_element(r.loads.begin(), r.loads.end());
    r.max_load_ratio = g.n ? static_cast<double>(r.max_load) * parts / static_cast<double>(g.n) : 0.0;
    r.capacity = capacity;
    r.capacity_ok = r.max_load <= capacity;
    return r;
}

QualityResult analyze_partition(gpart_ctx* ctx, const gpart_csr_view& g, const std::int32_t* labels, int parts) {
    if (parts < 0) throw std::invalid_argument("--parts must be non-negative");
    gpart_part_counters_t range{};
    GPART_CHECK(gpart_part_counters(ctx, &g, labels, 0, &range, nullptr));
    if (g.n > 0 && range.min_label < 0) throw std::invalid_argument("Partition id must be non-negative");
    const int inferred = g.n > 0 ? range.max_label + 1 : 0;
    int k = parts;
    if (parts == 0) k = inferred;
    else if (inferred > parts) throw std::invalid_argument("Partition file contains part id >= --parts");
    if (k <= 0) throw std::invalid_argument("Partition file is empty or has no parts");
    gpart_part_counters_t c{};
    QualityResult r;
    r.per_part.assign(static_cast<std::size_t>(7) * k, 0);
    GPART_CHECK(gpart_part_counters(ctx, &g, labels, k, &c, r.per_part.data()));
    if (c.first_invalid_neighbor >= 0) throw std::invalid_argument("CSR index out of vertex range");
    r.parts = k;
    r.vertices = g.n;
    r.directed_edges = g.m;
    r.edge_cut = c.directed_cut;
    r.weighted_edge_cut = c.weighted_directed_cut;
    r.edge_cut_ratio = g.m > 0 ? static_cast<double>(c.directed_cut) / static_cast<double>(g.m) : 0.0;
    const auto at = [&](int counter, int p) { return r.per_part[static_cast<std::size_t>(counter) * k + p]; };
    r.vertex_avg = static_cast<double>(g.n) / static_cast<double>(k);
    r.edge_set_avg = static_cast<double>(static_cast<std::uint64_t>(g.m) + c.directed_cut) / static_cast<double>(k);
    r.min_vertices = std::numeric_limits<std::int64_t>::max();
    r.min_edge_set = std::numeric_limits<std::int64_t>::max();
    double local_sum = 0.0, weighted_local_sum = 0.0;
    for (int p = 0; p < k; ++p) {
        const auto vertices = static_cast<std::int64_t>(at(0, p));
        const auto edge_set = static_cast<std::int64_t>(at(1, p) + at(2, p) + at(3, p));
        const std::uint64_t local = at(2, p) + at(3, p);
        const std::uint64_t weighted_local = at(5, p) + at(6, p);
        local_sum += static_cast<double>(local);
        weighted_local_sum += static_cast<double>(weighted_local);
        r.min_vertices = std::min(r.min_vertices, vertices);
        r.max_vertices = std::max(r.max_vertices, vertices);
        r.min_edge_set = std::min(r.min_edge_set, edge_set);
        r.max_edge_set = std::max(r.max_edge_set, edge_set);
        r.max_local_cut = std::max(r.max_local_cut, local);
        r.weighted_max_local_cut = std::max(r.weighted_max_local_cut, weighted_local);
    }
    r.avg_local_cut = local_sum / static_cast<double>(k);
    r.weighted_avg_local_cut = weighted_local_sum / static_cast<double>(k);
    r.vertex_imbalance = imbalance_ratio(static_cast<double>(r.max_vertices), r.vertex_avg);
    r.vertex_max_deviation = std::max(deviation_ratio(static_cast<double>(r.max_vertices), r.vertex_avg),
                                      deviation_ratio(static_cast<double>(r.min_vertices), r.vertex_avg));
    r.edge_set_imbalance = imbalance_ratio(static_cast<double>(r.max_edge_set), r.edge_set_avg);
    r.edge_set_max_deviation = std::max(deviation_ratio(static_cast<double>(r.max_edge_set), r.edge_set_avg),
                                        deviation_ratio(static_cast<double>(r.min_edge_set), r.edge_set_avg));
    return r;
}

}  // namespace gpart
