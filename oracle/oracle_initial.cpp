// oracle_initial.cpp -- CPU restatement of the GPU recursive spectral bisection.
//
// TEST INFRASTRUCTURE ONLY (parity checker / CPU baseline).
//
// Follows src/initial_partition_gpu.cu of the reference. FP64 results depend on
// the order of additions, so this file reproduces the reference's fixed order:
//  * per vertex, 32 lanes accumulate entries e = begin+lane, begin+lane+32, ...
//    in increasing e, then a shuffle-down tree (s = 16, 8, 4, 2, 1) combines the
//    lanes (:91-95, :117-125);
//  * per 8-vertex block (256 threads), vertex moments are combined by a 32-lane
//    tree per warp, then lanes 0..7 of warp 0 combine the 8 warp results with the
//    same 32-lane tree (block_sum :53-63);
//  * finish_moments sums the block partials with 256 threads striding by 256,
//    then block_sum again (:65-79).
// Multiply-add pairs are fused exactly where nvcc's default -fmad=true fuses
// them: the SpMV accumulation sum += w * x[u] (:122) and the variance
// squares/mass - mean*mean (:76). Everything else is separately rounded.
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace gpo {
namespace {

struct Moments {
    double mass, sum, squares;
};

// 32-lane __shfl_down_sync tree; returns lane 0. Lanes whose source is out of
// range keep their own value, which never reaches lane 0.
double tree32(double* a) {
    for (int s = 16; s; s /= 2)
        for (int i = 0; i + s < 32; ++i) a[i] = a[i] + a[i + s];
    return a[0];
}

Moments tree32(Moments* m) {
    double a[32], b[32], c[32];
    for (int i = 0; i < 32; ++i) {
        a[i] = m[i].mass;
        b[i] = m[i].sum;
        c[i] = m[i].squares;
    }
    return {tree32(a), tree32(b), tree32(c)};
}

// block_sum (:55-63): 256 thread values -> total.
Moments block_sum(const Moments* threads) {
    Moments warp_total[32];
    for (int w = 0; w < 8; ++w) {
        Moments lanes[32];
        for (int l = 0; l < 32; ++l) lanes[l] = threads[w * 32 + l];
        warp_total[w] = tree32(lanes);
    }
    for (int l = 8; l < 32; ++l) warp_total[l] = {0, 0, 0};
    return tree32(warp_total);
}

struct Normalization {
    double mean, scale;
};

Normalization finish_moments(const std::vector<Moments>& partials, int count) {
    Moments threads[256];
    for (int t = 0; t < 256; ++t) {
        Moments m{0, 0, 0};
        for (int i = t; i < count; i += 256) {
            m.mass += partials[i].mass;
            m.sum += partials[i].sum;
            m.squares += partials[i].squares;
        }
        threads[t] = m;
    }
    const Moments m = block_sum(threads);
    const double mean = m.sum / m.mass;
    const double variance = std::fmax(0.0, std::fma(-mean, mean, m.squares / m.mass));
    return {mean, variance > 1e-30 ? 1.0 / std::sqrt(variance) : 1.0};
}

// Partials of one launch: 8 vertices per block, vertex moment on lane 0 of its warp.
std::vector<Moments> block_partials(const std::vector<Moments>& vertex_moments, int count) {
    const int blocks = (count + 7) / 8;
    std::vector<Moments> partials(blocks);
    for (int b = 0; b < blocks; ++b) {
        Moments threads[256];
        for (int t = 0; t < 256; ++t) threads[t] = {0, 0, 0};
        for (int w = 0; w < 8; ++w) {
            const int i = b * 8 + w;
            if (i < count) threads[w * 32] = vertex_moments[i];
        }
        partials[b] = block_sum(threads);
    }
    return partials;
}

}  // namespace

std::vector<std::int32_t> initial_partition(const Graph& g, int parts, double imbalance,
                                            std::uint32_t seed, std::vector<double>* coords,
                                            int* task_count) {
    // src/initial_partition_gpu.cu:170-276
    const auto n64 = g.n();
    if (n64 <= 0 || n64 > std::numeric_limits<int>::max() - 255 || parts < 2 || parts > n64 ||
        parts > 32 || !std::isfinite(imbalance) || imbalance < 1.0)
        throw std::invalid_argument("invalid GPU initial partition input");
    const int n = static_cast<int>(n64);
    using U64 = unsigned long long;
    auto checked = [](U64 a, U64 b) { return ~U64{0} - a < b ? ~U64{0} : a + b; };
    U64 total = 0, edge_total = 0, max_weight = 0;
    for (auto w : g.vw) {
        total = checked(total, w);
        max_weight = std::max<U64>(max_weight, w);
    }
    for (auto w : g.ew) edge_total = checked(edge_total, w);
    if (!total || total == ~U64{0} || edge_total == ~U64{0})
        throw std::overflow_error("GPU initializer weight accumulation overflow");
    const U64 capacity = partition_capacity(total, parts, imbalance);
    if (max_weight > capacity) throw std::invalid_argument("a vertex exceeds partition capacity");

    std::vector<int> labels(n, 0), order(n);
    std::iota(order.begin(), order.end(), 0);
    std::vector<double> degree(n), x(n), y(n);
    if (coords) coords->assign(n, 0.0);
    struct Task {
        int begin, count, first_part, parts;
        U64 weight;
    };
    std::vector<Task> tasks{{0, n, 0, parts, total}};
    int splits = 0;
    while (!tasks.empty()) {
        Task task = tasks.back();
        tasks.pop_back();
        if (task.parts == 1) {
            if (task.weight > capacity) throw std::runtime_error("GPU initializer infeasible balance");
            continue;
        }
        ++splits;
        const int count = task.count;
        int* ids = order.data() + task.begin;
        const int group = task.first_part;
        const unsigned task_seed = seed ^ static_cast<unsigned>(task.first_part);
        std::vector<Moments> vm(count);
        // spectral_start (:81-107)
#pragma omp parallel for schedule(static)
        for (int i = 0; i < count; ++i) {
            const int v = ids[i];
            double lanes[32];
            for (int l = 0; l < 32; ++l) {
                double d = 0;
                for (std::int64_t e = g.off[v] + l; e < g.off[v + 1]; e += 32)
                    if (labels[g.adj[e]] == group) d += static_cast<double>(g.ew[e]);
                lanes[l] = d;
            }
            const double d = tree32(lanes);
            const double a = static_cast<double>(mix32(static_cast<unsigned>(v) ^ task_seed)) /
                                 4294967296.0 - 0.5;
            degree[v] = d;
            x[v] = a;
            const double mass = d > 0 ? d : 1.0;
            vm[i] = {mass, mass * a, mass * a * a};
        }
        Normalization norm = finish_moments(block_partials(vm, count), (count + 7) / 8);
        for (int iteration = 0; iteration <= 64; ++iteration) {
            // normalize (:138-142)
            for (int i = 0; i < count; ++i) {
                const int v = ids[i];
                x[v] = (x[v] - norm.mean) * norm.scale;
            }
            if (iteration == 64) break;
            // spectral_step (:109-136)
#pragma omp parallel for schedule(static)
            for (int i = 0; i < count; ++i) {
                const int v = ids[i];
                double lanes[32];
                for (int l = 0; l < 32; ++l) {
                    double s = 0;
                    for (std::int64_t e = g.off[v] + l; e < g.off[v + 1]; e += 32) {
                        const int u = g.adj[e];
                        if (labels[u] == group) s = std::fma(static_cast<double>(g.ew[e]), x[u], s);
                    }
                    lanes[l] = s;
                }
                const double sum = tree32(lanes);
                const double d = degree[v];
                const double a = d > 0 ? 0.5 * (x[v] + sum / d) : x[v];
                y[v] = a;
                const double mass = d > 0 ? d : 1.0;
                vm[i] = {mass, mass * a, mass * a * a};
            }
            norm = finish_moments(block_partials(vm, count), (count + 7) / 8);
            std::swap(x, y);
        }
        if (coords)
            for (int i = 0; i < count; ++i) (*coords)[ids[i]] = x[ids[i]];
        // SpectralOrder sort (:143-148, :230)
        std::sort(ids, ids + count, [&](int a, int b) { return x[a] < x[b] || (x[a] == x[b] && a < b); });
        std::vector<U64> prefix(count);
        U64 running = 0;
        for (int i = 0; i < count; ++i) {
            running += g.vw[ids[i]];
            prefix[i] = running;
        }
        // split selection (:233-248)
        const int left_parts = task.parts / 2, right_parts = task.parts - left_parts;
        const U64 target = static_cast<U64>(static_cast<long double>(task.weight) * left_parts / task.parts);
        const U64 left_limit = capacity > task.weight / left_parts ? task.weight : capacity * left_parts;
        const U64 right_limit = capacity > task.weight / right_parts ? task.weight : capacity * right_parts;
        const U64 min_left = task.weight > right_limit ? task.weight - right_limit : 0;
        int boundary = static_cast<int>(std::lower_bound(prefix.begin(), prefix.end(), target) -
                                        prefix.begin()) + 1;
        boundary = std::max(left_parts, std::min(count - right_parts, boundary));
        U64 left_weight = prefix[boundary - 1];
        if (left_weight > left_limit && boundary > left_parts) {
            --boundary;
            left_weight = prefix[boundary - 1];
        }
        if (left_weight < min_left && boundary < count - right_parts) {
            ++boundary;
            left_weight = prefix[boundary - 1];
        }
        if (left_weight > left_limit || left_weight < min_left)
            throw std::runtime_error("GPU spectral split cannot satisfy capacity");
        for (int i = 0; i < count; ++i)
            labels[ids[i]] = i < boundary ? task.first_part : task.first_part + left_parts;
        tasks.push_back({task.begin + boundary, count - boundary, task.first_part + left_parts,
                         right_parts, task.weight - left_weight});
        tasks.push_back({task.begin, boundary, task.first_part, left_parts, left_weight});
    }
    std::vector<U64> loads(parts, 0);
    for (int v = 0; v < n; ++v) loads[labels[v]] += g.vw[v];
    if (*std::max_element(loads.begin(), loads.end()) > capacity)
        throw std::runtime_error("GPU initializer balance validation failed");
    if (task_count) *task_count = splits;
    return std::vector<std::int32_t>(labels.begin(), labels.end());
}

}  // namespace gpo
