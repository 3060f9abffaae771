// cli_support.hpp -- what the two command-line tools share: the argument grammar
// (required positionals, optional positionals with defaults, trailing flags),
// the usage text derived from it, a clock with named marks and the exit-code
// convention (0 success, 1 a failure reported on stderr, 2 usage).
//
// The argument order, defaults, usage text, exit codes and report keys are the
// reference's CLI contract (app/partition_main.cu:20-142, app/coarsen_main.cu:13-79).
#pragma once

#include "../host/log.hpp"

#include <chrono>
#include <cstdint>
#include <exception>
#include <functional>
#include <iostream>
#include <map>
#include <optional>
#include <string>
#include <vector>

namespace gpart::cli {

enum ExitCode : int { kSuccess = 0, kFailure = 1, kUsage = 2 };

// The accepted command line: required positionals, then optional ones (shown as
// written in `optional`), then the flag. `repeated_flag` lets the flag appear any
// number of times at the end (the partition tool); otherwise at most once.
struct Grammar {
    std::vector<std::string> required;
    std::vector<std::string> optional;
    std::string flag;
    bool repeated_flag = false;
};

inline std::string usage_text(const char* program, const Grammar& g) {
    std::string text = std::string("Usage: ") + program;
    for (const auto& r : g.required) text += " <" + r + ">";
    for (const auto& o : g.optional) text += " [" + o + "]";
    text += " [" + g.flag + "]\n";
    return text;
}

class Arguments {
public:
    // nullopt when the count of positionals is outside the grammar.
    static std::optional<Arguments> parse(int argc, char** argv, const Grammar& g) {
        Arguments a;
        int end = argc;
        while (end > 1 && argv[end - 1] == g.flag) {
            a.flag_ = true;
            --end;
            if (!g.repeated_flag) break;
        }
        const std::size_t count = static_cast<std::size_t>(end - 1);
        if (count < g.required.size() || count > g.required.size() + g.optional.size()) return std::nullopt;
        a.values_.assign(argv + 1, argv + end);
        return a;
    }

    const std::string& operator[](std::size_t i) const { return values_.at(i); }
    bool flag() const { return flag_; }

    // Position i parsed by `convert` (std::stoi and kin: a malformed number throws
    // and ends the tool with exit code 1), or `fallback` when it was not given.
    template <typename T, typename Convert>
    T value_or(std::size_t i, T fallback, Convert convert) const {
        return i < values_.size() ? static_cast<T>(convert(values_[i])) : fallback;
    }

private:
    std::vector<std::string> values_;
    bool flag_ = false;
};

inline int to_int(const std::string& s) { return std::stoi(s); }
inline unsigned long to_ulong(const std::string& s) { return std::stoul(s); }
inline double to_double(const std::string& s) { return std::stod(s); }

// Named instants on the steady clock; span(a, b) is the seconds from a to b.
class Marks {
public:
    using clock = std::chrono::steady_clock;
    void mark(const std::string& name) { at_[name] = clock::now(); }
    double span(const std::string& from, const std::string& to) const {
        return std::chrono::duration<double>(at_.at(to) - at_.at(from)).count();
    }

private:
    std::map<std::string, clock::time_point> at_;
};

// Runs a tool body: its return value is the exit code; an exception prints
// "<tool>: <what>" on stderr and gives kFailure.
inline int run_tool(const char* tool, const std::function<int()>& body) {
    try {
        return body();
    } catch (const std::exception& e) {
        std::cerr << tool << ": " << e.what() << '\n';
        return kFailure;
    }
}

}  // namespace gpart::cli
