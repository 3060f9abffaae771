// log.hpp -- the report lines of the host library.
//
// The reference's drivers print their progress lines to stdout unconditionally.
// The CLIs here do the same; library callers (bindings, benchmarks) switch them
// off with gpart_set_log(0) so their own stdout stays parseable.
#pragma once

#include <iomanip>
#include <ostream>
#include <sstream>
#include <string>
#include <string_view>

namespace gpart::detail {
// stdout while logging is on, a discarding stream otherwise.
std::ostream& log();

// One report line: an optional leading word, then space-separated key=value
// fields; floating-point values carry 10 significant digits. Scripts find the
// values by key, so every key appears once per line.
class ReportLine {
public:
    explicit ReportLine(std::string_view head = {}) {
        text_ << std::setprecision(10) << head;
        empty_ = head.empty();
    }

    template <typename T>
    ReportLine& field(std::string_view key, const T& value) {
        if (!empty_) text_ << ' ';
        text_ << key << '=' << value;
        empty_ = false;
        return *this;
    }

    std::string str() const { return text_.str(); }

    // Writes the line and its newline in one insertion.
    void write(std::ostream& out) const { out << text_.str() + '\n'; }

private:
    std::ostringstream text_;
    bool empty_ = true;
};
}  // namespace gpart::detail
