// io::readMatrixMarket of the C++20 drop-in API (include/IO/MatrixMarket.cuh,
// src/IO/MatrixMarket.cu:35-122).  Accepts "coordinate" matrices with "real"
// or "integer" fields and "general" or "symmetric" storage; blank lines and
// '%' comment lines are skipped anywhere after the banner.  Rows keep the
// file's entry order (duplicates and unsorted columns included); symmetric
// off-diagonal entries are mirrored immediately after the stored one and must
// all lie in one triangle.  Every malformed input throws std::runtime_error.
#pragma once

#include <charconv>
#include <cmath>
#include <cstdint>
#include <istream>
#include <limits>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace thermofem::io {

struct HostCsr {
    int rows_{};
    int columns_{};
    std::vector<int> offsets_;
    std::vector<int> indices_;
    std::vector<double> values_;
};

namespace detail {

// Whitespace-separated fields of one line.
inline std::vector<std::string_view> fields(std::string_view line) {
    std::vector<std::string_view> out;
    std::size_t i = 0;
    while (true) {
        while (i < line.size() && (line[i] == ' ' || line[i] == '\t' || line[i] == '\r' || line[i] == '\v' ||
                                   line[i] == '\f'))
            ++i;
        if (i == line.size()) break;
        const std::size_t start = i;
        while (i < line.size() && !(line[i] == ' ' || line[i] == '\t' || line[i] == '\r' || line[i] == '\v' ||
                                    line[i] == '\f'))
            ++i;
        out.push_back(line.substr(start, i - start));
    }
    return out;
}

// Next line that holds data: its first character other than space, tab or
// carriage return exists and is not '%'.
inline bool dataLine(std::istream& in, std::string& line) {
    while (std::getline(in, line)) {
        const auto first = line.find_first_not_of(" \t\r");
        if (first != std::string::npos && line[first] != '%') return true;
    }
    return false;
}

template <class V>
bool parse(std::string_view token, V& value) {
    if (token.size() > 1 && token.front() == '+' && token[1] != '-') token.remove_prefix(1);
    const char* end = token.data() + token.size();
    std::from_chars_result r;
    if constexpr (std::is_floating_point_v<V>)
        r = std::from_chars(token.data(), end, value, std::chars_format::general);
    else
        r = std::from_chars(token.data(), end, value);
    return r.ec == std::errc() && r.ptr == end;
}

}  // namespace detail

inline HostCsr readMatrixMarket(std::istream& input) {
    using detail::fields;
    std::string line;
    if (!std::getline(input, line)) throw std::runtime_error("Missing Matrix Market header");
    const auto banner = fields(line);
    if (banner.size() != 5 || banner[0] != "%%MatrixMarket" || banner[1] != "matrix" || banner[2] != "coordinate" ||
        (banner[3] != "real" && banner[3] != "integer") || (banner[4] != "general" && banner[4] != "symmetric"))
        throw std::runtime_error("Unsupported Matrix Market header");
    const bool integerField = banner[3] == "integer";
    const bool symmetric = banner[4] == "symmetric";

    constexpr long long limit = std::numeric_limits<int>::max();
    if (!detail::dataLine(input, line)) throw std::runtime_error("Missing Matrix Market dimensions");
    const auto size = fields(line);
    long long rows = -1, columns = -1, count = -1;
    if (size.size() != 3 || !detail::parse(size[0], rows) || !detail::parse(size[1], columns) ||
        !detail::parse(size[2], count) || rows < 0 || columns < 0 || count < 0 || rows > limit || columns > limit ||
        count > limit || (symmetric && rows != columns))
        throw std::runtime_error("Invalid Matrix Market dimensions");

    // Entries in file order, mirrors right after their source entry.
    std::vector<int> entryRow, entryColumn;
    std::vector<double> entryValue;
    entryRow.reserve(std::size_t(count));
    entryColumn.reserve(std::size_t(count));
    entryValue.reserve(std::size_t(count));
    int triangle = 0;  // +1 lower, -1 upper, 0 not yet seen
    for (long long k = 0; k < count; ++k) {
        if (!detail::dataLine(input, line)) throw std::runtime_error("Truncated Matrix Market entries");
        const auto f = fields(line);
        long long r = 0, c = 0;
        double v = 0;
        if (f.size() < 3 || !detail::parse(f[0], r) || !detail::parse(f[1], c) || !detail::parse(f[2], v) || r < 1 ||
            r > rows || c < 1 || c > columns || !std::isfinite(v) || (integerField && std::trunc(v) != v))
            throw std::runtime_error("Invalid Matrix Market entry");
        if (f.size() != 3) throw std::runtime_error("Unexpected Matrix Market field");
        const bool mirror = symmetric && r != c;
        if (mirror) {
            const int side = r > c ? 1 : -1;
            if (triangle != 0 && triangle != side) throw std::runtime_error("Mixed triangles in symmetric storage");
            triangle = side;
        }
        if (entryRow.size() + (mirror ? 2u : 1u) > std::size_t(limit))
            throw std::runtime_error("Expanded CSR exceeds index range");
        entryRow.push_back(int(r - 1));
        entryColumn.push_back(int(c - 1));
        entryValue.push_back(v);
        if (mirror) {
            entryRow.push_back(int(c - 1));
            entryColumn.push_back(int(r - 1));
            entryValue.push_back(v);
        }
    }
    if (detail::dataLine(input, line)) throw std::runtime_error("Extra Matrix Market entries");
    if (input.bad()) throw std::runtime_error("Matrix Market input read failure");

    // Stable bucket sort by row.
    HostCsr out;
    out.rows_ = int(rows);
    out.columns_ = int(columns);
    out.offsets_.assign(std::size_t(rows) + 1, 0);
    for (int r : entryRow) ++out.offsets_[std::size_t(r) + 1];
    for (std::size_t r = 0; r < std::size_t(rows); ++r) out.offsets_[r + 1] += out.offsets_[r];
    out.indices_.resize(entryRow.size());
    out.values_.resize(entryRow.size());
    std::vector<int> cursor(out.offsets_.begin(), out.offsets_.end() - 1);
    for (std::size_t k = 0; k < entryRow.size(); ++k) {
        const auto slot = std::size_t(cursor[std::size_t(entryRow[k])]++);
        out.indices_[slot] = entryColumn[k];
        out.values_[slot] = entryValue[k];
    }
    return out;
}

}  // namespace thermofem::io
