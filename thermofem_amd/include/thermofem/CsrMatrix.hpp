// CsrMatrix<T, I> of the C++20 drop-in API (include/DataStructure/CsrMatrix.cuh:36-217):
// three buffers under one strategy, host-side validation for host input,
// device validation in fromBuffers (src/DataStructure/CsrMatrix.cu:12-58).
// Unsorted columns, duplicates and explicit zeros are allowed.
#pragma once

#include <cstdint>
#include <limits>
#include <utility>

#include "Memory.hpp"

namespace thermofem {

template <class T>
concept CsrIndex = std::signed_integral<T> && (sizeof(T) == 4 || sizeof(T) == 8);

template <BufferElement T, CsrIndex Index = std::int32_t>
class CsrMatrix {
public:
    using ValueType = T;
    CsrMatrix() = default;

    // Takes ownership of device buffers after validating them on the device;
    // on failure the inputs are left untouched (amg_native_test.cpp:125-128).
    static CsrMatrix fromBuffers(std::size_t rows, std::size_t columns, Buffer<Index>&& offsets,
                                 Buffer<Index>&& indices, Buffer<T>&& values) {
        if (!std::in_range<Index>(rows) || !std::in_range<Index>(columns) || !std::in_range<Index>(values.size()) ||
            rows == std::numeric_limits<std::size_t>::max())
            throw std::length_error("CSR dimensions exceed index range");
        if (offsets.size() != rows + 1 || indices.size() != values.size() ||
            offsets.memoryType() != MemoryType::Global || indices.memoryType() != MemoryType::Global ||
            values.memoryType() != MemoryType::Global ||
            (!indices.empty() && offsets.deviceId() != indices.deviceId()) ||
            (!values.empty() && offsets.deviceId() != values.deviceId()))
            throw std::invalid_argument("Invalid CSR device buffers");
        validateDevice(rows, columns, values.size(), offsets, indices);
        CsrMatrix result;
        result.rowCount_ = rows;
        result.columnCount_ = columns;
        result.rowOffsets_ = std::move(offsets);
        result.columnIndices_ = std::move(indices);
        result.values_ = std::move(values);
        return result;
    }

    CsrMatrix(std::size_t rows, std::size_t columns, std::span<const Index> rowOffsets,
              std::span<const Index> columnIndices, std::span<const T> values, MemoryStrategy strategy = GlobalMemory{})
        : rowCount_(rows), columnCount_(columns) {
        validateHost(rows, columns, rowOffsets, columnIndices, values.size());
        rowOffsets_ = Buffer<Index>(rowOffsets, strategy);
        columnIndices_ = Buffer<Index>(columnIndices, strategy);
        values_ = Buffer<T>(values, strategy);
    }

    void swap(CsrMatrix& other) noexcept {
        std::swap(rowCount_, other.rowCount_);
        std::swap(columnCount_, other.columnCount_);
        rowOffsets_.swap(other.rowOffsets_);
        columnIndices_.swap(other.columnIndices_);
        values_.swap(other.values_);
    }
    void changeMemoryType(MemoryStrategy strategy) {
        if (values_.memoryStrategy().index() == strategy.index()) return;
        CsrMatrix temporary(*this);
        temporary.rowOffsets_.changeMemoryType(strategy);
        temporary.columnIndices_.changeMemoryType(strategy);
        temporary.values_.changeMemoryType(strategy);
        swap(temporary);
    }
    void copyValuesFromHost(std::span<const T> values) { values_.copyFromHost(values); }
    [[nodiscard]] std::size_t rowCount() const noexcept { return rowCount_; }
    [[nodiscard]] std::size_t columnCount() const noexcept { return columnCount_; }
    [[nodiscard]] std::size_t nonZeroCount() const noexcept { return values_.size(); }
    [[nodiscard]] MemoryType memoryType() const noexcept { return values_.memoryType(); }
    [[nodiscard]] const Buffer<Index>& rowOffsets() const noexcept { return rowOffsets_; }
    [[nodiscard]] const Buffer<Index>& columnIndices() const noexcept { return columnIndices_; }
    [[nodiscard]] const Buffer<T>& values() const noexcept { return values_; }
    [[nodiscard]] T* deviceValues() { return values_.deviceData(); }
    [[nodiscard]] const T* deviceValues() const { return values_.deviceData(); }

private:
    static void validateHost(std::size_t rows, std::size_t columns, std::span<const Index> offsets,
                             std::span<const Index> indices, std::size_t count) {
        if (!std::in_range<Index>(rows) || !std::in_range<Index>(columns) || !std::in_range<Index>(count) ||
            rows == std::numeric_limits<std::size_t>::max())
            throw std::length_error("CSR dimensions exceed the index range");
        if (offsets.size() != rows + 1 || indices.size() != count || offsets.front() != 0 ||
            offsets.back() != static_cast<Index>(count))
            throw std::invalid_argument("Invalid CSR array sizes or endpoints");
        Index previous{};
        for (const auto offset : offsets) {
            if (offset < previous || offset > static_cast<Index>(count))
                throw std::invalid_argument("CSR row offsets must be nondecreasing and bounded");
            previous = offset;
        }
        for (const auto column : indices)
            if (column < 0 || static_cast<std::size_t>(column) >= columns)
                throw std::invalid_argument("CSR column index is out of range");
    }
    static void validateDevice(std::size_t rows, std::size_t columns, std::size_t count, const Buffer<Index>& offsets,
                               const Buffer<Index>& indices) {
        if constexpr (sizeof(Index) == 4) {
            tf_ctx* ctx = nullptr;
            detail::check(tf_ctx_create(offsets.deviceId(), &ctx), "CSR validation context");
            int bad = 0;
            const int rc = tf_csr_validate_structure(ctx, int(rows), int(columns), int64_t(count),
                                                     reinterpret_cast<const int*>(offsets.deviceData()),
                                                     reinterpret_cast<const int*>(indices.deviceData()), &bad);
            (void)tf_ctx_destroy(ctx);
            detail::check(rc, "CSR validation");
            if (bad) throw std::invalid_argument("Invalid device CSR structure");
        }
        else {
            // 64-bit indices are validated on a host copy.
            const auto o = offsets.hostCopy();
            const auto c = indices.hostCopy();
            bool bad = o.front() != 0 || std::size_t(o.back()) != count;
            for (std::size_t r = 0; r < rows && !bad; ++r) bad = o[r] < 0 || o[r + 1] < o[r] || std::size_t(o[r + 1]) > count;
            for (std::size_t j = 0; j < c.size() && !bad; ++j) bad = c[j] < 0 || std::size_t(c[j]) >= columns;
            if (bad) throw std::invalid_argument("Invalid device CSR structure");
        }
    }

    std::size_t rowCount_{};
    std::size_t columnCount_{};
    Buffer<Index> rowOffsets_;
    Buffer<Index> columnIndices_;
    Buffer<T> values_;
};

}  // namespace thermofem
