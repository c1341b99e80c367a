// Row-major dense matrix over a Buffer (include/DataStructure/DenseMatrix.cuh:
// the reference's class, same members, copy rules and errors).
#pragma once

#include <cstddef>
#include <span>
#include <stdexcept>
#include <utility>

#include "Memory.hpp"

namespace thermofem {

template <BufferElement T>
class DenseMatrix {
public:
    using ValueType = T;
    DenseMatrix() = default;
    DenseMatrix(std::size_t rows, std::size_t columns, MemoryStrategy strategy = GlobalMemory{})
        : rowCount_(rows), columnCount_(columns), values_(detail::checkedProduct(rows, columns), strategy) {}
    DenseMatrix(const DenseMatrix&) = default;
    DenseMatrix(DenseMatrix&& other) noexcept { swap(other); }
    DenseMatrix& operator=(const DenseMatrix& other) {
        if (this != &other) {
            DenseMatrix temporary(other);
            swap(temporary);
        }
        return *this;
    }
    DenseMatrix& operator=(DenseMatrix&& other) noexcept {
        if (this != &other) {
            DenseMatrix temporary(std::move(other));
            swap(temporary);
        }
        return *this;
    }
    void swap(DenseMatrix& other) noexcept {
        std::swap(rowCount_, other.rowCount_);
        std::swap(columnCount_, other.columnCount_);
        values_.swap(other.values_);
    }
    void copyDataFrom(const DenseMatrix& other) {
        if (rowCount_ != other.rowCount_ || columnCount_ != other.columnCount_)
            throw std::invalid_argument("Dense matrix shapes must match");
        values_.copyDataFrom(other.values_);
    }
    void changeMemoryType(MemoryStrategy strategy) { values_.changeMemoryType(strategy); }
    void copyFromHost(std::span<const T> values) { values_.copyFromHost(values); }
    void copyValuesFrom(const Buffer<T>& values) { values_.copyDataFrom(values); }
    void copyToHost(std::span<T> values) const { values_.copyToHost(values); }
    [[nodiscard]] std::size_t rowCount() const noexcept { return rowCount_; }
    [[nodiscard]] std::size_t columnCount() const noexcept { return columnCount_; }
    [[nodiscard]] const Buffer<T>& values() const noexcept { return values_; }
    [[nodiscard]] MemoryType memoryType() const noexcept { return values_.memoryType(); }
    [[nodiscard]] T* deviceData() { return values_.deviceData(); }
    [[nodiscard]] const T* deviceData() const { return values_.deviceData(); }
    // Element (row, column) through the host view: host-accessible strategies only.
    T& at(std::size_t row, std::size_t column) { return const_cast<T&>(std::as_const(*this).at(row, column)); }
    const T& at(std::size_t row, std::size_t column) const {
        if (row >= rowCount_ || column >= columnCount_) throw std::out_of_range("Dense matrix index is out of range");
        return values_.hostData()[row * columnCount_ + column];
    }

private:
    std::size_t rowCount_{};
    std::size_t columnCount_{};
    Buffer<T> values_;
};

}  // namespace thermofem
