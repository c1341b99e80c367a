// RAII device memory of the C++20 drop-in API: memory strategies, DeviceError,
// Buffer<T> and Vector<T>, over the C ABI in include/thermofem_hip.h.
// Mirrors include/DataStructure/{Memory,CudaBuffer,Vector}.cuh of the
// reference: deep copies adopt the source's strategy, copyDataFrom keeps the
// target's address and strategy, pinned memory has no device view and
// cross-device copies throw.
#pragma once

#include <thermofem_hip.h>

#include <cstddef>
#include <cstring>
#include <limits>
#include <span>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <variant>
#include <vector>

namespace thermofem {

// Runtime failure of the HIP layer (CudaError in the reference, Memory.cuh:52-66).
class DeviceError : public std::runtime_error {
public:
    DeviceError(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    [[nodiscard]] int code() const noexcept { return code_; }

private:
    int code_;
};

namespace detail {
// Translates a C ABI status into the reference's exception types.
inline void check(int rc, const char* operation) {
    if (rc == TF_OK) return;
    const std::string message = std::string(operation) + ": " + tf_last_error();
    switch (rc) {
    case TF_E_INVALID:
    case TF_E_NUMERIC: throw std::invalid_argument(message);
    case TF_E_CAPACITY: throw std::length_error(message);
    default: throw DeviceError(rc, message);
    }
}
inline int currentDevice() {
    int d = 0;
    check(tf_get_device(&d), "device query");
    return d;
}
// Switches to a device and restores the caller's (Memory.cu:26-48).
class DeviceScope {
public:
    explicit DeviceScope(int device) : previous_(currentDevice()) {
        if (device != previous_) check(tf_set_device(device), "device switch");
    }
    ~DeviceScope() { (void)tf_set_device(previous_); }
    DeviceScope(const DeviceScope&) = delete;
    DeviceScope& operator=(const DeviceScope&) = delete;

private:
    int previous_;
};
inline std::size_t checkedProduct(std::size_t count, std::size_t size) {
    if (size && count > std::numeric_limits<std::size_t>::max() / size) throw std::length_error("Buffer size overflow");
    return count * size;
}
}  // namespace detail

struct GlobalMemory {};
struct PinnedMemory {};
struct MappedMemory {};
struct ManagedMemory {};
using MemoryStrategy = std::variant<GlobalMemory, PinnedMemory, MappedMemory, ManagedMemory>;
enum class MemoryType { Global, Pinned, Mapped, Managed };

template <class T>
concept BufferElement = std::is_trivially_copyable_v<T> && std::is_trivially_destructible_v<T> &&
                        !std::is_const_v<T> && !std::is_volatile_v<T> && !std::is_pointer_v<T>;

// Typed buffer with value semantics (CudaBuffer.cuh:29-149).
template <BufferElement T>
class Buffer {
public:
    explicit Buffer(MemoryStrategy strategy = GlobalMemory{}) noexcept : strategy_(strategy) {}
    explicit Buffer(std::size_t count, MemoryStrategy strategy = GlobalMemory{}) : strategy_(strategy) {
        (void)detail::checkedProduct(count, sizeof(T));  // size errors come before any device query
        device_ = detail::currentDevice();
        allocate(count);
    }
    explicit Buffer(std::span<const T> values, MemoryStrategy strategy = GlobalMemory{})
        : Buffer(values.size(), strategy) {
        copyFromHost(values);
    }
    Buffer(const Buffer& other) : strategy_(other.strategy_), device_(other.device_) {
        detail::DeviceScope scope(device_);
        allocate(other.size_);
        copyDataFrom(other);
    }
    Buffer(Buffer&& other) noexcept { swap(other); }
    Buffer& operator=(const Buffer& other) {
        if (this != &other) {
            Buffer temporary(other);
            swap(temporary);
        }
        return *this;
    }
    Buffer& operator=(Buffer&& other) noexcept {
        if (this != &other) {
            Buffer temporary(std::move(other));
            swap(temporary);
        }
        return *this;
    }
    ~Buffer() { release(); }

    void swap(Buffer& other) noexcept {
        std::swap(data_, other.data_);
        std::swap(size_, other.size_);
        std::swap(strategy_, other.strategy_);
        std::swap(device_, other.device_);
    }
    void copyDataFrom(const Buffer& other) {
        if (size_ != other.size_) throw std::invalid_argument("Buffer sizes must match");
        if (!size_ || this == &other) return;
        if (device_ != other.device_) throw std::invalid_argument("Cross-device copies are not supported");
        detail::DeviceScope scope(device_);
        detail::check(tf_memcpy(data_, other.data_, sizeBytes()), "buffer copy");
    }
    void changeMemoryType(MemoryStrategy strategy) {
        if (strategy_.index() == strategy.index()) return;
        if (!size_) {
            strategy_ = strategy;
            return;
        }
        detail::DeviceScope scope(device_);
        Buffer temporary(size_, strategy);
        temporary.copyDataFrom(*this);
        swap(temporary);
    }
    void copyFromHost(std::span<const T> values) {
        if (values.size() != size_ || (size_ && !values.data())) throw std::invalid_argument("Host copy size mismatch");
        if (size_) detail::check(tf_memcpy(data_, values.data(), sizeBytes()), "host upload");
    }
    void copyToHost(std::span<T> values) const {
        if (values.size() != size_ || (size_ && !values.data())) throw std::invalid_argument("Host copy size mismatch");
        if (size_) detail::check(tf_memcpy(values.data(), data_, sizeBytes()), "host download");
    }
    [[nodiscard]] std::vector<T> hostCopy() const {
        std::vector<T> out(size_);
        copyToHost(out);
        return out;
    }
    [[nodiscard]] std::size_t size() const noexcept { return size_; }
    [[nodiscard]] bool empty() const noexcept { return size_ == 0; }
    [[nodiscard]] std::size_t sizeBytes() const noexcept { return size_ * sizeof(T); }
    [[nodiscard]] int deviceId() const noexcept { return device_; }
    [[nodiscard]] MemoryType memoryType() const noexcept { return static_cast<MemoryType>(strategy_.index()); }
    [[nodiscard]] const MemoryStrategy& memoryStrategy() const noexcept { return strategy_; }
    [[nodiscard]] T* data() noexcept { return data_; }
    [[nodiscard]] const T* data() const noexcept { return data_; }
    // Host view (Memory.cu:262-268): Global memory is not host accessible.
    [[nodiscard]] T* hostData() { return const_cast<T*>(std::as_const(*this).hostData()); }
    [[nodiscard]] const T* hostData() const {
        if (size_ && memoryType() == MemoryType::Global) throw std::logic_error("Global memory is not host accessible");
        return data_;
    }
    // Device view (Memory.cu:246-260): pinned memory has none, mapped memory is aliased.
    [[nodiscard]] T* deviceData() { return const_cast<T*>(std::as_const(*this).deviceData()); }
    [[nodiscard]] const T* deviceData() const {
        if (!size_) return nullptr;
        if (memoryType() == MemoryType::Pinned)
            throw std::logic_error("Pinned memory has no guaranteed device view; use MappedMemory");
        void* view = nullptr;
        detail::check(tf_device_view(int(strategy_.index()), data_, &view), "device view");
        return static_cast<const T*>(view);
    }

private:
    void allocate(std::size_t count) {
        const std::size_t bytes = detail::checkedProduct(count, sizeof(T));
        size_ = count;
        if (!bytes) return;
        void* p = nullptr;
        detail::check(tf_alloc(device_, int(strategy_.index()), bytes, &p), "allocation");
        data_ = static_cast<T*>(p);
    }
    void release() noexcept {
        if (data_) (void)tf_free(int(strategy_.index()), data_);
        data_ = nullptr;
        size_ = 0;
    }
    T* data_ = nullptr;
    std::size_t size_ = 0;
    MemoryStrategy strategy_{};
    int device_ = 0;
};

// A vector owns exactly one buffer (Vector.cuh:22-106).
template <BufferElement T>
class Vector {
public:
    using ValueType = T;
    explicit Vector(MemoryStrategy strategy = GlobalMemory{}) noexcept : values_(strategy) {}
    explicit Vector(std::size_t count, MemoryStrategy strategy = GlobalMemory{}) : values_(count, strategy) {}
    explicit Vector(std::span<const T> values, MemoryStrategy strategy = GlobalMemory{}) : values_(values, strategy) {}
    void copyDataFrom(const Vector& other) { values_.copyDataFrom(other.values_); }
    void copyDataFrom(const Buffer<T>& other) { values_.copyDataFrom(other); }
    void changeMemoryType(MemoryStrategy strategy) { values_.changeMemoryType(strategy); }
    void copyFromHost(std::span<const T> values) { values_.copyFromHost(values); }
    void copyToHost(std::span<T> values) const { values_.copyToHost(values); }
    [[nodiscard]] std::vector<T> hostCopy() const { return values_.hostCopy(); }
    void swap(Vector& other) noexcept { values_.swap(other.values_); }
    [[nodiscard]] std::size_t size() const noexcept { return values_.size(); }
    [[nodiscard]] bool empty() const noexcept { return values_.empty(); }
    [[nodiscard]] std::size_t sizeBytes() const noexcept { return values_.sizeBytes(); }
    [[nodiscard]] int deviceId() const noexcept { return values_.deviceId(); }
    [[nodiscard]] MemoryType memoryType() const noexcept { return values_.memoryType(); }
    [[nodiscard]] const MemoryStrategy& memoryStrategy() const noexcept { return values_.memoryStrategy(); }
    [[nodiscard]] const Buffer<T>& values() const noexcept { return values_; }
    [[nodiscard]] T* data() noexcept { return values_.data(); }
    [[nodiscard]] const T* data() const noexcept { return values_.data(); }
    [[nodiscard]] T* hostData() { return values_.hostData(); }
    [[nodiscard]] const T* hostData() const { return values_.hostData(); }
    [[nodiscard]] T* deviceData() { return values_.deviceData(); }
    [[nodiscard]] const T* deviceData() const { return values_.deviceData(); }

private:
    Buffer<T> values_;
};

}  // namespace thermofem
