// linalg::IterationContext<T, I> and linalg::inverseDiagonal of the C++20
// drop-in API (include/LinearAlgebra/IterationContext.cuh:25-84,
// include/LinearAlgebra/Diagonal.cuh:13-36), over the C ABI.  Every
// operation is synchronous, as in the reference.
#pragma once

#include <algorithm>
#include <cmath>
#include <memory>

#include "CsrMatrix.hpp"

namespace thermofem::linalg {

template <class T>
concept Scalar = std::same_as<T, float> || std::same_as<T, double>;

namespace detail {
using thermofem::detail::check;
// Device guard of every public call (Backend.cu:116-126).
using OperationScope = thermofem::detail::DeviceScope;

// Host copy of any vector, whatever its strategy (used by the reference's tests).
template <BufferElement T>
std::vector<T> hostCopy(const Vector<T>& v) {
    return v.hostCopy();
}
template <BufferElement T>
std::vector<T> hostCopy(const Buffer<T>& b) {
    return b.hostCopy();
}

// Owned execution context (stream + reduction scratch) on one device.
class Context {
public:
    explicit Context(int device) { check(tf_ctx_create(device, &ctx_), "context"); }
    ~Context() { (void)tf_ctx_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    tf_ctx* get() const noexcept { return ctx_; }
    void sync() const { check(tf_ctx_synchronize(ctx_), "synchronize"); }

private:
    tf_ctx* ctx_ = nullptr;
};

inline int librarySize(std::size_t n) {
    if (n > std::size_t(std::numeric_limits<int>::max())) throw std::length_error("Size exceeds the library's int range");
    return int(n);
}

// Canonical FP64/int32 device copy of any CsrMatrix (sorted columns,
// duplicates summed); the native form the kernels run on.
struct NativeCsr {
    int rows = 0;
    int64_t nnz = 0;
    Buffer<int> rp, ci;
    Buffer<double> v;
};

template <Scalar T, CsrIndex I>
NativeCsr toNative(const Context& ctx, const CsrMatrix<T, I>& a, bool canonicalize) {
    if (a.rowCount() >= std::size_t(std::numeric_limits<int>::max()) ||
        a.nonZeroCount() > std::size_t(std::numeric_limits<int>::max() / 2))
        throw std::length_error("AMG native index capacity exceeded");
    thermofem::detail::DeviceScope scope(a.rowOffsets().deviceId());
    const int n = int(a.rowCount());
    const std::size_t nnz = a.nonZeroCount();
    Buffer<int> rp(std::size_t(n) + 1), ci(nnz);
    Buffer<double> v(nnz);
    auto global = a;
    global.changeMemoryType(GlobalMemory{});
    if constexpr (sizeof(I) == 8) {
        check(tf_convert_i64_i32(ctx.get(), rp.size(), reinterpret_cast<const int64_t*>(global.rowOffsets().deviceData()),
                                 rp.data()),
              "index conversion");
        if (nnz)
            check(tf_convert_i64_i32(ctx.get(), nnz, reinterpret_cast<const int64_t*>(global.columnIndices().deviceData()),
                                     ci.data()),
                  "index conversion");
    }
    else {
        rp.copyDataFrom(reinterpret_cast<const Buffer<int>&>(global.rowOffsets()));
        if (nnz) ci.copyDataFrom(reinterpret_cast<const Buffer<int>&>(global.columnIndices()));
    }
    if (nnz) {
        if constexpr (std::same_as<T, float>)
            check(tf_convert_f32_f64(ctx.get(), nnz, global.values().deviceData(), v.data()), "value conversion");
        else
            v.copyDataFrom(global.values());
    }
    ctx.sync();
    NativeCsr out;
    out.rows = n;
    out.nnz = int64_t(nnz);
    if (!canonicalize) {
        out.rp = std::move(rp);
        out.ci = std::move(ci);
        out.v = std::move(v);
        return out;
    }
    int* oci = nullptr;
    double* ov = nullptr;
    int64_t onnz = 0;
    out.rp = Buffer<int>(std::size_t(n) + 1);
    check(tf_csr_canonicalize(ctx.get(), n, int64_t(nnz), rp.data(), ci.data(), v.data(), out.rp.data(), &oci, &ov,
                              &onnz),
          "canonical CSR");
    out.nnz = onnz;
    out.ci = Buffer<int>(std::size_t(onnz));
    out.v = Buffer<double>(std::size_t(onnz));
    if (onnz) {
        check(tf_memcpy(out.ci.data(), oci, std::size_t(onnz) * sizeof(int)), "canonical copy");
        check(tf_memcpy(out.v.data(), ov, std::size_t(onnz) * sizeof(double)), "canonical copy");
    }
    (void)tf_free(TF_MEM_GLOBAL, oci);
    (void)tf_free(TF_MEM_GLOBAL, ov);
    return out;
}

// Native FP64 CSR back to CsrMatrix<T, I> (fromNativeMatrix, GpuAggregation.cu:1133-1136).
template <Scalar T, CsrIndex I>
CsrMatrix<T, I> fromNative(const Context& ctx, const NativeCsr& a) {
    Buffer<I> rp(std::size_t(a.rows) + 1), ci(std::size_t(a.nnz));
    Buffer<T> v(std::size_t(a.nnz));
    if constexpr (sizeof(I) == 8) {
        check(tf_convert_i32_i64(ctx.get(), rp.size(), a.rp.data(), reinterpret_cast<int64_t*>(rp.data())),
              "index conversion");
        if (a.nnz)
            check(tf_convert_i32_i64(ctx.get(), ci.size(), a.ci.data(), reinterpret_cast<int64_t*>(ci.data())),
                  "index conversion");
    }
    else {
        rp.copyDataFrom(reinterpret_cast<const Buffer<I>&>(a.rp));
        if (a.nnz) ci.copyDataFrom(reinterpret_cast<const Buffer<I>&>(a.ci));
    }
    if (a.nnz) {
        if constexpr (std::same_as<T, float>)
            check(tf_convert_f64_f32(ctx.get(), v.size(), a.v.data(), v.data()), "value conversion");
        else
            v.copyDataFrom(a.v);
    }
    ctx.sync();
    return CsrMatrix<T, I>::fromBuffers(std::size_t(a.rows), std::size_t(a.rows), std::move(rp), std::move(ci),
                                        std::move(v));
}
}  // namespace detail

// Snapshot of the canonical matrix with persistent execution resources
// (IterationContext.cu:38-124).
template <Scalar T, CsrIndex I = int>
class IterationContext {
public:
    explicit IterationContext(const CsrMatrix<T, I>& matrix)
        : device_(matrix.rowOffsets().empty() ? thermofem::detail::currentDevice() : matrix.rowOffsets().deviceId()),
          ctx_(std::make_unique<detail::Context>(device_)) {
        detail::librarySize(matrix.rowCount());
        detail::librarySize(matrix.columnCount());
        if (matrix.rowCount() != matrix.columnCount())
            throw std::invalid_argument("IterationContext requires a square matrix");
        thermofem::detail::DeviceScope scope(device_);
        native_ = detail::toNative(*ctx_, matrix, true);
        int flags = 0;
        detail::check(tf_csr_check(ctx_->get(), native_.rows, native_.nnz, native_.rp.data(), native_.ci.data(),
                                   native_.v.data(), &flags),
                      "matrix check");
        if (flags & 2) throw std::invalid_argument("Matrix contains nonfinite entries");
        matrix_ = detail::fromNative<T, I>(*ctx_, native_);
        if constexpr (std::same_as<T, float>) {
            typed_ = Buffer<float>(std::size_t(native_.nnz));
            if (native_.nnz)
                detail::check(tf_convert_f64_f32(ctx_->get(), typed_.size(), native_.v.data(), typed_.data()),
                              "value conversion");
            ctx_->sync();
        }
    }
    IterationContext(const IterationContext&) = delete;
    IterationContext& operator=(const IterationContext&) = delete;

    [[nodiscard]] const CsrMatrix<T, I>& matrix() const noexcept { return matrix_; }

    void multiply(const Vector<T>& x, Vector<T>& y) {
        validate(x);
        validate(y);
        if (x.size() != matrix_.columnCount() || y.size() != matrix_.rowCount() || (x.size() && x.data() == y.data()))
            throw std::invalid_argument("SpMV shape or alias mismatch");
        thermofem::detail::DeviceScope scope(device_);
        if constexpr (std::same_as<T, double>)
            detail::check(tf_spmv_csr_f64(ctx_->get(), native_.rows, native_.nnz, native_.rp.data(), native_.ci.data(),
                                          values(), x.deviceData(), y.deviceData()),
                          "SpMV");
        else
            detail::check(tf_spmv_csr_f32(ctx_->get(), native_.rows, native_.nnz, native_.rp.data(), native_.ci.data(),
                                          values(), x.deviceData(), y.deviceData()),
                          "SpMV");
        ctx_->sync();
    }
    void residual(const Vector<T>& rhs, const Vector<T>& x, Vector<T>& out) {
        validate(rhs);
        validate(x);
        validate(out);
        if (rhs.size() != out.size() || (rhs.size() && (rhs.data() == out.data() || x.data() == out.data())) ||
            x.size() != matrix_.columnCount() || out.size() != matrix_.rowCount())
            throw std::invalid_argument("Residual shape or alias mismatch");
        thermofem::detail::DeviceScope scope(device_);
        if constexpr (std::same_as<T, double>)
            detail::check(tf_residual_csr_f64(ctx_->get(), native_.rows, native_.nnz, native_.rp.data(),
                                              native_.ci.data(), values(), rhs.deviceData(), x.deviceData(),
                                              out.deviceData()),
                          "residual");
        else
            detail::check(tf_residual_csr_f32(ctx_->get(), native_.rows, native_.nnz, native_.rp.data(),
                                              native_.ci.data(), values(), rhs.deviceData(), x.deviceData(),
                                              out.deviceData()),
                          "residual");
        ctx_->sync();
    }
    void clear(Vector<T>& x) {
        validate(x);
        thermofem::detail::DeviceScope scope(device_);
        detail::check(tf_fill_zero(ctx_->get(), x.deviceData(), x.sizeBytes()), "clear");
        ctx_->sync();
    }
    void axpy(T alpha, const Vector<T>& x, Vector<T>& y) {
        validate(x);
        validate(y);
        if (x.size() != y.size()) throw std::invalid_argument("AXPY shape mismatch");
        if (x.empty()) return;
        thermofem::detail::DeviceScope scope(device_);
        if constexpr (std::same_as<T, double>)
            detail::check(tf_axpy_f64(ctx_->get(), int(x.size()), alpha, x.deviceData(), y.deviceData()), "axpy");
        else
            detail::check(tf_axpy_f32(ctx_->get(), int(x.size()), alpha, x.deviceData(), y.deviceData()), "axpy");
        ctx_->sync();
    }
    void scale(T alpha, Vector<T>& x) {
        validate(x);
        if (x.empty()) return;
        thermofem::detail::DeviceScope scope(device_);
        if constexpr (std::same_as<T, double>)
            detail::check(tf_scal_f64(ctx_->get(), int(x.size()), alpha, x.deviceData()), "scal");
        else
            detail::check(tf_scal_f32(ctx_->get(), int(x.size()), alpha, x.deviceData()), "scal");
        ctx_->sync();
    }
    T dot(const Vector<T>& x, const Vector<T>& y) {
        validate(x);
        validate(y);
        if (x.size() != y.size()) throw std::invalid_argument("Dot shape mismatch");
        T result{};
        if (x.empty()) return result;
        thermofem::detail::DeviceScope scope(device_);
        if constexpr (std::same_as<T, double>)
            detail::check(tf_dot_f64(ctx_->get(), int(x.size()), x.deviceData(), y.deviceData(), &result), "dot");
        else
            detail::check(tf_dot_f32(ctx_->get(), int(x.size()), x.deviceData(), y.deviceData(), &result), "dot");
        return result;
    }
    T norm(const Vector<T>& x) {
        validate(x);
        T result{};
        if (x.empty()) return result;
        thermofem::detail::DeviceScope scope(device_);
        if constexpr (std::same_as<T, double>)
            detail::check(tf_nrm2_f64(ctx_->get(), int(x.size()), x.deviceData(), &result), "nrm2");
        else
            detail::check(tf_nrm2_f32(ctx_->get(), int(x.size()), x.deviceData(), &result), "nrm2");
        return result;
    }
    void jacobi(T weight, const Vector<T>& inverse, const Vector<T>& rhs, const Vector<T>& product, Vector<T>& x) {
        validate(inverse);
        validate(rhs);
        validate(product);
        validate(x);
        if (inverse.size() != x.size() || rhs.size() != x.size() || product.size() != x.size())
            throw std::invalid_argument("Jacobi shape mismatch");
        if (x.empty()) return;
        thermofem::detail::DeviceScope scope(device_);
        if constexpr (std::same_as<T, double>)
            detail::check(tf_jacobi_update_f64(ctx_->get(), int(x.size()), weight, inverse.deviceData(),
                                               rhs.deviceData(), product.deviceData(), x.deviceData()),
                          "jacobi");
        else
            detail::check(tf_jacobi_update_f32(ctx_->get(), int(x.size()), weight, inverse.deviceData(),
                                               rhs.deviceData(), product.deviceData(), x.deviceData()),
                          "jacobi");
        ctx_->sync();
    }

    // One fused weighted-Jacobi sweep, out = x + (w D^-1)(rhs - A x), in one
    // pass over the matrix (the library's addition to the reference's
    // multiply + jacobi pair, with the same arithmetic); out must not alias x.
    void sweep(T weight, const Vector<T>& inverse, const Vector<T>& rhs, const Vector<T>& x, Vector<T>& out) {
        validate(inverse);
        validate(rhs);
        validate(x);
        validate(out);
        if (inverse.size() != x.size() || rhs.size() != x.size() || out.size() != x.size() ||
            x.size() != matrix_.rowCount() || (x.size() && x.data() == out.data()))
            throw std::invalid_argument("Jacobi shape or alias mismatch");
        if (x.empty()) return;
        thermofem::detail::DeviceScope scope(device_);
        if constexpr (std::same_as<T, double>)
            detail::check(tf_jacobi_fused_f64(ctx_->get(), native_.rows, native_.nnz, native_.rp.data(),
                                              native_.ci.data(), values(), inverse.deviceData(), weight,
                                              rhs.deviceData(), x.deviceData(), out.deviceData()),
                          "jacobi sweep");
        else
            detail::check(tf_jacobi_fused_f32(ctx_->get(), native_.rows, native_.nnz, native_.rp.data(),
                                              native_.ci.data(), values(), inverse.deviceData(), weight,
                                              rhs.deviceData(), x.deviceData(), out.deviceData()),
                          "jacobi sweep");
        ctx_->sync();
    }

    // Library-internal views for the solvers built on this context.
    [[nodiscard]] const detail::NativeCsr& native() const noexcept { return native_; }
    [[nodiscard]] tf_ctx* handle() const noexcept { return ctx_->get(); }
    [[nodiscard]] int device() const noexcept { return device_; }

private:
    const T* values() const {
        if constexpr (std::same_as<T, double>)
            return native_.v.data();
        else
            return typed_.data();
    }
    void validate(const Vector<T>& v) const {
        detail::librarySize(v.size());
        if (v.memoryType() != MemoryType::Global || (v.size() && v.deviceId() != device_))
            throw std::invalid_argument("Iteration workspace requires same-device global vectors");
    }
    int device_;
    std::unique_ptr<detail::Context> ctx_;
    detail::NativeCsr native_;
    CsrMatrix<T, I> matrix_;
    Buffer<float> typed_;
};

// 1 / (sum of diagonal entries) per row; throws on a missing, zero or
// non-finite diagonal or reciprocal (Diagonal.cuh:13-36).
template <Scalar T, CsrIndex I>
Vector<T> inverseDiagonal(const CsrMatrix<T, I>& matrix) {
    if (!matrix.rowCount() || matrix.rowCount() != matrix.columnCount())
        throw std::invalid_argument("Inverse diagonal requires a nonempty square matrix");
    thermofem::detail::DeviceScope scope(matrix.rowOffsets().deviceId());
    detail::Context ctx(matrix.rowOffsets().deviceId());
    auto native = detail::toNative(ctx, matrix, false);
    Vector<T> inverse(matrix.rowCount());
    int bad = 0;
    if constexpr (std::same_as<T, double>) {
        detail::check(tf_inverse_diagonal_f64(ctx.get(), native.rows, native.rp.data(), native.ci.data(),
                                              native.v.data(), inverse.deviceData(), &bad),
                      "inverse diagonal");
    }
    else {
        Buffer<float> v(std::size_t(native.nnz));
        if (native.nnz)
            detail::check(tf_convert_f64_f32(ctx.get(), v.size(), native.v.data(), v.data()), "value conversion");
        detail::check(tf_inverse_diagonal_f32(ctx.get(), native.rows, native.rp.data(), native.ci.data(), v.data(),
                                              inverse.deviceData(), &bad),
                      "inverse diagonal");
    }
    if (bad) throw std::invalid_argument("Missing or invalid diagonal");
    return inverse;
}

// Host dense LU with partial pivoting in T (CoarseLu.cuh): factors at
// construction and throws std::invalid_argument when a pivot falls to
// epsilon(T) of the largest entry or a factor is not finite; apply() solves
// on the host (rhs and x may alias).
template <Scalar T>
class CoarseLu {
public:
    template <CsrIndex I>
    explicit CoarseLu(const CsrMatrix<T, I>& matrix) : n_(matrix.rowCount()) {
        if (!n_ || n_ != matrix.columnCount()) throw std::invalid_argument("Coarse LU requires a nonempty square matrix");
        a_.assign(thermofem::detail::checkedProduct(n_, n_), T(0));
        const auto rp = matrix.rowOffsets().hostCopy();
        const auto ci = matrix.columnIndices().hostCopy();
        const auto v = matrix.values().hostCopy();
        for (std::size_t i = 0; i < n_; ++i)
            for (auto k = rp[i]; k < rp[i + 1]; ++k) a_[i * n_ + std::size_t(ci[std::size_t(k)])] += v[std::size_t(k)];
        T largest = 0;
        for (T e : a_) largest = std::max(largest, std::abs(e));
        if (!std::isfinite(largest) || largest == 0) throw std::invalid_argument("Singular AMG coarse matrix");
        perm_.resize(n_);
        for (std::size_t k = 0; k < n_; ++k) {
            std::size_t best = k;
            for (std::size_t i = k + 1; i < n_; ++i)
                if (std::abs(at(i, k)) > std::abs(at(best, k))) best = i;
            if (std::abs(at(best, k)) / largest <= std::numeric_limits<T>::epsilon())
                throw std::invalid_argument("Singular or numerically singular AMG coarse matrix");
            perm_[k] = best;
            if (best != k) std::swap_ranges(&a_[k * n_], &a_[k * n_] + n_, &a_[best * n_]);
            for (std::size_t i = k + 1; i < n_; ++i) {
                const T l = at(i, k) /= at(k, k);
                for (std::size_t j = k + 1; j < n_; ++j) at(i, j) -= l * at(k, j);
            }
        }
        for (T e : a_)
            if (!std::isfinite(e)) throw std::invalid_argument("Nonfinite coarse LU factors");
    }
    void apply(const Vector<T>& rhs, Vector<T>& x) {
        if (rhs.size() != n_ || x.size() != n_) throw std::invalid_argument("Coarse LU vector shape mismatch");
        if (rhs.deviceId() != x.deviceId()) throw std::invalid_argument("Coarse LU vectors on different devices");
        std::vector<T> y = rhs.hostCopy();
        for (std::size_t k = 0; k < n_; ++k) std::swap(y[k], y[perm_[k]]);
        for (std::size_t i = 0; i < n_; ++i)
            for (std::size_t j = 0; j < i; ++j) y[i] -= at(i, j) * y[j];
        for (std::size_t i = n_; i-- > 0;) {
            for (std::size_t j = i + 1; j < n_; ++j) y[i] -= at(i, j) * y[j];
            y[i] /= at(i, i);
        }
        x.copyFromHost(y);
    }

private:
    T& at(std::size_t i, std::size_t j) { return a_[i * n_ + j]; }
    std::size_t n_;
    std::vector<T> a_;
    std::vector<std::size_t> perm_;
};

}  // namespace thermofem::linalg
