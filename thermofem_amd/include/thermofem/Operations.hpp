// The reference's one-shot linear-algebra functions (include/LinearAlgebra/
// Operations.cuh:160-525) and operators (Operators.cuh:8-125) over the C ABI.
// Each call runs on its own context of the operands' device; results keep
// the left operand's memory strategy.  Sparse operands are first brought to
// the canonical FP64 / int32 form (Operations.cuh:94-127 does it on the
// host); sparse results are canonical and, for float matrices, rounded from
// FP64.
#pragma once

#include <concepts>
#include <initializer_list>
#include <stdexcept>
#include <type_traits>

#include "CsrMatrix.hpp"
#include "DenseMatrix.hpp"
#include "LinearAlgebra.hpp"
#include "Memory.hpp"

namespace thermofem::linalg {

namespace detail {

// Operations.cuh:19-32
inline int commonDevice(std::initializer_list<int> devices) {
    int result = -1;
    for (const int device : devices) {
        if (device < 0) continue;
        if (result >= 0 && result != device) throw std::invalid_argument("Operands belong to different devices");
        result = device;
    }
    return result < 0 ? thermofem::detail::currentDevice() : result;
}

// Global-memory copy of any buffer on its device (Operations.cuh:54-59).
template <BufferElement T>
Buffer<T> deviceCopy(const Buffer<T>& b) {
    thermofem::detail::DeviceScope scope(b.deviceId());
    Buffer<T> r(b.size());
    r.copyDataFrom(b);
    return r;
}

inline int abiScal(tf_ctx* c, int n, double a, double* x) { return tf_scal_f64(c, n, a, x); }
inline int abiScal(tf_ctx* c, int n, float a, float* x) { return tf_scal_f32(c, n, a, x); }
inline int abiAxpy(tf_ctx* c, int n, double a, const double* x, double* y) { return tf_axpy_f64(c, n, a, x, y); }
inline int abiAxpy(tf_ctx* c, int n, float a, const float* x, float* y) { return tf_axpy_f32(c, n, a, x, y); }
inline int abiDot(tf_ctx* c, int n, const double* x, const double* y, double* r) { return tf_dot_f64(c, n, x, y, r); }
inline int abiDot(tf_ctx* c, int n, const float* x, const float* y, float* r) { return tf_dot_f32(c, n, x, y, r); }
inline int abiNrm2(tf_ctx* c, int n, const double* x, double* r) { return tf_nrm2_f64(c, n, x, r); }
inline int abiNrm2(tf_ctx* c, int n, const float* x, float* r) { return tf_nrm2_f32(c, n, x, r); }
inline int abiAsum(tf_ctx* c, int n, const double* x, double* r) { return tf_asum_f64(c, n, x, r); }
inline int abiAsum(tf_ctx* c, int n, const float* x, float* r) { return tf_asum_f32(c, n, x, r); }
inline int abiTranspose(tf_ctx* c, int r, int k, const double* a, double* t) { return tf_dense_transpose_f64(c, r, k, a, t); }
inline int abiTranspose(tf_ctx* c, int r, int k, const float* a, float* t) { return tf_dense_transpose_f32(c, r, k, a, t); }
inline int abiGemv(tf_ctx* c, int r, int k, const double* a, const double* x, double* y) { return tf_gemv_f64(c, r, k, a, x, y); }
inline int abiGemv(tf_ctx* c, int r, int k, const float* a, const float* x, float* y) { return tf_gemv_f32(c, r, k, a, x, y); }
inline int abiGemm(tf_ctx* c, int m, int n, int k, const double* a, const double* b, double* o) {
    return tf_gemm_f64(c, m, n, k, a, b, o);
}
inline int abiGemm(tf_ctx* c, int m, int n, int k, const float* a, const float* b, float* o) {
    return tf_gemm_f32(c, m, n, k, a, b, o);
}

// Result of a tf_csr_* call as a CsrMatrix<T, I> of rows x columns.
template <Scalar T, CsrIndex I>
CsrMatrix<T, I> fromAbi(const Context& ctx, std::size_t rows, std::size_t columns, Buffer<int>& rp, int* ci,
                        double* v, int64_t nnz, MemoryStrategy strategy) {
    NativeCsr n;
    n.rows = int(rows);
    n.nnz = nnz;
    n.rp = std::move(rp);
    n.ci = Buffer<int>(static_cast<std::size_t>(nnz));
    n.v = Buffer<double>(static_cast<std::size_t>(nnz));
    if (nnz) {
        check(tf_memcpy(n.ci.data(), ci, std::size_t(nnz) * sizeof(int)), "sparse result copy");
        check(tf_memcpy(n.v.data(), v, std::size_t(nnz) * sizeof(double)), "sparse result copy");
    }
    (void)tf_free(TF_MEM_GLOBAL, ci);
    (void)tf_free(TF_MEM_GLOBAL, v);
    Buffer<I> orp(rows + 1), oci(static_cast<std::size_t>(nnz));
    Buffer<T> ov(static_cast<std::size_t>(nnz));
    if constexpr (sizeof(I) == 8) {
        check(tf_convert_i32_i64(ctx.get(), orp.size(), n.rp.data(), reinterpret_cast<int64_t*>(orp.data())),
              "index conversion");
        if (nnz)
            check(tf_convert_i32_i64(ctx.get(), oci.size(), n.ci.data(), reinterpret_cast<int64_t*>(oci.data())),
                  "index conversion");
    }
    else {
        orp.copyDataFrom(reinterpret_cast<const Buffer<I>&>(n.rp));
        if (nnz) oci.copyDataFrom(reinterpret_cast<const Buffer<I>&>(n.ci));
    }
    if (nnz) {
        if constexpr (std::same_as<T, float>)
            check(tf_convert_f64_f32(ctx.get(), ov.size(), n.v.data(), ov.data()), "value conversion");
        else
            ov.copyDataFrom(n.v);
    }
    ctx.sync();
    auto out = CsrMatrix<T, I>::fromBuffers(rows, columns, std::move(orp), std::move(oci), std::move(ov));
    out.changeMemoryType(strategy);
    return out;
}

// Operations.cuh:131-145
template <Scalar T, CsrIndex I>
CsrMatrix<T, I> emptySparse(std::size_t rows, std::size_t columns, MemoryStrategy strategy) {
    if (!std::in_range<I>(rows) || !std::in_range<I>(columns)) throw std::length_error("CSR size exceeds index range");
    const std::vector<I> offsets(rows + 1, 0);
    return CsrMatrix<T, I>(rows, columns, offsets, {}, {}, strategy);
}

}  // namespace detail

// ------------------------------------------------------------ vectors

template <Scalar T>
Vector<T> scale(const Vector<T>& vector, std::type_identity_t<T> factor) {
    const int count = detail::librarySize(vector.size());
    detail::OperationScope scope(vector.deviceId());
    Vector<T> result(vector.size());
    result.copyDataFrom(vector);
    if (count) {
        detail::Context ctx(vector.deviceId());
        detail::check(detail::abiScal(ctx.get(), count, factor, result.data()), "scale");
        ctx.sync();
    }
    result.changeMemoryType(vector.memoryStrategy());
    return result;
}

template <Scalar T>
Vector<T> add(const Vector<T>& x, const Vector<T>& y, std::type_identity_t<T> beta = T{1}) {
    if (x.size() != y.size()) throw std::invalid_argument("Vector sizes must match");
    const int count = detail::librarySize(x.size());
    const int device = detail::commonDevice({x.deviceId(), y.deviceId()});
    detail::OperationScope scope(device);
    Vector<T> result(x.size());
    result.copyDataFrom(x);
    if (count) {
        auto input = detail::deviceCopy(y.values());
        detail::Context ctx(device);
        detail::check(detail::abiAxpy(ctx.get(), count, beta, input.data(), result.data()), "add");
        ctx.sync();
    }
    result.changeMemoryType(x.memoryStrategy());
    return result;
}

template <Scalar T>
Vector<T> subtract(const Vector<T>& x, const Vector<T>& y) {
    return add(x, y, T{-1});
}

template <Scalar T>
Vector<T> axpy(std::type_identity_t<T> alpha, const Vector<T>& x, const Vector<T>& y) {
    return add(y, x, alpha);
}

template <Scalar T>
T dot(const Vector<T>& x, const Vector<T>& y) {
    if (x.size() != y.size()) throw std::invalid_argument("Vector sizes must match");
    const int count = detail::librarySize(x.size());
    const int device = detail::commonDevice({x.deviceId(), y.deviceId()});
    if (!count) return T{};
    detail::OperationScope scope(device);
    auto left = detail::deviceCopy(x.values()), right = detail::deviceCopy(y.values());
    detail::Context ctx(device);
    T r{};
    detail::check(detail::abiDot(ctx.get(), count, left.data(), right.data(), &r), "dot");
    return r;
}

template <Scalar T>
T multiply(const Vector<T>& x, const Vector<T>& y) {
    return dot(x, y);
}

template <Scalar T>
T norm2(const Vector<T>& vector) {
    const int count = detail::librarySize(vector.size());
    if (!count) return T{};
    detail::OperationScope scope(vector.deviceId());
    auto input = detail::deviceCopy(vector.values());
    detail::Context ctx(vector.deviceId());
    T r{};
    detail::check(detail::abiNrm2(ctx.get(), count, input.data(), &r), "norm2");
    return r;
}

template <Scalar T>
T norm1(const Vector<T>& vector) {
    const int count = detail::librarySize(vector.size());
    if (!count) return T{};
    detail::OperationScope scope(vector.deviceId());
    auto input = detail::deviceCopy(vector.values());
    detail::Context ctx(vector.deviceId());
    T r{};
    detail::check(detail::abiAsum(ctx.get(), count, input.data(), &r), "norm1");
    return r;
}

// ------------------------------------------------------- dense matrices

template <Scalar T>
DenseMatrix<T> scale(const DenseMatrix<T>& matrix, std::type_identity_t<T> factor) {
    const int count = detail::librarySize(matrix.values().size());
    detail::OperationScope scope(matrix.values().deviceId());
    DenseMatrix<T> result(matrix.rowCount(), matrix.columnCount());
    result.copyValuesFrom(matrix.values());
    if (count) {
        detail::Context ctx(matrix.values().deviceId());
        detail::check(detail::abiScal(ctx.get(), count, factor, result.deviceData()), "scale");
        ctx.sync();
    }
    result.changeMemoryType(matrix.values().memoryStrategy());
    return result;
}

template <Scalar T>
DenseMatrix<T> add(const DenseMatrix<T>& a, const DenseMatrix<T>& b, std::type_identity_t<T> beta = T{1}) {
    if (a.rowCount() != b.rowCount() || a.columnCount() != b.columnCount())
        throw std::invalid_argument("Dense matrix shapes must match");
    const int count = detail::librarySize(a.values().size());
    const int device = detail::commonDevice({a.values().deviceId(), b.values().deviceId()});
    detail::OperationScope scope(device);
    DenseMatrix<T> result(a.rowCount(), a.columnCount());
    result.copyValuesFrom(a.values());
    if (count) {
        auto input = detail::deviceCopy(b.values());
        detail::Context ctx(device);
        detail::check(detail::abiAxpy(ctx.get(), count, beta, input.data(), result.deviceData()), "add");
        ctx.sync();
    }
    result.changeMemoryType(a.values().memoryStrategy());
    return result;
}

template <Scalar T>
DenseMatrix<T> subtract(const DenseMatrix<T>& a, const DenseMatrix<T>& b) {
    return add(a, b, T{-1});
}

template <Scalar T>
DenseMatrix<T> transpose(const DenseMatrix<T>& matrix) {
    const int rows = detail::librarySize(matrix.rowCount()), columns = detail::librarySize(matrix.columnCount());
    detail::OperationScope scope(matrix.values().deviceId());
    DenseMatrix<T> result(matrix.columnCount(), matrix.rowCount());
    if (rows && columns) {
        auto input = detail::deviceCopy(matrix.values());
        detail::Context ctx(matrix.values().deviceId());
        detail::check(detail::abiTranspose(ctx.get(), rows, columns, input.data(), result.deviceData()), "transpose");
        ctx.sync();
    }
    result.changeMemoryType(matrix.values().memoryStrategy());
    return result;
}

template <Scalar T>
Vector<T> multiply(const DenseMatrix<T>& matrix, const Vector<T>& vector) {
    if (matrix.columnCount() != vector.size()) throw std::invalid_argument("Matrix-vector inner dimensions must match");
    const int rows = detail::librarySize(matrix.rowCount()), columns = detail::librarySize(matrix.columnCount());
    const int device = detail::commonDevice({matrix.values().deviceId(), vector.deviceId()});
    detail::OperationScope scope(device);
    Vector<T> result(matrix.rowCount());
    if (rows) {
        auto a = detail::deviceCopy(matrix.values());
        auto x = detail::deviceCopy(vector.values());
        detail::Context ctx(device);
        detail::check(detail::abiGemv(ctx.get(), rows, columns, a.data(), x.data(), result.data()), "gemv");
        ctx.sync();
    }
    result.changeMemoryType(matrix.values().memoryStrategy());
    return result;
}

template <Scalar T>
DenseMatrix<T> multiply(const DenseMatrix<T>& a, const DenseMatrix<T>& b) {
    if (a.columnCount() != b.rowCount()) throw std::invalid_argument("Matrix inner dimensions must match");
    const int rows = detail::librarySize(a.rowCount()), columns = detail::librarySize(b.columnCount()),
              inner = detail::librarySize(a.columnCount());
    const int device = detail::commonDevice({a.values().deviceId(), b.values().deviceId()});
    detail::OperationScope scope(device);
    DenseMatrix<T> result(a.rowCount(), b.columnCount());
    if (rows && columns) {
        auto left = detail::deviceCopy(a.values()), right = detail::deviceCopy(b.values());
        detail::Context ctx(device);
        detail::check(detail::abiGemm(ctx.get(), rows, columns, inner, left.data(), right.data(), result.deviceData()),
                      "gemm");
        ctx.sync();
    }
    result.changeMemoryType(a.values().memoryStrategy());
    return result;
}

template <Scalar T>
T frobeniusNorm(const DenseMatrix<T>& matrix) {
    const int count = detail::librarySize(matrix.values().size());
    if (!count) return T{};
    detail::OperationScope scope(matrix.values().deviceId());
    auto input = detail::deviceCopy(matrix.values());
    detail::Context ctx(matrix.values().deviceId());
    T r{};
    detail::check(detail::abiNrm2(ctx.get(), count, input.data(), &r), "norm2");
    return r;
}

// ----------------------------------------------------- sparse matrices

template <Scalar T, CsrIndex I>
CsrMatrix<T, I> scale(const CsrMatrix<T, I>& matrix, std::type_identity_t<T> factor) {
    const int count = detail::librarySize(matrix.nonZeroCount());
    detail::OperationScope scope(matrix.rowOffsets().deviceId());
    auto result = matrix;
    if (count) {
        auto input = detail::deviceCopy(matrix.values());
        detail::Context ctx(matrix.rowOffsets().deviceId());
        detail::check(detail::abiScal(ctx.get(), count, factor, input.data()), "scale");
        ctx.sync();
        result.copyValuesFromHost(input.hostCopy());
    }
    return result;
}

template <Scalar T, CsrIndex I>
Vector<T> multiply(const CsrMatrix<T, I>& matrix, const Vector<T>& vector) {
    if (matrix.columnCount() != vector.size())
        throw std::invalid_argument("Sparse matrix-vector inner dimensions must match");
    const int device = detail::commonDevice({matrix.rowOffsets().deviceId(), vector.deviceId()});
    detail::OperationScope scope(device);
    Vector<T> result(matrix.rowCount());
    detail::Context ctx(device);
    if (matrix.nonZeroCount() && matrix.rowCount()) {
        auto a = detail::toNative(ctx, matrix, false);  // rows summed in stored order
        Buffer<double> x(vector.size()), y(matrix.rowCount());
        auto xin = detail::deviceCopy(vector.values());
        if constexpr (std::same_as<T, float>)
            detail::check(tf_convert_f32_f64(ctx.get(), x.size(), xin.data(), x.data()), "value conversion");
        else
            x.copyDataFrom(xin);
        detail::check(tf_spmv_csr_f64(ctx.get(), a.rows, a.nnz, a.rp.data(), a.ci.data(), a.v.data(), x.data(),
                                      y.data()),
                      "spmv");
        // the library stream finishes before a plain copy reads its output
        ctx.sync();
        if constexpr (std::same_as<T, float>)
            detail::check(tf_convert_f64_f32(ctx.get(), y.size(), y.data(), result.data()), "value conversion");
        else
            result.copyDataFrom(y);
        ctx.sync();
    }
    else if (result.size()) {
        std::vector<T> zeros(result.size());
        result.copyFromHost(zeros);
    }
    result.changeMemoryType(matrix.values().memoryStrategy());
    return result;
}

template <Scalar T, CsrIndex I>
DenseMatrix<T> multiply(const CsrMatrix<T, I>& matrix, const DenseMatrix<T>& dense) {
    if (matrix.columnCount() != dense.rowCount()) throw std::invalid_argument("Sparse-dense inner dimensions must match");
    const int device = detail::commonDevice({matrix.rowOffsets().deviceId(), dense.values().deviceId()});
    detail::OperationScope scope(device);
    DenseMatrix<T> result(matrix.rowCount(), dense.columnCount());
    const int k = detail::librarySize(dense.columnCount());
    detail::Context ctx(device);
    if (matrix.nonZeroCount() && k && matrix.rowCount()) {
        auto a = detail::toNative(ctx, matrix, false);
        Buffer<double> X(dense.values().size()), Y(result.values().size());
        auto xin = detail::deviceCopy(dense.values());
        if constexpr (std::same_as<T, float>)
            detail::check(tf_convert_f32_f64(ctx.get(), X.size(), xin.data(), X.data()), "value conversion");
        else
            X.copyDataFrom(xin);
        detail::check(tf_spmm_csr_f64(ctx.get(), a.rows, a.nnz, a.rp.data(), a.ci.data(), a.v.data(), k, X.data(),
                                      Y.data()),
                      "spmm");
        ctx.sync();
        if constexpr (std::same_as<T, float>)
            detail::check(tf_convert_f64_f32(ctx.get(), Y.size(), Y.data(), result.deviceData()), "value conversion");
        else
            result.copyValuesFrom(Y);
        ctx.sync();
    }
    else if (result.values().size()) {
        std::vector<T> zeros(result.values().size());
        result.copyFromHost(zeros);
    }
    result.changeMemoryType(matrix.values().memoryStrategy());
    return result;
}

template <Scalar T, CsrIndex I>
CsrMatrix<T, I> add(const CsrMatrix<T, I>& a, const CsrMatrix<T, I>& b, std::type_identity_t<T> beta = T{1}) {
    if (a.rowCount() != b.rowCount() || a.columnCount() != b.columnCount())
        throw std::invalid_argument("CSR shapes must match");
    detail::librarySize(a.rowCount());
    detail::librarySize(a.columnCount());
    detail::librarySize(a.nonZeroCount());
    detail::librarySize(b.nonZeroCount());
    const int device = detail::commonDevice({a.rowOffsets().deviceId(), b.rowOffsets().deviceId()});
    detail::OperationScope scope(device);
    if ((!a.nonZeroCount() && !b.nonZeroCount()) || !a.rowCount() || !a.columnCount())
        return detail::emptySparse<T, I>(a.rowCount(), a.columnCount(), a.values().memoryStrategy());
    detail::Context ctx(device);
    auto l = detail::toNative(ctx, a, false), r = detail::toNative(ctx, b, false);
    Buffer<int> rp(a.rowCount() + 1);
    int* ci = nullptr;
    double* v = nullptr;
    int64_t nnz = 0;
    detail::check(tf_csr_add_f64(ctx.get(), l.rows, int(a.columnCount()), l.nnz, l.rp.data(), l.ci.data(), l.v.data(),
                                 r.nnz, r.rp.data(), r.ci.data(), r.v.data(), double(beta), rp.data(), &ci, &v, &nnz),
                  "sparse add");
    return detail::fromAbi<T, I>(ctx, a.rowCount(), a.columnCount(), rp, ci, v, nnz, a.values().memoryStrategy());
}

template <Scalar T, CsrIndex I>
CsrMatrix<T, I> subtract(const CsrMatrix<T, I>& a, const CsrMatrix<T, I>& b) {
    return add(a, b, T{-1});
}

template <Scalar T, CsrIndex I>
CsrMatrix<T, I> transpose(const CsrMatrix<T, I>& matrix) {
    detail::librarySize(matrix.rowCount());
    detail::librarySize(matrix.columnCount());
    detail::librarySize(matrix.nonZeroCount());
    detail::OperationScope scope(matrix.rowOffsets().deviceId());
    if (!matrix.nonZeroCount() || !matrix.rowCount() || !matrix.columnCount())
        return detail::emptySparse<T, I>(matrix.columnCount(), matrix.rowCount(), matrix.values().memoryStrategy());
    detail::Context ctx(matrix.rowOffsets().deviceId());
    auto a = detail::toNative(ctx, matrix, false);
    Buffer<int> rp(matrix.columnCount() + 1);
    int* ci = nullptr;
    double* v = nullptr;
    detail::check(tf_csr_transpose_f64(ctx.get(), a.rows, int(matrix.columnCount()), a.nnz, a.rp.data(), a.ci.data(),
                                       a.v.data(), rp.data(), &ci, &v),
                  "sparse transpose");
    std::vector<int> hrp = rp.hostCopy();
    return detail::fromAbi<T, I>(ctx, matrix.columnCount(), matrix.rowCount(), rp, ci, v, int64_t(hrp.back()),
                                 matrix.values().memoryStrategy());
}

template <Scalar T, CsrIndex I>
CsrMatrix<T, I> multiply(const CsrMatrix<T, I>& a, const CsrMatrix<T, I>& b) {
    if (a.columnCount() != b.rowCount()) throw std::invalid_argument("Sparse matrix inner dimensions must match");
    detail::librarySize(a.rowCount());
    detail::librarySize(a.columnCount());
    detail::librarySize(b.columnCount());
    detail::librarySize(a.nonZeroCount());
    detail::librarySize(b.nonZeroCount());
    const int device = detail::commonDevice({a.rowOffsets().deviceId(), b.rowOffsets().deviceId()});
    detail::OperationScope scope(device);
    if (!a.nonZeroCount() || !b.nonZeroCount() || !a.rowCount() || !b.columnCount())
        return detail::emptySparse<T, I>(a.rowCount(), b.columnCount(), a.values().memoryStrategy());
    detail::Context ctx(device);
    auto l = detail::toNative(ctx, a, false), r = detail::toNative(ctx, b, false);
    Buffer<int> rp(a.rowCount() + 1);
    int* ci = nullptr;
    double* v = nullptr;
    int64_t nnz = 0;
    detail::check(tf_csr_multiply_f64(ctx.get(), l.rows, int(a.columnCount()), int(b.columnCount()), l.nnz,
                                      l.rp.data(), l.ci.data(), l.v.data(), r.nnz, r.rp.data(), r.ci.data(),
                                      r.v.data(), rp.data(), &ci, &v, &nnz),
                  "sparse product");
    return detail::fromAbi<T, I>(ctx, a.rowCount(), b.columnCount(), rp, ci, v, nnz, a.values().memoryStrategy());
}

template <Scalar T, CsrIndex I>
T frobeniusNorm(const CsrMatrix<T, I>& matrix) {
    if (!matrix.nonZeroCount() || !matrix.rowCount() || !matrix.columnCount()) return T{};
    // duplicates are summed first (the canonical matrix's norm, Operations.cuh:516-523)
    const auto merged = add(matrix, detail::emptySparse<T, I>(matrix.rowCount(), matrix.columnCount(), GlobalMemory{}));
    detail::OperationScope scope(matrix.rowOffsets().deviceId());
    detail::Context ctx(matrix.rowOffsets().deviceId());
    auto values = detail::deviceCopy(merged.values());
    T r{};
    detail::check(detail::abiNrm2(ctx.get(), detail::librarySize(values.size()), values.data(), &r), "norm2");
    return r;
}

}  // namespace thermofem::linalg

namespace thermofem {

// Operators.cuh:8-125: objects that linalg::scale accepts.
template <class T>
concept LinearObject = requires(const T& value) {
    typename T::ValueType;
    requires linalg::Scalar<typename T::ValueType>;
    { linalg::scale(value, typename T::ValueType{1}) } -> std::same_as<T>;
};

template <LinearObject Object>
Object operator+(const Object& left, const Object& right) {
    return linalg::add(left, right);
}
template <LinearObject Object>
Object operator-(const Object& left, const Object& right) {
    return linalg::subtract(left, right);
}
template <LinearObject Object>
Object operator+(const Object& value) {
    return value;
}
template <LinearObject Object>
Object operator-(const Object& value) {
    return linalg::scale(value, typename Object::ValueType{-1});
}
template <LinearObject Object>
Object operator*(const Object& value, typename Object::ValueType factor) {
    return linalg::scale(value, factor);
}
template <LinearObject Object>
Object operator*(typename Object::ValueType factor, const Object& value) {
    return linalg::scale(value, factor);
}
template <LinearObject Object>
Object operator/(const Object& value, typename Object::ValueType divisor) {
    if (divisor == typename Object::ValueType{}) throw std::domain_error("Division by zero");
    return linalg::scale(value, typename Object::ValueType{1} / divisor);
}
template <LinearObject Left, LinearObject Right>
    requires requires(const Left& left, const Right& right) { linalg::multiply(left, right); }
auto operator*(const Left& left, const Right& right) {
    return linalg::multiply(left, right);
}
template <LinearObject Object>
Object& operator+=(Object& left, const Object& right) {
    left = left + right;
    return left;
}
template <LinearObject Object>
Object& operator-=(Object& left, const Object& right) {
    left = left - right;
    return left;
}
template <LinearObject Object>
Object& operator*=(Object& value, typename Object::ValueType factor) {
    value = value * factor;
    return value;
}
template <LinearObject Object>
Object& operator/=(Object& value, typename Object::ValueType divisor) {
    value = value / divisor;
    return value;
}
template <LinearObject Object>
    requires requires(const Object& left, const Object& right) {
        { linalg::multiply(left, right) } -> std::same_as<Object>;
    }
Object& operator*=(Object& left, const Object& right) {
    left = left * right;
    return left;
}

}  // namespace thermofem
