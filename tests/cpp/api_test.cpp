// Tests of the C++20 drop-in API over libthermofem_hip.so.  Each case follows
// one of the reference's GoogleTest cases (file:line cited per case) with the
// same inputs and expectations; typed cases run for float and double.
// Usage: api_test [--cpu-only] [name-filter]
#include <thermofem/thermofem.hpp>

#include <cstdint>
#include <limits>
#include <sstream>

#include "harness.hpp"

using namespace thermofem;

namespace {

template <class T, class I = int>
CsrMatrix<T, I> poisson(int n, bool nonsymmetric = false, T diagonal = T(2)) {
    std::vector<I> offsets{0}, columns;
    std::vector<T> values;
    for (int row = 0; row < n; ++row) {
        if (row) {
            columns.push_back(I(row - 1));
            values.push_back(T(-1));
        }
        columns.push_back(I(row));
        values.push_back(nonsymmetric ? T(3) : diagonal);
        if (row + 1 < n) {
            columns.push_back(I(row + 1));
            values.push_back(nonsymmetric ? T(-0.5) : T(-1));
        }
        offsets.push_back(I(values.size()));
    }
    return CsrMatrix<T, I>(std::size_t(n), std::size_t(n), offsets, columns, values);
}

template <class T>
Vector<T> vec(const std::vector<T>& v, MemoryStrategy s = GlobalMemory{}) {
    return Vector<T>(std::span<const T>(v), s);
}

std::vector<MemoryStrategy> strategies() { return {GlobalMemory{}, PinnedMemory{}, MappedMemory{}, ManagedMemory{}}; }

}  // namespace

// ---- options and host-side validation (no device) ----

// cg_test.cpp:186-207, solver_test.cpp:310-312, :176-178, :76-79
TEST_CASE(optionsRejectedWithoutDevice, false) {
    CgOptions<double> cg;
    cg.maxIterations_ = -1;
    EXPECT_THROW((Cg<double>(cg)), std::invalid_argument);
    cg = {};
    cg.residualRefreshInterval_ = 0;
    EXPECT_THROW((Cg<double>(cg)), std::invalid_argument);
    for (double value : {-1.0, 1.0, std::numeric_limits<double>::infinity(), std::numeric_limits<double>::quiet_NaN()}) {
        cg = {};
        cg.relativeTolerance_ = value;
        EXPECT_THROW((Cg<double>(cg)), std::invalid_argument);
    }
    for (double value : {-1.0, std::numeric_limits<double>::infinity(), std::numeric_limits<double>::quiet_NaN()}) {
        cg = {};
        cg.absoluteTolerance_ = value;
        EXPECT_THROW((Cg<double>(cg)), std::invalid_argument);
    }
    BiCgStabOptions<double> bi;
    bi.relativeTolerance_ = -1;
    EXPECT_THROW((BiCgStab<double>(bi)), std::invalid_argument);
    AmgOptions<double> amg;
    amg.maxLevels_ = 0;
    EXPECT_THROW((Amg<double>(amg)), std::invalid_argument);
    amg = {};
    amg.coarseRows_ = 1025;
    EXPECT_THROW((Amg<double>(amg)), std::invalid_argument);
    amg = {};
    amg.jacobiWeight_ = 2;
    EXPECT_THROW((Amg<double>(amg)), std::invalid_argument);
    EXPECT_THROW((Jacobi<double>(0)), std::invalid_argument);
    EXPECT_THROW((Jacobi<float>(1, std::numeric_limits<float>::quiet_NaN())), std::invalid_argument);
}

// data_structure_test.cpp:147-191
TEST_CASE(csrHostValidationWithoutDevice, false) {
    const std::vector<int> offsets{0, 1, 2}, columns{0, 1};
    const std::vector<double> values{1, 2};
    EXPECT_THROW((CsrMatrix<double>(2, 2, {}, columns, values)), std::invalid_argument);
    EXPECT_THROW((CsrMatrix<double>(3, 2, offsets, columns, values)), std::invalid_argument);
    EXPECT_THROW((CsrMatrix<double>(2, 2, offsets, columns, {})), std::invalid_argument);
    const std::vector<int> badStart{1, 1, 2}, badEnd{0, 1, 3};
    EXPECT_THROW((CsrMatrix<double>(2, 2, badStart, columns, values)), std::invalid_argument);
    EXPECT_THROW((CsrMatrix<double>(2, 2, badEnd, columns, values)), std::invalid_argument);
    const std::vector<int> decreasing{0, 2, 1, 2}, negative{0, -1, 1, 2}, outOfBounds{0, 1, 3, 2};
    EXPECT_THROW((CsrMatrix<double>(3, 2, decreasing, columns, values)), std::invalid_argument);
    EXPECT_THROW((CsrMatrix<double>(3, 2, negative, columns, values)), std::invalid_argument);
    EXPECT_THROW((CsrMatrix<double>(3, 2, outOfBounds, columns, values)), std::invalid_argument);
    const std::vector<int> badColumns{0, 2};
    EXPECT_THROW((CsrMatrix<double>(2, 2, offsets, badColumns, values)), std::invalid_argument);
    EXPECT_THROW((CsrMatrix<double>(std::numeric_limits<std::size_t>::max(), 2, {}, {}, {})), std::length_error);
    EXPECT_THROW((Buffer<double>(std::numeric_limits<std::size_t>::max() / 4)), std::length_error);
}

// ---- Matrix Market reader (no device) ----

// matrix_market_test.cpp:8-16
TEST_CASE(matrixMarketGroupsUnsortedCoordinatesAndPreservesDuplicates, false) {
    std::istringstream input(
        "%%MatrixMarket matrix coordinate real general\n% comment\n3 3 4\n3 1 2\n1 2 4\n1 2 -1\n3 3 5\n");
    const auto m = io::readMatrixMarket(input);
    EXPECT_TRUE(m.offsets_ == (std::vector<int>{0, 2, 2, 4}));
    EXPECT_TRUE(m.indices_ == (std::vector<int>{1, 1, 0, 2}));
    EXPECT_TRUE(m.values_ == (std::vector<double>{4, -1, 2, 5}));
}

// matrix_market_test.cpp:17-25
TEST_CASE(matrixMarketExpandsSymmetricTriangleAndAcceptsEmpty, false) {
    std::istringstream input("%%MatrixMarket matrix coordinate integer symmetric\n2 2 3\n1 1 2\n2 1 -1\n2 2 3\n");
    const auto m = io::readMatrixMarket(input);
    EXPECT_TRUE(m.offsets_ == (std::vector<int>{0, 2, 4}));
    EXPECT_TRUE(m.indices_ == (std::vector<int>{0, 1, 0, 1}));
    EXPECT_TRUE(m.values_ == (std::vector<double>{2, -1, -1, 3}));
    std::istringstream empty("%%MatrixMarket matrix coordinate real general\n0 0 0\n");
    EXPECT_TRUE(io::readMatrixMarket(empty).offsets_ == (std::vector<int>{0}));
}

// matrix_market_test.cpp:26-44
TEST_CASE(matrixMarketRejectsMalformedUnsupportedAndNonfiniteInputs, false) {
    const std::vector<std::string> inputs{"",
                                          "%%MatrixMarket matrix array real general\n2 2\n",
                                          "%%MatrixMarket matrix coordinate complex general\n2 2 0\n",
                                          "%%MatrixMarket matrix coordinate real symmetric\n2 3 0\n",
                                          "%%MatrixMarket matrix coordinate real general\n-1 2 0\n",
                                          "%%MatrixMarket matrix coordinate real general\n2147483648 2 0\n",
                                          "%%MatrixMarket matrix coordinate real general\n2 2 1\n",
                                          "%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 2\n",
                                          "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 3 2\n",
                                          "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 nan\n",
                                          "%%MatrixMarket matrix coordinate integer general\n2 2 1\n1 1 1.5\n",
                                          "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 2 extra\n",
                                          "%%MatrixMarket matrix coordinate real general\n2 2 0\n1 1 2\n",
                                          "%%MatrixMarket matrix coordinate real symmetric\n2 2 2\n1 2 1\n2 1 1\n"};
    for (const auto& text : inputs) {
        std::istringstream input(text);
        EXPECT_THROW(io::readMatrixMarket(input), std::runtime_error);
    }
}

// ---- memory and matrices ----

// data_structure_test.cpp:219-284, :296-345
TEST_CASE(bufferStrategiesCopiesAndViews, true) {
    const std::vector<double> expected{1, 2, 3};
    for (const auto& sourceStrategy : strategies()) {
        for (const auto& targetStrategy : strategies()) {
            Buffer<double> source(std::span<const double>(expected), sourceStrategy);
            Buffer<double> target(3, targetStrategy);
            double* address = target.data();
            target.copyDataFrom(source);
            EXPECT_EQ(target.data(), address);
            EXPECT_EQ(int(target.memoryType()), int(targetStrategy.index()));
            EXPECT_TRUE(target.hostCopy() == expected);
            target = source;
            EXPECT_EQ(int(target.memoryType()), int(sourceStrategy.index()));
            EXPECT_NE(target.data(), source.data());
            EXPECT_TRUE(target.hostCopy() == expected);
            Buffer<double> wrong(2, targetStrategy);
            EXPECT_THROW(wrong.copyDataFrom(source), std::invalid_argument);
            source.changeMemoryType(targetStrategy);
            EXPECT_EQ(int(source.memoryType()), int(targetStrategy.index()));
            EXPECT_TRUE(source.hostCopy() == expected);
            if (source.memoryType() == MemoryType::Global)
                EXPECT_THROW((void)source.hostData(), std::logic_error);
            else
                EXPECT_EQ(source.hostData()[2], 3.0);
            if (source.memoryType() == MemoryType::Pinned)
                EXPECT_THROW((void)source.deviceData(), std::logic_error);
            else
                EXPECT_NE(source.deviceData(), nullptr);
            double* before = source.data();
            Buffer<double> moved(std::move(source));
            EXPECT_EQ(moved.data(), before);
            EXPECT_TRUE(source.empty());
            Vector<double> v(3, targetStrategy);
            v.copyDataFrom(Vector<double>(std::span<const double>(expected), sourceStrategy));
            EXPECT_TRUE(v.hostCopy() == expected);
        }
    }
    Buffer<double> empty(MappedMemory{});
    Buffer<double> other(std::move(empty));
    EXPECT_EQ(int(other.memoryType()), int(MemoryType::Mapped));
    EXPECT_EQ(other.deviceData(), nullptr);
}

// data_structure_test.cpp:381-416, amg_native_test.cpp:115-131
TEST_CASE(csrStorageCanonicalAndDeviceValidation, true) {
    const std::vector<int> offsets{0, 2, 3}, columns{0, 1, 1};
    const std::vector<double> values{1, 2, 3};
    CsrMatrix<double> source(2, 2, offsets, columns, values, PinnedMemory{});
    auto copy = source;
    copy.changeMemoryType(GlobalMemory{});
    EXPECT_EQ(int(copy.rowOffsets().memoryType()), int(MemoryType::Global));
    EXPECT_TRUE(copy.rowOffsets().hostCopy() == offsets);
    EXPECT_TRUE(copy.columnIndices().hostCopy() == columns);
    EXPECT_TRUE(copy.values().hostCopy() == values);
    const std::vector<int> rp{0, 3, 6}, ci{1, 0, 0, 1, 0, 1};
    const std::vector<double> dup{-1, 1, 1, 1, -1, 1};
    CsrMatrix<double> a(2, 2, rp, ci, dup);
    multigrid::GpuAggregation operations;
    const auto canonical = operations.canonical(a);
    EXPECT_TRUE(canonical.values().hostCopy() == (std::vector<double>{2, -1, -1, 2}));
    const std::vector<int> badOffsets{0, 7, 6};
    Buffer<int> o{std::span<const int>(badOffsets)}, i{std::span<const int>(ci)};
    Buffer<double> e{std::span<const double>(dup)};
    EXPECT_THROW(CsrMatrix<double>::fromBuffers(2, 2, std::move(o), std::move(i), std::move(e)), std::invalid_argument);
    EXPECT_EQ(o.size(), 3u);
    EXPECT_EQ(e.size(), 6u);
    Buffer<int> o2{std::span<const int>(rp)}, i2{std::span<const int>(ci)};
    Buffer<double> e2{std::span<const double>(dup)};
    auto owned = CsrMatrix<double>::fromBuffers(2, 2, std::move(o2), std::move(i2), std::move(e2));
    EXPECT_EQ(owned.nonZeroCount(), 6u);
    EXPECT_TRUE(o2.empty());
    EXPECT_THROW(operations.strength(a, -.1), std::invalid_argument);
    const auto graph = operations.strength(canonical, .25);
    EXPECT_THROW(operations.aggregate(graph, 0, true), std::invalid_argument);
}

// ---- linear algebra and multigrid operations ----

// multigrid_operations_test.cpp:33-64
TYPED_CASE(diagonalResidualAndGalerkinMatchHandCalculation) {
    const std::vector<int> offsets{0, 2, 5, 8, 10}, columns{0, 1, 0, 1, 2, 1, 2, 3, 2, 3};
    const std::vector<T> values{2, -1, -1, 2, -1, -1, 2, -1, -1, 2};
    const CsrMatrix<T> a(4, 4, offsets, columns, values);
    const auto inverse = linalg::detail::hostCopy(linalg::inverseDiagonal(a));
    for (auto value : inverse) EXPECT_EQ(value, T(0.5));
    multigrid::GpuAggregation setup;
    const auto native = multigrid::toNativeMatrix(a);
    const auto graph = setup.strength(native, .25);
    const auto transfer = setup.aggregate(graph, 8, true);
    const auto coarse = setup.galerkin(native, transfer);
    ASSERT_EQ(coarse.rowCount(), 1u);
    EXPECT_TRUE(linalg::detail::hostCopy(coarse.values()) == (std::vector<double>{2}));
    const std::vector<T> ones(4, 1), zeros(4, 0);
    Vector<T> x = vec(ones), rhs = vec(zeros), residual(4);
    linalg::IterationContext<T> context(a);
    context.residual(rhs, x, residual);
    EXPECT_TRUE(linalg::detail::hostCopy(residual) == (std::vector<T>{-1, 0, 0, -1}));
    Vector<T> diagonal = vec(inverse);
    context.clear(x);
    context.clear(residual);
    rhs.copyFromHost(ones);
    context.jacobi(T(1), diagonal, rhs, residual, x);
    for (T value : linalg::detail::hostCopy(x)) EXPECT_EQ(value, T(0.5));
    EXPECT_THROW(context.residual(rhs, x, rhs), std::invalid_argument);
    EXPECT_THROW(setup.strength(native, -1), std::invalid_argument);
}

// multigrid_operations_test.cpp:66-83
TYPED_CASE(coarseLuPivotAndDuplicateDiagonal) {
    const std::vector<int> offsets{0, 2, 4}, columns{0, 0, 0, 1};
    const std::vector<T> values{1, 1, 4, 3}, b{2, 7};
    const CsrMatrix<T> a(2, 2, offsets, columns, values);
    const auto diagonal = linalg::detail::hostCopy(linalg::inverseDiagonal(a));
    EXPECT_NEAR(diagonal[0], 0.5, 1e-6);
    EXPECT_NEAR(diagonal[1], 1.0 / 3, 1e-6);
    linalg::CoarseLu<T> lu(a);
    Vector<T> rhs = vec(b), x(2);
    lu.apply(rhs, x);
    const auto result = linalg::detail::hostCopy(x);
    EXPECT_NEAR(result[0], 1, 1e-6);
    EXPECT_NEAR(result[1], 1, 1e-6);
    Vector<T> wrong(3);
    EXPECT_THROW(lu.apply(wrong, x), std::invalid_argument);
    const std::vector<T> invalid{0, 0, 4, 3};
    const CsrMatrix<T> bad(2, 2, offsets, columns, invalid);
    EXPECT_THROW(linalg::inverseDiagonal(bad), std::invalid_argument);
    EXPECT_THROW((linalg::CoarseLu<T>(bad)), std::invalid_argument);
}

// linear_algebra_test.cpp:112-116 (dot -24, norm sqrt 14) on the iteration context
TYPED_CASE(iterationContextReductions) {
    auto a = poisson<T>(3);
    linalg::IterationContext<T> context(a);
    Vector<T> x = vec(std::vector<T>{1, -2, 3}), y = vec(std::vector<T>{4, 5, -6});
    EXPECT_NEAR(context.dot(x, y), -24, 1e-5);
    EXPECT_NEAR(context.norm(x), std::sqrt(14.0), 1e-5);
    context.axpy(T(2), x, y);
    EXPECT_TRUE(linalg::detail::hostCopy(y) == (std::vector<T>{6, 1, 0}));
    context.scale(T(-1), y);
    EXPECT_TRUE(linalg::detail::hostCopy(y) == (std::vector<T>{-6, -1, -0.0}));
    Vector<T> product(3);
    context.multiply(x, product);
    EXPECT_TRUE(linalg::detail::hostCopy(product) == (std::vector<T>{4, -8, 8}));
}

// solver_test.cpp:96-109
TYPED_CASE(contextOwnsSnapshotAndSupportsWideIndices) {
    auto a = poisson<T, std::int64_t>(2);
    linalg::IterationContext<T, std::int64_t> context(a);
    const std::vector<T> zeros(4, 0), input{1, 2};
    a.copyValuesFromHost(zeros);
    const Vector<T> x = vec(input);
    Vector<T> y(2);
    context.multiply(x, y);
    const auto output = linalg::detail::hostCopy(y);
    EXPECT_NEAR(output[0], 0, 1e-6);
    EXPECT_NEAR(output[1], 3, 1e-6);
    EXPECT_THROW(context.multiply(x, const_cast<Vector<T>&>(x)), std::invalid_argument);
}

// amg_native_test.cpp:133-153
TEST_CASE(gpuCoarsePivotScaleAndSingularDetection, true) {
    const std::vector<int> rp{0, 2, 4}, ci{0, 1, 0, 1};
    for (double scale : {1e-100, 1.0, 1e100}) {
        const std::vector<double> v{scale, 2 * scale, 4 * scale, 3 * scale}, b{5 * scale, 10 * scale};
        CsrMatrix<double> a(2, 2, rp, ci, v);
        auto factor = multigrid::factorCoarse(a);
        ASSERT_TRUE(factor.factored_);
        Vector<double> rhs = vec(b), x(2);
        multigrid::solveCoarse(factor, rhs, x);
        const auto actual = linalg::detail::hostCopy(x);
        EXPECT_NEAR(actual[0], 1, 1e-12);
        EXPECT_NEAR(actual[1], 2, 1e-12);
        multigrid::solveCoarse(factor, rhs, rhs);
        EXPECT_TRUE(linalg::detail::hostCopy(rhs) == actual);
    }
    const std::vector<double> singular{1, 1, 1, 1};
    CsrMatrix<double> a(2, 2, rp, ci, singular);
    const auto factor = multigrid::factorCoarse(a);
    EXPECT_TRUE(factor.singular_);
    EXPECT_FALSE(factor.factored_);
}

// amg_native_test.cpp:96-107: restricting ones gives aggregate sizes; prolong adds
TEST_CASE(transfersRestrictSizesAndProlongAdds, true) {
    auto a = poisson<double>(61);
    multigrid::GpuAggregation operations;
    const auto native = operations.canonical(a);
    const auto graph = operations.strength(native, .25);
    const auto t = operations.aggregate(graph, 8, true);
    const std::vector<double> ones(61, 1);
    Vector<double> fine = vec(ones), restricted(std::size_t(t.coarseRows_));
    multigrid::restrictAggregates(t, fine, restricted);
    const auto sums = linalg::detail::hostCopy(restricted);
    const auto offsets = t.offsets_.hostCopy(), map = t.aggregate_.hostCopy();
    int histogramTotal = 0;
    for (int k = 0; k < 4; ++k) histogramTotal += t.histogram_[std::size_t(k)];
    EXPECT_EQ(histogramTotal, t.coarseRows_);
    for (int c = 0; c < t.coarseRows_; ++c) EXPECT_EQ(sums[std::size_t(c)], double(offsets[std::size_t(c) + 1] - offsets[std::size_t(c)]));
    multigrid::prolongAggregates(t, restricted, fine);
    const auto actual = linalg::detail::hostCopy(fine);
    for (int i = 0; i < 61; ++i) EXPECT_EQ(actual[std::size_t(i)], 1 + sums[std::size_t(map[std::size_t(i)])]);
    const auto coarse = operations.galerkin(native, t);
    EXPECT_LE(coarse.nonZeroCount(), native.nonZeroCount());
}

// ---- smoothers, preconditioners, solvers ----

// solver_test.cpp:56-74
TYPED_CASE(jacobiMatchesAnalyticSweepsAndPreservesPolicy) {
    const auto a = poisson<T>(2);
    std::unique_ptr<Smoother<T>> smoother = std::make_unique<Jacobi<T>>(2, T(1));
    smoother->setup(a);
    const std::vector<T> b{1, 0}, initial{0, 0};
    Vector<T> rhs = vec(b, PinnedMemory{}), x = vec(initial, PinnedMemory{});
    smoother->apply(rhs, x);
    auto values = linalg::detail::hostCopy(x);
    EXPECT_EQ(int(x.memoryType()), int(MemoryType::Pinned));
    EXPECT_NEAR(values[0], 0.5, 1e-6);
    EXPECT_NEAR(values[1], 0.25, 1e-6);
    rhs.copyFromHost(b);
    smoother->apply(rhs, rhs);
    values = linalg::detail::hostCopy(rhs);
    EXPECT_NEAR(values[0], 0.75, 1e-6);
    EXPECT_NEAR(values[1], 0.25, 1e-6);
}

// solver_test.cpp:75-95
TYPED_CASE(jacobiSetupValidationAndReuse) {
    Jacobi<T> smoother;
    Vector<T> rhs(2), x(2);
    EXPECT_THROW(smoother.apply(rhs, x), std::logic_error);
    smoother.setup(poisson<T>(2));
    const std::vector<int> offsets{0, 1, 2}, columns{1, 0};
    const std::vector<T> values{1, 1};
    const CsrMatrix<T> bad(2, 2, offsets, columns, values);
    EXPECT_THROW(smoother.setup(bad), std::invalid_argument);
    rhs.copyFromHost(std::vector<T>{1, 1});
    x.copyFromHost(std::vector<T>{0, 0});
    EXPECT_NO_THROW(smoother.apply(rhs, x));
    Vector<T> wrong(3);
    EXPECT_THROW(smoother.apply(wrong, x), std::invalid_argument);
    smoother.setup(poisson<T>(3));
    EXPECT_THROW(smoother.apply(rhs, x), std::invalid_argument);
}

// solver_test.cpp:110-148
TYPED_CASE(amgCycleReducesErrorAndIsLinear) {
    const auto a = poisson<T>(127);
    AmgOptions<T> amgOptions;
    amgOptions.coarseRows_ = 8;
    Amg<T> amg(amgOptions);
    Vector<T> uninitialized(127);
    EXPECT_THROW(amg.apply(uninitialized, uninitialized), std::logic_error);
    amg.setup(a);
    ASSERT_GE(amg.levelCount(), 3u);
    const std::vector<T> ones(127, 1), zeros(127, 0);
    Vector<T> rhs = vec(ones), x = vec(zeros);
    linalg::IterationContext<T> context(a);
    amg.apply(rhs, x);
    const auto first = linalg::detail::hostCopy(x);
    T errorSquared{}, exactSquared{};
    for (std::size_t i = 0; i < first.size(); ++i) {
        const T exact = T(i + 1) * T(127 - i) / T(2);
        errorSquared += (first[i] - exact) * (first[i] - exact);
        exactSquared += exact * exact;
    }
    EXPECT_LT(std::sqrt(errorSquared / exactSquared), T(1));
    context.scale(T(2), rhs);
    amg.apply(rhs, x);
    const auto second = linalg::detail::hostCopy(x);
    const T tolerance = std::same_as<T, float> ? T(1e-2) : T(1e-9);
    for (std::size_t i = 0; i < first.size(); ++i) EXPECT_NEAR(second[i], 2 * first[i], tolerance);
    amg.apply(rhs, rhs);
    const auto aliased = linalg::detail::hostCopy(rhs);
    for (std::size_t i = 0; i < first.size(); ++i) EXPECT_NEAR(aliased[i], second[i], tolerance);
}

// solver_test.cpp:149-179
TYPED_CASE(amgCoarsePivotAndFailurePreservesState) {
    const std::vector<int> offsets{0, 2, 4}, columns{0, 1, 0, 1};
    const std::vector<T> values{1, 2, 4, 3}, b{5, 10};
    const CsrMatrix<T> a(2, 2, offsets, columns, values);
    Amg<T> amg;
    amg.setup(a);
    Vector<T> rhs = vec(b, PinnedMemory{}), x(2, PinnedMemory{});
    amg.apply(rhs, x);
    auto result = linalg::detail::hostCopy(x);
    EXPECT_NEAR(result[0], 1, 1e-5);
    EXPECT_NEAR(result[1], 2, 1e-5);
    const std::vector<T> invalid{0, 1, 1, 1};
    EXPECT_THROW(amg.setup(CsrMatrix<T>(2, 2, offsets, columns, invalid)), std::invalid_argument);
    EXPECT_NO_THROW(amg.apply(rhs, x));
    result = linalg::detail::hostCopy(x);
    EXPECT_NEAR(result[0], 1, 1e-5);
    EXPECT_EQ(int(x.memoryType()), int(MemoryType::Pinned));
    const std::vector<T> singular{1, 1, 1, 1};
    EXPECT_NO_THROW(amg.setup(CsrMatrix<T>(2, 2, offsets, columns, singular)));
    EXPECT_EQ(amg.coarseSolver(), std::string("Jacobi (singular LU)"));
    EXPECT_NO_THROW(amg.apply(rhs, x));
    AmgOptions<T> options;
    options.maxLevels_ = 1;
    options.coarseRows_ = 2;
    Amg<T> bounded(options);
    EXPECT_NO_THROW(bounded.setup(poisson<T>(100)));
    EXPECT_EQ(bounded.levelCount(), 1u);
    EXPECT_EQ(bounded.coarseSolver(), std::string("Jacobi (coarsening stopped)"));
    EXPECT_EQ(bounded.stopReason(), std::string("level limit"));
}

// solver_test.cpp:180-213
TYPED_CASE(bicgstabAmgSolvesPoissonAndReportsTrueResidual) {
    const int n = 63;
    const auto a = poisson<T>(n);
    BiCgStabOptions<T> options;
    options.relativeTolerance_ = std::same_as<T, float> ? T(2e-4) : T(1e-9);
    BiCgStab<T> solver(options);
    solver.setup(a);
    const std::vector<T> b(n, 1), initial(n, 0);
    Vector<T> rhs = vec(b, PinnedMemory{}), x = vec(initial, PinnedMemory{});
    auto* address = x.data();
    const auto result = solver.apply(rhs, x);
    ASSERT_TRUE(result.status_ == SolveStatus::Converged);
    EXPECT_LT(result.iterations_, 100);
    EXPECT_EQ(x.data(), address);
    EXPECT_EQ(int(x.memoryType()), int(MemoryType::Pinned));
    const auto actual = linalg::detail::hostCopy(x);
    double residualSquared{};
    for (int i = 0; i < n; ++i) {
        const double exact = double(i + 1) * double(n - i) / 2;
        EXPECT_NEAR(actual[std::size_t(i)], exact, exact * (std::same_as<T, float> ? 1e-3 : 1e-8));
        const double r = 1 - 2 * double(actual[std::size_t(i)]) + (i ? double(actual[std::size_t(i) - 1]) : 0) +
                         (i + 1 < n ? double(actual[std::size_t(i) + 1]) : 0);
        residualSquared += r * r;
    }
    EXPECT_NEAR(result.residualNorm_, std::sqrt(residualSquared), (std::same_as<T, float> ? 1e-3 : 1e-10));
    EXPECT_LE(result.relativeResidual_, options.relativeTolerance_);
    const auto repeated = solver.apply(rhs, x);
    EXPECT_TRUE(repeated.status_ == SolveStatus::Converged);
    EXPECT_EQ(repeated.iterations_, 0);
    rhs.copyFromHost(b);
    EXPECT_TRUE(solver.apply(rhs, rhs).status_ == SolveStatus::Converged);
}

// The hierarchy a device-side BiCGStab builds is visible through the Amg it
// owns (example/solve-matrix.cu:122-140 queries it after solver.setup).
TEST_CASE(solverBuiltHierarchyIsQueryable, true) {
    auto amg = std::make_unique<Amg<double>>(AmgOptions<double>{});
    Amg<double>* hierarchy = amg.get();
    BiCgStab<double> solver({}, std::move(amg));
    solver.setup(poisson<double>(300));
    EXPECT_GE(hierarchy->levelCount(), 2u);
    EXPECT_EQ(hierarchy->stopReason(), std::string("coarse size"));
    EXPECT_EQ(hierarchy->coarseSolver(), std::string("pivoted GPU LU"));
    EXPECT_EQ(hierarchy->levelStatistics().front().rows_, 300u);
    Vector<double> rhs = vec(std::vector<double>(300, 1.0)), z(300);
    hierarchy->apply(rhs, z);
    EXPECT_GT(linalg::detail::hostCopy(z)[150], 0.0);
}

// solver_test.cpp:214-237
TYPED_CASE(bicgstabIdentitySolvesNonsymmetricWithResidualRefresh) {
    const int n = 80;
    const auto a = poisson<T>(n, true);
    std::vector<T> exact(n), b(n), initial(n, T(0.2));
    for (int i = 0; i < n; ++i) exact[std::size_t(i)] = T(1) + T(i % 7) / T(7);
    for (int i = 0; i < n; ++i)
        b[std::size_t(i)] = 3 * exact[std::size_t(i)] - (i ? exact[std::size_t(i) - 1] : 0) -
                            T(0.5) * (i + 1 < n ? exact[std::size_t(i) + 1] : 0);
    BiCgStabOptions<T> options;
    options.residualRefreshInterval_ = 3;
    options.relativeTolerance_ = std::same_as<T, float> ? T(1e-5) : T(1e-11);
    std::unique_ptr<Solver<T>> solver = std::make_unique<BiCgStab<T>>(options, std::make_unique<Identity<T>>());
    solver->setup(a);
    Vector<T> rhs = vec(b), x = vec(initial);
    const auto report = solver->apply(rhs, x);
    ASSERT_TRUE(report.status_ == SolveStatus::Converged);
    const auto actual = linalg::detail::hostCopy(x);
    for (int i = 0; i < n; ++i)
        EXPECT_NEAR(actual[std::size_t(i)], exact[std::size_t(i)], (std::same_as<T, float> ? 1e-4 : 1e-9));
}

// solver_test.cpp:238-254
TYPED_CASE(bicgstabZeroRhsAndInitialGuess) {
    BiCgStabOptions<T> options;
    options.absoluteTolerance_ = std::same_as<T, float> ? T(1e-5) : T(1e-12);
    BiCgStab<T> solver(options, std::make_unique<Identity<T>>());
    solver.setup(poisson<T>(5));
    const std::vector<T> zeros(5, 0), ones(5, 1);
    Vector<T> rhs = vec(zeros), x = vec(zeros);
    auto report = solver.apply(rhs, x);
    EXPECT_TRUE(report.status_ == SolveStatus::Converged);
    EXPECT_EQ(report.iterations_, 0);
    x.copyFromHost(ones);
    report = solver.apply(rhs, x);
    EXPECT_TRUE(report.status_ == SolveStatus::Converged);
    EXPECT_LE(report.residualNorm_, options.absoluteTolerance_);
    EXPECT_EQ(report.relativeResidual_, report.residualNorm_);
}

// solver_test.cpp:255-287
TYPED_CASE(bicgstabReportsLimitBreakdownAndNonfinite) {
    const auto a = poisson<T>(3, true);
    const std::vector<T> b{1, 2, 3}, zeros(3, 0);
    Vector<T> rhs = vec(b), x = vec(zeros);
    BiCgStabOptions<T> options;
    options.maxIterations_ = 0;
    BiCgStab<T> limited(options, std::make_unique<Identity<T>>());
    limited.setup(a);
    auto result = limited.apply(rhs, x);
    EXPECT_TRUE(result.status_ == SolveStatus::MaxIterations);
    EXPECT_EQ(result.iterations_, 0);
    options.maxIterations_ = 1;
    options.relativeTolerance_ = 0;
    BiCgStab<T> oneStep(options, std::make_unique<Identity<T>>());
    oneStep.setup(a);
    result = oneStep.apply(rhs, x);
    EXPECT_TRUE(result.status_ == SolveStatus::MaxIterations);
    EXPECT_EQ(result.iterations_, 1);
    const std::vector<int> offsets{0, 0, 0, 0}, columns;
    const std::vector<T> values;
    const CsrMatrix<T> zeroMatrix(3, 3, offsets, columns, values);
    BiCgStab<T> solver({}, std::make_unique<Identity<T>>());
    solver.setup(zeroMatrix);
    x.copyFromHost(zeros);
    result = solver.apply(rhs, x);
    EXPECT_TRUE(result.status_ == SolveStatus::Breakdown);
    EXPECT_NEAR(result.residualNorm_, std::sqrt(14.0), 1e-5);
    auto invalid = b;
    invalid[1] = std::numeric_limits<T>::quiet_NaN();
    rhs.copyFromHost(invalid);
    EXPECT_TRUE(solver.apply(rhs, x).status_ == SolveStatus::NonFinite);
}

// solver_test.cpp:288-313
TYPED_CASE(bicgstabLifecycleWideIndicesAndFailedSetup) {
    BiCgStab<T, std::int64_t> solver;
    const std::vector<T> b{1, 0}, initial{0, 0};
    Vector<T> rhs = vec(b), x = vec(initial);
    EXPECT_THROW(solver.apply(rhs, x), std::logic_error);
    solver.setup(poisson<T, std::int64_t>(2));
    EXPECT_TRUE(solver.apply(rhs, x).status_ == SolveStatus::Converged);
    Vector<T> wrong(3);
    EXPECT_THROW(solver.apply(wrong, x), std::invalid_argument);
    EXPECT_THROW(solver.setup(CsrMatrix<T, std::int64_t>{}), std::invalid_argument);
    EXPECT_THROW(solver.apply(rhs, x), std::logic_error);
    solver.setup(poisson<T, std::int64_t>(2));
    EXPECT_TRUE(solver.apply(rhs, x).status_ == SolveStatus::Converged);
}

// solver_test.cpp:314-324
TYPED_CASE(amgAcceptsIndependentSmootherFactory) {
    int created{};
    Amg<T> amg({}, [&created]() -> std::unique_ptr<Smoother<T>> {
        ++created;
        return std::make_unique<Jacobi<T>>(3, T(0.6));
    });
    amg.setup(poisson<T>(255));
    EXPECT_EQ(std::size_t(created), 2 * (amg.levelCount() - 1));
    EXPECT_GE(amg.levelCount(), 2u);
    const std::vector<T> ones(255, 1), zeros(255, 0);
    Vector<T> rhs = vec(ones), x = vec(zeros);
    amg.apply(rhs, x);
    for (auto v : amg.lastCycleVisits()) EXPECT_EQ(v, 1u);
    BiCgStabOptions<T> options;
    options.relativeTolerance_ = std::same_as<T, float> ? T(1e-4) : T(1e-9);
    BiCgStab<T> solver(options, std::make_unique<Amg<T>>(AmgOptions<T>{}, [] {
                           return std::unique_ptr<Smoother<T>>(std::make_unique<Jacobi<T>>(3, T(0.6)));
                       }));
    solver.setup(poisson<T>(255));
    EXPECT_TRUE(solver.apply(rhs, x).status_ == SolveStatus::Converged);
    Amg<T> invalid({}, []() -> std::unique_ptr<Smoother<T>> { return {}; });
    EXPECT_THROW(invalid.setup(poisson<T>(255)), std::invalid_argument);
}

// solver_test.cpp:325-345
TYPED_CASE(solverPreservesAllSupportedMemoryPolicies) {
    BiCgStab<T> solver;
    solver.setup(poisson<T>(3));
    const std::vector<T> b{1, 0, 1}, initial(3, 0);
    for (const auto& inputPolicy : strategies()) {
        const Vector<T> rhs = vec(b, inputPolicy);
        for (const auto& outputPolicy : strategies()) {
            Vector<T> x = vec(initial, outputPolicy);
            const auto type = x.memoryType();
            auto* address = x.data();
            ASSERT_TRUE(solver.apply(rhs, x).status_ == SolveStatus::Converged);
            EXPECT_EQ(int(x.memoryType()), int(type));
            EXPECT_EQ(x.data(), address);
            for (T value : linalg::detail::hostCopy(x)) EXPECT_NEAR(value, 1, 1e-5);
        }
    }
}

// cg_test.cpp:50-78
TYPED_CASE(cgIdentitySolvesPoissonWithIndependentCpuResidual) {
    constexpr int n = 31;
    CgOptions<T> options;
    options.relativeTolerance_ = std::same_as<T, float> ? T(2e-5) : T(1e-11);
    std::unique_ptr<Solver<T>> solver = std::make_unique<Cg<T>>(options, std::make_unique<Identity<T>>());
    solver->setup(poisson<T>(n));
    const std::vector<T> ones(n, 1), zeros(n, 0);
    Vector<T> rhs = vec(ones), x = vec(zeros);
    const auto report = solver->apply(rhs, x);
    ASSERT_TRUE(report.status_ == SolveStatus::Converged);
    EXPECT_GT(report.iterations_, 0);
    const auto s = linalg::detail::hostCopy(x);
    double residualSquared = 0;
    for (int i = 0; i < n; ++i) {
        const double exact = double(i + 1) * double(n - i) / 2;
        EXPECT_NEAR(s[std::size_t(i)], exact, (std::same_as<T, float> ? 0.01 : 1e-8));
        const double r = 1 - (2 * double(s[std::size_t(i)]) - (i ? double(s[std::size_t(i) - 1]) : 0) -
                              (i + 1 < n ? double(s[std::size_t(i) + 1]) : 0));
        residualSquared += r * r;
    }
    EXPECT_NEAR(report.residualNorm_, std::sqrt(residualSquared), (std::same_as<T, float> ? 1e-4 : 1e-12));
    EXPECT_LE(std::sqrt(residualSquared), double(options.relativeTolerance_) * std::sqrt(double(n)));
}

// cg_test.cpp:80-99
TYPED_CASE(cgDefaultIdentityPreservesPinnedStorageSupportsAliasAndReuse) {
    Cg<T, std::int64_t> solver;
    auto matrix = poisson<T, std::int64_t>(2);
    solver.setup(matrix);
    matrix.copyValuesFromHost(std::vector<T>(4, 0));
    const std::vector<T> b{1, 0}, initial{T(0.1), T(-0.2)};
    Vector<T> rhs = vec(b, PinnedMemory{}), x = vec(initial, PinnedMemory{});
    auto* address = x.data();
    ASSERT_TRUE(solver.apply(rhs, x).status_ == SolveStatus::Converged);
    EXPECT_EQ(address, x.data());
    EXPECT_EQ(int(x.memoryType()), int(MemoryType::Pinned));
    auto solution = linalg::detail::hostCopy(x);
    EXPECT_NEAR(solution[0], 2.0 / 3, 1e-5);
    EXPECT_NEAR(solution[1], 1.0 / 3, 1e-5);
    ASSERT_TRUE(solver.apply(rhs, rhs).status_ == SolveStatus::Converged);
    solution = linalg::detail::hostCopy(rhs);
    EXPECT_NEAR(solution[0], 2.0 / 3, 1e-5);
    EXPECT_NEAR(solution[1], 1.0 / 3, 1e-5);
}

// cg_test.cpp:100-133
TYPED_CASE(cgInitialConvergenceZeroRhsAndIterationLimits) {
    CgOptions<T> options;
    options.maxIterations_ = 0;
    Cg<T> initialOnly(options);
    initialOnly.setup(poisson<T>(3));
    const std::vector<T> zeros(3, 0), b{1, 0, 0};
    Vector<T> rhs = vec(zeros), x = vec(zeros);
    auto report = initialOnly.apply(rhs, x);
    EXPECT_TRUE(report.status_ == SolveStatus::Converged);
    EXPECT_EQ(report.iterations_, 0);
    EXPECT_EQ(report.relativeResidual_, T(0));
    rhs.copyFromHost(b);
    report = initialOnly.apply(rhs, x);
    EXPECT_TRUE(report.status_ == SolveStatus::MaxIterations);
    EXPECT_EQ(report.iterations_, 0);
    EXPECT_NEAR(report.residualNorm_, 1, 1e-6);
    options.maxIterations_ = 1;
    Cg<T> bounded(options);
    bounded.setup(poisson<T>(3));
    report = bounded.apply(rhs, x);
    EXPECT_TRUE(report.status_ == SolveStatus::MaxIterations);
    EXPECT_EQ(report.iterations_, 1);
    EXPECT_NEAR(report.residualNorm_, 0.5, 1e-6);
    CgOptions<T> zeroOptions;
    zeroOptions.absoluteTolerance_ = T(1e-6);
    Cg<T> zeroSolver(zeroOptions);
    zeroSolver.setup(poisson<T>(3));
    rhs.copyFromHost(zeros);
    report = zeroSolver.apply(rhs, x);
    EXPECT_TRUE(report.status_ == SolveStatus::Converged);
    EXPECT_EQ(report.relativeResidual_, report.residualNorm_);
}

// cg_test.cpp:135-155
TYPED_CASE(cgRefreshRestartsAndExactNonzeroInitialGuess) {
    CgOptions<T> options;
    options.residualRefreshInterval_ = 1;
    options.maxIterations_ = 500;
    options.relativeTolerance_ = std::same_as<T, float> ? T(1e-5) : T(1e-10);
    Cg<T> solver(options);
    solver.setup(poisson<T>(5, false, T(4)));
    const std::vector<T> b{3, 2, 2, 2, 3}, zero(5, 0), exact(5, 1);
    Vector<T> rhs = vec(b), x = vec(zero);
    auto report = solver.apply(rhs, x);
    ASSERT_TRUE(report.status_ == SolveStatus::Converged);
    EXPECT_GT(report.iterations_, 1);
    for (auto value : linalg::detail::hostCopy(x)) EXPECT_NEAR(value, 1, (std::same_as<T, float> ? 1e-4 : 1e-9));
    x.copyFromHost(exact);
    report = solver.apply(rhs, x);
    EXPECT_TRUE(report.status_ == SolveStatus::Converged);
    EXPECT_EQ(report.iterations_, 0);
}

// cg_test.cpp:157-170
TYPED_CASE(cgNonpositiveCurvatureAndNonfiniteRhsAreReported) {
    const std::vector<T> b{1}, zero{0};
    Vector<T> rhs = vec(b), x = vec(zero);
    Cg<T> solver;
    solver.setup(poisson<T>(1, false, T(-1)));
    EXPECT_TRUE(solver.apply(rhs, x).status_ == SolveStatus::Breakdown);
    solver.setup(poisson<T>(1, false, T(0)));
    EXPECT_TRUE(solver.apply(rhs, x).status_ == SolveStatus::Breakdown);
    solver.setup(poisson<T>(1));
    rhs.copyFromHost(std::vector<T>{std::numeric_limits<T>::quiet_NaN()});
    EXPECT_TRUE(solver.apply(rhs, x).status_ == SolveStatus::NonFinite);
}

// cg_test.cpp:172-185
TYPED_CASE(cgSetupAndVectorContracts) {
    Cg<T> solver;
    Vector<T> rhs(2), x(2), wrong(3);
    EXPECT_THROW(solver.apply(rhs, x), std::logic_error);
    solver.setup(poisson<T>(2));
    EXPECT_THROW(solver.apply(wrong, x), std::invalid_argument);
    EXPECT_THROW(solver.apply(rhs, wrong), std::invalid_argument);
    const std::vector<int> offsets{0, 1, 2}, columns{0, 1};
    const std::vector<T> invalid{1, std::numeric_limits<T>::infinity()};
    EXPECT_THROW(solver.setup(CsrMatrix<T>(2, 2, offsets, columns, invalid)), std::invalid_argument);
    EXPECT_THROW(solver.apply(rhs, x), std::logic_error);
    EXPECT_NO_THROW(solver.setup(poisson<T>(2)));
    EXPECT_THROW(solver.setup(CsrMatrix<T>{}), std::invalid_argument);
    EXPECT_THROW(solver.apply(rhs, x), std::logic_error);
}

// amg_native_test.cpp:329-395 (visit counts, linearity, aliasing, BiCGStab+AMG per cycle)
template <class T, class I>
static void checkCycles() {
    const auto a = poisson<T, I>(127);
    std::vector<T> b(127), c(127), sum(127);
    for (int i = 0; i < 127; ++i) {
        b[std::size_t(i)] = T(std::sin(.13 * i) + .4);
        c[std::size_t(i)] = T(std::cos(.07 * i));
        sum[std::size_t(i)] = b[std::size_t(i)] + c[std::size_t(i)];
    }
    for (auto kind : {multigrid::Cycle::V, multigrid::Cycle::W, multigrid::Cycle::F}) {
        AmgOptions<T> options;
        options.coarseRows_ = 4;
        options.cycle_ = kind;
        Amg<T, I> amg(options);
        amg.setup(a);
        ASSERT_GE(amg.levelCount(), 3u);
        Vector<T> rhs = vec(b), x(127), other(127), combined(127);
        amg.apply(rhs, x);
        const auto actual = linalg::detail::hostCopy(x);
        const double tol = std::same_as<T, float> ? 2e-3 : 2e-10;
        const auto visits = amg.lastCycleVisits();
        for (std::size_t d = 0; d < visits.size(); ++d) {
            if (kind == multigrid::Cycle::V) EXPECT_EQ(visits[d], 1u);
            if (kind == multigrid::Cycle::W) EXPECT_EQ(visits[d], std::size_t(1) << d);
            if (kind == multigrid::Cycle::F) EXPECT_EQ(visits[d], d + 1 == visits.size() ? visits.size() : d + 2);
        }
        rhs.copyFromHost(c);
        amg.apply(rhs, other);
        rhs.copyFromHost(sum);
        amg.apply(rhs, combined);
        const auto xc = linalg::detail::hostCopy(other), xs = linalg::detail::hostCopy(combined);
        for (std::size_t i = 0; i < 127; ++i) EXPECT_NEAR(xs[i], actual[i] + xc[i], tol);
        rhs.copyFromHost(b);
        amg.apply(rhs, rhs);
        const auto alias = linalg::detail::hostCopy(rhs);
        for (std::size_t i = 0; i < 127; ++i) EXPECT_NEAR(alias[i], actual[i], tol);
        BiCgStabOptions<T> solveOptions;
        solveOptions.relativeTolerance_ = std::same_as<T, float> ? T(5e-4) : T(1e-9);
        BiCgStab<T, I> solver(solveOptions, std::make_unique<Amg<T, I>>(options));
        solver.setup(a);
        rhs.copyFromHost(b);
        x.copyFromHost(std::vector<T>(127, 0));
        const auto report = solver.apply(rhs, x);
        EXPECT_TRUE(report.status_ == SolveStatus::Converged);
        EXPECT_LE(report.relativeResidual_, solveOptions.relativeTolerance_);
    }
}
TEST_CASE(cyclesDouble, true) { checkCycles<double, int>(); }
TEST_CASE(cyclesFloatWide, true) { checkCycles<float, std::int64_t>(); }

// amg_native_test.cpp:398-424
TEST_CASE(fallbackMatchesFixedJacobiAndZeroSweeps, true) {
    const auto a = poisson<double>(100);
    std::vector<double> b(100, 1), zero(100, 0);
    Vector<double> rhs = vec(b), x(100), expected = vec(zero);
    AmgOptions<double> o;
    o.maxLevels_ = 1;
    o.coarseSweeps_ = 8;
    Amg<double> amg(o);
    amg.setup(a);
    amg.apply(rhs, x);
    Jacobi<double> jacobi(8, .75);
    jacobi.setup(a);
    jacobi.apply(rhs, expected);
    EXPECT_TRUE(linalg::detail::hostCopy(x) == linalg::detail::hostCopy(expected));
    o.maxLevels_ = 20;
    o.coarseRows_ = 4;
    o.preSweeps_ = o.postSweeps_ = 0;
    for (auto kind : {multigrid::Cycle::V, multigrid::Cycle::W, multigrid::Cycle::F}) {
        o.cycle_ = kind;
        Amg<double> noSmoothing(o);
        noSmoothing.setup(a);
        noSmoothing.apply(rhs, x);
        auto first = linalg::detail::hostCopy(x);
        noSmoothing.apply(rhs, x);
        EXPECT_TRUE(first == linalg::detail::hostCopy(x));
    }
}

// Engine and host-driven paths agree: Cg over the engine Amg and over a
// factory-built hierarchy with the default Jacobi settings.
TEST_CASE(engineAndFactoryHierarchiesAgree, true) {
    const auto a = poisson<double>(500);
    AmgOptions<double> o;
    o.coarseRows_ = 8;
    Amg<double> engine(o);
    Amg<double> factory(o, [] { return std::unique_ptr<Smoother<double>>(std::make_unique<Jacobi<double>>(1, 0.75)); });
    engine.setup(a);
    factory.setup(a);
    EXPECT_EQ(engine.levelCount(), factory.levelCount());
    const auto se = engine.levelStatistics(), sf = factory.levelStatistics();
    for (std::size_t d = 0; d < se.size() && d < sf.size(); ++d) {
        EXPECT_EQ(se[d].rows_, sf[d].rows_);
        EXPECT_EQ(se[d].nonzeros_, sf[d].nonzeros_);
        EXPECT_TRUE(se[d].aggregateSizes_ == sf[d].aggregateSizes_);
    }
    EXPECT_EQ(engine.stopReason(), factory.stopReason());
    EXPECT_EQ(engine.coarseSolver(), factory.coarseSolver());
    std::vector<double> b(500);
    for (int i = 0; i < 500; ++i) b[std::size_t(i)] = std::sin(0.01 * i) + 1;
    Vector<double> rhs = vec(b), xe(500), xf(500);
    engine.apply(rhs, xe);
    factory.apply(rhs, xf);
    const auto ve = xe.hostCopy(), vf = xf.hostCopy();
    double diff = 0, norm = 0;
    for (std::size_t i = 0; i < ve.size(); ++i) {
        diff = std::max(diff, std::abs(ve[i] - vf[i]));
        norm = std::max(norm, std::abs(vf[i]));
    }
    EXPECT_LE(diff, 1e-12 * norm);
}

// ---- linalg free functions and operators (linear_algebra_test.cpp) ----

namespace {
namespace la = thermofem::linalg;

template <class T>
void expectBuffer(const Buffer<T>& b, std::initializer_list<T> expected) {
    const auto actual = b.hostCopy();
    ASSERT_EQ(actual.size(), expected.size());
    const double tolerance = std::same_as<T, float> ? 1e-5 : 1e-12;
    std::size_t i = 0;
    for (const auto value : expected) EXPECT_NEAR(double(actual[i++]), double(value), tolerance);
}
template <class T>
void expectVector(const Vector<T>& v, std::initializer_list<T> expected) {
    expectBuffer(v.values(), expected);
}
// the mathematical matrix of a CSR (duplicates summed), row-major
template <class T, class I>
void expectSparse(const CsrMatrix<T, I>& m, std::initializer_list<T> expected) {
    const auto rp = m.rowOffsets().hostCopy();
    const auto ci = m.columnIndices().hostCopy();
    const auto v = m.values().hostCopy();
    std::vector<double> dense(m.rowCount() * m.columnCount(), 0.0);
    for (std::size_t r = 0; r < m.rowCount(); ++r)
        for (auto j = rp[r]; j < rp[r + 1]; ++j) dense[r * m.columnCount() + std::size_t(ci[std::size_t(j)])] += v[std::size_t(j)];
    ASSERT_EQ(dense.size(), expected.size());
    const double tolerance = std::same_as<T, float> ? 1e-5 : 1e-12;
    std::size_t i = 0;
    for (const auto value : expected) EXPECT_NEAR(dense[i++], double(value), tolerance);
}

// linear_algebra_test.cpp:97-131
template <class T>
void run_vectorArithmetic() {
    for (const auto& ls : strategies())
        for (const auto& rs : strategies()) {
            Vector<T> x(std::span<const T>(std::vector<T>{1, -2, 3}), ls), y(std::span<const T>(std::vector<T>{4, 5, -6}), rs);
            auto sum = x + y;
            EXPECT_EQ(sum.memoryStrategy().index(), ls.index());
            expectVector<T>(sum, {5, 3, -3});
            expectVector<T>(x - y, {-3, -7, 9});
            expectVector<T>(-x, {-1, 2, -3});
            expectVector<T>(+x, {1, -2, 3});
            expectVector<T>(T(2) * x, {2, -4, 6});
            expectVector<T>(x * T(2), {2, -4, 6});
            expectVector<T>(x / T(2), {T(0.5), -1, T(1.5)});
            EXPECT_NEAR(double(x * y), -24.0, 1e-5);
            EXPECT_NEAR(double(la::dot(x, y)), -24.0, 1e-5);
            EXPECT_NEAR(double(la::norm1(x)), 6.0, 1e-5);
            EXPECT_NEAR(double(la::norm2(x)), std::sqrt(14.0), 1e-5);
            expectVector<T>(la::axpy(T(2), x, y), {6, 1, 0});
            auto copy = x;
            copy += copy;
            copy -= x;
            copy *= T(2);
            copy /= T(2);
            expectVector<T>(copy, {1, -2, 3});
            EXPECT_EQ(copy.memoryStrategy().index(), ls.index());
            EXPECT_THROW(x / T(0), std::domain_error);
            expectVector<T>(x, {1, -2, 3});
        }
}

// linear_algebra_test.cpp:133-170
template <class T>
void run_denseArithmetic() {
    for (const auto& strategy : strategies()) {
        DenseMatrix<T> a(2, 3, strategy), b(3, 2, PinnedMemory{});
        a.copyFromHost(std::vector<T>{1, 2, 3, 4, 5, 6});
        b.copyFromHost(std::vector<T>{7, 8, 9, 10, 11, 12});
        Vector<T> x(std::span<const T>(std::vector<T>{1, 2, 3}), PinnedMemory{});
        expectVector<T>(a * x, {14, 32});
        auto product = a * b;
        EXPECT_EQ(product.rowCount(), 2u);
        EXPECT_EQ(product.columnCount(), 2u);
        EXPECT_EQ(product.values().memoryStrategy().index(), strategy.index());
        expectBuffer<T>(product.values(), {58, 64, 139, 154});
        auto transposed = la::transpose(a);
        EXPECT_EQ(transposed.rowCount(), 3u);
        EXPECT_EQ(transposed.columnCount(), 2u);
        expectBuffer<T>(transposed.values(), {1, 4, 2, 5, 3, 6});
        expectBuffer<T>((a + a).values(), {2, 4, 6, 8, 10, 12});
        expectBuffer<T>((a - a).values(), {0, 0, 0, 0, 0, 0});
        expectBuffer<T>((-a).values(), {-1, -2, -3, -4, -5, -6});
        auto copy = a;
        copy += a;
        copy -= a;
        copy *= T(2);
        copy /= T(2);
        expectBuffer<T>(copy.values(), {1, 2, 3, 4, 5, 6});
        copy *= b;
        expectBuffer<T>(copy.values(), {58, 64, 139, 154});
        EXPECT_NEAR(double(la::frobeniusNorm(a)), std::sqrt(91.0), 1e-5);
        EXPECT_THROW(a / T(0), std::domain_error);
    }
}

// linear_algebra_test.cpp:176-220 (duplicate column 2 in row 0 of a)
template <class T, class I>
void checkSparse(MemoryStrategy strategy) {
    CsrMatrix<T, I> a(2, 3, std::vector<I>{0, 3, 4}, std::vector<I>{2, 0, 2, 1}, std::vector<T>{1, 2, 3, 5}, strategy);
    CsrMatrix<T, I> b(3, 2, std::vector<I>{0, 1, 2, 4}, std::vector<I>{0, 1, 1, 0}, std::vector<T>{1, 2, 4, 3},
                      PinnedMemory{});
    Vector<T> x(std::span<const T>(std::vector<T>{1, 2, 3}), PinnedMemory{});
    expectVector<T>(a * x, {14, 10});
    DenseMatrix<T> dense(3, 2, PinnedMemory{});
    dense.copyFromHost(std::vector<T>{1, 0, 0, 2, 3, 4});
    expectBuffer<T>((a * dense).values(), {14, 16, 0, 10});
    auto product = a * b;
    expectSparse<T, I>(product, {14, 16, 0, 10});
    EXPECT_EQ(product.values().memoryStrategy().index(), strategy.index());
    expectSparse<T, I>(a + a, {4, 0, 8, 0, 10, 0});
    expectSparse<T, I>(a - a, {0, 0, 0, 0, 0, 0});
    auto transposed = la::transpose(a);
    expectSparse<T, I>(transposed, {2, 0, 0, 5, 4, 0});
    expectSparse<T, I>(-a, {-2, 0, -4, 0, -5, 0});
    expectSparse<T, I>(a * T(2), {4, 0, 8, 0, 10, 0});
    EXPECT_NEAR(double(la::frobeniusNorm(a)), std::sqrt(45.0), 1e-5);
    auto copy = a;
    copy += copy;
    copy -= a;
    copy *= T(2);
    copy /= T(2);
    expectSparse<T, I>(copy, {2, 0, 4, 0, 5, 0});
    copy *= b;
    expectSparse<T, I>(copy, {14, 16, 0, 10});
    EXPECT_THROW(a / T(0), std::domain_error);
}

// linear_algebra_test.cpp:222-254
template <class T>
void run_zeroDimensions() {
    DenseMatrix<T> a(2, 0, PinnedMemory{}), b(0, 3);
    expectBuffer<T>((a * b).values(), {0, 0, 0, 0, 0, 0});
    Vector<T> empty;
    expectVector<T>(a * empty, {0, 0});
    EXPECT_EQ(double(la::dot(empty, empty)), 0.0);
    EXPECT_EQ(double(la::norm1(empty)), 0.0);
    EXPECT_EQ(double(la::norm2(empty)), 0.0);
    EXPECT_EQ(la::transpose(a).rowCount(), 0u);
    CsrMatrix<T> sparse(2, 3, std::vector<int>{0, 0, 0}, std::vector<int>{}, std::vector<T>{}, PinnedMemory{});
    Vector<T> x(std::span<const T>(std::vector<T>{1, 2, 3}));
    expectVector<T>(sparse * x, {0, 0});
    expectSparse<T, int>(sparse + sparse, {0, 0, 0, 0, 0, 0});
    expectSparse<T, int>(la::transpose(sparse), {0, 0, 0, 0, 0, 0});
    CsrMatrix<T> right(3, 2, std::vector<int>{0, 0, 0, 0}, std::vector<int>{}, std::vector<T>{});
    expectSparse<T, int>(sparse * right, {0, 0, 0, 0});
    CsrMatrix<T> defaultEmpty;
    EXPECT_EQ((defaultEmpty + defaultEmpty).rowCount(), 0u);
    EXPECT_EQ(la::transpose(defaultEmpty).rowCount(), 0u);
    CsrMatrix<T> one(2, 3, std::vector<int>{0, 1, 1}, std::vector<int>{1}, std::vector<T>{4});
    expectSparse<T, int>(sparse + one, {0, 4, 0, 0, 0, 0});
    expectSparse<T, int>(one + sparse, {0, 4, 0, 0, 0, 0});
    // two nonzero operands whose product has no structural nonzero
    CsrMatrix<T> disjoint(3, 2, std::vector<int>{0, 1, 1, 1}, std::vector<int>{0}, std::vector<T>{2});
    expectSparse<T, int>(one * disjoint, {0, 0, 0, 0});
}
}  // namespace

TEST_CASE(vectorArithmetic, true) {
    run_vectorArithmetic<float>();
    run_vectorArithmetic<double>();
}
TEST_CASE(denseArithmetic, true) {
    run_denseArithmetic<float>();
    run_denseArithmetic<double>();
}
TEST_CASE(sparseArithmetic, true) {
    for (const auto& s : strategies()) {
        checkSparse<float, std::int32_t>(s);
        checkSparse<float, std::int64_t>(s);
        checkSparse<double, std::int32_t>(s);
        checkSparse<double, std::int64_t>(s);
    }
}
TEST_CASE(zeroDimensions, true) {
    run_zeroDimensions<float>();
    run_zeroDimensions<double>();
}
// linear_algebra_test.cpp:256-269 (the matrices are allocated on the device)
TEST_CASE(linalgInvalidShapes, true) {
    Vector<double> x;
    EXPECT_THROW(x / 0.0, std::domain_error);
    DenseMatrix<double> a(2, 0), b(3, 0);
    EXPECT_THROW(a + b, std::invalid_argument);
    EXPECT_THROW(a * b, std::invalid_argument);
    EXPECT_THROW(b * DenseMatrix<double>(1, 0), std::invalid_argument);
    EXPECT_THROW(la::detail::librarySize(static_cast<std::size_t>(std::numeric_limits<int>::max()) + 1),
                 std::length_error);
    CsrMatrix<double> sparse;
    EXPECT_THROW(sparse * b, std::invalid_argument);
}

int main(int argc, char** argv) { return harness::run(argc, argv); }
