// The C++20 drop-in API over libthermofem_hip.so, checked suite by suite:
// host-side contracts (options, CSR validation, Matrix Market), storage,
// the linear-algebra and multigrid operations, smoothers and the AMG
// hierarchy, CG and BiCGStab scenarios, and the linalg free functions.  The
// inputs and expected values are those the reference's tests fix (cited per
// suite as file:line); typed suites run for float and double.
// Usage: api_test [--cpu-only] [name-filter]
#include <thermofem/thermofem.hpp>

#include <cstdint>
#include <limits>
#include <optional>
#include <sstream>

#include "probe.hpp"

using namespace thermofem;
using probe::Probe;

namespace {

constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();
constexpr double kInf = std::numeric_limits<double>::infinity();

template <class T>
constexpr bool isFloat = std::same_as<T, float>;

// Tridiagonal [-1, d, -1] (or the nonsymmetric [-1, 3, -0.5]) of order n.
template <class T, class I = int>
CsrMatrix<T, I> tridiagonal(int n, T diagonal = T(2), bool skewed = false) {
    std::vector<I> rp(1, I(0)), ci;
    std::vector<T> v;
    auto put = [&](int c, T x) {
        ci.push_back(I(c));
        v.push_back(x);
    };
    for (int r = 0; r < n; ++r) {
        if (r > 0) put(r - 1, T(-1));
        put(r, skewed ? T(3) : diagonal);
        if (r + 1 < n) put(r + 1, skewed ? T(-0.5) : T(-1));
        rp.push_back(I(v.size()));
    }
    return CsrMatrix<T, I>(std::size_t(n), std::size_t(n), rp, ci, v);
}

template <class T>
Vector<T> onDevice(const std::vector<T>& values, MemoryStrategy where = GlobalMemory{}) {
    return Vector<T>(std::span<const T>(values), where);
}

template <class T>
std::vector<T> host(const Vector<T>& v) {
    return v.hostCopy();
}

const std::vector<MemoryStrategy>& allStrategies() {
    static const std::vector<MemoryStrategy> s{GlobalMemory{}, PinnedMemory{}, MappedMemory{}, ManagedMemory{}};
    return s;
}

// x_i = (i + 1)(n - i) / 2 solves tridiagonal(n) x = 1.
double tridiagonalOnesSolution(int n, int i) { return double(i + 1) * double(n - i) / 2; }

// |1 - (2 x_i - x_{i-1} - x_{i+1})| summed in squares: the residual of x for
// tridiagonal(n) x = 1, computed on the host.
template <class T>
double tridiagonalOnesResidual(const std::vector<T>& x) {
    const int n = int(x.size());
    double sum = 0;
    for (int i = 0; i < n; ++i) {
        const double left = i ? double(x[std::size_t(i) - 1]) : 0, right = i + 1 < n ? double(x[std::size_t(i) + 1]) : 0;
        const double r = 1 - (2 * double(x[std::size_t(i)]) - left - right);
        sum += r * r;
    }
    return std::sqrt(sum);
}

// ---------------------------------------------------------------- host side

// Option ranges of Cg.cuh, BiCgStab.cuh, Amg.cuh:66-75 and Jacobi.cuh
// (cg_test.cpp:186-207, solver_test.cpp:76-79,176-178,310-312).
void optionContracts(Probe& p) {
    auto rejectsCg = [&](auto edit, const char* what) {
        CgOptions<double> o;
        edit(o);
        p.raises<std::invalid_argument>([&] { Cg<double> s(o); }, what);
    };
    rejectsCg([](auto& o) { o.maxIterations_ = -1; }, "CG negative iteration limit");
    rejectsCg([](auto& o) { o.residualRefreshInterval_ = 0; }, "CG zero refresh interval");
    for (double r : {-1.0, 1.0, kInf, kNaN}) rejectsCg([r](auto& o) { o.relativeTolerance_ = r; }, "CG relative tolerance");
    for (double a : {-1.0, kInf, kNaN}) rejectsCg([a](auto& o) { o.absoluteTolerance_ = a; }, "CG absolute tolerance");
    BiCgStabOptions<double> bi;
    bi.relativeTolerance_ = -1;
    p.raises<std::invalid_argument>([&] { BiCgStab<double> s(bi); }, "BiCGStab negative tolerance");
    auto rejectsAmg = [&](auto edit, const char* what) {
        AmgOptions<double> o;
        edit(o);
        p.raises<std::invalid_argument>([&] { Amg<double> m(o); }, what);
    };
    rejectsAmg([](auto& o) { o.maxLevels_ = 0; }, "AMG zero levels");
    rejectsAmg([](auto& o) { o.coarseRows_ = 1025; }, "AMG coarse rows past 1024");
    rejectsAmg([](auto& o) { o.jacobiWeight_ = 2; }, "AMG Jacobi weight 2");
    p.raises<std::invalid_argument>([] { Jacobi<double> j(0); }, "Jacobi zero sweeps");
    p.raises<std::invalid_argument>([] { Jacobi<float> j(1, std::numeric_limits<float>::quiet_NaN()); }, "Jacobi NaN");
}

// CsrMatrix construction checks on the host (data_structure_test.cpp:147-191).
void csrContracts(Probe& p) {
    using M = CsrMatrix<double>;
    const std::vector<int> rp{0, 1, 2}, ci{0, 1};
    const std::vector<double> v{1, 2};
    struct Bad {
        const char* what;
        std::size_t rows;
        std::vector<int> rp, ci;
        std::vector<double> v;
    };
    const std::vector<Bad> cases{{"missing offsets", 2, {}, ci, v},
                                 {"offsets of the wrong length", 3, rp, ci, v},
                                 {"values missing", 2, rp, ci, {}},
                                 {"first offset not zero", 2, {1, 1, 2}, ci, v},
                                 {"last offset past the entries", 2, {0, 1, 3}, ci, v},
                                 {"decreasing offsets", 3, {0, 2, 1, 2}, ci, v},
                                 {"negative offset", 3, {0, -1, 1, 2}, ci, v},
                                 {"offset out of bounds", 3, {0, 1, 3, 2}, ci, v},
                                 {"column out of range", 2, rp, {0, 2}, v}};
    for (const auto& c : cases) p.raises<std::invalid_argument>([&] { M m(c.rows, 2, c.rp, c.ci, c.v); }, c.what);
    p.raises<std::length_error>([] { M m(std::numeric_limits<std::size_t>::max(), 2, {}, {}, {}); }, "row count");
    p.raises<std::length_error>([] { Buffer<double> b(std::numeric_limits<std::size_t>::max() / 4); }, "buffer size");
}

// io::readMatrixMarket (matrix_market_test.cpp:8-44).
void matrixMarket(Probe& p) {
    auto read = [](const std::string& text) {
        std::istringstream in(text);
        return io::readMatrixMarket(in);
    };
    // unsorted coordinates are grouped by row, duplicates kept in file order
    auto m = read("%%MatrixMarket matrix coordinate real general\n% comment\n3 3 4\n3 1 2\n1 2 4\n1 2 -1\n3 3 5\n");
    p.expect(m.offsets_ == std::vector<int>{0, 2, 2, 4}, "general: offsets");
    p.expect(m.indices_ == std::vector<int>{1, 1, 0, 2}, "general: columns");
    p.expect(m.values_ == std::vector<double>{4, -1, 2, 5}, "general: values");
    // a symmetric lower triangle is mirrored; integers are accepted
    m = read("%%MatrixMarket matrix coordinate integer symmetric\n2 2 3\n1 1 2\n2 1 -1\n2 2 3\n");
    p.expect(m.offsets_ == std::vector<int>{0, 2, 4}, "symmetric: offsets");
    p.expect(m.indices_ == std::vector<int>{0, 1, 0, 1}, "symmetric: columns");
    p.expect(m.values_ == std::vector<double>{2, -1, -1, 3}, "symmetric: values");
    p.expect(read("%%MatrixMarket matrix coordinate real general\n0 0 0\n").offsets_ == std::vector<int>{0}, "empty");
    const char* rejected[] = {"",
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
    for (const char* text : rejected) p.raises<std::runtime_error>([&] { (void)read(text); }, std::string("reject: ") + text);
}

// ---------------------------------------------------------------- storage

// Buffer / Vector over the four strategies: copy rules, views, moves
// (data_structure_test.cpp:219-345).
void memoryStrategies(Probe& p) {
    const std::vector<double> data{1, 2, 3};
    for (const auto& from : allStrategies())
        for (const auto& to : allStrategies()) {
            Buffer<double> src(std::span<const double>(data), from);
            Buffer<double> dst(3, to);
            double* const kept = dst.data();
            dst.copyDataFrom(src);  // keeps the target's address and strategy
            p.expect(dst.data() == kept && int(dst.memoryType()) == int(to.index()) && dst.hostCopy() == data,
                     "copyDataFrom keeps the target");
            dst = src;  // deep copy adopting the source's strategy
            p.expect(int(dst.memoryType()) == int(from.index()) && dst.data() != src.data() && dst.hostCopy() == data,
                     "assignment copies with the source strategy");
            Buffer<double> small(2, to);
            p.raises<std::invalid_argument>([&] { small.copyDataFrom(src); }, "size mismatch");
            src.changeMemoryType(to);
            p.expect(int(src.memoryType()) == int(to.index()) && src.hostCopy() == data, "changeMemoryType");
            if (src.memoryType() == MemoryType::Global)
                p.raises<std::logic_error>([&] { (void)src.hostData(); }, "no host view of device memory");
            else
                p.expect(src.hostData()[2] == 3.0, "host view");
            if (src.memoryType() == MemoryType::Pinned)
                p.raises<std::logic_error>([&] { (void)src.deviceData(); }, "no device view of pinned memory");
            else
                p.expect(src.deviceData() != nullptr, "device view");
            double* const addr = src.data();
            Buffer<double> moved(std::move(src));
            p.expect(moved.data() == addr && src.empty(), "move transfers the allocation");
            Vector<double> v(3, to);
            v.copyDataFrom(Vector<double>(std::span<const double>(data), from));
            p.expect(v.hostCopy() == data, "vector copy");
        }
    Buffer<double> none(MappedMemory{});
    Buffer<double> moved(std::move(none));
    p.expect(int(moved.memoryType()) == int(MemoryType::Mapped) && moved.deviceData() == nullptr, "empty mapped move");
}

// CsrMatrix copies, canonical form, fromBuffers (data_structure_test.cpp:381-416,
// amg_native_test.cpp:115-131).
void csrStorage(Probe& p) {
    const std::vector<int> rp{0, 2, 3}, ci{0, 1, 1};
    const std::vector<double> v{1, 2, 3};
    CsrMatrix<double> pinned(2, 2, rp, ci, v, PinnedMemory{});
    auto copy = pinned;
    copy.changeMemoryType(GlobalMemory{});
    p.expect(int(copy.rowOffsets().memoryType()) == int(MemoryType::Global), "moved to device");
    p.expect(copy.rowOffsets().hostCopy() == rp && copy.columnIndices().hostCopy() == ci && copy.values().hostCopy() == v,
             "copy contents");
    const std::vector<int> drp{0, 3, 6}, dci{1, 0, 0, 1, 0, 1};
    const std::vector<double> dv{-1, 1, 1, 1, -1, 1};
    CsrMatrix<double> dup(2, 2, drp, dci, dv);
    multigrid::GpuAggregation ops;
    const auto canonical = ops.canonical(dup);
    p.expect(canonical.values().hostCopy() == std::vector<double>{2, -1, -1, 2}, "duplicates summed in order");
    const std::vector<int> broken{0, 7, 6};
    Buffer<int> bo{std::span<const int>(broken)}, bi{std::span<const int>(dci)};
    Buffer<double> bv{std::span<const double>(dv)};
    p.raises<std::invalid_argument>(
        [&] { (void)CsrMatrix<double>::fromBuffers(2, 2, std::move(bo), std::move(bi), std::move(bv)); },
        "fromBuffers rejects bad offsets");
    p.expect(bo.size() == 3u && bv.size() == 6u, "rejected buffers stay with the caller");
    Buffer<int> go{std::span<const int>(drp)}, gi{std::span<const int>(dci)};
    Buffer<double> gv{std::span<const double>(dv)};
    const auto owned = CsrMatrix<double>::fromBuffers(2, 2, std::move(go), std::move(gi), std::move(gv));
    p.expect(owned.nonZeroCount() == 6u && go.empty(), "fromBuffers takes the buffers");
    p.raises<std::invalid_argument>([&] { (void)ops.strength(dup, -.1); }, "negative strength threshold");
    const auto graph = ops.strength(canonical, .25);
    p.raises<std::invalid_argument>([&] { (void)ops.aggregate(graph, 0, true); }, "zero matching rounds");
}

// ---------------------------------------------------------------- operations

// On tridiagonal(4): D^-1 = 1/2, residual of ones, Galerkin [2], one Jacobi
// update to 1/2 (multigrid_operations_test.cpp:33-64).
template <class T>
void operationsOnTridiagonal4(Probe& p) {
    const auto a = tridiagonal<T>(4);
    const auto dinv = host(linalg::inverseDiagonal(a));
    for (T d : dinv) p.expect(d == T(0.5), "inverse diagonal 1/2");
    multigrid::GpuAggregation ops;
    const auto native = multigrid::toNativeMatrix(a);
    const auto coarse = ops.galerkin(native, ops.aggregate(ops.strength(native, .25), 8, true));
    p.need(coarse.rowCount() == 1u, "one aggregate");
    p.expect(coarse.values().hostCopy() == std::vector<double>{2}, "Galerkin [2]");
    linalg::IterationContext<T> op(a);
    Vector<T> x = onDevice(std::vector<T>(4, 1)), b = onDevice(std::vector<T>(4, 0)), r(4);
    op.residual(b, x, r);
    p.expect(host(r) == std::vector<T>{-1, 0, 0, -1}, "residual of ones");
    Vector<T> d = onDevice(dinv);
    op.clear(x);
    op.clear(r);  // the product A x of the zero iterate
    b.copyFromHost(std::vector<T>(4, 1));
    op.jacobi(T(1), d, b, r, x);
    for (T xi : host(x)) p.expect(xi == T(0.5), "Jacobi update");
    p.raises<std::invalid_argument>([&] { op.residual(b, x, b); }, "residual aliasing");
    p.raises<std::invalid_argument>([&] { (void)ops.strength(native, -1); }, "threshold");
}

// Duplicate diagonal entries and the pivoting coarse LU
// (multigrid_operations_test.cpp:66-83).
template <class T>
void duplicateDiagonalAndCoarseLu(Probe& p) {
    const std::vector<int> rp{0, 2, 4}, ci{0, 0, 0, 1};
    const CsrMatrix<T> a(2, 2, rp, ci, std::vector<T>{1, 1, 4, 3});
    const auto d = host(linalg::inverseDiagonal(a));
    p.near(d[0], 0.5, 1e-6, "summed diagonal 2");
    p.near(d[1], 1.0 / 3, 1e-6, "diagonal 3");
    linalg::CoarseLu<T> lu(a);
    Vector<T> b = onDevice(std::vector<T>{2, 7}), x(2);
    lu.apply(b, x);
    for (T xi : host(x)) p.near(xi, 1, 1e-6, "LU solution ones");
    Vector<T> wrong(3);
    p.raises<std::invalid_argument>([&] { lu.apply(wrong, x); }, "LU shape");
    const CsrMatrix<T> zeroDiagonal(2, 2, rp, ci, std::vector<T>{0, 0, 4, 3});
    p.raises<std::invalid_argument>([&] { (void)linalg::inverseDiagonal(zeroDiagonal); }, "zero diagonal");
    p.raises<std::invalid_argument>([&] { linalg::CoarseLu<T> bad(zeroDiagonal); }, "singular LU input");
}

// dot -24, norm sqrt(14) (linear_algebra_test.cpp:112-116), axpy, scale, SpMV.
template <class T>
void contextReductions(Probe& p) {
    linalg::IterationContext<T> op(tridiagonal<T>(3));
    Vector<T> x = onDevice(std::vector<T>{1, -2, 3}), y = onDevice(std::vector<T>{4, 5, -6});
    p.near(op.dot(x, y), -24, 1e-5, "dot");
    p.near(op.norm(x), std::sqrt(14.0), 1e-5, "norm");
    op.axpy(T(2), x, y);
    p.expect(host(y) == std::vector<T>{6, 1, 0}, "axpy");
    op.scale(T(-1), y);
    p.expect(host(y) == std::vector<T>{-6, -1, -0.0}, "scale");
    Vector<T> ax(3);
    op.multiply(x, ax);
    p.expect(host(ax) == std::vector<T>{4, -8, 8}, "SpMV");
}

// The context keeps its own copy of the matrix; 64-bit indices
// (solver_test.cpp:96-109).
template <class T>
void contextSnapshot(Probe& p) {
    auto a = tridiagonal<T, std::int64_t>(2);
    linalg::IterationContext<T, std::int64_t> op(a);
    a.copyValuesFromHost(std::vector<T>(4, 0));
    const Vector<T> x = onDevice(std::vector<T>{1, 2});
    Vector<T> y(2);
    op.multiply(x, y);
    const auto out = host(y);
    p.near(out[0], 0, 1e-6, "snapshot row 0");
    p.near(out[1], 3, 1e-6, "snapshot row 1");
    p.raises<std::invalid_argument>([&] { op.multiply(x, const_cast<Vector<T>&>(x)); }, "SpMV aliasing");
}

// A = s [1 2; 4 3], b = s [5, 10] -> [1, 2] for s = 1e-100, 1, 1e100; aliasing;
// [1 1; 1 1] singular (amg_native_test.cpp:133-153).
void coarseFactorScales(Probe& p) {
    const std::vector<int> rp{0, 2, 4}, ci{0, 1, 0, 1};
    for (double s : {1e-100, 1.0, 1e100}) {
        const CsrMatrix<double> a(2, 2, rp, ci, std::vector<double>{s, 2 * s, 4 * s, 3 * s});
        const auto f = multigrid::factorCoarse(a);
        p.need(f.factored_, "factored at scale");
        Vector<double> b = onDevice(std::vector<double>{5 * s, 10 * s}), x(2);
        multigrid::solveCoarse(f, b, x);
        const auto got = host(x);
        p.near(got[0], 1, 1e-12, "x0");
        p.near(got[1], 2, 1e-12, "x1");
        multigrid::solveCoarse(f, b, b);
        p.expect(host(b) == got, "aliased solve");
    }
    const auto singular = multigrid::factorCoarse(CsrMatrix<double>(2, 2, rp, ci, std::vector<double>{1, 1, 1, 1}));
    p.expect(singular.singular_ && !singular.factored_, "singular detected");
}

// Restricting ones counts each aggregate's members; prolongation adds
// (amg_native_test.cpp:96-107).
void transfers(Probe& p) {
    const int n = 61;
    multigrid::GpuAggregation ops;
    const auto native = ops.canonical(tridiagonal<double>(n));
    const auto t = ops.aggregate(ops.strength(native, .25), 8, true);
    Vector<double> fine = onDevice(std::vector<double>(n, 1.0)), coarse(std::size_t(t.coarseRows_));
    multigrid::restrictAggregates(t, fine, coarse);
    const auto sums = host(coarse);
    const auto off = t.offsets_.hostCopy(), map = t.aggregate_.hostCopy();
    int counted = 0;
    for (int k : t.histogram_) counted += k;
    p.expect(counted == t.coarseRows_, "histogram counts every aggregate");
    for (int c = 0; c < t.coarseRows_; ++c)
        p.expect(sums[std::size_t(c)] == double(off[std::size_t(c) + 1] - off[std::size_t(c)]), "restricted ones");
    multigrid::prolongAggregates(t, coarse, fine);
    const auto added = host(fine);
    for (int i = 0; i < n; ++i)
        p.expect(added[std::size_t(i)] == 1 + sums[std::size_t(map[std::size_t(i)])], "prolongation adds");
    p.expect(ops.galerkin(native, t).nonZeroCount() <= native.nonZeroCount(), "Galerkin does not grow");
}

// ---------------------------------------------------------------- smoothers and AMG

// Jacobi(2 sweeps, weight 1) on tridiagonal(2), b = [1, 0]: [0.5, 0.25];
// aliased, from x = b: [0.75, 0.25]; pinned storage kept (solver_test.cpp:56-74).
template <class T>
void jacobiSweeps(Probe& p) {
    std::unique_ptr<Smoother<T>> j = std::make_unique<Jacobi<T>>(2, T(1));
    j->setup(tridiagonal<T>(2));
    Vector<T> b = onDevice(std::vector<T>{1, 0}, PinnedMemory{}), x = onDevice(std::vector<T>{0, 0}, PinnedMemory{});
    j->apply(b, x);
    auto got = host(x);
    p.expect(int(x.memoryType()) == int(MemoryType::Pinned), "strategy kept");
    p.near(got[0], 0.5, 1e-6, "sweep x0");
    p.near(got[1], 0.25, 1e-6, "sweep x1");
    j->apply(b, b);
    got = host(b);
    p.near(got[0], 0.75, 1e-6, "aliased x0");
    p.near(got[1], 0.25, 1e-6, "aliased x1");
}

// Jacobi's setup / shape contracts (solver_test.cpp:75-95).
template <class T>
void jacobiContracts(Probe& p) {
    Jacobi<T> j;
    Vector<T> b(2), x(2), three(3);
    p.raises<std::logic_error>([&] { j.apply(b, x); }, "apply before setup");
    j.setup(tridiagonal<T>(2));
    const CsrMatrix<T> offDiagonalOnly(2, 2, std::vector<int>{0, 1, 2}, std::vector<int>{1, 0}, std::vector<T>{1, 1});
    p.raises<std::invalid_argument>([&] { j.setup(offDiagonalOnly); }, "missing diagonal");
    b.copyFromHost(std::vector<T>{1, 1});
    x.copyFromHost(std::vector<T>{0, 0});
    p.quiet([&] { j.apply(b, x); }, "earlier setup still usable");
    p.raises<std::invalid_argument>([&] { j.apply(three, x); }, "shape");
    j.setup(tridiagonal<T>(3));
    p.raises<std::invalid_argument>([&] { j.apply(b, x); }, "shape after a new setup");
}

// One V-cycle on tridiagonal(127), b = 1, reduces the error; scaling b scales
// the result; aliasing (solver_test.cpp:110-148).
template <class T>
void amgCycleLinearity(Probe& p) {
    const int n = 127;
    const auto a = tridiagonal<T>(n);
    AmgOptions<T> o;
    o.coarseRows_ = 8;
    Amg<T> amg(o);
    Vector<T> unset(n);
    p.raises<std::logic_error>([&] { amg.apply(unset, unset); }, "apply before setup");
    amg.setup(a);
    p.need(amg.levelCount() >= 3u, "at least three levels");
    Vector<T> b = onDevice(std::vector<T>(n, 1)), x = onDevice(std::vector<T>(n, 0));
    amg.apply(b, x);
    const auto once = host(x);
    double e2 = 0, x2 = 0;
    for (int i = 0; i < n; ++i) {
        const double exact = tridiagonalOnesSolution(n, i);
        e2 += (double(once[std::size_t(i)]) - exact) * (double(once[std::size_t(i)]) - exact);
        x2 += exact * exact;
    }
    p.expect(std::sqrt(e2 / x2) < 1, "error reduced");
    linalg::IterationContext<T>(a).scale(T(2), b);
    amg.apply(b, x);
    const auto twice = host(x);
    const double tol = isFloat<T> ? 1e-2 : 1e-9;
    for (int i = 0; i < n; ++i) p.near(twice[std::size_t(i)], 2 * once[std::size_t(i)], tol, "linear in b");
    amg.apply(b, b);
    const auto aliased = host(b);
    for (int i = 0; i < n; ++i) p.near(aliased[std::size_t(i)], twice[std::size_t(i)], tol, "aliased apply");
}

// Two-row coarsest level by LU ([1 2; 4 3] x = [5, 10] -> [1, 2]); a failed
// setup keeps the previous hierarchy; singular and level-limited coarse
// Jacobi (solver_test.cpp:149-179).
template <class T>
void amgCoarseChoices(Probe& p) {
    const std::vector<int> rp{0, 2, 4}, ci{0, 1, 0, 1};
    Amg<T> amg;
    amg.setup(CsrMatrix<T>(2, 2, rp, ci, std::vector<T>{1, 2, 4, 3}));
    Vector<T> b = onDevice(std::vector<T>{5, 10}, PinnedMemory{}), x(2, PinnedMemory{});
    amg.apply(b, x);
    auto got = host(x);
    p.near(got[0], 1, 1e-5, "LU x0");
    p.near(got[1], 2, 1e-5, "LU x1");
    p.raises<std::invalid_argument>([&] { amg.setup(CsrMatrix<T>(2, 2, rp, ci, std::vector<T>{0, 1, 1, 1})); },
                                    "zero diagonal rejected");
    p.quiet([&] { amg.apply(b, x); }, "previous hierarchy kept");
    p.near(host(x)[0], 1, 1e-5, "previous hierarchy's answer");
    p.expect(int(x.memoryType()) == int(MemoryType::Pinned), "strategy kept");
    p.quiet([&] { amg.setup(CsrMatrix<T>(2, 2, rp, ci, std::vector<T>{1, 1, 1, 1})); }, "singular accepted");
    p.expect(amg.coarseSolver() == "Jacobi (singular LU)", "singular coarse solver");
    p.quiet([&] { amg.apply(b, x); }, "singular apply");
    AmgOptions<T> o;
    o.maxLevels_ = 1;
    o.coarseRows_ = 2;
    Amg<T> oneLevel(o);
    p.quiet([&] { oneLevel.setup(tridiagonal<T>(100)); }, "level-limited setup");
    p.expect(oneLevel.levelCount() == 1u && oneLevel.coarseSolver() == "Jacobi (coarsening stopped)" &&
                 oneLevel.stopReason() == "level limit",
             "level-limited hierarchy");
}

// The hierarchy a BiCGStab builds on the device can be queried through the
// Amg it owns, as the solve-matrix program does after setup.
void ownedHierarchyQueries(Probe& p) {
    auto owned = std::make_unique<Amg<double>>(AmgOptions<double>{});
    Amg<double>* amg = owned.get();
    BiCgStab<double> solver({}, std::move(owned));
    solver.setup(tridiagonal<double>(300));
    p.expect(amg->levelCount() >= 2u && amg->stopReason() == "coarse size" && amg->coarseSolver() == "pivoted GPU LU",
             "hierarchy description");
    p.expect(amg->levelStatistics().front().rows_ == 300u, "finest rows");
    Vector<double> b = onDevice(std::vector<double>(300, 1.0)), z(300);
    amg->apply(b, z);
    p.expect(host(z)[150] > 0.0, "owned hierarchy applies");
}

// Factory smoothers: two per non-coarsest level, V-visits once each, usable
// under BiCGStab; a factory returning null is rejected (solver_test.cpp:314-324).
template <class T>
void smootherFactory(Probe& p) {
    int made = 0;
    Amg<T> amg({}, [&made]() -> std::unique_ptr<Smoother<T>> {
        ++made;
        return std::make_unique<Jacobi<T>>(3, T(0.6));
    });
    const auto a = tridiagonal<T>(255);
    amg.setup(a);
    p.expect(amg.levelCount() >= 2u && std::size_t(made) == 2 * (amg.levelCount() - 1), "two smoothers a level");
    Vector<T> b = onDevice(std::vector<T>(255, 1)), x = onDevice(std::vector<T>(255, 0));
    amg.apply(b, x);
    for (auto v : amg.lastCycleVisits()) p.expect(v == 1u, "V visits");
    BiCgStabOptions<T> o;
    o.relativeTolerance_ = isFloat<T> ? T(1e-4) : T(1e-9);
    BiCgStab<T> solver(o, std::make_unique<Amg<T>>(AmgOptions<T>{}, [] {
                           return std::unique_ptr<Smoother<T>>(std::make_unique<Jacobi<T>>(3, T(0.6)));
                       }));
    solver.setup(a);
    p.expect(solver.apply(b, x).status_ == SolveStatus::Converged, "factory hierarchy converges");
    Amg<T> broken({}, []() -> std::unique_ptr<Smoother<T>> { return {}; });
    p.raises<std::invalid_argument>([&] { broken.setup(a); }, "null smoother");
}

// V/W/F on tridiagonal(127) with coarse rows 4: visit counts, linearity to
// the precision's bound, aliasing, and BiCGStab + AMG per cycle
// (amg_native_test.cpp:329-395).
template <class T, class I>
void cycleKinds(Probe& p) {
    const int n = 127;
    const auto a = tridiagonal<T, I>(n);
    std::vector<T> b(n), c(n), bc(n);
    for (int i = 0; i < n; ++i) {
        b[std::size_t(i)] = T(std::sin(.13 * i) + .4);
        c[std::size_t(i)] = T(std::cos(.07 * i));
        bc[std::size_t(i)] = b[std::size_t(i)] + c[std::size_t(i)];
    }
    const double tol = isFloat<T> ? 2e-3 : 2e-10;
    for (auto kind : {multigrid::Cycle::V, multigrid::Cycle::W, multigrid::Cycle::F}) {
        AmgOptions<T> o;
        o.coarseRows_ = 4;
        o.cycle_ = kind;
        Amg<T, I> amg(o);
        amg.setup(a);
        p.need(amg.levelCount() >= 3u, "three levels");
        auto applied = [&](const std::vector<T>& rhs) {
            Vector<T> in = onDevice(rhs), out(n);
            amg.apply(in, out);
            return host(out);
        };
        const auto xb = applied(b), xc = applied(c), xbc = applied(bc);
        const auto visits = amg.lastCycleVisits();
        for (std::size_t d = 0; d < visits.size(); ++d) {
            const std::size_t want = kind == multigrid::Cycle::V   ? 1
                                     : kind == multigrid::Cycle::W ? std::size_t(1) << d
                                                                   : (d + 1 == visits.size() ? visits.size() : d + 2);
            p.expect(visits[d] == want, "visit count");
        }
        for (int i = 0; i < n; ++i) p.near(xbc[std::size_t(i)], xb[std::size_t(i)] + xc[std::size_t(i)], tol, "linear");
        Vector<T> alias = onDevice(b);
        amg.apply(alias, alias);
        const auto xa = host(alias);
        for (int i = 0; i < n; ++i) p.near(xa[std::size_t(i)], xb[std::size_t(i)], tol, "aliased");
        BiCgStabOptions<T> so;
        so.relativeTolerance_ = isFloat<T> ? T(5e-4) : T(1e-9);
        BiCgStab<T, I> solver(so, std::make_unique<Amg<T, I>>(o));
        solver.setup(a);
        Vector<T> rhs = onDevice(b), x = onDevice(std::vector<T>(n, 0));
        const auto r = solver.apply(rhs, x);
        p.expect(r.status_ == SolveStatus::Converged && r.relativeResidual_ <= so.relativeTolerance_,
                 "BiCGStab with this cycle");
    }
}

// One level only: the coarse solve is Jacobi(8, 0.75) exactly; cycles
// without smoothing repeat bit for bit (amg_native_test.cpp:398-424).
void fallbackAndRepeatability(Probe& p) {
    const auto a = tridiagonal<double>(100);
    Vector<double> b = onDevice(std::vector<double>(100, 1.0)), x(100), ref = onDevice(std::vector<double>(100, 0.0));
    AmgOptions<double> o;
    o.maxLevels_ = 1;
    o.coarseSweeps_ = 8;
    Amg<double> amg(o);
    amg.setup(a);
    amg.apply(b, x);
    Jacobi<double> j(8, .75);
    j.setup(a);
    j.apply(b, ref);
    p.expect(host(x) == host(ref), "coarse Jacobi equals Jacobi(8, .75)");
    o.maxLevels_ = 20;
    o.coarseRows_ = 4;
    o.preSweeps_ = o.postSweeps_ = 0;
    for (auto kind : {multigrid::Cycle::V, multigrid::Cycle::W, multigrid::Cycle::F}) {
        o.cycle_ = kind;
        Amg<double> bare(o);
        bare.setup(a);
        bare.apply(b, x);
        const auto first = host(x);
        bare.apply(b, x);
        p.expect(first == host(x), "repeatable without smoothing");
    }
}

// The device engine and the host-driven hierarchy (a factory of
// Jacobi(1, 0.75)) build the same levels and cycle to 1e-12.
void engineMatchesHostPath(Probe& p) {
    const int n = 500;
    const auto a = tridiagonal<double>(n);
    AmgOptions<double> o;
    o.coarseRows_ = 8;
    Amg<double> engine(o);
    Amg<double> hosted(o, [] { return std::unique_ptr<Smoother<double>>(std::make_unique<Jacobi<double>>(1, 0.75)); });
    engine.setup(a);
    hosted.setup(a);
    p.need(engine.levelCount() == hosted.levelCount(), "level count");
    const auto se = engine.levelStatistics(), sh = hosted.levelStatistics();
    for (std::size_t d = 0; d < se.size(); ++d)
        p.expect(se[d].rows_ == sh[d].rows_ && se[d].nonzeros_ == sh[d].nonzeros_ &&
                     se[d].aggregateSizes_ == sh[d].aggregateSizes_,
                 "level statistics");
    p.expect(engine.stopReason() == hosted.stopReason() && engine.coarseSolver() == hosted.coarseSolver(), "choices");
    std::vector<double> rhs(n);
    for (int i = 0; i < n; ++i) rhs[std::size_t(i)] = std::sin(0.01 * i) + 1;
    Vector<double> b = onDevice(rhs), xe(n), xh(n);
    engine.apply(b, xe);
    hosted.apply(b, xh);
    const auto ve = host(xe), vh = host(xh);
    double diff = 0, scale = 0;
    for (std::size_t i = 0; i < std::size_t(n); ++i) {
        diff = std::max(diff, std::abs(ve[i] - vh[i]));
        scale = std::max(scale, std::abs(vh[i]));
    }
    p.expect(diff <= 1e-12 * scale, "same cycle");
}

// ---------------------------------------------------------------- Krylov scenarios

// A solve scenario: a matrix, right-hand side and start, the options, the
// status to reach, and a check of the report and solution.
template <class T, class I>
struct Scenario {
    const char* what;
    std::function<CsrMatrix<T, I>()> matrix;
    std::vector<T> rhs, start;
    std::function<void(CgOptions<T>&)> edit;
    SolveStatus status;
    std::function<void(Probe&, const SolveResult<T>&, const std::vector<T>&)> verify;
};

template <class T, class I, class Solver>
SolveResult<T> solveOnce(Solver& s, const CsrMatrix<T, I>& a, const std::vector<T>& rhs, std::vector<T>& x) {
    s.setup(a);
    Vector<T> b = onDevice(rhs), xv = onDevice(x);
    const auto r = s.apply(b, xv);
    x = host(xv);
    return r;
}

// CG (cg_test.cpp:50-185).
template <class T>
void cgScenarios(Probe& p) {
    const int n31 = 31;
    using S = Scenario<T, int>;
    const std::vector<S> cases{
        {"identity CG on tridiagonal(31), b = 1, against the closed form and an independent residual",
         [] { return tridiagonal<T>(31); }, std::vector<T>(n31, 1), std::vector<T>(n31, 0),
         [](auto& o) { o.relativeTolerance_ = isFloat<T> ? T(2e-5) : T(1e-11); }, SolveStatus::Converged,
         [](Probe& q, const SolveResult<T>& r, const std::vector<T>& x) {
             q.expect(r.iterations_ > 0, "iterated");
             for (int i = 0; i < 31; ++i)
                 q.near(x[std::size_t(i)], tridiagonalOnesSolution(31, i), isFloat<T> ? 0.01 : 1e-8, "solution");
             const double res = tridiagonalOnesResidual(x);
             q.near(r.residualNorm_, res, isFloat<T> ? 1e-4 : 1e-12, "reported residual");
             q.expect(res <= (isFloat<T> ? 2e-5 : 1e-11) * std::sqrt(31.0), "residual within tolerance");
         }},
        {"no iterations allowed, b = 0: converged at once", [] { return tridiagonal<T>(3); }, std::vector<T>(3, 0),
         std::vector<T>(3, 0), [](auto& o) { o.maxIterations_ = 0; }, SolveStatus::Converged,
         [](Probe& q, const SolveResult<T>& r, const std::vector<T>&) {
             q.expect(r.iterations_ == 0 && r.relativeResidual_ == T(0), "zero iterations, zero residual");
         }},
        {"no iterations allowed, b = e0: residual 1", [] { return tridiagonal<T>(3); }, std::vector<T>{1, 0, 0},
         std::vector<T>(3, 0), [](auto& o) { o.maxIterations_ = 0; }, SolveStatus::MaxIterations,
         [](Probe& q, const SolveResult<T>& r, const std::vector<T>&) {
             q.expect(r.iterations_ == 0, "no iteration");
             q.near(r.residualNorm_, 1, 1e-6, "residual 1");
         }},
        {"one iteration, b = e0: residual 1/2", [] { return tridiagonal<T>(3); }, std::vector<T>{1, 0, 0},
         std::vector<T>(3, 0), [](auto& o) { o.maxIterations_ = 1; }, SolveStatus::MaxIterations,
         [](Probe& q, const SolveResult<T>& r, const std::vector<T>&) {
             q.expect(r.iterations_ == 1, "one iteration");
             q.near(r.residualNorm_, 0.5, 1e-6, "residual 1/2");
         }},
        {"b = 0 from the one-iteration iterate, absolute tolerance: relative equals absolute",
         [] { return tridiagonal<T>(3); }, std::vector<T>(3, 0), std::vector<T>{T(0.5), 0, 0},
         [](auto& o) { o.absoluteTolerance_ = T(1e-6); },
         SolveStatus::Converged,
         [](Probe& q, const SolveResult<T>& r, const std::vector<T>&) {
             q.expect(r.relativeResidual_ == r.residualNorm_, "relative = absolute");
         }},
        {"refresh every iteration on tridiag(-1, 4, -1), b = A 1", [] { return tridiagonal<T>(5, T(4)); },
         std::vector<T>{3, 2, 2, 2, 3}, std::vector<T>(5, 0),
         [](auto& o) {
             o.residualRefreshInterval_ = 1;
             o.maxIterations_ = 500;
             o.relativeTolerance_ = isFloat<T> ? T(1e-5) : T(1e-10);
         },
         SolveStatus::Converged,
         [](Probe& q, const SolveResult<T>& r, const std::vector<T>& x) {
             q.expect(r.iterations_ > 1, "several iterations");
             for (T xi : x) q.near(xi, 1, isFloat<T> ? 1e-4 : 1e-9, "ones");
         }},
        {"the exact solution as start: converged at once", [] { return tridiagonal<T>(5, T(4)); },
         std::vector<T>{3, 2, 2, 2, 3}, std::vector<T>(5, 1),
         [](auto& o) { o.relativeTolerance_ = isFloat<T> ? T(1e-5) : T(1e-10); }, SolveStatus::Converged,
         [](Probe& q, const SolveResult<T>& r, const std::vector<T>&) { q.expect(r.iterations_ == 0, "none"); }},
        {"negative curvature: breakdown", [] { return tridiagonal<T>(1, T(-1)); }, std::vector<T>{1},
         std::vector<T>{0}, [](auto&) {}, SolveStatus::Breakdown, nullptr},
        {"zero curvature: breakdown", [] { return tridiagonal<T>(1, T(0)); }, std::vector<T>{1}, std::vector<T>{0},
         [](auto&) {}, SolveStatus::Breakdown, nullptr},
        {"NaN right-hand side: non-finite", [] { return tridiagonal<T>(1); },
         std::vector<T>{std::numeric_limits<T>::quiet_NaN()}, std::vector<T>{0}, [](auto&) {}, SolveStatus::NonFinite,
         nullptr},
    };
    for (const auto& c : cases) {
        CgOptions<T> o;
        c.edit(o);
        Cg<T> solver(o);
        std::vector<T> x = c.start;
        const auto r = solveOnce<T, int, Cg<T>>(solver, c.matrix(), c.rhs, x);
        if (p.expect(r.status_ == c.status, std::string(c.what) + ": status") && c.verify) c.verify(p, r, x);
    }
}

// CG object contracts: storage kept, aliasing, snapshot, 64-bit indices,
// setup order (cg_test.cpp:80-99,172-185).
template <class T>
void cgContracts(Probe& p) {
    Cg<T, std::int64_t> wide;
    auto m = tridiagonal<T, std::int64_t>(2);
    wide.setup(m);
    m.copyValuesFromHost(std::vector<T>(4, 0));  // the solver keeps its snapshot
    Vector<T> b = onDevice(std::vector<T>{1, 0}, PinnedMemory{}), x = onDevice(std::vector<T>{T(0.1), T(-0.2)}, PinnedMemory{});
    T* const addr = x.data();
    p.need(wide.apply(b, x).status_ == SolveStatus::Converged, "2x2 solve");
    p.expect(x.data() == addr && int(x.memoryType()) == int(MemoryType::Pinned), "x kept in place");
    const auto got = host(x);
    p.near(got[0], 2.0 / 3, 1e-5, "x0");
    p.near(got[1], 1.0 / 3, 1e-5, "x1");
    p.need(wide.apply(b, b).status_ == SolveStatus::Converged, "aliased solve");
    p.near(host(b)[0], 2.0 / 3, 1e-5, "aliased x0");
    p.near(host(b)[1], 1.0 / 3, 1e-5, "aliased x1");
    Cg<T> cg;
    Vector<T> r2(2), x2(2), three(3);
    p.raises<std::logic_error>([&] { (void)cg.apply(r2, x2); }, "apply before setup");
    cg.setup(tridiagonal<T>(2));
    p.raises<std::invalid_argument>([&] { (void)cg.apply(three, x2); }, "rhs shape");
    p.raises<std::invalid_argument>([&] { (void)cg.apply(r2, three); }, "x shape");
    const CsrMatrix<T> infinite(2, 2, std::vector<int>{0, 1, 2}, std::vector<int>{0, 1},
                                std::vector<T>{1, std::numeric_limits<T>::infinity()});
    p.raises<std::invalid_argument>([&] { cg.setup(infinite); }, "non-finite matrix");
    p.raises<std::logic_error>([&] { (void)cg.apply(r2, x2); }, "failed setup leaves no solver");
    p.quiet([&] { cg.setup(tridiagonal<T>(2)); }, "setup again");
    p.raises<std::invalid_argument>([&] { cg.setup(CsrMatrix<T>{}); }, "empty matrix");
    p.raises<std::logic_error>([&] { (void)cg.apply(r2, x2); }, "after a rejected setup");
}

// BiCGStab (solver_test.cpp:180-345).
template <class T>
void bicgstabScenarios(Probe& p) {
    {  // AMG default, tridiagonal(63), b = 1: closed form, true residual, re-apply, aliasing
        const int n = 63;
        BiCgStabOptions<T> o;
        o.relativeTolerance_ = isFloat<T> ? T(2e-4) : T(1e-9);
        BiCgStab<T> s(o);
        s.setup(tridiagonal<T>(n));
        const std::vector<T> ones(n, 1);
        Vector<T> b = onDevice(ones, PinnedMemory{}), x = onDevice(std::vector<T>(n, 0), PinnedMemory{});
        T* const addr = x.data();
        const auto r = s.apply(b, x);
        p.need(r.status_ == SolveStatus::Converged && r.iterations_ < 100, "AMG BiCGStab converges");
        p.expect(x.data() == addr && int(x.memoryType()) == int(MemoryType::Pinned), "x kept in place");
        const auto got = host(x);
        for (int i = 0; i < n; ++i) {
            const double e = tridiagonalOnesSolution(n, i);
            p.near(got[std::size_t(i)], e, e * (isFloat<T> ? 1e-3 : 1e-8), "closed form");
        }
        p.near(r.residualNorm_, tridiagonalOnesResidual(got), isFloat<T> ? 1e-3 : 1e-10, "true residual reported");
        p.expect(r.relativeResidual_ <= o.relativeTolerance_, "within tolerance");
        const auto again = s.apply(b, x);
        p.expect(again.status_ == SolveStatus::Converged && again.iterations_ == 0, "converged start");
        b.copyFromHost(ones);
        p.expect(s.apply(b, b).status_ == SolveStatus::Converged, "aliased");
    }
    {  // identity, nonsymmetric tridiag(-1, 3, -0.5) of order 80, refresh every 3
        const int n = 80;
        std::vector<T> exact(n), b(n);
        for (int i = 0; i < n; ++i) exact[std::size_t(i)] = T(1) + T(i % 7) / T(7);
        for (int i = 0; i < n; ++i)
            b[std::size_t(i)] = 3 * exact[std::size_t(i)] - (i ? exact[std::size_t(i) - 1] : 0) -
                                T(0.5) * (i + 1 < n ? exact[std::size_t(i) + 1] : 0);
        BiCgStabOptions<T> o;
        o.residualRefreshInterval_ = 3;
        o.relativeTolerance_ = isFloat<T> ? T(1e-5) : T(1e-11);
        std::unique_ptr<Solver<T>> s = std::make_unique<BiCgStab<T>>(o, std::make_unique<Identity<T>>());
        s->setup(tridiagonal<T>(n, T(3), true));
        Vector<T> rhs = onDevice(b), x = onDevice(std::vector<T>(n, T(0.2)));
        p.need(s->apply(rhs, x).status_ == SolveStatus::Converged, "nonsymmetric converges");
        const auto got = host(x);
        for (int i = 0; i < n; ++i) p.near(got[std::size_t(i)], exact[std::size_t(i)], isFloat<T> ? 1e-4 : 1e-9, "x");
    }
    {  // zero right-hand side, then a start that is not zero, absolute tolerance
        BiCgStabOptions<T> o;
        o.absoluteTolerance_ = isFloat<T> ? T(1e-5) : T(1e-12);
        BiCgStab<T> s(o, std::make_unique<Identity<T>>());
        s.setup(tridiagonal<T>(5));
        Vector<T> b = onDevice(std::vector<T>(5, 0)), x = onDevice(std::vector<T>(5, 0));
        auto r = s.apply(b, x);
        p.expect(r.status_ == SolveStatus::Converged && r.iterations_ == 0, "b = 0");
        x.copyFromHost(std::vector<T>(5, 1));
        r = s.apply(b, x);
        p.expect(r.status_ == SolveStatus::Converged && r.residualNorm_ <= o.absoluteTolerance_ &&
                     r.relativeResidual_ == r.residualNorm_,
                 "b = 0 from ones");
    }
    {  // iteration limits, breakdown on the zero matrix, NaN right-hand side
        const auto a = tridiagonal<T>(3, T(3), true);
        const std::vector<T> b{1, 2, 3}, zeros(3, 0);
        Vector<T> rhs = onDevice(b), x = onDevice(zeros);
        BiCgStabOptions<T> o;
        o.maxIterations_ = 0;
        BiCgStab<T> none(o, std::make_unique<Identity<T>>());
        none.setup(a);
        auto r = none.apply(rhs, x);
        p.expect(r.status_ == SolveStatus::MaxIterations && r.iterations_ == 0, "no iterations");
        o.maxIterations_ = 1;
        o.relativeTolerance_ = 0;
        BiCgStab<T> one(o, std::make_unique<Identity<T>>());
        one.setup(a);
        r = one.apply(rhs, x);
        p.expect(r.status_ == SolveStatus::MaxIterations && r.iterations_ == 1, "one iteration");
        BiCgStab<T> s({}, std::make_unique<Identity<T>>());
        s.setup(CsrMatrix<T>(3, 3, std::vector<int>{0, 0, 0, 0}, std::vector<int>{}, std::vector<T>{}));
        x.copyFromHost(zeros);
        r = s.apply(rhs, x);
        p.expect(r.status_ == SolveStatus::Breakdown, "zero matrix breaks down");
        p.near(r.residualNorm_, std::sqrt(14.0), 1e-5, "residual |b|");
        auto nan = b;
        nan[1] = std::numeric_limits<T>::quiet_NaN();
        rhs.copyFromHost(nan);
        p.expect(s.apply(rhs, x).status_ == SolveStatus::NonFinite, "NaN right-hand side");
    }
    {  // life cycle with 64-bit indices and a failed setup
        BiCgStab<T, std::int64_t> s;
        Vector<T> b = onDevice(std::vector<T>{1, 0}), x = onDevice(std::vector<T>{0, 0}), three(3);
        p.raises<std::logic_error>([&] { (void)s.apply(b, x); }, "apply before setup");
        s.setup(tridiagonal<T, std::int64_t>(2));
        p.expect(s.apply(b, x).status_ == SolveStatus::Converged, "2x2");
        p.raises<std::invalid_argument>([&] { (void)s.apply(three, x); }, "shape");
        p.raises<std::invalid_argument>([&] { s.setup(CsrMatrix<T, std::int64_t>{}); }, "empty matrix");
        p.raises<std::logic_error>([&] { (void)s.apply(b, x); }, "after a rejected setup");
        s.setup(tridiagonal<T, std::int64_t>(2));
        p.expect(s.apply(b, x).status_ == SolveStatus::Converged, "set up again");
    }
    {  // every memory strategy in and out
        BiCgStab<T> s;
        s.setup(tridiagonal<T>(3));
        const std::vector<T> b{1, 0, 1};
        for (const auto& in : allStrategies()) {
            const Vector<T> rhs = onDevice(b, in);
            for (const auto& out : allStrategies()) {
                Vector<T> x = onDevice(std::vector<T>(3, 0), out);
                const auto type = x.memoryType();
                T* const addr = x.data();
                p.need(s.apply(rhs, x).status_ == SolveStatus::Converged, "strategy pair");
                p.expect(x.memoryType() == type && x.data() == addr, "x kept");
                for (T v : host(x)) p.near(v, 1, 1e-5, "ones");
            }
        }
    }
}

// ---------------------------------------------------------------- linalg free functions

namespace la = thermofem::linalg;

template <class T>
double tolerance() {
    return isFloat<T> ? 1e-5 : 1e-12;
}
template <class T>
void sameValues(Probe& p, const std::vector<T>& got, std::initializer_list<double> want, const char* what) {
    if (!p.expect(got.size() == want.size(), std::string(what) + ": size")) return;
    std::size_t i = 0;
    for (double w : want) p.near(double(got[i++]), w, tolerance<T>(), what);
}
// the matrix a CSR stands for (duplicates summed), row-major
template <class T, class I>
std::vector<double> dense(const CsrMatrix<T, I>& m) {
    const auto rp = m.rowOffsets().hostCopy();
    const auto ci = m.columnIndices().hostCopy();
    const auto v = m.values().hostCopy();
    std::vector<double> d(m.rowCount() * m.columnCount(), 0.0);
    for (std::size_t r = 0; r < m.rowCount(); ++r)
        for (auto k = rp[r]; k < rp[r + 1]; ++k) d[r * m.columnCount() + std::size_t(ci[std::size_t(k)])] += v[std::size_t(k)];
    return d;
}
template <class T, class I>
void sameMatrix(Probe& p, const CsrMatrix<T, I>& m, std::initializer_list<double> want, const char* what) {
    sameValues<double>(p, dense(m), want, what);
}

// Vector operators and reductions over every strategy pair
// (linear_algebra_test.cpp:97-131).
template <class T>
void vectorAlgebra(Probe& p) {
    const std::vector<T> xv{1, -2, 3}, yv{4, 5, -6};
    for (const auto& ls : allStrategies())
        for (const auto& rs : allStrategies()) {
            const Vector<T> x(std::span<const T>(xv), ls), y(std::span<const T>(yv), rs);
            const auto sum = x + y;
            p.expect(sum.memoryStrategy().index() == ls.index(), "left operand's strategy");
            sameValues<T>(p, host(sum), {5, 3, -3}, "x + y");
            sameValues<T>(p, host(x - y), {-3, -7, 9}, "x - y");
            sameValues<T>(p, host(-x), {-1, 2, -3}, "-x");
            sameValues<T>(p, host(+x), {1, -2, 3}, "+x");
            sameValues<T>(p, host(T(2) * x), {2, -4, 6}, "2 x");
            sameValues<T>(p, host(x * T(2)), {2, -4, 6}, "x 2");
            sameValues<T>(p, host(x / T(2)), {0.5, -1, 1.5}, "x / 2");
            p.near(double(x * y), -24, 1e-5, "x . y");
            p.near(double(la::dot(x, y)), -24, 1e-5, "dot");
            p.near(double(la::norm1(x)), 6, 1e-5, "norm1");
            p.near(double(la::norm2(x)), std::sqrt(14.0), 1e-5, "norm2");
            sameValues<T>(p, host(la::axpy(T(2), x, y)), {6, 1, 0}, "axpy");
            auto acc = x;
            acc += acc;
            acc -= x;
            acc *= T(2);
            acc /= T(2);
            sameValues<T>(p, host(acc), {1, -2, 3}, "compound assignments");
            p.expect(acc.memoryStrategy().index() == ls.index(), "strategy kept");
            p.raises<std::domain_error>([&] { (void)(x / T(0)); }, "division by zero");
            sameValues<T>(p, host(x), {1, -2, 3}, "operand unchanged");
        }
}

// Dense products, transpose, sums, norm (linear_algebra_test.cpp:133-170).
template <class T>
void denseAlgebra(Probe& p) {
    for (const auto& s : allStrategies()) {
        DenseMatrix<T> a(2, 3, s), b(3, 2, PinnedMemory{});
        a.copyFromHost(std::vector<T>{1, 2, 3, 4, 5, 6});
        b.copyFromHost(std::vector<T>{7, 8, 9, 10, 11, 12});
        const Vector<T> x(std::span<const T>(std::vector<T>{1, 2, 3}), PinnedMemory{});
        sameValues<T>(p, host(a * x), {14, 32}, "A x");
        const auto ab = a * b;
        p.expect(ab.rowCount() == 2u && ab.columnCount() == 2u && ab.values().memoryStrategy().index() == s.index(),
                 "A B shape and strategy");
        sameValues<T>(p, ab.values().hostCopy(), {58, 64, 139, 154}, "A B");
        const auto at = la::transpose(a);
        p.expect(at.rowCount() == 3u && at.columnCount() == 2u, "transpose shape");
        sameValues<T>(p, at.values().hostCopy(), {1, 4, 2, 5, 3, 6}, "transpose");
        sameValues<T>(p, (a + a).values().hostCopy(), {2, 4, 6, 8, 10, 12}, "A + A");
        sameValues<T>(p, (a - a).values().hostCopy(), {0, 0, 0, 0, 0, 0}, "A - A");
        sameValues<T>(p, (-a).values().hostCopy(), {-1, -2, -3, -4, -5, -6}, "-A");
        auto c = a;
        c += a;
        c -= a;
        c *= T(2);
        c /= T(2);
        sameValues<T>(p, c.values().hostCopy(), {1, 2, 3, 4, 5, 6}, "compound assignments");
        c *= b;
        sameValues<T>(p, c.values().hostCopy(), {58, 64, 139, 154}, "in-place product");
        p.near(double(la::frobeniusNorm(a)), std::sqrt(91.0), 1e-5, "Frobenius norm");
        p.raises<std::domain_error>([&] { (void)(a / T(0)); }, "division by zero");
    }
}

// Sparse products, sums, transpose with a duplicate column in row 0
// (linear_algebra_test.cpp:176-220).
template <class T, class I>
void sparseAlgebra(Probe& p, const MemoryStrategy& s) {
    const CsrMatrix<T, I> a(2, 3, std::vector<I>{0, 3, 4}, std::vector<I>{2, 0, 2, 1}, std::vector<T>{1, 2, 3, 5}, s);
    const CsrMatrix<T, I> b(3, 2, std::vector<I>{0, 1, 2, 4}, std::vector<I>{0, 1, 1, 0}, std::vector<T>{1, 2, 4, 3},
                            PinnedMemory{});
    const Vector<T> x(std::span<const T>(std::vector<T>{1, 2, 3}), PinnedMemory{});
    sameValues<T>(p, host(a * x), {14, 10}, "A x");
    DenseMatrix<T> d(3, 2, PinnedMemory{});
    d.copyFromHost(std::vector<T>{1, 0, 0, 2, 3, 4});
    sameValues<T>(p, (a * d).values().hostCopy(), {14, 16, 0, 10}, "A D");
    const auto ab = a * b;
    sameMatrix(p, ab, {14, 16, 0, 10}, "A B");
    p.expect(ab.values().memoryStrategy().index() == s.index(), "product strategy");
    sameMatrix(p, a + a, {4, 0, 8, 0, 10, 0}, "A + A");
    sameMatrix(p, a - a, {0, 0, 0, 0, 0, 0}, "A - A");
    sameMatrix(p, la::transpose(a), {2, 0, 0, 5, 4, 0}, "transpose");
    sameMatrix(p, -a, {-2, 0, -4, 0, -5, 0}, "-A");
    sameMatrix(p, a * T(2), {4, 0, 8, 0, 10, 0}, "2 A");
    p.near(double(la::frobeniusNorm(a)), std::sqrt(45.0), 1e-5, "Frobenius norm");
    auto c = a;
    c += c;
    c -= a;
    c *= T(2);
    c /= T(2);
    sameMatrix(p, c, {2, 0, 4, 0, 5, 0}, "compound assignments");
    c *= b;
    sameMatrix(p, c, {14, 16, 0, 10}, "in-place product");
    p.raises<std::domain_error>([&] { (void)(a / T(0)); }, "division by zero");
}

// Zero-size operands (linear_algebra_test.cpp:222-254).
template <class T>
void emptyOperands(Probe& p) {
    DenseMatrix<T> a(2, 0, PinnedMemory{}), b(0, 3);
    sameValues<T>(p, (a * b).values().hostCopy(), {0, 0, 0, 0, 0, 0}, "2x0 times 0x3");
    const Vector<T> empty;
    sameValues<T>(p, host(a * empty), {0, 0}, "2x0 times empty");
    p.expect(double(la::dot(empty, empty)) == 0 && double(la::norm1(empty)) == 0 && double(la::norm2(empty)) == 0,
             "empty reductions");
    p.expect(la::transpose(a).rowCount() == 0u, "transpose of 2x0");
    const CsrMatrix<T> none(2, 3, std::vector<int>{0, 0, 0}, std::vector<int>{}, std::vector<T>{}, PinnedMemory{});
    const Vector<T> x(std::span<const T>(std::vector<T>{1, 2, 3}));
    sameValues<T>(p, host(none * x), {0, 0}, "no entries times x");
    sameMatrix(p, none + none, {0, 0, 0, 0, 0, 0}, "sum without entries");
    sameMatrix(p, la::transpose(none), {0, 0, 0, 0, 0, 0}, "transpose without entries");
    const CsrMatrix<T> right(3, 2, std::vector<int>{0, 0, 0, 0}, std::vector<int>{}, std::vector<T>{});
    sameMatrix(p, none * right, {0, 0, 0, 0}, "product without entries");
    const CsrMatrix<T> dflt;
    p.expect((dflt + dflt).rowCount() == 0u && la::transpose(dflt).rowCount() == 0u, "default matrices");
    const CsrMatrix<T> one(2, 3, std::vector<int>{0, 1, 1}, std::vector<int>{1}, std::vector<T>{4});
    sameMatrix(p, none + one, {0, 4, 0, 0, 0, 0}, "empty + one");
    sameMatrix(p, one + none, {0, 4, 0, 0, 0, 0}, "one + empty");
    const CsrMatrix<T> disjoint(3, 2, std::vector<int>{0, 1, 1, 1}, std::vector<int>{0}, std::vector<T>{2});
    sameMatrix(p, one * disjoint, {0, 0, 0, 0}, "product with no structural nonzero");
}

// Shape errors of the free functions (linear_algebra_test.cpp:256-269).
void algebraShapeErrors(Probe& p) {
    const Vector<double> x;
    p.raises<std::domain_error>([&] { (void)(x / 0.0); }, "empty vector / 0");
    const DenseMatrix<double> a(2, 0), b(3, 0);
    p.raises<std::invalid_argument>([&] { (void)(a + b); }, "dense sum shape");
    p.raises<std::invalid_argument>([&] { (void)(a * b); }, "dense product shape");
    p.raises<std::invalid_argument>([&] { (void)(b * DenseMatrix<double>(1, 0)); }, "dense product inner size");
    p.raises<std::length_error>(
        [] { la::detail::librarySize(static_cast<std::size_t>(std::numeric_limits<int>::max()) + 1); },
        "size past the library's int range");
    const CsrMatrix<double> sparse;
    p.raises<std::invalid_argument>([&] { (void)(sparse * b); }, "sparse times dense shape");
}

template <template <class> class F>
void bothPrecisions(Probe& p) {
    F<float>::run(p);
    F<double>::run(p);
}
#define TF_TYPED(name)                               \
    template <class T>                               \
    struct name##_ {                                 \
        static void run(Probe& p) { name<T>(p); }    \
    };
TF_TYPED(operationsOnTridiagonal4)
TF_TYPED(duplicateDiagonalAndCoarseLu)
TF_TYPED(contextReductions)
TF_TYPED(contextSnapshot)
TF_TYPED(jacobiSweeps)
TF_TYPED(jacobiContracts)
TF_TYPED(amgCycleLinearity)
TF_TYPED(amgCoarseChoices)
TF_TYPED(smootherFactory)
TF_TYPED(cgScenarios)
TF_TYPED(cgContracts)
TF_TYPED(bicgstabScenarios)
TF_TYPED(vectorAlgebra)
TF_TYPED(denseAlgebra)
TF_TYPED(emptyOperands)
#undef TF_TYPED

}  // namespace

int main(int argc, char** argv) {
    const std::vector<probe::Suite> suites{
        {"option_contracts", false, optionContracts},
        {"csr_contracts", false, csrContracts},
        {"matrix_market", false, matrixMarket},
        {"memory_strategies", true, memoryStrategies},
        {"csr_storage", true, csrStorage},
        {"operations_on_tridiagonal4", true, bothPrecisions<operationsOnTridiagonal4_>},
        {"duplicate_diagonal_and_coarse_lu", true, bothPrecisions<duplicateDiagonalAndCoarseLu_>},
        {"context_reductions", true, bothPrecisions<contextReductions_>},
        {"context_snapshot", true, bothPrecisions<contextSnapshot_>},
        {"coarse_factor_scales", true, coarseFactorScales},
        {"transfers", true, transfers},
        {"jacobi_sweeps", true, bothPrecisions<jacobiSweeps_>},
        {"jacobi_contracts", true, bothPrecisions<jacobiContracts_>},
        {"amg_cycle_linearity", true, bothPrecisions<amgCycleLinearity_>},
        {"amg_coarse_choices", true, bothPrecisions<amgCoarseChoices_>},
        {"owned_hierarchy_queries", true, ownedHierarchyQueries},
        {"smoother_factory", true, bothPrecisions<smootherFactory_>},
        {"cycle_kinds_double", true, cycleKinds<double, int>},
        {"cycle_kinds_float_wide", true, cycleKinds<float, std::int64_t>},
        {"fallback_and_repeatability", true, fallbackAndRepeatability},
        {"engine_matches_host_path", true, engineMatchesHostPath},
        {"cg_scenarios", true, bothPrecisions<cgScenarios_>},
        {"cg_contracts", true, bothPrecisions<cgContracts_>},
        {"bicgstab_scenarios", true, bothPrecisions<bicgstabScenarios_>},
        {"vector_algebra", true, bothPrecisions<vectorAlgebra_>},
        {"dense_algebra", true, bothPrecisions<denseAlgebra_>},
        {"sparse_algebra", true,
         [](Probe& p) {
             for (const auto& s : allStrategies()) {
                 sparseAlgebra<float, std::int32_t>(p, s);
                 sparseAlgebra<float, std::int64_t>(p, s);
                 sparseAlgebra<double, std::int32_t>(p, s);
                 sparseAlgebra<double, std::int64_t>(p, s);
             }
         }},
        {"empty_operands", true, bothPrecisions<emptyOperands_>},
        {"algebra_shape_errors", true, algebraShapeErrors},
    };
    return probe::run(suites, argc, argv);
}
