// multigrid:: pieces of the C++20 drop-in API: the cycle scheduler
// (include/Multigrid/Cycle.cuh), GpuAggregation and the aggregate transfers /
// coarse LU (include/Multigrid/GpuAggregation.cuh:13-101), over the C ABI.
#pragma once

#include <array>
#include <concepts>

#include "LinearAlgebra.hpp"

namespace thermofem::multigrid {

enum class Cycle { V, W, F };

template <class O>
concept CycleOperations = requires(O& operations, std::size_t level) {
    { operations.isCoarsest(level) } -> std::convertible_to<bool>;
    operations.begin(level);
    operations.clear(level);
    operations.coarseSolve(level);
    operations.preSmooth(level);
    operations.postSmooth(level);
    operations.restrictRhs(level);
    operations.restrictResidual(level);
    operations.prolong(level);
};

// V: clear (zero guess), pre-smooth, restrict the residual, recurse with a
// zero guess, prolong, post-smooth.  W visits the coarse level twice, the
// second time keeping the first result.  F: clear, restrict the rhs, F on the
// coarse level, prolong, then one V at this level without clearing.  The
// coarsest level only solves.  (Cycle.cuh:35-63.)
template <CycleOperations O>
void executeCycle(Cycle kind, std::size_t level, O& operations, bool zeroInitial = true) {
    if (kind != Cycle::V && kind != Cycle::W && kind != Cycle::F) throw std::invalid_argument("Unknown AMG cycle");
    operations.begin(level);
    if (zeroInitial || kind == Cycle::F) operations.clear(level);
    if (operations.isCoarsest(level)) {
        operations.coarseSolve(level);
        return;
    }
    if (kind == Cycle::F) {
        operations.restrictRhs(level);
        executeCycle(Cycle::F, level + 1, operations, true);
        operations.prolong(level);
        executeCycle(Cycle::V, level, operations, false);
        return;
    }
    operations.preSmooth(level);
    operations.restrictResidual(level);
    executeCycle(kind, level + 1, operations, true);
    if (kind == Cycle::W) executeCycle(kind, level + 1, operations, false);
    operations.prolong(level);
    operations.postSmooth(level);
}

// Fine-to-coarse map and member CSR (GpuAggregation.cuh:13-17).
struct AggregateTransfer {
    Buffer<int> aggregate_, offsets_, members_;
    int fineRows_{}, coarseRows_{};
    std::array<int, 4> histogram_{};
};

// Dense pivoted LU of the coarsest matrix (GpuAggregation.cuh:20-25), plus
// the explicit inverse the cycle applies.
struct CoarseFactor {
    Buffer<double> lu_, inverse_;
    Buffer<int> pivots_;
    int rows_{};
    bool factored_{}, singular_{};
};

namespace detail {
using thermofem::detail::check;
inline Buffer<int> adoptInt(int* p, std::size_t n) {
    Buffer<int> b(n);
    if (n) check(tf_memcpy(b.data(), p, n * sizeof(int)), "adopt");
    (void)tf_free(TF_MEM_GLOBAL, p);
    return b;
}
inline Buffer<double> adoptDouble(double* p, std::size_t n) {
    Buffer<double> b(n);
    if (n) check(tf_memcpy(b.data(), p, n * sizeof(double)), "adopt");
    (void)tf_free(TF_MEM_GLOBAL, p);
    return b;
}
inline CsrMatrix<double> adoptCsr(int rows, int* rp, int* ci, double* v, int64_t nnz) {
    return CsrMatrix<double>::fromBuffers(std::size_t(rows), std::size_t(rows), adoptInt(rp, std::size_t(rows) + 1),
                                          adoptInt(ci, std::size_t(nnz)), adoptDouble(v, std::size_t(nnz)));
}
}  // namespace detail

// The aggregation setup stages (GpuAggregation.cuh:34-69).  compactKeys is
// accepted for interface compatibility; merges here use no sort keys.
class GpuAggregation {
public:
    explicit GpuAggregation(bool compactKeys = false)
        : compact_(compactKeys), device_(thermofem::detail::currentDevice()), ctx_(device_) {}

    CsrMatrix<double> canonical(const CsrMatrix<double>& a) {
        require(a);
        int* ci = nullptr;
        double* v = nullptr;
        int64_t nnz = 0;
        Buffer<int> rp(a.rowCount() + 1);
        detail::check(tf_csr_canonicalize(ctx_.get(), int(a.rowCount()), int64_t(a.nonZeroCount()),
                                          a.rowOffsets().deviceData(), a.columnIndices().deviceData(),
                                          a.values().deviceData(), rp.data(), &ci, &v, &nnz),
                      "AMG canonical");
        return CsrMatrix<double>::fromBuffers(a.rowCount(), a.rowCount(), std::move(rp),
                                              detail::adoptInt(ci, std::size_t(nnz)),
                                              detail::adoptDouble(v, std::size_t(nnz)));
    }
    void validate(const CsrMatrix<double>& a) {
        require(a);
        int bad = 0;
        detail::check(tf_agg_validate_rowmax(ctx_.get(), int(a.rowCount()), a.rowOffsets().deviceData(),
                                             a.columnIndices().deviceData(), a.values().deviceData(), nullptr, &bad),
                      "AMG validation");
        if (bad)
            throw std::invalid_argument("Native AMG: missing/zero diagonal, non-finite value or reciprocal in hierarchy");
    }
    CsrMatrix<double> strength(const CsrMatrix<double>& a, double threshold) {
        if (!std::isfinite(threshold) || threshold < 0 || threshold > 1)
            throw std::invalid_argument("Invalid strength threshold");
        require(a);
        int *rp = nullptr, *ci = nullptr;
        double* w = nullptr;
        int64_t nnz = 0;
        detail::check(tf_agg_strength(ctx_.get(), int(a.rowCount()), int64_t(a.nonZeroCount()),
                                      a.rowOffsets().deviceData(), a.columnIndices().deviceData(),
                                      a.values().deviceData(), threshold, &rp, &ci, &w, &nnz),
                      "AMG strength");
        return detail::adoptCsr(int(a.rowCount()), rp, ci, w, nnz);
    }
    AggregateTransfer aggregate(const CsrMatrix<double>& graph, int rounds, bool absorbSingletons) {
        if (rounds < 1 || rounds > 8 || !graph.rowCount()) throw std::invalid_argument("Invalid matching configuration");
        int *map = nullptr, *offsets = nullptr, *members = nullptr;
        int coarse = 0;
        int64_t hist[4] = {0, 0, 0, 0};
        const int n = int(graph.rowCount());
        detail::check(tf_agg_aggregate(ctx_.get(), n, graph.rowOffsets().deviceData(),
                                       graph.columnIndices().deviceData(), graph.values().deviceData(), rounds,
                                       absorbSingletons ? 1 : 0, &map, &offsets, &members, &coarse, hist),
                      "AMG aggregation");
        AggregateTransfer t;
        t.fineRows_ = n;
        t.coarseRows_ = coarse;
        t.aggregate_ = detail::adoptInt(map, std::size_t(n));
        t.offsets_ = detail::adoptInt(offsets, std::size_t(coarse) + 1);
        t.members_ = detail::adoptInt(members, std::size_t(n));
        for (int k = 0; k < 4; ++k) t.histogram_[std::size_t(k)] = int(hist[k]);
        return t;
    }
    CsrMatrix<double> galerkin(const CsrMatrix<double>& a, const AggregateTransfer& t) {
        if (a.rowCount() != std::size_t(t.fineRows_) || t.coarseRows_ < 1)
            throw std::invalid_argument("AMG transfer dimension mismatch");
        int *rp = nullptr, *ci = nullptr;
        double* v = nullptr;
        int64_t nnz = 0;
        detail::check(tf_galerkin(ctx_.get(), int(a.rowCount()), a.rowOffsets().deviceData(),
                                  a.columnIndices().deviceData(), a.values().deviceData(), t.aggregate_.data(),
                                  t.coarseRows_, t.offsets_.data(), t.members_.data(), &rp, &ci, &v, &nnz),
                      "AMG Galerkin");
        return detail::adoptCsr(t.coarseRows_, rp, ci, v, nnz);
    }
    [[nodiscard]] bool compactSortKeys() const noexcept { return compact_; }

private:
    void require(const CsrMatrix<double>& a) const {
        if (!a.rowCount() || a.rowCount() >= std::size_t(std::numeric_limits<int>::max()) ||
            a.rowCount() != a.columnCount() || a.nonZeroCount() > std::size_t(std::numeric_limits<int>::max() / 2) ||
            a.rowOffsets().deviceId() != device_)
            throw std::invalid_argument("AMG requires nonempty square int32-capacity CSR on workspace device");
    }
    bool compact_;
    int device_;
    linalg::detail::Context ctx_;
};

template <linalg::Scalar T>
void validateTransfer(const AggregateTransfer& t, const Vector<T>& f, const Vector<T>& c) {
    if (f.size() != std::size_t(t.fineRows_) || c.size() != std::size_t(t.coarseRows_) ||
        f.deviceId() != t.aggregate_.deviceId() || c.deviceId() != t.aggregate_.deviceId() ||
        f.memoryType() != MemoryType::Global || c.memoryType() != MemoryType::Global || f.data() == c.data())
        throw std::invalid_argument("Invalid aggregate transfer vectors");
}

// coarse[c] = sum of fine over the members of c, ascending (GpuAggregation.cu:950-960,1050-1058).
template <linalg::Scalar T>
void restrictAggregates(const AggregateTransfer& t, const Vector<T>& fine, Vector<T>& coarse) {
    validateTransfer(t, fine, coarse);
    thermofem::detail::DeviceScope scope(t.aggregate_.deviceId());
    linalg::detail::Context ctx(t.aggregate_.deviceId());
    if constexpr (std::same_as<T, double>)
        detail::check(tf_restrict_f64(ctx.get(), t.coarseRows_, t.offsets_.data(), t.members_.data(),
                                      fine.deviceData(), coarse.deviceData()),
                      "AMG restriction");
    else
        detail::check(tf_restrict_f32(ctx.get(), t.coarseRows_, t.offsets_.data(), t.members_.data(),
                                      fine.deviceData(), coarse.deviceData()),
                      "AMG restriction");
    ctx.sync();
}

// fine[i] += coarse[map[i]] (GpuAggregation.cu:970-976,1059-1067).
template <linalg::Scalar T>
void prolongAggregates(const AggregateTransfer& t, const Vector<T>& coarse, Vector<T>& fine) {
    validateTransfer(t, fine, coarse);
    thermofem::detail::DeviceScope scope(t.aggregate_.deviceId());
    linalg::detail::Context ctx(t.aggregate_.deviceId());
    if constexpr (std::same_as<T, double>)
        detail::check(tf_prolong_f64(ctx.get(), t.fineRows_, t.aggregate_.data(), coarse.deviceData(),
                                     fine.deviceData()),
                      "AMG prolongation");
    else
        detail::check(tf_prolong_f32(ctx.get(), t.fineRows_, t.aggregate_.data(), coarse.deviceData(),
                                     fine.deviceData()),
                      "AMG prolongation");
    ctx.sync();
}

// Pivoted dense LU (GpuAggregation.cu:827-893,1068-1086); order 1..1024.
inline CoarseFactor factorCoarse(const CsrMatrix<double>& a) {
    if (!a.rowCount() || a.rowCount() > 1024 || a.rowCount() != a.columnCount())
        throw std::invalid_argument("GPU LU requires order 1..1024");
    thermofem::detail::DeviceScope scope(a.rowOffsets().deviceId());
    linalg::detail::Context ctx(a.rowOffsets().deviceId());
    CoarseFactor f;
    f.rows_ = int(a.rowCount());
    f.lu_ = Buffer<double>(a.rowCount() * a.rowCount());
    f.inverse_ = Buffer<double>(a.rowCount() * a.rowCount());
    f.pivots_ = Buffer<int>(a.rowCount());
    int singular = 0;
    detail::check(tf_coarse_lu_factor_f64(ctx.get(), f.rows_, a.rowOffsets().deviceData(),
                                          a.columnIndices().deviceData(), a.values().deviceData(), f.lu_.data(),
                                          f.pivots_.data(), f.inverse_.data(), &singular),
                  "AMG coarse factor");
    f.singular_ = singular != 0;
    f.factored_ = !f.singular_;
    return f;
}

// Permuted forward/backward substitution in FP64; rhs and x may alias
// (GpuAggregation.cu:904-939,1087-1098).
template <linalg::Scalar T>
void solveCoarse(const CoarseFactor& factor, const Vector<T>& rhs, Vector<T>& x) {
    if (!factor.factored_ || rhs.size() != std::size_t(factor.rows_) || x.size() != rhs.size() ||
        rhs.deviceId() != factor.lu_.deviceId() || x.deviceId() != factor.lu_.deviceId() ||
        rhs.memoryType() != MemoryType::Global || x.memoryType() != MemoryType::Global)
        throw std::invalid_argument("Invalid GPU coarse solve");
    thermofem::detail::DeviceScope scope(factor.lu_.deviceId());
    linalg::detail::Context ctx(factor.lu_.deviceId());
    if constexpr (std::same_as<T, double>)
        detail::check(tf_coarse_lu_solve_f64(ctx.get(), factor.rows_, factor.lu_.data(), factor.pivots_.data(),
                                             rhs.deviceData(), x.deviceData()),
                      "AMG coarse solve");
    else
        detail::check(tf_coarse_lu_solve_f32(ctx.get(), factor.rows_, factor.lu_.data(), factor.pivots_.data(),
                                             rhs.deviceData(), x.deviceData()),
                      "AMG coarse solve");
    ctx.sync();
}

// {T, I} CSR <-> the native {double, int} CSR (GpuAggregation.cu:1101-1136).
template <linalg::Scalar T, CsrIndex I>
CsrMatrix<double> toNativeMatrix(const CsrMatrix<T, I>& a) {
    linalg::detail::Context ctx(a.rowOffsets().deviceId());
    auto n = linalg::detail::toNative(ctx, a, false);
    return CsrMatrix<double>::fromBuffers(std::size_t(n.rows), std::size_t(n.rows), std::move(n.rp), std::move(n.ci),
                                          std::move(n.v));
}
template <linalg::Scalar T, CsrIndex I>
CsrMatrix<T, I> fromNativeMatrix(const CsrMatrix<double>& a) {
    linalg::detail::Context ctx(a.rowOffsets().deviceId());
    linalg::detail::NativeCsr n;
    n.rows = int(a.rowCount());
    n.nnz = int64_t(a.nonZeroCount());
    n.rp = a.rowOffsets();
    n.ci = a.columnIndices();
    n.v = a.values();
    return linalg::detail::fromNative<T, I>(ctx, n);
}

}  // namespace thermofem::multigrid
