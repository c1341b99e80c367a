// Smoother / Preconditioner / Solver interfaces, Jacobi, Identity, Amg, Cg and
// BiCgStab of the C++20 drop-in API (include/Solver/Interfaces.cuh,
// include/Smoother/Jacobi.cuh, include/Preconditioner/{Amg,Identity}.cuh,
// include/Solver/{Cg,BiCgStab}.cuh).
//
// Amg without a smoother factory, and Cg / BiCgStab over double / int32
// with that Amg or with Identity, run on the library's engine: fused gfx950
// kernels, device-resident Krylov scalars and a hipGraph per iteration.  Any
// other combination (user smoothers or preconditioners, float, 64-bit
// indices) runs the reference's host-driven algorithm over the same kernels.
#pragma once

#include <algorithm>
#include <functional>
#include <memory>
#include <optional>
#include <string>

#include "Multigrid.hpp"

namespace thermofem {

enum class SolveStatus { Converged, MaxIterations, Breakdown, NonFinite };

template <linalg::Scalar T>
struct SolveResult {
    SolveStatus status_{SolveStatus::MaxIterations};
    int iterations_{};
    T residualNorm_{};
    T relativeResidual_{};
};

template <linalg::Scalar T, CsrIndex I = int>
class Smoother {
public:
    virtual ~Smoother() = default;
    virtual void setup(const CsrMatrix<T, I>& matrix) = 0;
    virtual void apply(const Vector<T>& rhs, Vector<T>& x) = 0;
};

template <linalg::Scalar T, CsrIndex I = int>
class Preconditioner {
public:
    virtual ~Preconditioner() = default;
    virtual void setup(const CsrMatrix<T, I>& matrix) = 0;
    virtual void apply(const Vector<T>& rhs, Vector<T>& x) = 0;
};

template <linalg::Scalar T, CsrIndex I = int>
class Solver {
public:
    virtual ~Solver() = default;
    virtual void setup(const CsrMatrix<T, I>& matrix) = 0;
    virtual SolveResult<T> apply(const Vector<T>& rhs, Vector<T>& x) = 0;
};

namespace iteration {
template <linalg::Scalar T, CsrIndex I>
void validateMatrix(const CsrMatrix<T, I>& matrix) {
    if (!matrix.rowCount() || matrix.rowCount() != matrix.columnCount())
        throw std::invalid_argument("Solver requires a nonempty square matrix");
}
template <linalg::Scalar T>
void validateVectors(std::size_t n, int device, const Vector<T>& rhs, const Vector<T>& x) {
    if (rhs.size() != n || x.size() != n) throw std::invalid_argument("Solver vector shape mismatch");
    if ((rhs.size() && rhs.deviceId() != device) || (x.size() && x.deviceId() != device))
        throw std::invalid_argument("Solver vectors must live on the solver's device");
}
}  // namespace iteration

// Weighted Jacobi, x += w D^-1 (b - A x) per sweep (Jacobi.cuh:26-93).  Each
// sweep is one fused pass (IterationContext::sweep) between two iterate
// buffers; the right-hand side is copied first, so rhs and x may alias.
template <linalg::Scalar T, CsrIndex I = int>
class Jacobi final : public Smoother<T, I> {
public:
    explicit Jacobi(int sweeps = 2, T weight = T(2) / T(3)) : sweeps_(sweeps), weight_(weight) {
        if (sweeps < 1 || !std::isfinite(weight) || weight <= 0 || weight >= 2)
            throw std::invalid_argument("Invalid Jacobi parameters");
    }
    void setup(const CsrMatrix<T, I>& matrix) override {
        iteration::validateMatrix(matrix);
        detail::DeviceScope scope(matrix.rowOffsets().deviceId());
        auto w = std::make_unique<Work>(matrix);
        w->dinv = linalg::inverseDiagonal(w->op.matrix());
        work_ = std::move(w);
    }
    void apply(const Vector<T>& rhs, Vector<T>& x) override {
        if (!work_) throw std::logic_error("Jacobi requires setup");
        Work& w = *work_;
        iteration::validateVectors(w.op.matrix().rowCount(), w.op.device(), rhs, x);
        w.b.copyDataFrom(rhs);
        w.iterate[0].copyDataFrom(x);
        int at = 0;
        for (int k = 0; k < sweeps_; ++k, at ^= 1) w.op.sweep(weight_, w.dinv, w.b, w.iterate[at], w.iterate[at ^ 1]);
        x.copyDataFrom(w.iterate[at]);
    }

private:
    struct Work {
        linalg::IterationContext<T, I> op;
        Vector<T> dinv, b, iterate[2];
        explicit Work(const CsrMatrix<T, I>& m)
            : op(m), dinv(m.rowCount()), b(m.rowCount()), iterate{Vector<T>(m.rowCount()), Vector<T>(m.rowCount())} {}
    };
    int sweeps_;
    T weight_;
    std::unique_ptr<Work> work_;
};

// z = r (Identity.cuh:11-41).
template <linalg::Scalar T, CsrIndex I = int>
class Identity final : public Preconditioner<T, I> {
public:
    void setup(const CsrMatrix<T, I>& matrix) override {
        iteration::validateMatrix(matrix);
        rows_ = matrix.rowCount();
        device_ = matrix.rowOffsets().deviceId();
    }
    void apply(const Vector<T>& rhs, Vector<T>& x) override {
        if (!rows_) throw std::logic_error("Identity requires setup");
        iteration::validateVectors(rows_, device_, rhs, x);
        x.copyDataFrom(rhs);
    }

private:
    std::size_t rows_{};
    int device_{};
};

template <linalg::Scalar T>
struct AmgOptions {
    int maxLevels_{20};
    int coarseRows_{64};  // 1..1024 (the reference allows 1..64)
    int matchingRounds_{8};
    int preSweeps_{1}, postSweeps_{1};
    int coarseSweeps_{8};
    T jacobiWeight_{T(0.75)};
    T strengthThreshold_{T(0.25)};
    bool absorbSingletons_{true};
    bool compactSortKeys_{false};
    multigrid::Cycle cycle_{multigrid::Cycle::V};
};

struct AmgLevelStatistics {
    std::size_t rows_{}, nonzeros_{};
    std::array<int, 4> aggregateSizes_{};
};

// Aggregation AMG preconditioner (Amg.cuh:13-278).
template <linalg::Scalar T, CsrIndex I = int>
class Amg final : public Preconditioner<T, I> {
public:
    using SmootherFactory = std::function<std::unique_ptr<Smoother<T, I>>()>;
    explicit Amg(AmgOptions<T> options = {}, SmootherFactory factory = {})
        : options_(options), factory_(std::move(factory)) {
        if (options.maxLevels_ < 1 || options.maxLevels_ > 20 || options.coarseRows_ < 1 ||
            options.coarseRows_ > 1024 || options.matchingRounds_ < 1 || options.matchingRounds_ > 8 ||
            options.preSweeps_ < 0 || options.postSweeps_ < 0 || options.coarseSweeps_ < 1 ||
            !std::isfinite(options.strengthThreshold_) || options.strengthThreshold_ < 0 ||
            options.strengthThreshold_ > 1 || !std::isfinite(options.jacobiWeight_) || options.jacobiWeight_ <= 0 ||
            options.jacobiWeight_ >= 2 ||
            (options.cycle_ != multigrid::Cycle::V && options.cycle_ != multigrid::Cycle::W &&
             options.cycle_ != multigrid::Cycle::F))
            throw std::invalid_argument("Invalid AMG options");
    }
    ~Amg() override {
        if (engine_) (void)tf_amg_destroy(engine_);
    }
    Amg(const Amg&) = delete;
    Amg& operator=(const Amg&) = delete;

    void setup(const CsrMatrix<T, I>& matrix) override {
        iteration::validateMatrix(matrix);
        detail::DeviceScope scope(matrix.rowOffsets().deviceId());
        if (!factory_) {
            setupEngine(matrix);
            return;
        }
        setupGeneric(matrix);
    }
    void apply(const Vector<T>& rhs, Vector<T>& x) override {
        if (factory_ ? !host_ : !engineReady_) throw std::logic_error("AMG requires setup");
        iteration::validateVectors(rows_, device_, rhs, x);
        detail::DeviceScope scope(device_);
        if (!factory_) {
            applyEngine(rhs, x);
            return;
        }
        host_->run(options_.cycle_, rhs, x);
    }
    [[nodiscard]] std::size_t levelCount() const noexcept {
        if (factory_) return host_ ? host_->size() : 0;
        return engineReady_ ? std::size_t(tf_amg_level_count(engine_)) : 0;
    }
    [[nodiscard]] std::vector<AmgLevelStatistics> levelStatistics() const {
        require();
        if (factory_) return host_->stats;
        std::vector<AmgLevelStatistics> out(levelCount());
        for (std::size_t d = 0; d < out.size(); ++d) {
            int64_t rows = 0, nnz = 0, hist[4] = {0, 0, 0, 0};
            detail::check(tf_amg_level_info(engine_, int(d), &rows, &nnz, hist), "AMG statistics");
            out[d].rows_ = std::size_t(rows);
            out[d].nonzeros_ = std::size_t(nnz);
            for (int k = 0; k < 4; ++k) out[d].aggregateSizes_[std::size_t(k)] = int(hist[k]);
        }
        return out;
    }
    [[nodiscard]] std::string stopReason() const {
        require();
        return factory_ ? host_->stop : std::string(tf_amg_stop_reason(engine_));
    }
    [[nodiscard]] std::string coarseSolver() const {
        require();
        return factory_ ? host_->coarseSolver : std::string(tf_amg_coarse_solver(engine_));
    }
    [[nodiscard]] std::vector<std::size_t> lastCycleVisits() const {
        require();
        if (factory_) return host_->visits;
        std::vector<int64_t> v(levelCount());
        detail::check(tf_amg_visits(engine_, v.data()), "AMG visits");
        return std::vector<std::size_t>(v.begin(), v.end());
    }
    // Engine handle for solvers that run the preconditioned iteration on the
    // device (nullptr with a smoother factory).
    [[nodiscard]] tf_amg* engine() {
        if (factory_) return nullptr;
        ensureEngine(detail::currentDevice());
        return engine_;
    }
    // Called by such a solver after its setup has built this hierarchy on the
    // engine, so that the queries and apply() see it.
    void engineSetUp(std::size_t rows, int device) {
        engineReady_ = true;
        rows_ = rows;
        device_ = device;
        work_ = Buffer<double>(2 * rows_);
    }

private:
    tf_amg_options engineOptions() const {
        tf_amg_options o{};
        o.max_levels = options_.maxLevels_;
        o.coarse_rows = options_.coarseRows_;
        o.matching_rounds = options_.matchingRounds_;
        o.pre_sweeps = options_.preSweeps_;
        o.post_sweeps = options_.postSweeps_;
        o.coarse_sweeps = options_.coarseSweeps_;
        o.jacobi_weight = double(options_.jacobiWeight_);
        o.strength_threshold = double(options_.strengthThreshold_);
        o.absorb_singletons = options_.absorbSingletons_ ? 1 : 0;
        o.cycle = int(options_.cycle_);
        o.level_precision = std::same_as<T, float> ? 32 : 64;
        o.coarse_tail = 0;
        return o;
    }
    void ensureEngine(int device) {
        if (engine_ && engineDevice_ == device) return;
        if (engine_) (void)tf_amg_destroy(engine_);
        engine_ = nullptr;
        engineReady_ = false;
        const tf_amg_options o = engineOptions();
        detail::check(tf_amg_create(device, &o, &engine_), "AMG create");
        engineDevice_ = device;
    }
    void setupEngine(const CsrMatrix<T, I>& matrix) {
        const int device = matrix.rowOffsets().deviceId();
        ensureEngine(device);
        linalg::detail::Context ctx(device);
        auto native = linalg::detail::toNative(ctx, matrix, false);
        detail::check(tf_amg_setup(engine_, native.rows, native.nnz, native.rp.data(), native.ci.data(),
                                   native.v.data(), 0),
                      "AMG setup");
        engineReady_ = true;
        rows_ = matrix.rowCount();
        device_ = device;
        work_ = Buffer<double>(2 * rows_);
    }
    void applyEngine(const Vector<T>& rhs, Vector<T>& x) {
        double* in = work_.data();
        double* out = work_.data() + rows_;
        linalg::detail::Context ctx(device_);
        Vector<T> staged(rows_);
        staged.copyDataFrom(rhs);
        if constexpr (std::same_as<T, double>)
            detail::check(tf_memcpy(in, staged.deviceData(), rows_ * sizeof(double)), "AMG input");
        else
            detail::check(tf_convert_f32_f64(ctx.get(), rows_, staged.deviceData(), in), "AMG input");
        ctx.sync();
        detail::check(tf_amg_apply(engine_, in, out, 0), "AMG apply");
        if constexpr (std::same_as<T, double>)
            detail::check(tf_memcpy(staged.deviceData(), out, rows_ * sizeof(double)), "AMG output");
        else
            detail::check(tf_convert_f64_f32(ctx.get(), rows_, out, staged.deviceData()), "AMG output");
        ctx.sync();
        x.copyDataFrom(staged);
    }

    // The host-driven hierarchy (injected smoothers, float levels, 64-bit
    // indices): per-level operators, vectors, transfers and smoothers, kept as
    // parallel arrays indexed by level, and the V/W/F schedule of
    // multigrid::executeCycle run over them (Amg.cuh:81-143,197-257).  The
    // coarsest level is solved by the pivoted LU or, where it could not be
    // factored, by Jacobi(coarseSweeps, weight).
    struct HostLevels {
        struct Vectors {
            Vector<T> b, x, r;
            explicit Vectors(std::size_t n) : b(n), x(n), r(n) {}
        };
        std::vector<std::unique_ptr<linalg::IterationContext<T, I>>> op;
        std::vector<std::unique_ptr<Vectors>> vec;
        std::vector<multigrid::AggregateTransfer> down;  // level d -> d + 1
        std::vector<std::unique_ptr<Smoother<T, I>>> pre, post;
        std::vector<AmgLevelStatistics> stats;
        std::vector<std::size_t> visits;
        multigrid::CoarseFactor factor;
        std::unique_ptr<Smoother<T, I>> coarseJacobi;
        std::string stop, coarseSolver;

        std::size_t size() const { return op.size(); }
        void add(const CsrMatrix<T, I>& a, std::size_t nonzeros) {
            op.push_back(std::make_unique<linalg::IterationContext<T, I>>(a));
            vec.push_back(std::make_unique<Vectors>(a.rowCount()));
            pre.emplace_back();
            post.emplace_back();
            stats.push_back({a.rowCount(), nonzeros, {}});
        }
        void run(multigrid::Cycle kind, const Vector<T>& rhs, Vector<T>& x) {
            vec.front()->b.copyDataFrom(rhs);
            visits.assign(size(), 0);
            multigrid::executeCycle(kind, 0, *this, true);
            x.copyDataFrom(vec.front()->x);
        }
        // multigrid::CycleOperations over level d
        bool isCoarsest(std::size_t d) const { return d + 1 == size(); }
        void begin(std::size_t d) { ++visits[d]; }
        void clear(std::size_t d) { op[d]->clear(vec[d]->x); }
        void preSmooth(std::size_t d) {
            if (pre[d]) pre[d]->apply(vec[d]->b, vec[d]->x);
        }
        void postSmooth(std::size_t d) {
            if (post[d]) post[d]->apply(vec[d]->b, vec[d]->x);
        }
        void restrictRhs(std::size_t d) { multigrid::restrictAggregates(down[d], vec[d]->b, vec[d + 1]->b); }
        void restrictResidual(std::size_t d) {
            op[d]->residual(vec[d]->b, vec[d]->x, vec[d]->r);
            multigrid::restrictAggregates(down[d], vec[d]->r, vec[d + 1]->b);
        }
        void prolong(std::size_t d) { multigrid::prolongAggregates(down[d], vec[d + 1]->x, vec[d]->x); }
        void coarseSolve(std::size_t d) {
            if (factor.factored_)
                multigrid::solveCoarse(factor, vec[d]->b, vec[d]->x);
            else
                coarseJacobi->apply(vec[d]->b, vec[d]->x);
        }
    };
    // Why coarsening ends at a level of `rows` rows that is the `count`-th
    // (empty: it goes on), before the strength graph is built.
    std::string stopBefore(std::size_t rows, std::size_t count) const {
        if (rows <= std::size_t(options_.coarseRows_)) return "coarse size";
        if (count == std::size_t(options_.maxLevels_)) return "level limit";
        return {};
    }
    void setupGeneric(const CsrMatrix<T, I>& matrix) {
        auto h = std::make_unique<HostLevels>();
        multigrid::GpuAggregation aggregation(options_.compactSortKeys_);
        auto level = aggregation.canonical(multigrid::toNativeMatrix(matrix));
        for (;;) {
            aggregation.validate(level);
            auto typed = multigrid::fromNativeMatrix<T, I>(level);
            (void)linalg::inverseDiagonal(typed);  // rejects a zero or non-finite diagonal
            h->add(typed, level.nonZeroCount());
            h->stop = stopBefore(level.rowCount(), h->size());
            if (!h->stop.empty()) break;
            auto graph = aggregation.strength(level, double(options_.strengthThreshold_));
            if (!graph.nonZeroCount()) {
                h->stop = "empty strength graph";
                break;
            }
            auto transfer = aggregation.aggregate(graph, options_.matchingRounds_, options_.absorbSingletons_);
            if (transfer.coarseRows_ >= transfer.fineRows_) {
                h->stop = "no contraction";
                break;
            }
            auto next = aggregation.galerkin(level, transfer);
            const std::size_t d = h->size() - 1;
            h->stats[d].aggregateSizes_ = transfer.histogram_;
            if (options_.preSweeps_) h->pre[d] = makeSmoother(typed);
            if (options_.postSweeps_) h->post[d] = makeSmoother(typed);
            h->down.push_back(std::move(transfer));
            level = std::move(next);
        }
        if (level.rowCount() <= std::size_t(std::max(64, options_.coarseRows_)))
            h->factor = multigrid::factorCoarse(level);
        if (h->factor.factored_) {
            h->coarseSolver = "pivoted GPU LU";
        }
        else {
            h->coarseSolver = h->factor.singular_ ? "Jacobi (singular LU)" : "Jacobi (coarsening stopped)";
            h->coarseJacobi = std::make_unique<Jacobi<T, I>>(options_.coarseSweeps_, options_.jacobiWeight_);
            h->coarseJacobi->setup(h->op.back()->matrix());
        }
        h->visits.assign(h->size(), 0);
        host_ = std::move(h);
        rows_ = matrix.rowCount();
        device_ = matrix.rowOffsets().deviceId();
    }
    std::unique_ptr<Smoother<T, I>> makeSmoother(const CsrMatrix<T, I>& a) {
        auto smoother = factory_();
        if (!smoother) throw std::invalid_argument("AMG smoother factory returned null");
        smoother->setup(a);
        return smoother;
    }
    void require() const {
        if (factory_ ? !host_ : !engineReady_) throw std::logic_error("AMG requires setup");
    }

    AmgOptions<T> options_;
    SmootherFactory factory_;
    std::unique_ptr<HostLevels> host_;
    tf_amg* engine_ = nullptr;
    int engineDevice_ = -1;
    bool engineReady_ = false;
    std::size_t rows_{};
    int device_{};
    Buffer<double> work_;
};

template <linalg::Scalar T>
struct CgOptions {
    int maxIterations_{1000};
    T relativeTolerance_{std::same_as<T, float> ? T(1e-5) : T(1e-8)};
    T absoluteTolerance_{0};
    int residualRefreshInterval_{50};
};
template <linalg::Scalar T>
struct BiCgStabOptions {
    int maxIterations_{1000};
    T relativeTolerance_{std::same_as<T, float> ? T(1e-5) : T(1e-8)};
    T absoluteTolerance_{0};
    int residualRefreshInterval_{50};
};

namespace detail {
// The scalars of a Krylov recurrence and the tests each reduction's value
// goes through, as one state machine: the host twin of the control block the
// library's device-resident iteration keeps (a step either continues or
// records the status the recurrence stops with; Cg.cuh:123-189,
// BiCgStab.cuh:144-247).  count is the iteration under way (1-based).
template <linalg::Scalar T>
struct Recurrence {
    T normB{}, tol{}, res{}, rho{}, rhoPrev{}, alpha{1}, beta{}, omega{1};
    int count = 0, limit, refreshEvery;
    bool restart = true;  // the next direction starts afresh from the (preconditioned) residual
    std::optional<SolveStatus> stop;

    Recurrence(int maxIterations, int refreshInterval) : limit(maxIterations), refreshEvery(refreshInterval) {}
    bool running() const { return !stop && count < limit; }
    bool halt(SolveStatus why) {
        stop = why;
        return false;
    }
    bool finite(T v) { return std::isfinite(v) || halt(SolveStatus::NonFinite); }
    int iterations() const { return stop ? count : limit; }
    // status from the final true residual
    SolveStatus classify(T residual) const {
        const SolveStatus s = stop.value_or(SolveStatus::MaxIterations);
        if (!std::isfinite(residual) || !std::isfinite(normB)) return SolveStatus::NonFinite;
        if (residual <= tol) return SolveStatus::Converged;
        return s == SolveStatus::Converged ? SolveStatus::MaxIterations : s;
    }
    // CG: rho = r.z; breakdown unless positive; beta for a continued direction
    bool cgRho(T value) {
        ++count;
        rho = value;
        if (!finite(rho)) return false;
        if (rho <= 0) return halt(SolveStatus::Breakdown);
        if (restart) return true;
        beta = rho / rhoPrev;
        return finite(beta);
    }
    // CG: alpha = rho / p.Ap; breakdown unless the curvature is positive
    bool cgCurvature(T value) {
        if (!finite(value)) return false;
        if (value <= 0) return halt(SolveStatus::Breakdown);
        alpha = rho / value;
        return finite(alpha);
    }
    // BiCGStab: rho = r_hat.r; breakdown at zero; beta for a continued direction
    bool biRho(T value) {
        ++count;
        rho = value;
        if (!finite(rho)) return false;
        if (rho == 0) return halt(SolveStatus::Breakdown);
        if (restart) return true;
        beta = (rho / rhoPrev) * (alpha / omega);
        return finite(beta);
    }
    // BiCGStab: alpha = rho / r_hat.v
    bool biAlpha(T denominator) {
        if (!finite(denominator)) return false;
        if (denominator == 0) return halt(SolveStatus::Breakdown);
        alpha = rho / denominator;
        return finite(alpha);
    }
    // BiCGStab: omega = t.s / t.t
    bool biOmega(T tt, T ts) {
        if (!finite(tt) || !finite(ts)) return false;
        if (tt == 0) return halt(SolveStatus::Breakdown);
        omega = ts / tt;
        if (!finite(omega)) return false;
        return omega != 0 || halt(SolveStatus::Breakdown);
    }
};

// Shared machinery of Cg and BiCgStab: option checks, the state snapshot, the
// engine path, the start of a host-driven solve, its residual refresh and
// its report.
template <linalg::Scalar T, CsrIndex I, class Options>
class KrylovBase : public Solver<T, I> {
protected:
    KrylovBase(Options options, std::unique_ptr<Preconditioner<T, I>> preconditioner, int kind, const char* name)
        : options_(options), preconditioner_(std::move(preconditioner)), kind_(kind), name_(name) {
        if (options.maxIterations_ < 0 || options.residualRefreshInterval_ < 1 ||
            !std::isfinite(options.relativeTolerance_) || options.relativeTolerance_ < 0 ||
            options.relativeTolerance_ >= 1 || !std::isfinite(options.absoluteTolerance_) ||
            options.absoluteTolerance_ < 0)
            throw std::invalid_argument(std::string("Invalid ") + name + " options");
    }
    ~KrylovBase() override {
        if (engine_) (void)tf_solver_destroy(engine_);
    }

public:
    void setup(const CsrMatrix<T, I>& matrix) override {
        state_.reset();
        if (engine_) (void)tf_solver_destroy(engine_);
        engine_ = nullptr;
        iteration::validateMatrix(matrix);
        DeviceScope scope(matrix.rowOffsets().deviceId());
        auto next = std::make_unique<State>(matrix, kind_);
        if (useEngine()) {
            tf_krylov_options ko{options_.maxIterations_, double(options_.relativeTolerance_),
                                 double(options_.absoluteTolerance_), options_.residualRefreshInterval_, 1};
            tf_solver* s = nullptr;
            if (auto* amg = dynamic_cast<Amg<double, int>*>(preconditioner_.get()))
                check(tf_solver_create_with_amg(kind_, &ko, amg->engine(), &s), "solver create");
            else
                check(tf_solver_create(next->context_.device(), kind_, &ko, nullptr, &s), "solver create");
            const auto& nat = next->context_.native();
            const int rc = tf_solver_setup(s, nat.rows, nat.nnz, nat.rp.data(), nat.ci.data(), nat.v.data(), 0);
            if (rc != TF_OK) {
                (void)tf_solver_destroy(s);
                check(rc, "solver setup");
            }
            if (auto* amg = dynamic_cast<Amg<double, int>*>(preconditioner_.get()))
                amg->engineSetUp(matrix.rowCount(), next->context_.device());
            engine_ = s;
        }
        else {
            preconditioner_->setup(next->context_.matrix());
        }
        state_ = std::move(next);
    }
    SolveResult<T> apply(const Vector<T>& rhs, Vector<T>& x) override {
        if (!state_) throw std::logic_error(std::string(name_) + " requires setup");
        auto& s = *state_;
        iteration::validateVectors(s.x_.size(), s.context_.device(), rhs, x);
        DeviceScope scope(s.context_.device());
        s.rhs_.copyDataFrom(rhs);
        s.x_.copyDataFrom(x);
        SolveResult<T> report;
        if (engine_) {
            tf_solve_result r{};
            if constexpr (std::same_as<T, double>)
                check(tf_solver_apply(engine_, s.rhs_.deviceData(), s.x_.deviceData(), 0, &r), "solve");
            report = {SolveStatus(r.status), r.iterations, T(r.residual_norm), T(r.relative_residual)};
        }
        else {
            report = solve();
        }
        x.copyDataFrom(s.x_);
        return report;
    }

protected:
    struct State {
        linalg::IterationContext<T, I> context_;
        Vector<T> rhs_, x_, r_, z_, p_, product_, shadow_, s_, t_, ph_, sh_;
        State(const CsrMatrix<T, I>& m, int kind)
            : context_(m), rhs_(m.rowCount()), x_(m.rowCount()), r_(m.rowCount()), z_(m.rowCount()),
              p_(m.rowCount()), product_(m.rowCount()), shadow_(kind ? m.rowCount() : 0),
              s_(kind ? m.rowCount() : 0), t_(kind ? m.rowCount() : 0), ph_(kind ? m.rowCount() : 0),
              sh_(kind ? m.rowCount() : 0) {}
    };
    virtual SolveResult<T> solve() = 0;
    bool useEngine() const {
        if constexpr (std::same_as<T, double> && std::same_as<I, int>) {
            if (dynamic_cast<Identity<double, int>*>(preconditioner_.get())) return true;
            if (auto* amg = dynamic_cast<Amg<double, int>*>(preconditioner_.get())) return amg->engine() != nullptr;
        }
        return false;
    }
    // The report of a host-driven solve: the true residual b - A x decides
    // the status (Cg.cuh:105-119, BiCgStab.cuh:122-136).
    SolveResult<T> report(const Recurrence<T>& k) {
        auto& s = *state_;
        s.context_.residual(s.rhs_, s.x_, s.r_);
        const T res = s.context_.norm(s.r_);
        return {k.classify(res), k.iterations(), res, k.normB > 0 ? res / k.normB : res};
    }
    // normB, the tolerance and the initial residual r = b - A x: a recurrence
    // that has already stopped when either is non-finite or x already meets
    // the tolerance.
    Recurrence<T> start() {
        auto& s = *state_;
        const auto& o = options_;
        Recurrence<T> k(o.maxIterations_, o.residualRefreshInterval_);
        k.normB = s.context_.norm(s.rhs_);
        k.tol = std::max(o.absoluteTolerance_, o.relativeTolerance_ * k.normB);
        s.context_.residual(s.rhs_, s.x_, s.r_);
        k.res = s.context_.norm(s.r_);
        if (!std::isfinite(k.normB) || !std::isfinite(k.res))
            k.halt(SolveStatus::NonFinite);
        else if (k.res <= k.tol)
            k.halt(SolveStatus::Converged);
        return k;
    }
    // |r| after an update: stops on a non-finite value; when the residual meets
    // the tolerance or a refresh is due, recomputes r = b - A x and stops if
    // that one meets it (or is non-finite), otherwise restarts the direction.
    // True when the recurrence continues.
    bool refresh(Recurrence<T>& k, T updatedNorm, Vector<T>* shadow = nullptr) {
        auto& s = *state_;
        if (!k.finite(updatedNorm)) return false;
        k.res = updatedNorm;
        if (k.res > k.tol && k.count % k.refreshEvery != 0) return true;
        s.context_.residual(s.rhs_, s.x_, s.r_);
        k.res = s.context_.norm(s.r_);
        if (!std::isfinite(k.res) || k.res <= k.tol) return k.halt(SolveStatus::MaxIterations);
        if (shadow) shadow->copyDataFrom(s.r_);
        k.restart = true;
        return true;
    }
    Options options_;
    std::unique_ptr<Preconditioner<T, I>> preconditioner_;
    std::unique_ptr<State> state_;
    tf_solver* engine_ = nullptr;
    int kind_;
    const char* name_;
};
}  // namespace detail

// Preconditioned CG (Cg.cuh:10-193); default preconditioner Identity.
template <linalg::Scalar T, CsrIndex I = int>
class Cg final : public detail::KrylovBase<T, I, CgOptions<T>> {
    using Base = detail::KrylovBase<T, I, CgOptions<T>>;

public:
    explicit Cg(CgOptions<T> options = {}, std::unique_ptr<Preconditioner<T, I>> preconditioner = {})
        : Base(options, preconditioner ? std::move(preconditioner) : std::make_unique<Identity<T, I>>(), 0, "CG") {}

private:
    // z = M r, rho, direction, q = A p, alpha, x and r updated, |r| with the
    // refresh test -- each reduction checked by the recurrence.
    SolveResult<T> solve() override {
        auto& v = *this->state_;
        auto& A = v.context_;
        auto k = this->start();
        while (k.running()) {
            this->preconditioner_->apply(v.r_, v.z_);
            if (!k.cgRho(A.dot(v.r_, v.z_))) break;
            if (k.restart) {
                v.p_.copyDataFrom(v.z_);
                k.restart = false;
            }
            else {
                A.scale(k.beta, v.p_);
                A.axpy(T(1), v.z_, v.p_);
            }
            A.multiply(v.p_, v.product_);
            if (!k.cgCurvature(A.dot(v.p_, v.product_))) break;
            A.axpy(k.alpha, v.p_, v.x_);
            A.axpy(-k.alpha, v.product_, v.r_);
            if (!this->refresh(k, A.norm(v.r_))) break;
            k.rhoPrev = k.rho;
        }
        return this->report(k);
    }
};

// Right-preconditioned BiCGStab (BiCgStab.cuh:11-251); default preconditioner Amg.
template <linalg::Scalar T, CsrIndex I = int>
class BiCgStab final : public detail::KrylovBase<T, I, BiCgStabOptions<T>> {
    using Base = detail::KrylovBase<T, I, BiCgStabOptions<T>>;

public:
    explicit BiCgStab(BiCgStabOptions<T> options = {}, std::unique_ptr<Preconditioner<T, I>> preconditioner = {})
        : Base(options, preconditioner ? std::move(preconditioner) : std::make_unique<Amg<T, I>>(), 1, "BiCGStab") {}

private:
    // Right-preconditioned: p_hat = M p, v = A p_hat, alpha, s, and either the
    // early exit on |s| (refresh as a restart) or s_hat = M s, t = A s_hat,
    // omega, x and r updated, |r| with the refresh test.
    SolveResult<T> solve() override {
        auto& w = *this->state_;
        auto& A = w.context_;
        auto& M = *this->preconditioner_;
        Vector<T>& v = w.product_;
        auto k = this->start();
        if (k.running()) w.shadow_.copyDataFrom(w.r_);
        while (k.running()) {
            if (!k.biRho(A.dot(w.shadow_, w.r_))) break;
            if (k.restart) {
                w.p_.copyDataFrom(w.r_);
                k.restart = false;
            }
            else {
                A.axpy(-k.omega, v, w.p_);
                A.scale(k.beta, w.p_);
                A.axpy(T(1), w.r_, w.p_);
            }
            M.apply(w.p_, w.ph_);
            A.multiply(w.ph_, v);
            if (!k.biAlpha(A.dot(w.shadow_, v))) break;
            w.s_.copyDataFrom(w.r_);
            A.axpy(-k.alpha, v, w.s_);
            A.axpy(k.alpha, w.ph_, w.x_);
            const T normS = A.norm(w.s_);
            if (!k.finite(normS)) break;
            if (normS <= k.tol) {
                // s already meets the tolerance: the recomputed residual decides
                // (a restart keeps rhoPrev)
                A.residual(w.rhs_, w.x_, w.r_);
                k.res = A.norm(w.r_);
                if (k.res <= k.tol || !std::isfinite(k.res)) {
                    k.halt(SolveStatus::MaxIterations);
                    break;
                }
                w.shadow_.copyDataFrom(w.r_);
                k.restart = true;
                continue;
            }
            M.apply(w.s_, w.sh_);
            A.multiply(w.sh_, w.t_);
            if (!k.biOmega(A.dot(w.t_, w.t_), A.dot(w.t_, w.s_))) break;
            A.axpy(k.omega, w.sh_, w.x_);
            w.r_.copyDataFrom(w.s_);
            A.axpy(-k.omega, w.t_, w.r_);
            if (!this->refresh(k, A.norm(w.r_), &w.shadow_)) break;
            k.rhoPrev = k.rho;
        }
        return this->report(k);
    }
};

}  // namespace thermofem
