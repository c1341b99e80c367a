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

// Weighted Jacobi, x += w D^-1 (b - A x) per sweep; rhs and x may alias (Jacobi.cuh:26-93).
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
        auto next = std::make_unique<State>(matrix);
        next->inverse_ = linalg::inverseDiagonal(next->context_.matrix());
        state_ = std::move(next);
    }
    void apply(const Vector<T>& rhs, Vector<T>& x) override {
        if (!state_) throw std::logic_error("Jacobi requires setup");
        auto& s = *state_;
        iteration::validateVectors(s.context_.matrix().rowCount(), s.context_.device(), rhs, x);
        s.rhs_.copyDataFrom(rhs);
        s.x_.copyDataFrom(x);
        for (int sweep = 0; sweep < sweeps_; ++sweep) {
            s.context_.multiply(s.x_, s.product_);
            s.context_.jacobi(weight_, s.inverse_, s.rhs_, s.product_, s.x_);
        }
        x.copyDataFrom(s.x_);
    }

private:
    struct State {
        linalg::IterationContext<T, I> context_;
        Vector<T> inverse_, rhs_, x_, product_;
        explicit State(const CsrMatrix<T, I>& m)
            : context_(m), inverse_(m.rowCount()), rhs_(m.rowCount()), x_(m.rowCount()), product_(m.rowCount()) {}
    };
    int sweeps_;
    T weight_;
    std::unique_ptr<State> state_;
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
        if (factory_ ? !state_ : !engineReady_) throw std::logic_error("AMG requires setup");
        iteration::validateVectors(rows_, device_, rhs, x);
        detail::DeviceScope scope(device_);
        if (!factory_) {
            applyEngine(rhs, x);
            return;
        }
        auto& first = *state_->levels_.front();
        first.rhs_.copyDataFrom(rhs);
        std::fill(state_->visits_.begin(), state_->visits_.end(), 0);
        CycleAdapter adapter{*state_};
        multigrid::executeCycle(options_.cycle_, 0, adapter, true);
        x.copyDataFrom(first.x_);
    }
    [[nodiscard]] std::size_t levelCount() const noexcept {
        if (factory_) return state_ ? state_->levels_.size() : 0;
        return engineReady_ ? std::size_t(tf_amg_level_count(engine_)) : 0;
    }
    [[nodiscard]] std::vector<AmgLevelStatistics> levelStatistics() const {
        require();
        if (factory_) return state_->statistics_;
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
        return factory_ ? state_->stopReason_ : std::string(tf_amg_stop_reason(engine_));
    }
    [[nodiscard]] std::string coarseSolver() const {
        require();
        return factory_ ? state_->coarseSolver_ : std::string(tf_amg_coarse_solver(engine_));
    }
    [[nodiscard]] std::vector<std::size_t> lastCycleVisits() const {
        require();
        if (factory_) return state_->visits_;
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

    // Host-driven hierarchy with injected smoothers (Amg.cuh:81-143,197-257).
    struct Level {
        linalg::IterationContext<T, I> context_;
        Vector<T> rhs_, x_, residual_;
        std::unique_ptr<Smoother<T, I>> pre_, post_;
        std::unique_ptr<multigrid::AggregateTransfer> transfer_;
        explicit Level(const CsrMatrix<T, I>& a)
            : context_(a), rhs_(a.rowCount()), x_(a.rowCount()), residual_(a.rowCount()) {}
    };
    struct State {
        std::vector<std::unique_ptr<Level>> levels_;
        std::vector<AmgLevelStatistics> statistics_;
        std::vector<std::size_t> visits_;
        multigrid::CoarseFactor factor_;
        std::unique_ptr<Smoother<T, I>> fallback_;
        std::string stopReason_, coarseSolver_;
    };
    struct CycleAdapter {
        State& state_;
        bool isCoarsest(std::size_t d) const { return d + 1 == state_.levels_.size(); }
        void begin(std::size_t d) { ++state_.visits_[d]; }
        void clear(std::size_t d) {
            auto& l = *state_.levels_[d];
            l.context_.clear(l.x_);
        }
        void coarseSolve(std::size_t d) {
            auto& l = *state_.levels_[d];
            if (state_.factor_.factored_)
                multigrid::solveCoarse(state_.factor_, l.rhs_, l.x_);
            else
                state_.fallback_->apply(l.rhs_, l.x_);
        }
        void preSmooth(std::size_t d) {
            auto& l = *state_.levels_[d];
            if (l.pre_) l.pre_->apply(l.rhs_, l.x_);
        }
        void postSmooth(std::size_t d) {
            auto& l = *state_.levels_[d];
            if (l.post_) l.post_->apply(l.rhs_, l.x_);
        }
        void restrictRhs(std::size_t d) {
            auto& l = *state_.levels_[d];
            multigrid::restrictAggregates(*l.transfer_, l.rhs_, state_.levels_[d + 1]->rhs_);
        }
        void restrictResidual(std::size_t d) {
            auto& l = *state_.levels_[d];
            l.context_.residual(l.rhs_, l.x_, l.residual_);
            multigrid::restrictAggregates(*l.transfer_, l.residual_, state_.levels_[d + 1]->rhs_);
        }
        void prolong(std::size_t d) {
            auto& l = *state_.levels_[d];
            multigrid::prolongAggregates(*l.transfer_, state_.levels_[d + 1]->x_, l.x_);
        }
    };
    void setupGeneric(const CsrMatrix<T, I>& matrix) {
        auto next = std::make_unique<State>();
        multigrid::GpuAggregation operations(options_.compactSortKeys_);
        auto native = operations.canonical(multigrid::toNativeMatrix(matrix));
        for (int depth = 0;; ++depth) {
            operations.validate(native);
            auto typed = multigrid::fromNativeMatrix<T, I>(native);
            (void)linalg::inverseDiagonal(typed);
            auto level = std::make_unique<Level>(typed);
            next->statistics_.push_back({native.rowCount(), native.nonZeroCount(), {}});
            if (native.rowCount() <= std::size_t(options_.coarseRows_)) {
                next->stopReason_ = "coarse size";
                next->levels_.push_back(std::move(level));
                break;
            }
            if (depth + 1 == options_.maxLevels_) {
                next->stopReason_ = "level limit";
                next->levels_.push_back(std::move(level));
                break;
            }
            auto graph = operations.strength(native, double(options_.strengthThreshold_));
            if (!graph.nonZeroCount()) {
                next->stopReason_ = "empty strength graph";
                next->levels_.push_back(std::move(level));
                break;
            }
            auto transfer = operations.aggregate(graph, options_.matchingRounds_, options_.absorbSingletons_);
            if (transfer.coarseRows_ >= transfer.fineRows_) {
                next->stopReason_ = "no contraction";
                next->levels_.push_back(std::move(level));
                break;
            }
            auto coarse = operations.galerkin(native, transfer);
            next->statistics_.back().aggregateSizes_ = transfer.histogram_;
            level->transfer_ = std::make_unique<multigrid::AggregateTransfer>(std::move(transfer));
            if (options_.preSweeps_) level->pre_ = makeSmoother(typed);
            if (options_.postSweeps_) level->post_ = makeSmoother(typed);
            next->levels_.push_back(std::move(level));
            native = std::move(coarse);
        }
        if (native.rowCount() <= std::size_t(std::max(64, options_.coarseRows_)))
            next->factor_ = multigrid::factorCoarse(native);
        if (next->factor_.factored_) {
            next->coarseSolver_ = "pivoted GPU LU";
        }
        else {
            next->coarseSolver_ = next->factor_.singular_ ? "Jacobi (singular LU)" : "Jacobi (coarsening stopped)";
            next->fallback_ = std::make_unique<Jacobi<T, I>>(options_.coarseSweeps_, options_.jacobiWeight_);
            next->fallback_->setup(next->levels_.back()->context_.matrix());
        }
        next->visits_.resize(next->levels_.size());
        state_ = std::move(next);
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
        if (factory_ ? !state_ : !engineReady_) throw std::logic_error("AMG requires setup");
    }

    AmgOptions<T> options_;
    SmootherFactory factory_;
    std::unique_ptr<State> state_;
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
// Shared machinery of Cg and BiCgStab: option checks, the state snapshot, the
// engine path and finish() (Cg.cuh:105-119, BiCgStab.cuh:122-136).
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
    SolveResult<T> finish(SolveStatus status, int iterations, T normB, T tolerance) {
        auto& s = *state_;
        s.context_.residual(s.rhs_, s.x_, s.r_);
        const T residual = s.context_.norm(s.r_);
        if (!std::isfinite(residual) || !std::isfinite(normB))
            status = SolveStatus::NonFinite;
        else if (residual <= tolerance)
            status = SolveStatus::Converged;
        else if (status == SolveStatus::Converged)
            status = SolveStatus::MaxIterations;
        return {status, iterations, residual, normB > 0 ? residual / normB : residual};
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
    SolveResult<T> solve() override {
        auto& s = *this->state_;
        auto& op = s.context_;
        const auto& o = this->options_;
        const T normB = op.norm(s.rhs_);
        const T tol = std::max(o.absoluteTolerance_, o.relativeTolerance_ * normB);
        op.residual(s.rhs_, s.x_, s.r_);
        T residual = op.norm(s.r_);
        if (!std::isfinite(normB) || !std::isfinite(residual)) return this->finish(SolveStatus::NonFinite, 0, normB, tol);
        if (residual <= tol) return this->finish(SolveStatus::Converged, 0, normB, tol);
        T previousRho = 0;
        bool fresh = true;
        for (int k = 0; k < o.maxIterations_; ++k) {
            const int it = k + 1;
            this->preconditioner_->apply(s.r_, s.z_);
            const T rho = op.dot(s.r_, s.z_);
            if (!std::isfinite(rho)) return this->finish(SolveStatus::NonFinite, it, normB, tol);
            if (rho <= 0) return this->finish(SolveStatus::Breakdown, it, normB, tol);
            if (fresh) {
                s.p_.copyDataFrom(s.z_);
                fresh = false;
            }
            else {
                const T beta = rho / previousRho;
                if (!std::isfinite(beta)) return this->finish(SolveStatus::NonFinite, it, normB, tol);
                op.scale(beta, s.p_);
                op.axpy(T(1), s.z_, s.p_);
            }
            op.multiply(s.p_, s.product_);
            const T curvature = op.dot(s.p_, s.product_);
            if (!std::isfinite(curvature)) return this->finish(SolveStatus::NonFinite, it, normB, tol);
            if (curvature <= 0) return this->finish(SolveStatus::Breakdown, it, normB, tol);
            const T alpha = rho / curvature;
            if (!std::isfinite(alpha)) return this->finish(SolveStatus::NonFinite, it, normB, tol);
            op.axpy(alpha, s.p_, s.x_);
            op.axpy(-alpha, s.product_, s.r_);
            residual = op.norm(s.r_);
            if (!std::isfinite(residual)) return this->finish(SolveStatus::NonFinite, it, normB, tol);
            if (residual <= tol || it % o.residualRefreshInterval_ == 0) {
                op.residual(s.rhs_, s.x_, s.r_);
                residual = op.norm(s.r_);
                if (!std::isfinite(residual) || residual <= tol)
                    return this->finish(SolveStatus::MaxIterations, it, normB, tol);
                fresh = true;
            }
            previousRho = rho;
        }
        return this->finish(SolveStatus::MaxIterations, o.maxIterations_, normB, tol);
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
    SolveResult<T> solve() override {
        auto& st = *this->state_;
        auto& op = st.context_;
        const auto& o = this->options_;
        auto& M = *this->preconditioner_;
        const T normB = op.norm(st.rhs_);
        const T tol = std::max(o.absoluteTolerance_, o.relativeTolerance_ * normB);
        op.residual(st.rhs_, st.x_, st.r_);
        T residual = op.norm(st.r_);
        if (!std::isfinite(normB) || !std::isfinite(residual)) return this->finish(SolveStatus::NonFinite, 0, normB, tol);
        if (residual <= tol) return this->finish(SolveStatus::Converged, 0, normB, tol);
        st.shadow_.copyDataFrom(st.r_);
        T rhoPrevious = 1, alpha = 1, omega = 1;
        bool fresh = true;
        Vector<T>& v = st.product_;
        for (int k = 0; k < o.maxIterations_; ++k) {
            const int it = k + 1;
            const T rho = op.dot(st.shadow_, st.r_);
            if (!std::isfinite(rho)) return this->finish(SolveStatus::NonFinite, it, normB, tol);
            if (rho == 0) return this->finish(SolveStatus::Breakdown, it, normB, tol);
            if (fresh) {
                st.p_.copyDataFrom(st.r_);
                fresh = false;
            }
            else {
                const T beta = (rho / rhoPrevious) * (alpha / omega);
                if (!std::isfinite(beta)) return this->finish(SolveStatus::NonFinite, it, normB, tol);
                op.axpy(-omega, v, st.p_);
                op.scale(beta, st.p_);
                op.axpy(T(1), st.r_, st.p_);
            }
            M.apply(st.p_, st.ph_);
            op.multiply(st.ph_, v);
            const T denominator = op.dot(st.shadow_, v);
            if (!std::isfinite(denominator)) return this->finish(SolveStatus::NonFinite, it, normB, tol);
            if (denominator == 0) return this->finish(SolveStatus::Breakdown, it, normB, tol);
            alpha = rho / denominator;
            if (!std::isfinite(alpha)) return this->finish(SolveStatus::NonFinite, it, normB, tol);
            st.s_.copyDataFrom(st.r_);
            op.axpy(-alpha, v, st.s_);
            op.axpy(alpha, st.ph_, st.x_);
            const T normS = op.norm(st.s_);
            if (!std::isfinite(normS)) return this->finish(SolveStatus::NonFinite, it, normB, tol);
            if (normS <= tol) {
                op.residual(st.rhs_, st.x_, st.r_);
                residual = op.norm(st.r_);
                if (residual <= tol || !std::isfinite(residual))
                    return this->finish(SolveStatus::MaxIterations, it, normB, tol);
                st.shadow_.copyDataFrom(st.r_);
                fresh = true;
                continue;
            }
            M.apply(st.s_, st.sh_);
            op.multiply(st.sh_, st.t_);
            const T tt = op.dot(st.t_, st.t_), ts = op.dot(st.t_, st.s_);
            if (!std::isfinite(tt) || !std::isfinite(ts)) return this->finish(SolveStatus::NonFinite, it, normB, tol);
            if (tt == 0) return this->finish(SolveStatus::Breakdown, it, normB, tol);
            omega = ts / tt;
            if (!std::isfinite(omega)) return this->finish(SolveStatus::NonFinite, it, normB, tol);
            if (omega == 0) return this->finish(SolveStatus::Breakdown, it, normB, tol);
            op.axpy(omega, st.sh_, st.x_);
            st.r_.copyDataFrom(st.s_);
            op.axpy(-omega, st.t_, st.r_);
            residual = op.norm(st.r_);
            if (!std::isfinite(residual)) return this->finish(SolveStatus::NonFinite, it, normB, tol);
            if (residual <= tol || it % o.residualRefreshInterval_ == 0) {
                op.residual(st.rhs_, st.x_, st.r_);
                residual = op.norm(st.r_);
                if (residual <= tol || !std::isfinite(residual))
                    return this->finish(SolveStatus::MaxIterations, it, normB, tol);
                st.shadow_.copyDataFrom(st.r_);
                fresh = true;
            }
            rhoPrevious = rho;
        }
        return this->finish(SolveStatus::MaxIterations, o.maxIterations_, normB, tol);
    }
};

}  // namespace thermofem
