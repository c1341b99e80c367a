// solve-matrix: reads a Matrix Market matrix, solves A x = A x_exact with
// BiCGStab (FP64, rtol 1e-8, atol 0, at most 1000 iterations, true residual
// refreshed every 50) preconditioned by aggregation AMG or the identity, and
// prints key=value lines.  The command line, the keys, the manufactured
// solution and the exit codes are the contract of example/solve-matrix.cu
// and docs/matrix-market.md:40-70:
//   usage: solve-matrix path [amg|identity] [v|w|f]
//   exit 0: converged with host relative residual <= 1.01e-8 and relative
//           solution error <= 1e-4; 2: solved but that check failed;
//   1: bad arguments, unreadable or invalid matrix, or a runtime error.
//
// The program is organised as a request (parsed arguments), a problem (the
// matrix with its manufactured right-hand side), the solve on the library,
// and an independent host check of the answer with compensated sums.
#include <thermofem/MatrixMarket.hpp>
#include <thermofem/thermofem.hpp>

#include <chrono>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    std::string path;
    bool amg = true;
    std::string cycleName = "v";
    thermofem::multigrid::Cycle cycle = thermofem::multigrid::Cycle::V;

    static Request parse(int argc, char** argv) {
        if (argc < 2 || argc > 4) throw std::invalid_argument("Usage: solve-matrix path [amg|identity] [v|w|f]");
        static const std::map<std::string, thermofem::multigrid::Cycle> cycles{
            {"v", thermofem::multigrid::Cycle::V},
            {"w", thermofem::multigrid::Cycle::W},
            {"f", thermofem::multigrid::Cycle::F}};
        Request q;
        q.path = argv[1];
        const std::string pre = argc >= 3 ? argv[2] : "amg";
        if (argc == 4) q.cycleName = argv[3];
        const auto c = cycles.find(q.cycleName);
        if (c == cycles.end()) throw std::invalid_argument("Unknown cycle");
        if (pre != "amg" && pre != "identity") throw std::invalid_argument("Unknown preconditioner");
        q.cycle = c->second;
        q.amg = pre == "amg";
        return q;
    }
};

// Neumaier's compensated summation: FP64 sums with an error term, accurate
// enough for the host check without extended precision.
struct Sum {
    double s = 0, c = 0;
    void add(double v) {
        const double t = s + v;
        c += std::fabs(s) >= std::fabs(v) ? (s - t) + v : (v - t) + s;
        s = t;
    }
    double value() const { return s + c; }
};

// y = A x with each row summed by compensated addition of the products.
std::vector<double> rowProducts(const thermofem::io::HostCsr& a, const std::vector<double>& x) {
    std::vector<double> y(std::size_t(a.rows_));
    for (std::size_t i = 0; i < y.size(); ++i) {
        Sum row;
        for (int k = a.offsets_[i]; k < a.offsets_[i + 1]; ++k)
            row.add(a.values_[std::size_t(k)] * x[std::size_t(a.indices_[std::size_t(k)])]);
        y[i] = row.value();
    }
    return y;
}

struct Problem {
    thermofem::io::HostCsr host;
    std::vector<double> exact, rhs;
    double readSeconds = 0;

    explicit Problem(const std::string& path) {
        const auto t0 = Clock::now();
        std::ifstream in(path);
        if (!in) throw std::runtime_error("Cannot open matrix file");
        host = thermofem::io::readMatrixMarket(in);
        if (host.rows_ <= 0 || host.rows_ != host.columns_)
            throw std::invalid_argument("Expected nonempty square matrix");
        readSeconds = std::chrono::duration<double>(Clock::now() - t0).count();
        // manufactured solution (docs/matrix-market.md:46-50)
        exact.resize(std::size_t(host.rows_));
        for (std::size_t i = 0; i < exact.size(); ++i) {
            const double t = double(i);
            exact[i] = 1.0 + 0.25 * std::sin(0.013 * t) + 0.125 * std::cos(0.007 * t);
        }
        rhs = rowProducts(host, exact);
    }
    std::size_t rows() const { return exact.size(); }
};

// The host check of a solution: relative residual |b - A x| / |b| and
// relative error against the manufactured solution.
struct Check {
    double residual = 0, error = 0;
    bool finite = true;
    Check(const Problem& p, const std::vector<double>& x) {
        const auto ax = rowProducts(p.host, x);
        Sum rr, bb, ee, xx;
        for (std::size_t i = 0; i < x.size(); ++i) {
            finite = finite && std::isfinite(x[i]) && std::isfinite(ax[i]);
            const double r = p.rhs[i] - ax[i], e = x[i] - p.exact[i];
            rr.add(r * r);
            bb.add(p.rhs[i] * p.rhs[i]);
            ee.add(e * e);
            xx.add(p.exact[i] * p.exact[i]);
        }
        residual = std::sqrt(rr.value() / (bb.value() > 0 ? bb.value() : 1.0));
        error = std::sqrt(ee.value() / xx.value());
    }
};

const char* statusName(thermofem::SolveStatus s) {
    static const char* const names[] = {"Converged", "MaxIterations", "Breakdown", "NonFinite"};
    const auto k = std::size_t(s);
    return k < 4 ? names[k] : "Unknown";
}

template <class V>
void put(const char* key, const V& value) {
    std::cout << key << '=' << value << '\n';
}

void printHierarchy(const thermofem::Amg<double>& amg, const Request& q, const Problem& p) {
    put("cycle", q.cycleName);
    put("levels", amg.levelCount());
    put("stop_reason", amg.stopReason());
    put("coarse_solver", amg.coarseSolver());
    std::size_t rows = 0, nonzeros = 0, d = 0;
    for (const auto& level : amg.levelStatistics()) {
        rows += level.rows_;
        nonzeros += level.nonzeros_;
        std::cout << "level=" << d++ << " rows=" << level.rows_ << " nnz=" << level.nonzeros_ << " aggregates=";
        for (int count : level.aggregateSizes_) std::cout << count << ',';
        std::cout << '\n';
    }
    put("grid_complexity", double(rows) / double(p.rows()));
    put("operator_complexity", double(nonzeros) / double(p.host.values_.size()));
}

int run(const Request& q) {
    using namespace thermofem;
    const Problem p(q.path);
    put("matrix", q.path);
    put("rows", p.rows());
    put("nnz", p.host.values_.size());
    put("preconditioner", q.amg ? "amg" : "identity");
    put("read_seconds", p.readSeconds);

    const int device = detail::currentDevice();
    char gpu[256] = {0};
    int runtime = 0;
    detail::check(tf_device_info(device, gpu, sizeof gpu, &runtime), "device info");
    put("device", device);
    put("gpu", gpu);
    put("hip_runtime", runtime);
    BiCgStabOptions<double> options;
    options.relativeTolerance_ = 1e-8;
    put("rtol", options.relativeTolerance_);
    put("atol", options.absoluteTolerance_);
    put("max_iterations", options.maxIterations_);
    put("refresh_interval", options.residualRefreshInterval_);

    const std::size_t n = p.rows();
    const CsrMatrix<double> A(n, n, p.host.offsets_, p.host.indices_, p.host.values_);
    const Vector<double> b{std::span<const double>(p.rhs)};
    Vector<double> x(n);
    x.copyFromHost(std::vector<double>(n, 0.0));
    AmgOptions<double> amgOptions;
    amgOptions.cycle_ = q.cycle;
    Amg<double>* hierarchy = nullptr;
    std::unique_ptr<Preconditioner<double>> M;
    if (q.amg) {
        auto amg = std::make_unique<Amg<double>>(amgOptions);
        hierarchy = amg.get();  // owned by the solver below; read after its setup
        M = std::move(amg);
    }
    else {
        M = std::make_unique<Identity<double>>();
    }
    BiCgStab<double> solver(options, std::move(M));

    put("phase", "setup");
    const auto s0 = Clock::now();
    solver.setup(A);
    const auto s1 = Clock::now();
    put("setup_seconds", std::chrono::duration<double>(s1 - s0).count());
    if (hierarchy) printHierarchy(*hierarchy, q, p);

    put("phase", "solve");
    const auto result = solver.apply(b, x);
    const auto s2 = Clock::now();
    const Check check(p, x.hostCopy());
    put("status", statusName(result.status_));
    put("iterations", result.iterations_);
    put("solve_seconds", std::chrono::duration<double>(s2 - s1).count());
    put("reported_relative_residual", result.relativeResidual_);
    put("host_relative_residual", check.residual);
    put("relative_solution_error", check.error);
    const bool pass = check.finite && result.status_ == SolveStatus::Converged && check.residual <= 1.01e-8 &&
                      check.error <= 1e-4;
    put("validation", pass ? "PASS" : "FAIL");
    return pass ? 0 : 2;
}

}  // namespace

int main(int argc, char** argv) {
    std::cout << std::unitbuf << std::setprecision(12);
    try {
        return run(Request::parse(argc, argv));
    }
    catch (const std::exception& e) {
        std::cerr << "error=" << e.what() << '\n';
        return 1;
    }
}
