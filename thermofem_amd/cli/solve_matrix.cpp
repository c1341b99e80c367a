// solve-matrix: reads a Matrix Market matrix, solves A x = A x_exact with
// BiCGStab (FP64, rtol 1e-8, atol 0, at most 1000 iterations, true residual
// refreshed every 50) preconditioned by aggregation AMG or the identity, and
// prints key=value lines (example/solve-matrix.cu:54-180,
// docs/matrix-market.md:56-70).
//   usage: solve-matrix path [amg|identity] [v|w|f]
//   exit 0: converged with host relative residual <= 1.01e-8 and relative
//           solution error <= 1e-4; 2: solved but that check failed;
//   1: bad arguments, unreadable or invalid matrix, or a runtime error.
#include <thermofem/MatrixMarket.hpp>
#include <thermofem/thermofem.hpp>

#include <chrono>
#include <fstream>
#include <iomanip>
#include <iostream>

namespace {

const char* statusName(thermofem::SolveStatus s) {
    switch (s) {
    case thermofem::SolveStatus::Converged: return "Converged";
    case thermofem::SolveStatus::MaxIterations: return "MaxIterations";
    case thermofem::SolveStatus::Breakdown: return "Breakdown";
    case thermofem::SolveStatus::NonFinite: return "NonFinite";
    }
    return "Unknown";
}

// A x in long double, row by row in stored order.
std::vector<double> hostProduct(const thermofem::io::HostCsr& a, const std::vector<double>& x) {
    std::vector<double> y(std::size_t(a.rows_));
    for (int i = 0; i < a.rows_; ++i) {
        long double acc = 0;
        for (int k = a.offsets_[std::size_t(i)]; k < a.offsets_[std::size_t(i) + 1]; ++k)
            acc += (long double)a.values_[std::size_t(k)] * x[std::size_t(a.indices_[std::size_t(k)])];
        y[std::size_t(i)] = double(acc);
    }
    return y;
}

double seconds(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
}

}  // namespace

int main(int argc, char** argv) {
    using namespace thermofem;
    using Clock = std::chrono::steady_clock;
    std::cout << std::unitbuf << std::setprecision(12);
    try {
        if (argc < 2 || argc > 4) throw std::invalid_argument("Usage: solve-matrix path [amg|identity] [v|w|f]");
        const std::string mode = argc >= 3 ? argv[2] : "amg";
        const std::string cycle = argc == 4 ? argv[3] : "v";
        if (cycle != "v" && cycle != "w" && cycle != "f") throw std::invalid_argument("Unknown cycle");
        if (mode != "amg" && mode != "identity") throw std::invalid_argument("Unknown preconditioner");

        const auto t0 = Clock::now();
        std::ifstream file(argv[1]);
        if (!file) throw std::runtime_error("Cannot open matrix file");
        const auto host = io::readMatrixMarket(file);
        if (host.rows_ <= 0 || host.rows_ != host.columns_) throw std::invalid_argument("Expected nonempty square matrix");
        const auto t1 = Clock::now();
        const auto n = std::size_t(host.rows_);
        std::cout << "matrix=" << argv[1] << "\nrows=" << n << "\nnnz=" << host.values_.size()
                  << "\npreconditioner=" << mode << "\nread_seconds=" << seconds(t0, t1) << '\n';

        std::vector<double> exact(n), initial(n, 0.0);
        for (std::size_t i = 0; i < n; ++i)
            exact[i] = 1.0 + 0.25 * std::sin(double(i) * 0.013) + 0.125 * std::cos(double(i) * 0.007);
        const auto b = hostProduct(host, exact);

        const int device = detail::currentDevice();
        char name[256] = {0};
        int runtime = 0;
        detail::check(tf_device_info(device, name, sizeof name, &runtime), "device info");
        std::cout << "device=" << device << "\ngpu=" << name << "\nhip_runtime=" << runtime << "\nrtol=" << 1e-8
                  << "\natol=0\nmax_iterations=1000\nrefresh_interval=50\n";

        const CsrMatrix<double> matrix(n, n, host.offsets_, host.indices_, host.values_);
        const Vector<double> rhs{std::span<const double>(b)};
        Vector<double> x{std::span<const double>(initial)};
        AmgOptions<double> amgOptions;
        amgOptions.cycle_ = cycle == "v" ? multigrid::Cycle::V : cycle == "w" ? multigrid::Cycle::W : multigrid::Cycle::F;
        auto amg = std::make_unique<Amg<double>>(amgOptions);
        Amg<double>* hierarchy = amg.get();
        std::unique_ptr<Preconditioner<double>> preconditioner;
        if (mode == "amg")
            preconditioner = std::move(amg);
        else
            preconditioner = std::make_unique<Identity<double>>();
        BiCgStabOptions<double> options;
        options.relativeTolerance_ = 1e-8;
        BiCgStab<double> solver(options, std::move(preconditioner));

        std::cout << "phase=setup\n";
        const auto t2 = Clock::now();
        solver.setup(matrix);
        const auto t3 = Clock::now();
        std::cout << "setup_seconds=" << seconds(t2, t3) << '\n';
        if (mode == "amg") {
            std::cout << "cycle=" << cycle << "\nlevels=" << hierarchy->levelCount()
                      << "\nstop_reason=" << hierarchy->stopReason() << "\ncoarse_solver=" << hierarchy->coarseSolver()
                      << '\n';
            std::size_t rowsTotal = 0, nnzTotal = 0;
            const auto stats = hierarchy->levelStatistics();
            for (std::size_t d = 0; d < stats.size(); ++d) {
                rowsTotal += stats[d].rows_;
                nnzTotal += stats[d].nonzeros_;
                std::cout << "level=" << d << " rows=" << stats[d].rows_ << " nnz=" << stats[d].nonzeros_
                          << " aggregates=";
                for (int c : stats[d].aggregateSizes_) std::cout << c << ',';
                std::cout << '\n';
            }
            std::cout << "grid_complexity=" << double(rowsTotal) / double(n)
                      << "\noperator_complexity=" << double(nnzTotal) / double(host.values_.size()) << '\n';
        }

        std::cout << "phase=solve\n";
        const auto report = solver.apply(rhs, x);
        const auto t4 = Clock::now();
        const auto solution = x.hostCopy();
        const auto product = hostProduct(host, solution);
        long double rr = 0, bb = 0, ee = 0, xx = 0;
        bool finite = true;
        for (std::size_t i = 0; i < n; ++i) {
            finite = finite && std::isfinite(solution[i]) && std::isfinite(product[i]);
            const long double r = (long double)b[i] - product[i], e = (long double)solution[i] - exact[i];
            rr += r * r;
            bb += (long double)b[i] * b[i];
            ee += e * e;
            xx += (long double)exact[i] * exact[i];
        }
        const double hostResidual = double(std::sqrt(rr / (bb > 0 ? bb : 1)));
        const double solutionError = double(std::sqrt(ee / xx));
        std::cout << "status=" << statusName(report.status_) << "\niterations=" << report.iterations_
                  << "\nsolve_seconds=" << seconds(t3, t4) << "\nreported_relative_residual=" << report.relativeResidual_
                  << "\nhost_relative_residual=" << hostResidual << "\nrelative_solution_error=" << solutionError << '\n';
        const bool pass =
            finite && report.status_ == SolveStatus::Converged && hostResidual <= 1.01e-8 && solutionError <= 1e-4;
        std::cout << "validation=" << (pass ? "PASS" : "FAIL") << '\n';
        return pass ? 0 : 2;
    }
    catch (const std::exception& e) {
        std::cerr << "error=" << e.what() << '\n';
        return 1;
    }
}
