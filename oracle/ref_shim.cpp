// C entry points over the reference's own host aggregation oracle
// (test/support/NativeAggregationOracle.{cuh,cu}) and Matrix Market reader
// (src/IO/MatrixMarket.cu), compiled from the reference tree by oracle/Makefile
// into oracle/_ref/.  TEST INFRASTRUCTURE ONLY: used to pin the restated
// oracle (tf_oracle.cpp) and to write golden fixtures; never shipped.
#include <NativeAggregationOracle.cuh>

#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>

namespace {
thread_local std::string g_error;
using thermofem::io::HostCsr;
namespace to = thermofem::testoracle;

HostCsr make(int n, const int* rp, const int* ci, const double* v) {
    HostCsr a;
    a.rows_ = a.columns_ = n;
    a.offsets_.assign(rp, rp + n + 1);
    a.indices_.assign(ci, ci + rp[n]);
    a.values_.assign(v, v + rp[n]);
    return a;
}
int64_t put(const HostCsr& a, int* rp, int* ci, double* v) {
    std::memcpy(rp, a.offsets_.data(), a.offsets_.size() * sizeof(int));
    std::memcpy(ci, a.indices_.data(), a.indices_.size() * sizeof(int));
    std::memcpy(v, a.values_.data(), a.values_.size() * sizeof(double));
    return int64_t(a.values_.size());
}
template <class F>
int64_t guarded(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}
}  // namespace

extern "C" {
const char* tfr_last_error(void) { return g_error.c_str(); }

int64_t tfr_canonical(int n, const int* rp, const int* ci, const double* v, int* orp, int* oci, double* ov) {
    return guarded([&] { return put(to::canonicalMatrix(make(n, rp, ci, v)), orp, oci, ov); });
}
int64_t tfr_strength(int n, const int* rp, const int* ci, const double* v, double theta, int* orp, int* oci,
                     double* ov) {
    return guarded([&] { return put(to::strengthGraph(make(n, rp, ci, v), theta).edges_, orp, oci, ov); });
}
int64_t tfr_match(int n, const int* rp, const int* ci, const double* v, int rounds, int* map) {
    return guarded([&] {
        auto m = to::matchPairs({make(n, rp, ci, v)}, rounds);
        std::memcpy(map, m.data(), m.size() * sizeof(int));
        int64_t c = 0;
        for (int x : m) c = x + 1 > c ? x + 1 : c;
        return c;
    });
}
int64_t tfr_contract(int n, const int* rp, const int* ci, const double* v, const int* map, int* orp, int* oci,
                     double* ov) {
    return guarded([&] {
        std::vector<int> m(map, map + n);
        return put(to::contractGraph({make(n, rp, ci, v)}, m).edges_, orp, oci, ov);
    });
}
int64_t tfr_aggregate(int n, const int* rp, const int* ci, const double* v, int rounds, int absorb, int* map,
                      int* offsets, int* members, int64_t* histogram) {
    return guarded([&] {
        auto t = to::aggregate({make(n, rp, ci, v)}, rounds, absorb != 0);
        std::memcpy(map, t.fineToCoarse_.data(), t.fineToCoarse_.size() * sizeof(int));
        std::memcpy(offsets, t.offsets_.data(), t.offsets_.size() * sizeof(int));
        std::memcpy(members, t.members_.data(), t.members_.size() * sizeof(int));
        for (int k = 0; k < 4; ++k) histogram[k] = int64_t(t.histogram_[k]);
        return int64_t(t.coarseRows());
    });
}
int64_t tfr_galerkin(int n, const int* rp, const int* ci, const double* v, int nc, const int* map,
                     const int* offsets, const int* members, int* orp, int* oci, double* ov) {
    return guarded([&] {
        to::AggregateMap t;
        t.fineToCoarse_.assign(map, map + n);
        t.offsets_.assign(offsets, offsets + nc + 1);
        t.members_.assign(members, members + n);
        return put(to::aggregateGalerkin(make(n, rp, ci, v), t), orp, oci, ov);
    });
}
/* Reads a Matrix Market file; first call with null arrays returns nnz and rows. */
int64_t tfr_read_matrix_market(const char* path, int* rows, int* cols, int* rp, int* ci, double* v) {
    return guarded([&] {
        std::ifstream file(path);
        auto a = thermofem::io::readMatrixMarket(file);
        *rows = a.rows_;
        *cols = a.columns_;
        if (rp) put(a, rp, ci, v);
        return int64_t(a.values_.size());
    });
}
}
