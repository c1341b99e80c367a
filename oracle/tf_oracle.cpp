// CPU oracle: host FP64 re-execution of ThermoFEM's aggregation-AMG Krylov path.
//
// TEST INFRASTRUCTURE ONLY. It is the parity checker for the HIP library and
// the timed CPU baseline; the product library never links or calls it.
//
// Every function restates the reference algorithm and cites the reference
// file:line it follows.  Summation orders are fixed so that results are
// reproducible:
//  * coordinate merges (canonical / strength / contraction / Galerkin) sum or
//    max-merge duplicates sequentially in original entry order, exactly like
//    the stable_sort + sequential merge of test/support/NativeAggregationOracle.cu:43-75;
//  * SpMV rows sum sequentially in stored order;
//  * dot products sum fixed 4096-element blocks sequentially, then the block
//    sums sequentially (independent of the OpenMP thread count).
// Build with -ffp-contract=off so no multiply-add is fused.
#include "tf_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_error;

struct Csr {
    int n = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
    int64_t nnz() const { return int64_t(ci.size()); }
};

struct Entry {
    int r, c;
    double v;
};

Csr fromArrays(int n, const int* rp, const int* ci, const double* v) {
    Csr a;
    a.n = n;
    a.rp.assign(rp, rp + n + 1);
    const int64_t nnz = rp[n];
    a.ci.assign(ci, ci + nnz);
    a.v.assign(v, v + nnz);
    return a;
}

void toArrays(const Csr& a, int* rp, int* ci, double* v) {
    std::copy(a.rp.begin(), a.rp.end(), rp);
    std::copy(a.ci.begin(), a.ci.end(), ci);
    std::copy(a.v.begin(), a.v.end(), v);
}

// Shape/finiteness check of NativeAggregationOracle.cu:21-37 and the device
// check of GpuAggregation.cu:283-304 (flag 2 conditions).
void validateCsr(const Csr& a) {
    if (a.n < 1 || int64_t(a.rp.size()) != int64_t(a.n) + 1 || a.rp.front() != 0 ||
        int64_t(a.rp.back()) != a.nnz() || a.nnz() > std::numeric_limits<int>::max() / 2) {
        throw std::invalid_argument("AMG invalid CSR or nonfinite coefficient");
    }
    for (int i = 0; i < a.n; ++i) {
        if (a.rp[i] < 0 || a.rp[i] > a.rp[i + 1]) throw std::invalid_argument("AMG invalid CSR or nonfinite coefficient");
    }
    for (int64_t j = 0; j < a.nnz(); ++j) {
        if (a.ci[j] < 0 || a.ci[j] >= a.n || !std::isfinite(a.v[j]))
            throw std::invalid_argument("AMG invalid CSR or nonfinite coefficient");
    }
}

// Coordinate merge: NativeAggregationOracle.cu:43-75 (stable_sort by (row,col),
// then sequential sum/max of equal keys, non-finite merge rejected) and the
// device compress of GpuAggregation.cu:178-221.  A stable counting sort by
// row followed by a stable sort by column inside each row is the same
// permutation as one stable sort by (row,col), so results are identical.
Csr compress(int n, const std::vector<Entry>& e, bool maximum) {
    const int64_t m = int64_t(e.size());
    if (m > std::numeric_limits<int>::max()) throw std::length_error("Aggregate setup index overflow");
    std::vector<int64_t> start(static_cast<size_t>(n) + 1, 0);
    for (const auto& x : e) start[static_cast<size_t>(x.r) + 1]++;
    for (int i = 0; i < n; ++i) start[i + 1] += start[i];
    std::vector<Entry> s(static_cast<size_t>(m));
    {
        std::vector<int64_t> pos(start.begin(), start.end() - 1);
        for (const auto& x : e) s[static_cast<size_t>(pos[x.r]++)] = x;
    }
    std::vector<int> unique(static_cast<size_t>(n), 0);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 2048) reduction(| : bad)
    for (int r = 0; r < n; ++r) {
        Entry* b = s.data() + start[r];
        Entry* en = s.data() + start[r + 1];
        const int64_t len = en - b;
        if (len <= 48) {
            for (int64_t i = 1; i < len; ++i) {  // stable insertion sort
                Entry x = b[i];
                int64_t j = i;
                while (j > 0 && b[j - 1].c > x.c) {
                    b[j] = b[j - 1];
                    --j;
                }
                b[j] = x;
            }
        }
        else {
            std::stable_sort(b, en, [](const Entry& p, const Entry& q) { return p.c < q.c; });
        }
        int u = 0;
        for (Entry* it = b; it != en; ++it) {
            if (u > 0 && b[u - 1].c == it->c) {
                double& val = b[u - 1].v;
                val = maximum ? std::max(val, it->v) : val + it->v;
                if (!std::isfinite(val)) bad |= 1;
            }
            else {
                b[u++] = *it;
            }
        }
        unique[r] = u;
    }
    if (bad) throw std::invalid_argument("Nonfinite aggregate reduction");
    Csr out;
    out.n = n;
    out.rp.assign(static_cast<size_t>(n) + 1, 0);
    for (int r = 0; r < n; ++r) out.rp[r + 1] = out.rp[r] + unique[r];
    out.ci.resize(static_cast<size_t>(out.rp[n]));
    out.v.resize(static_cast<size_t>(out.rp[n]));
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; ++r) {
        const Entry* b = s.data() + start[r];
        for (int k = 0; k < unique[r]; ++k) {
            out.ci[static_cast<size_t>(out.rp[r] + k)] = b[k].c;
            out.v[static_cast<size_t>(out.rp[r] + k)] = b[k].v;
        }
    }
    return out;
}

// NativeAggregationOracle.cu:119-129 / GpuAggregation.cu:991-1007.
Csr canonical(const Csr& a) {
    validateCsr(a);
    std::vector<Entry> e;
    e.reserve(static_cast<size_t>(a.nnz()));
    for (int r = 0; r < a.n; ++r)
        for (int j = a.rp[r]; j < a.rp[r + 1]; ++j) e.push_back({r, a.ci[j], a.v[j]});
    return compress(a.n, e, false);
}

// GpuAggregation.cu:247-271 (validate kernel): finite values, nonzero finite
// diagonal with finite reciprocal; rowmax = max |a_ij| over j != i.
void validateRowmax(const Csr& a, std::vector<double>* rowmax) {
    int bad = 0;
    if (rowmax) rowmax->assign(static_cast<size_t>(a.n), 0.0);
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int i = 0; i < a.n; ++i) {
        double diag = 0, mx = 0;
        for (int j = a.rp[i]; j < a.rp[i + 1]; ++j) {
            if (!std::isfinite(a.v[j])) bad |= 1;
            if (a.ci[j] == i)
                diag = a.v[j];
            else
                mx = std::fmax(mx, std::fabs(a.v[j]));
        }
        if (diag == 0 || !std::isfinite(diag) || !std::isfinite(1.0 / diag)) bad |= 1;
        if (rowmax) (*rowmax)[i] = mx;
    }
    if (bad) throw std::invalid_argument("Native AMG: missing/zero diagonal, non-finite value or reciprocal in hierarchy");
}

// NativeAggregationOracle.cu:130-160 / GpuAggregation.cu:339-390.
Csr strength(const Csr& a, double threshold) {
    if (!std::isfinite(threshold) || threshold < 0 || threshold > 1)
        throw std::invalid_argument("Invalid strength threshold");
    if (a.nnz() > int64_t(std::numeric_limits<int>::max() / 2))
        throw std::overflow_error("Native AMG: strength graph exceeds 32-bit capacity");
    std::vector<Entry> e;
    e.reserve(static_cast<size_t>(a.nnz()) * 2);
    for (int r = 0; r < a.n; ++r) {
        double maximum = 0;
        for (int j = a.rp[r]; j < a.rp[r + 1]; ++j)
            if (a.ci[j] != r) maximum = std::max(maximum, std::abs(a.v[j]));
        if (maximum == 0) continue;
        for (int j = a.rp[r]; j < a.rp[r + 1]; ++j) {
            const int c = a.ci[j];
            const double w = std::abs(a.v[j]) / maximum;
            if (c != r && w > 0 && w >= threshold) {
                e.push_back({r, c, w});
                e.push_back({c, r, w});
            }
        }
    }
    return compress(a.n, e, true);
}

// SplitMix64 finalizer over the ordered endpoint pair:
// GpuAggregation.cu:49-54, NativeAggregationOracle.cu:100-106.
uint64_t edgePriority(int i, int j) {
    uint64_t h = (uint64_t(unsigned(std::min(i, j))) << 32) | unsigned(std::max(i, j));
    h = (h ^ (h >> 30)) * 0xbf58476d1ce4e5b9ULL;
    h = (h ^ (h >> 27)) * 0x94d049bb133111ebULL;
    return h ^ (h >> 31);
}

// Synchronous mutual-proposal matching: NativeAggregationOracle.cu:161-206,
// GpuAggregation.cu:402-505.  Returns the compact map; count via out param.
std::vector<int> matchPairs(const Csr& g, int rounds, int* countOut) {
    if (rounds < 1 || rounds > 8) throw std::invalid_argument("Invalid matching configuration");
    const int n = g.n;
    std::vector<int> mate(static_cast<size_t>(n), -1), proposal(static_cast<size_t>(n), -1);
    for (int round = 0; round < rounds; ++round) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            int best = -1;
            double score = 0;
            uint64_t priority = 0;
            if (mate[i] < 0) {
                for (int j = g.rp[i]; j < g.rp[i + 1]; ++j) {
                    const int c = g.ci[j];
                    const double w = g.v[j];
                    if (c == i || mate[c] >= 0 || w <= 0 || w < score) continue;
                    const uint64_t cand = edgePriority(i, c);
                    if (w > score || (w == score && (best < 0 || cand > priority || (cand == priority && c < best)))) {
                        best = c;
                        score = w;
                        priority = cand;
                    }
                }
            }
            proposal[i] = best;
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            const int p = proposal[i];
            if (p >= 0 && proposal[p] == i) mate[i] = p;
        }
    }
    std::vector<int> prefix(static_cast<size_t>(n) + 1, 0), map(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) prefix[i + 1] = prefix[i] + ((mate[i] < 0 || i < mate[i]) ? 1 : 0);
    for (int i = 0; i < n; ++i) map[i] = prefix[mate[i] < 0 ? i : std::min(i, mate[i])];
    *countOut = prefix[n];
    return map;
}

// NativeAggregationOracle.cu:76-98 (mapSize): the map must be compact.
int mapSize(const std::vector<int>& map, int n) {
    if (int64_t(map.size()) != n || map.empty()) throw std::invalid_argument("Invalid aggregate map size");
    const int largest = *std::max_element(map.begin(), map.end());
    if (largest < 0 || largest >= n) throw std::invalid_argument("Invalid aggregate map range");
    std::vector<char> used(static_cast<size_t>(largest) + 1, 0);
    for (int value : map) {
        if (value < 0 || value > largest) throw std::invalid_argument("Invalid aggregate map entry");
        used[static_cast<size_t>(value)] = 1;
    }
    if (std::find(used.begin(), used.end(), 0) != used.end()) throw std::invalid_argument("Noncompact aggregate map");
    return largest + 1;
}

// NativeAggregationOracle.cu:207-221 / GpuAggregation.cu:534-576.
Csr contractGraph(const Csr& g, const std::vector<int>& map) {
    const int count = mapSize(map, g.n);
    std::vector<Entry> e;
    e.reserve(static_cast<size_t>(g.nnz()));
    for (int r = 0; r < g.n; ++r)
        for (int j = g.rp[r]; j < g.rp[r + 1]; ++j)
            if (map[r] != map[g.ci[j]] && g.v[j] > 0) e.push_back({map[r], map[g.ci[j]], g.v[j]});
    return compress(count, e, true);
}

struct Aggregates {
    std::vector<int> map, offsets, members;
    int64_t histogram[4] = {0, 0, 0, 0};
    int coarse() const { return int(offsets.size()) - 1; }
};

std::vector<std::vector<int>> memberGroups(const std::vector<int>& map, int count) {
    std::vector<std::vector<int>> groups(static_cast<size_t>(count));
    for (size_t i = 0; i < map.size(); ++i) groups[static_cast<size_t>(map[i])].push_back(int(i));
    return groups;
}

// Two-phase matching, composition, two snapshot rounds of singleton absorption
// and compaction: NativeAggregationOracle.cu:222-305, GpuAggregation.cu:766-818.
Aggregates aggregate(const Csr& g, int rounds, bool absorb) {
    if (rounds < 1 || rounds > 8 || g.n < 1) throw std::invalid_argument("Invalid matching configuration");
    int first = 0, second = 0;
    auto map = matchPairs(g, rounds, &first);
    const Csr contracted = contractGraph(g, map);
    const auto map2 = matchPairs(contracted, rounds, &second);
    for (auto& value : map) value = map2[static_cast<size_t>(value)];
    const int count = mapSize(map, g.n);
    if (absorb) {
        for (int round = 0; round < 2; ++round) {
            const auto groups = memberGroups(map, count);
            std::vector<int> proposal(static_cast<size_t>(g.n), -1);
            std::vector<double> scores(static_cast<size_t>(g.n), 0);
            for (int i = 0; i < g.n; ++i) {
                if (groups[static_cast<size_t>(map[i])].size() != 1) continue;
                for (int j = g.rp[i]; j < g.rp[i + 1]; ++j) {
                    const int c = map[static_cast<size_t>(g.ci[j])];
                    const double w = g.v[j];
                    if (groups[static_cast<size_t>(c)].size() < 2 || groups[static_cast<size_t>(c)].size() >= 4 || w <= 0) continue;
                    if (w > scores[i] || (w == scores[i] && (proposal[i] < 0 || c < proposal[i]))) {
                        proposal[i] = c;
                        scores[i] = w;
                    }
                }
            }
            auto next = map;
            for (int c = 0; c < count; ++c) {
                const auto& grp = groups[static_cast<size_t>(c)];
                if (grp.size() < 2 || grp.size() >= 4) continue;
                std::vector<int> cand;
                for (int row : grp)
                    for (int j = g.rp[row]; j < g.rp[row + 1]; ++j)
                        if (proposal[static_cast<size_t>(g.ci[j])] == c) cand.push_back(g.ci[j]);
                std::sort(cand.begin(), cand.end());
                cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
                std::sort(cand.begin(), cand.end(), [&scores](int a, int b) {
                    return scores[a] > scores[b] || (scores[a] == scores[b] && a < b);
                });
                const size_t accepted = std::min(cand.size(), static_cast<size_t>(4) - grp.size());
                for (size_t k = 0; k < accepted; ++k) next[static_cast<size_t>(cand[k])] = c;
            }
            map.swap(next);
        }
    }
    const auto groups = memberGroups(map, count);
    std::vector<int> compact(static_cast<size_t>(count), -1);
    int used = 0;
    for (int c = 0; c < count; ++c)
        if (!groups[static_cast<size_t>(c)].empty()) compact[static_cast<size_t>(c)] = used++;
    for (auto& value : map) value = compact[static_cast<size_t>(value)];
    Aggregates out;
    out.map = std::move(map);
    out.offsets.push_back(0);
    for (const auto& grp : groups) {
        if (grp.empty()) continue;
        if (grp.size() > 4) throw std::logic_error("Aggregate capacity exceeded");
        out.histogram[grp.size() - 1]++;
        out.members.insert(out.members.end(), grp.begin(), grp.end());
        out.offsets.push_back(int(out.members.size()));
    }
    return out;
}

// NativeAggregationOracle.cu:306-318 / GpuAggregation.cu:136-146,227-234.
Csr galerkin(const Csr& a, const std::vector<int>& map) {
    const int count = mapSize(map, a.n);
    std::vector<Entry> e;
    e.reserve(static_cast<size_t>(a.nnz()));
    for (int r = 0; r < a.n; ++r)
        for (int j = a.rp[r]; j < a.rp[r + 1]; ++j) e.push_back({map[r], map[static_cast<size_t>(a.ci[j])], a.v[j]});
    return compress(count, e, false);
}

// Dense partial-pivoting LU: GpuAggregation.cu:827-893.  The pivot threshold is
// 64*eps*max|A| (:861); the first maximal |m[i][k]| wins (:856-860).
bool luFactor(const Csr& a, std::vector<double>& m, std::vector<int>& piv) {
    const int n = a.n;
    m.assign(static_cast<size_t>(n) * n, 0.0);
    piv.assign(static_cast<size_t>(n), 0);
    double scale = 0;
    for (int r = 0; r < n; ++r)
        for (int j = a.rp[r]; j < a.rp[r + 1]; ++j) {
            m[static_cast<size_t>(r) * n + a.ci[j]] = a.v[j];
            scale = std::fmax(scale, std::fabs(a.v[j]));
        }
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(m[static_cast<size_t>(i) * n + k]) > std::fabs(m[static_cast<size_t>(p) * n + k])) p = i;
        const double pv = m[static_cast<size_t>(p) * n + k];
        if (!std::isfinite(pv) || std::fabs(pv) <= 64 * 2.2204460492503131e-16 * scale) return false;
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(m[static_cast<size_t>(p) * n + j], m[static_cast<size_t>(k) * n + j]);
        for (int r = k + 1; r < n; ++r) {
            m[static_cast<size_t>(r) * n + k] /= m[static_cast<size_t>(k) * n + k];
            for (int j = k + 1; j < n; ++j) m[static_cast<size_t>(r) * n + j] -= m[static_cast<size_t>(r) * n + k] * m[static_cast<size_t>(k) * n + j];
        }
    }
    for (double x : m)
        if (!std::isfinite(x)) return false;
    return true;
}

// GpuAggregation.cu:904-939: permute, forward, backward, all in double.
void luSolve(int n, const double* lu, const int* piv, std::vector<double>& v) {
    for (int k = 0; k < n; ++k) std::swap(v[k], v[static_cast<size_t>(piv[k])]);
    for (int k = 0; k < n; ++k)
        for (int r = k + 1; r < n; ++r) v[r] -= lu[static_cast<size_t>(r) * n + k] * v[k];
    for (int k = n - 1; k >= 0; --k) {
        v[k] /= lu[static_cast<size_t>(k) * n + k];
        for (int r = 0; r < k; ++r) v[r] -= lu[static_cast<size_t>(r) * n + k] * v[k];
    }
}

template <class T>
struct TypedCsr {
    int n = 0;
    std::vector<int> rp, ci;
    std::vector<T> v;
};

template <class T>
TypedCsr<T> typed(const Csr& a) {
    TypedCsr<T> t;
    t.n = a.n;
    t.rp = a.rp;
    t.ci = a.ci;
    t.v.resize(a.v.size());
    for (size_t j = 0; j < a.v.size(); ++j) t.v[j] = T(a.v[j]);
    return t;
}

// Sequential row sums in stored order (the oracle's fixed order).
template <class T>
void spmv(const TypedCsr<T>& a, const T* x, T* y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < a.n; ++i) {
        T acc = 0;
        for (int j = a.rp[i]; j < a.rp[i + 1]; ++j) acc += a.v[j] * x[a.ci[j]];
        y[i] = acc;
    }
}

// Diagonal.cuh:13-36: duplicate diagonal entries summed in T, then 1/d in T.
template <class T>
std::vector<T> inverseDiagonal(const TypedCsr<T>& a) {
    std::vector<T> inv(static_cast<size_t>(a.n));
    int bad = 0;
    for (int r = 0; r < a.n; ++r) {
        T d{};
        for (int j = a.rp[r]; j < a.rp[r + 1]; ++j)
            if (a.ci[j] == r) d += a.v[j];
        if (!std::isfinite(d) || d == 0 || !std::isfinite(T(1) / d)) bad = 1;
        inv[r] = T(1) / d;
    }
    if (bad) throw std::invalid_argument("Missing or invalid diagonal");
    return inv;
}

// Weighted Jacobi sweeps: Jacobi.cuh:60-74 with the update of
// IterationContext.cu:26-32, x += (w*d)*(b - Ax).
template <class T>
void jacobiSweeps(const TypedCsr<T>& a, const std::vector<T>& inv, int sweeps, T w, const T* b, T* x,
                  std::vector<T>& product) {
    product.resize(static_cast<size_t>(a.n));
    for (int s = 0; s < sweeps; ++s) {
        spmv(a, x, product.data());
#pragma omp parallel for schedule(static)
        for (int i = 0; i < a.n; ++i) x[i] += w * inv[i] * (b[i] - product[i]);
    }
}

double blockedDot(const double* x, const double* y, int64_t n) {
    const int64_t B = 4096;
    const int64_t nb = (n + B - 1) / B;
    std::vector<double> part(static_cast<size_t>(std::max<int64_t>(nb, 1)), 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; ++b) {
        double s = 0;
        const int64_t e = std::min(n, (b + 1) * B);
        for (int64_t i = b * B; i < e; ++i) s += x[i] * y[i];
        part[static_cast<size_t>(b)] = s;
    }
    double s = 0;
    for (int64_t b = 0; b < nb; ++b) s += part[static_cast<size_t>(b)];
    return s;
}

struct Options {
    int maxLevels = 20, coarseRows = 64, rounds = 8, pre = 1, post = 1, coarseSweeps = 8;
    double omega = 0.75, theta = 0.25;
    bool absorb = true;
    int cycle = 0;
    int precision = 64;
};

struct AmgBase {
    virtual ~AmgBase() = default;
    Options opt;
    std::vector<Csr> natives;
    std::vector<Aggregates> transfers;
    std::vector<int64_t> rows, nnz;
    std::vector<std::array<int64_t, 4>> hist;
    std::vector<int64_t> visits;
    std::string stopReason, coarseSolver;
    bool factored = false, singular = false;
    std::vector<double> lu;
    std::vector<int> piv;
    virtual void buildTyped() = 0;
    virtual void apply(const double* rhs, double* x) = 0;

    // Amg.cuh:81-143.  The LU is attempted at rows <= max(64, coarseRows)
    // (the reference caps coarseRows at 64, so its bound is 64).
    void setup(const Csr& input) {
        natives.clear();
        transfers.clear();
        rows.clear();
        nnz.clear();
        hist.clear();
        factored = singular = false;
        Csr native = canonical(input);
        for (int depth = 0;; ++depth) {
            validateRowmax(native, nullptr);
            rows.push_back(native.n);
            nnz.push_back(native.nnz());
            hist.push_back({0, 0, 0, 0});
            if (native.n <= opt.coarseRows) {
                stopReason = "coarse size";
                natives.push_back(std::move(native));
                break;
            }
            if (depth + 1 == opt.maxLevels) {
                stopReason = "level limit";
                natives.push_back(std::move(native));
                break;
            }
            Csr graph = strength(native, opt.theta);
            if (graph.nnz() == 0) {
                stopReason = "empty strength graph";
                natives.push_back(std::move(native));
                break;
            }
            Aggregates t = aggregate(graph, opt.rounds, opt.absorb);
            if (t.coarse() >= native.n) {
                stopReason = "no contraction";
                natives.push_back(std::move(native));
                break;
            }
            Csr coarse = galerkin(native, t.map);
            for (int k = 0; k < 4; ++k) hist.back()[k] = t.histogram[k];
            transfers.push_back(std::move(t));
            natives.push_back(std::move(native));
            native = std::move(coarse);
        }
        const Csr& last = natives.back();
        if (last.n <= std::max(64, opt.coarseRows)) {
            factored = luFactor(last, lu, piv);
            singular = !factored;
        }
        coarseSolver = factored ? "pivoted GPU LU" : (singular ? "Jacobi (singular LU)" : "Jacobi (coarsening stopped)");
        visits.assign(natives.size(), 0);
        buildTyped();
    }
};

template <class L>
struct Amg final : AmgBase {
    std::vector<TypedCsr<L>> A;
    std::vector<std::vector<L>> inv, b, x, r, product;

    void buildTyped() override {
        const size_t nl = natives.size();
        A.clear();
        inv.clear();
        for (size_t d = 0; d < nl; ++d) {
            A.push_back(typed<L>(natives[d]));
            inv.push_back(inverseDiagonal(A.back()));
        }
        b.assign(nl, {});
        x.assign(nl, {});
        r.assign(nl, {});
        product.assign(nl, {});
        for (size_t d = 0; d < nl; ++d) {
            b[d].assign(static_cast<size_t>(A[d].n), L(0));
            x[d].assign(static_cast<size_t>(A[d].n), L(0));
            r[d].assign(static_cast<size_t>(A[d].n), L(0));
        }
    }
    bool coarsest(size_t d) const { return d + 1 == natives.size(); }
    void smooth(size_t d, int sweeps, int /*unused*/) {
        if (sweeps > 0) jacobiSweeps(A[d], inv[d], sweeps, L(opt.omega), b[d].data(), x[d].data(), product[d]);
    }
    void restrictVector(size_t d, const std::vector<L>& fine) {  // GpuAggregation.cu:950-960
        const auto& t = transfers[d];
        auto& coarse = b[d + 1];
        for (int c = 0; c < t.coarse(); ++c) {
            L s = 0;
            for (int k = t.offsets[c]; k < t.offsets[c + 1]; ++k) s += fine[static_cast<size_t>(t.members[k])];
            coarse[c] = s;
        }
    }
    void prolong(size_t d) {  // GpuAggregation.cu:970-976
        const auto& t = transfers[d];
        for (size_t i = 0; i < x[d].size(); ++i) x[d][i] += x[d + 1][static_cast<size_t>(t.map[i])];
    }
    void coarseSolve(size_t d) {  // Amg.cuh:208-216
        if (factored) {
            std::vector<double> v(b[d].begin(), b[d].end());
            luSolve(A[d].n, lu.data(), piv.data(), v);
            for (size_t i = 0; i < v.size(); ++i) x[d][i] = L(v[i]);
        }
        else {
            jacobiSweeps(A[d], inv[d], opt.coarseSweeps, L(opt.omega), b[d].data(), x[d].data(), product[d]);
        }
    }
    // Cycle.cuh:35-63.
    void cycle(int kind, size_t d, bool zero) {
        ++visits[d];
        if (zero || kind == 2) std::fill(x[d].begin(), x[d].end(), L(0));
        if (coarsest(d)) {
            coarseSolve(d);
            return;
        }
        if (kind == 2) {
            restrictVector(d, b[d]);
            cycle(2, d + 1, true);
            prolong(d);
            cycle(0, d, false);
            return;
        }
        smooth(d, opt.pre, 0);
        spmv(A[d], x[d].data(), r[d].data());  // residual: IterationContext.cu:166-175
        for (size_t i = 0; i < r[d].size(); ++i) r[d][i] = b[d][i] - r[d][i];
        restrictVector(d, r[d]);
        cycle(kind, d + 1, true);
        if (kind == 1) cycle(kind, d + 1, false);
        prolong(d);
        smooth(d, opt.post, 0);
    }
    // Amg.cuh:149-161.
    void apply(const double* rhs, double* out) override {
        for (size_t i = 0; i < b[0].size(); ++i) b[0][i] = L(rhs[i]);
        std::fill(visits.begin(), visits.end(), 0);
        cycle(opt.cycle, 0, true);
        for (size_t i = 0; i < x[0].size(); ++i) out[i] = double(x[0][i]);
    }
};

struct Result {
    int status = 1, iterations = 0;
    double residual = 0, relative = 0;
};

struct Krylov {
    TypedCsr<double> A;
    AmgBase* M;
    int n;
    std::vector<double> rhs, x;
    int maxit, refresh;
    double rtol, atol;

    void precond(const double* in, double* out) {
        if (M)
            M->apply(in, out);
        else
            std::copy(in, in + n, out);
    }
    double dot(const std::vector<double>& a, const std::vector<double>& b) { return blockedDot(a.data(), b.data(), n); }
    double norm(const std::vector<double>& a) { return std::sqrt(dot(a, a)); }
    void residual(std::vector<double>& r) {
        spmv(A, x.data(), r.data());
        for (int i = 0; i < n; ++i) r[i] = rhs[i] - r[i];
    }
    // Cg.cuh:105-119 / BiCgStab.cuh:122-136.
    Result finish(int status, int iterations, double normB, double tol) {
        std::vector<double> r(static_cast<size_t>(n));
        residual(r);
        const double res = norm(r);
        if (!std::isfinite(res) || !std::isfinite(normB))
            status = 3;
        else if (res <= tol)
            status = 0;
        else if (status == 0)
            status = 1;
        return {status, iterations, res, normB > 0 ? res / normB : res};
    }
    // Cg.cuh:123-189.
    Result cg() {
        const size_t N = static_cast<size_t>(n);
        std::vector<double> r(N), z(N), p(N), q(N);
        const double normB = norm(rhs);
        const double tol = std::max(atol, rtol * normB);
        residual(r);
        double res = norm(r);
        if (!std::isfinite(normB) || !std::isfinite(res)) return finish(3, 0, normB, tol);
        if (res <= tol) return finish(0, 0, normB, tol);
        double prevRho = 0;
        bool fresh = true;
        for (int k = 0; k < maxit; ++k) {
            const int it = k + 1;
            precond(r.data(), z.data());
            const double rho = dot(r, z);
            if (!std::isfinite(rho)) return finish(3, it, normB, tol);
            if (rho <= 0) return finish(2, it, normB, tol);
            if (fresh) {
                p = z;
                fresh = false;
            }
            else {
                const double beta = rho / prevRho;
                if (!std::isfinite(beta)) return finish(3, it, normB, tol);
                for (int i = 0; i < n; ++i) p[i] = beta * p[i];
                for (int i = 0; i < n; ++i) p[i] = 1.0 * z[i] + p[i];
            }
            spmv(A, p.data(), q.data());
            const double curv = dot(p, q);
            if (!std::isfinite(curv)) return finish(3, it, normB, tol);
            if (curv <= 0) return finish(2, it, normB, tol);
            const double alpha = rho / curv;
            if (!std::isfinite(alpha)) return finish(3, it, normB, tol);
            for (int i = 0; i < n; ++i) x[i] = alpha * p[i] + x[i];
            for (int i = 0; i < n; ++i) r[i] = -alpha * q[i] + r[i];
            res = norm(r);
            if (!std::isfinite(res)) return finish(3, it, normB, tol);
            if (res <= tol || it % refresh == 0) {
                residual(r);
                res = norm(r);
                if (!std::isfinite(res) || res <= tol) return finish(1, it, normB, tol);
                fresh = true;
            }
            prevRho = rho;
        }
        return finish(1, maxit, normB, tol);
    }
    // BiCgStab.cuh:144-247.
    Result bicgstab() {
        const size_t N = static_cast<size_t>(n);
        std::vector<double> r(N), sh(N), p(N), v(N), s(N), t(N), ph(N), shat(N);
        const double normB = norm(rhs);
        const double tol = std::max(atol, rtol * normB);
        residual(r);
        double res = norm(r);
        if (!std::isfinite(normB) || !std::isfinite(res)) return finish(3, 0, normB, tol);
        if (res <= tol) return finish(0, 0, normB, tol);
        sh = r;
        double rhoPrev = 1, alpha = 1, omega = 1;
        bool fresh = true;
        for (int k = 0; k < maxit; ++k) {
            const int it = k + 1;
            const double rho = dot(sh, r);
            if (!std::isfinite(rho)) return finish(3, it, normB, tol);
            if (rho == 0) return finish(2, it, normB, tol);
            if (fresh) {
                p = r;
                fresh = false;
            }
            else {
                const double beta = (rho / rhoPrev) * (alpha / omega);
                if (!std::isfinite(beta)) return finish(3, it, normB, tol);
                for (int i = 0; i < n; ++i) p[i] = -omega * v[i] + p[i];
                for (int i = 0; i < n; ++i) p[i] = beta * p[i];
                for (int i = 0; i < n; ++i) p[i] = 1.0 * r[i] + p[i];
            }
            precond(p.data(), ph.data());
            spmv(A, ph.data(), v.data());
            const double den = dot(sh, v);
            if (!std::isfinite(den)) return finish(3, it, normB, tol);
            if (den == 0) return finish(2, it, normB, tol);
            alpha = rho / den;
            if (!std::isfinite(alpha)) return finish(3, it, normB, tol);
            s = r;
            for (int i = 0; i < n; ++i) s[i] = -alpha * v[i] + s[i];
            for (int i = 0; i < n; ++i) x[i] = alpha * ph[i] + x[i];
            const double ns = norm(s);
            if (!std::isfinite(ns)) return finish(3, it, normB, tol);
            if (ns <= tol) {
                residual(r);
                res = norm(r);
                if (res <= tol || !std::isfinite(res)) return finish(1, it, normB, tol);
                sh = r;
                fresh = true;
                continue;
            }
            precond(s.data(), shat.data());
            spmv(A, shat.data(), t.data());
            const double tt = dot(t, t), ts = dot(t, s);
            if (!std::isfinite(tt) || !std::isfinite(ts)) return finish(3, it, normB, tol);
            if (tt == 0) return finish(2, it, normB, tol);
            omega = ts / tt;
            if (!std::isfinite(omega)) return finish(3, it, normB, tol);
            if (omega == 0) return finish(2, it, normB, tol);
            for (int i = 0; i < n; ++i) x[i] = omega * shat[i] + x[i];
            r = s;
            for (int i = 0; i < n; ++i) r[i] = -omega * t[i] + r[i];
            res = norm(r);
            if (!std::isfinite(res)) return finish(3, it, normB, tol);
            if (res <= tol || it % refresh == 0) {
                residual(r);
                res = norm(r);
                if (res <= tol || !std::isfinite(res)) return finish(1, it, normB, tol);
                sh = r;
                fresh = true;
            }
            rhoPrev = rho;
        }
        return finish(1, maxit, normB, tol);
    }
};

template <class F>
auto guarded(F&& f, decltype(f()) failure) -> decltype(f()) {
    try {
        return f();
    }
    catch (const std::exception& e) {
        g_error = e.what();
        return failure;
    }
}

}  // namespace

extern "C" {

const char* tfo_last_error(void) { return g_error.c_str(); }

int tfo_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int64_t tfo_canonical(int n, const int* rp, const int* ci, const double* v, int* out_rp, int* out_ci, double* out_v) {
    return guarded(
        [&]() -> int64_t {
            Csr c = canonical(fromArrays(n, rp, ci, v));
            toArrays(c, out_rp, out_ci, out_v);
            return c.nnz();
        },
        int64_t(-1));
}

int tfo_validate_rowmax(int n, const int* rp, const int* ci, const double* v, double* rowmax) {
    return guarded(
        [&]() -> int {
            std::vector<double> m;
            validateRowmax(fromArrays(n, rp, ci, v), &m);
            std::copy(m.begin(), m.end(), rowmax);
            return 0;
        },
        -1);
}

int64_t tfo_strength(int n, const int* rp, const int* ci, const double* v, double theta, int* g_rp, int* g_ci,
                     double* g_w) {
    return guarded(
        [&]() -> int64_t {
            Csr a = fromArrays(n, rp, ci, v);
            validateCsr(a);
            Csr g = strength(a, theta);
            toArrays(g, g_rp, g_ci, g_w);
            return g.nnz();
        },
        int64_t(-1));
}

int tfo_match(int n, const int* g_rp, const int* g_ci, const double* g_w, int rounds, int* map) {
    return guarded(
        [&]() -> int {
            int count = 0;
            auto m = matchPairs(fromArrays(n, g_rp, g_ci, g_w), rounds, &count);
            std::copy(m.begin(), m.end(), map);
            return count;
        },
        -1);
}

int64_t tfo_contract(int n, const int* g_rp, const int* g_ci, const double* g_w, const int* map, int* out_rp,
                     int* out_ci, double* out_w) {
    return guarded(
        [&]() -> int64_t {
            std::vector<int> m(map, map + n);
            Csr c = contractGraph(fromArrays(n, g_rp, g_ci, g_w), m);
            toArrays(c, out_rp, out_ci, out_w);
            return c.nnz();
        },
        int64_t(-1));
}

int tfo_aggregate(int n, const int* g_rp, const int* g_ci, const double* g_w, int rounds, int absorb, int* map,
                  int* offsets, int* members, int64_t* histogram) {
    return guarded(
        [&]() -> int {
            Aggregates t = aggregate(fromArrays(n, g_rp, g_ci, g_w), rounds, absorb != 0);
            std::copy(t.map.begin(), t.map.end(), map);
            std::copy(t.offsets.begin(), t.offsets.end(), offsets);
            std::copy(t.members.begin(), t.members.end(), members);
            for (int k = 0; k < 4; ++k) histogram[k] = t.histogram[k];
            return t.coarse();
        },
        -1);
}

int64_t tfo_galerkin(int n, const int* rp, const int* ci, const double* v, const int* map, int* out_rp, int* out_ci,
                     double* out_v) {
    return guarded(
        [&]() -> int64_t {
            std::vector<int> m(map, map + n);
            Csr c = galerkin(fromArrays(n, rp, ci, v), m);
            toArrays(c, out_rp, out_ci, out_v);
            return c.nnz();
        },
        int64_t(-1));
}

int tfo_lu_factor(int n, const int* rp, const int* ci, const double* v, double* lu, int* pivots) {
    return guarded(
        [&]() -> int {
            std::vector<double> m;
            std::vector<int> p;
            const bool ok = luFactor(fromArrays(n, rp, ci, v), m, p);
            std::copy(m.begin(), m.end(), lu);
            std::copy(p.begin(), p.end(), pivots);
            return ok ? 1 : 0;
        },
        -1);
}

void tfo_lu_solve(int n, const double* lu, const int* pivots, const double* b, double* x) {
    std::vector<double> v(b, b + n);
    luSolve(n, lu, pivots, v);
    std::copy(v.begin(), v.end(), x);
}

void tfo_spmv(int n, const int* rp, const int* ci, const double* v, const double* x, double* y) {
    TypedCsr<double> a;
    a.n = n;
    a.rp.assign(rp, rp + n + 1);
    a.ci.assign(ci, ci + rp[n]);
    a.v.assign(v, v + rp[n]);
    spmv(a, x, y);
}

void tfo_spmv_f32(int n, const int* rp, const int* ci, const float* v, const float* x, float* y) {
    TypedCsr<float> a;
    a.n = n;
    a.rp.assign(rp, rp + n + 1);
    a.ci.assign(ci, ci + rp[n]);
    a.v.assign(v, v + rp[n]);
    spmv(a, x, y);
}

void* tfo_amg_create(const tfo_amg_options* o) {
    return guarded(
        [&]() -> void* {
            Options opt;
            if (o) {
                opt.maxLevels = o->max_levels;
                opt.coarseRows = o->coarse_rows;
                opt.rounds = o->matching_rounds;
                opt.pre = o->pre_sweeps;
                opt.post = o->post_sweeps;
                opt.coarseSweeps = o->coarse_sweeps;
                opt.omega = o->jacobi_weight;
                opt.theta = o->strength_threshold;
                opt.absorb = o->absorb_singletons != 0;
                opt.cycle = o->cycle;
                opt.precision = o->level_precision;
            }
            // Amg.cuh:66-75, with coarse_rows widened to 1..1024.
            if (opt.maxLevels < 1 || opt.maxLevels > 20 || opt.coarseRows < 1 || opt.coarseRows > 1024 ||
                opt.rounds < 1 || opt.rounds > 8 || opt.pre < 0 || opt.post < 0 || opt.coarseSweeps < 1 ||
                !std::isfinite(opt.theta) || opt.theta < 0 || opt.theta > 1 || !std::isfinite(opt.omega) ||
                opt.omega <= 0 || opt.omega >= 2 || opt.cycle < 0 || opt.cycle > 2 ||
                (opt.precision != 32 && opt.precision != 64))
                throw std::invalid_argument("Invalid AMG options");
            AmgBase* a = opt.precision == 32 ? static_cast<AmgBase*>(new Amg<float>()) : new Amg<double>();
            a->opt = opt;
            return a;
        },
        nullptr);
}

void tfo_amg_destroy(void* amg) { delete static_cast<AmgBase*>(amg); }

int tfo_amg_setup(void* amg, int n, const int* rp, const int* ci, const double* v) {
    return guarded(
        [&]() -> int {
            static_cast<AmgBase*>(amg)->setup(fromArrays(n, rp, ci, v));
            return 0;
        },
        -1);
}

int tfo_amg_levels(void* amg) { return int(static_cast<AmgBase*>(amg)->natives.size()); }

int tfo_amg_level_info(void* amg, int level, int64_t* rows, int64_t* nnz, int64_t* histogram) {
    auto* a = static_cast<AmgBase*>(amg);
    if (level < 0 || level >= int(a->natives.size())) return -1;
    *rows = a->rows[static_cast<size_t>(level)];
    *nnz = a->nnz[static_cast<size_t>(level)];
    for (int k = 0; k < 4; ++k) histogram[k] = a->hist[static_cast<size_t>(level)][static_cast<size_t>(k)];
    return 0;
}

int tfo_amg_level_matrix(void* amg, int level, int* rp, int* ci, double* v) {
    auto* a = static_cast<AmgBase*>(amg);
    if (level < 0 || level >= int(a->natives.size())) return -1;
    toArrays(a->natives[static_cast<size_t>(level)], rp, ci, v);
    return 0;
}

int tfo_amg_level_transfer(void* amg, int level, int* map, int* offsets, int* members) {
    auto* a = static_cast<AmgBase*>(amg);
    if (level < 0 || level >= int(a->transfers.size())) return -1;
    const auto& t = a->transfers[static_cast<size_t>(level)];
    std::copy(t.map.begin(), t.map.end(), map);
    std::copy(t.offsets.begin(), t.offsets.end(), offsets);
    std::copy(t.members.begin(), t.members.end(), members);
    return t.coarse();
}

const char* tfo_amg_stop_reason(void* amg) { return static_cast<AmgBase*>(amg)->stopReason.c_str(); }
const char* tfo_amg_coarse_solver(void* amg) { return static_cast<AmgBase*>(amg)->coarseSolver.c_str(); }

int tfo_amg_apply(void* amg, const double* rhs, double* x) {
    return guarded(
        [&]() -> int {
            static_cast<AmgBase*>(amg)->apply(rhs, x);
            return 0;
        },
        -1);
}

int tfo_amg_visits(void* amg, int64_t* visits) {
    auto* a = static_cast<AmgBase*>(amg);
    std::copy(a->visits.begin(), a->visits.end(), visits);
    return int(a->visits.size());
}

int tfo_krylov_solve(int kind, void* amg, int n, const int* rp, const int* ci, const double* v, const double* b,
                     double* x, int max_iterations, double rtol, double atol, int refresh, tfo_solve_result* result) {
    return guarded(
        [&]() -> int {
            if (max_iterations < 0 || refresh < 1 || !std::isfinite(rtol) || rtol < 0 || rtol >= 1 ||
                !std::isfinite(atol) || atol < 0)
                throw std::invalid_argument("Invalid solver options");
            Krylov k;
            Csr c = canonical(fromArrays(n, rp, ci, v));
            k.A.n = c.n;
            k.A.rp = c.rp;
            k.A.ci = c.ci;
            k.A.v = c.v;
            k.M = static_cast<AmgBase*>(amg);
            k.n = n;
            k.rhs.assign(b, b + n);
            k.x.assign(x, x + n);
            k.maxit = max_iterations;
            k.rtol = rtol;
            k.atol = atol;
            k.refresh = refresh;
            Result r = kind == 0 ? k.cg() : k.bicgstab();
            std::copy(k.x.begin(), k.x.end(), x);
            result->status = r.status;
            result->iterations = r.iterations;
            result->residual_norm = r.residual;
            result->relative_residual = r.relative;
            return 0;
        },
        -1);
}

int tfo_jacobi(int n, const int* rp, const int* ci, const double* v, int sweeps, double weight, const double* b,
               double* x) {
    return guarded(
        [&]() -> int {
            Csr c = canonical(fromArrays(n, rp, ci, v));
            TypedCsr<double> a;
            a.n = c.n;
            a.rp = c.rp;
            a.ci = c.ci;
            a.v = c.v;
            auto inv = inverseDiagonal(a);
            std::vector<double> product;
            jacobiSweeps(a, inv, sweeps, weight, b, x, product);
            return 0;
        },
        -1);
}

}  // extern "C"
