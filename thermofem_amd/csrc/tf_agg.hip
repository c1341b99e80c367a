// Aggregation-AMG setup on gfx950: CSR checks, row validation with the row
// maximum, the symmetric strength graph, synchronous pairwise matching,
// graph contraction, singleton absorption, member tables and the Galerkin
// product for piecewise-constant prolongation.
//
// Coordinate merges do not use a global sort.  Every merge here is local to
// one output row (a CSR row, or the union of the <= 4 member rows of an
// aggregate), so entries are staged per output row with key
// (column << 32 | position-in-row), each row segment is sorted on its own
// (insertion sort for short rows, LDS bitonic for long ones) and duplicate
// columns are merged sequentially.  Staging in encounter order (members
// ascending, entries in stored order) makes every sum run in original entry
// order, which is the order of the stable sort + sequential merge the
// reference's host oracle defines (test/support/NativeAggregationOracle.cu:43-75),
// so structures and values are bit-identical to it.
#include <algorithm>

#include "tf_device.hpp"
#include "tf_internal.hpp"

namespace tf {
namespace {

using u64 = unsigned long long;

enum Slot {
    kKeys = 0,
    kVals = 1,
    kCount = 2,
    kStart = 3,
    kFill = 4,
    kMate = 5,
    kProposal = 6,
    kFlags = 7,
    kPrefix = 8,
    kSizes = 9,
    kSlots = 10,
    kNextMap = 11,
    kPropWeight = 12,
    kRowmax = 13,
    kLarge = 14,
    kSecond = 15,
    kMemOffsets = 16,
    kMembers = 17,
    kHist = 18,
    kUnique = 19,
    kChanged = 20,
    kStaging = 21,
    kRawLen = 22,
    kMaxPart = 23,
    kSizes2 = 24,   // member table of the second pairing (the first's stays in kSizes / kSlots)
    kSlots2 = 25,
    kOldScan = 26,
};

inline dim3 grid1(int64_t n) { return dim3(unsigned((n + kBlock - 1) / kBlock > 0 ? (n + kBlock - 1) / kBlock : 1)); }

__device__ __forceinline__ u64 pack_key(int col, int pos) { return (u64(unsigned(col)) << 32) | unsigned(pos); }
__device__ __forceinline__ int key_col(u64 k) { return int(k >> 32); }

// SplitMix64 finalizer over the ordered pair (GpuAggregation.cu:49-54).
__device__ __forceinline__ u64 edge_priority(int i, int j) {
    u64 h = (u64(unsigned(min(i, j))) << 32) | unsigned(max(i, j));
    h = (h ^ (h >> 30)) * 0xbf58476d1ce4e5b9ULL;
    h = (h ^ (h >> 27)) * 0x94d049bb133111ebULL;
    return h ^ (h >> 31);
}

__global__ void k_fill_i32(int n, int value, int* p) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i < n) p[i] = value;
}

// checkCanonical (GpuAggregation.cu:283-304): bit 0 unsorted/duplicate, bit 1 invalid.
__global__ void k_csr_check(int n, int nnz, const int* __restrict__ rp, const int* __restrict__ ci,
                            const double* __restrict__ v, int* flags) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= n) return;
    if (rp[0] != 0 || rp[n] != nnz || rp[i] < 0 || rp[i] > rp[i + 1] || rp[i + 1] > nnz) {
        atomicOr(flags, 2);
        return;
    }
    // entries in chunks of 8 whose loads go out together (indices clamped
    // into the row, extra slots ignored); one flag update per thread
    int previous = -1, bad = 0;
    const int e = rp[i + 1];
    for (int j0 = rp[i]; j0 < e; j0 += 8) {
        int cc[8];
        double vv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = min(j0 + u, e - 1);
            cc[u] = ci[j];
            vv[u] = v[j];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (j0 + u >= e) break;
            if (cc[u] < 0 || cc[u] >= n || !isfinite(vv[u])) bad |= 2;
            if (cc[u] <= previous) bad |= 1;
            previous = cc[u];
        }
    }
    if (bad) atomicOr(flags, bad);
}

// Device CSR structure check of CsrMatrix::fromBuffers (CsrMatrix.cu:12-31).
__global__ void k_csr_structure(int rows, int cols, int64_t count, const int* __restrict__ rp,
                                const int* __restrict__ ci, int* bad) {
    const int r = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (r == 0 && (rp[0] != 0 || int64_t(rp[rows]) != count)) atomicOr(bad, 1);
    if (r >= rows) return;
    if (rp[r] < 0 || rp[r + 1] < rp[r] || int64_t(rp[r + 1]) > count) {
        atomicOr(bad, 1);
        return;
    }
    for (int j = rp[r]; j < rp[r + 1]; ++j)
        if (ci[j] < 0 || ci[j] >= cols) atomicOr(bad, 1);
}

// validate (GpuAggregation.cu:247-271).
__global__ void k_validate(int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ a,
                           int* bad, double* __restrict__ rowmax) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= n) return;
    double diag = 0, mx = 0;
    bool invalid = false;
    // chunks of 8 entries whose loads go out together (clamped indices)
    const int e = rp[i + 1];
    for (int j0 = rp[i]; j0 < e; j0 += 8) {
        int cc[8];
        double aa[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = min(j0 + u, e - 1);
            cc[u] = ci[j];
            aa[u] = a[j];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (j0 + u >= e) break;
            invalid |= !isfinite(aa[u]);
            if (cc[u] == i)
                diag = aa[u];
            else
                mx = fmax(mx, fabs(aa[u]));
        }
    }
    if (invalid || diag == 0 || !isfinite(diag) || !isfinite(1.0 / diag)) atomicOr(bad, 1);
    if (rowmax) rowmax[i] = mx;
}

// k_validate fused with the level's inverse diagonal in its precision T
// (canonical rows: the diagonal is one entry, inv = 1 / T(a_ii)), raising
// bad2 as k_inverse_diagonal would.  G lanes per row, one entry per lane
// (coalesced loads); the row maximum is a max (exact in any order) and the
// diagonal comes from the lane that holds it.
template <class T, int G>
__global__ void k_validate_inv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                               const double* __restrict__ a, int* bad, double* __restrict__ rowmax,
                               T* __restrict__ inv, int* bad2) {
    const int lane = int(threadIdx.x) % G;
    const int groups = int(gridDim.x) * (kBlock / G);
    for (int i = int((int64_t(blockIdx.x) * kBlock + threadIdx.x) / G); i < n; i += groups) {
        double diag = 0, mx = 0;
        int invalid = 0, found = 0;
        const int e = rp[i + 1];
        for (int j = rp[i] + lane; j < e; j += G) {
            const int c = ci[j];
            const double av = a[j];
            invalid |= isfinite(av) ? 0 : 1;
            if (c == i) {
                diag = av;
                found = 1;
            }
            else {
                mx = fmax(mx, fabs(av));
            }
        }
        int who = found ? lane : G;
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) {
            mx = fmax(mx, __shfl_xor(mx, o, G));
            invalid |= __shfl_xor(invalid, o, G);
            who = min(who, __shfl_xor(who, o, G));
        }
        const int base = (int(threadIdx.x) & (kWave - 1)) / G * G;
        diag = __shfl(diag, base + (who < G ? who : 0), kWave);
        if (who == G) diag = 0;
        if (lane != 0) continue;
        if (invalid || diag == 0 || !isfinite(diag) || !isfinite(1.0 / diag)) atomicOr(bad, 1);
        rowmax[i] = mx;
        const T d = T(diag);
        const T r = T(1) / d;
        if (!isfinite(d) || d == T(0) || !isfinite(r)) atomicOr(bad2, 1);
        inv[i] = r;
    }
}

// Strength edges (GpuAggregation.cu:339-364): w = |a_ij| / rowmax_i for
// j != i when rowmax_i > 0, w > 0 and w >= theta; both directions.
__global__ void k_strength_count(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                 const double* __restrict__ a, const double* __restrict__ rowmax, double theta,
                                 int* cnt) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= n) return;
    const double mx = rowmax[i];
    if (!(mx > 0)) return;
    int out = 0;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
        const int c = ci[j];
        if (c == i) continue;
        const double w = fabs(a[j]) / mx;
        if (w >= theta && w > 0) {
            ++out;
            atomicAdd(cnt + c, 1);
        }
    }
    if (out) atomicAdd(cnt + i, out);
}

__global__ void k_strength_fill(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                const double* __restrict__ a, const double* __restrict__ rowmax, double theta,
                                const int* __restrict__ start, int* fill, u64* __restrict__ keys,
                                double* __restrict__ vals) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= n) return;
    const double mx = rowmax[i];
    if (!(mx > 0)) return;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
        const int c = ci[j];
        if (c == i) continue;
        const double w = fabs(a[j]) / mx;
        if (w >= theta && w > 0) {
            const int p = atomicAdd(fill + i, 1);
            keys[start[i] + p] = pack_key(c, p);
            vals[start[i] + p] = w;
            const int q = atomicAdd(fill + c, 1);
            keys[start[c] + q] = pack_key(i, q);
            vals[start[c] + q] = w;
        }
    }
}


// Flags a canonical matrix whose pattern is not symmetric.
__global__ void k_struct_sym(int n, const int* __restrict__ rp, const int* __restrict__ ci, int* flag) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= n) return;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
        const int c = ci[j];
        if (c != i && find_sorted_col(ci, rp[c], rp[c + 1], i) < 0) {
            atomicOr(flag, 1);
            return;
        }
    }
}

// Strength graph of a structurally symmetric canonical matrix without any
// merge: row i of the graph keeps the pattern of row i, with weight
// max(q_ij, q_ji) where q_ij = |a_ij|/rowmax_i when that qualifies
// (rowmax_i > 0, q > 0, q >= theta) and 0 otherwise; entries with weight 0
// are dropped.  This is exactly the max-merge of both emitted directions.
// One pass writes the weights in the layout of A (column -1 = dropped) and
// the kept count per row; k_compact_rows packs them.
// VS: the values are symmetric bit for bit, so a_ji is read as a_ij and the
// transpose is not looked up.
template <int G, bool VS = false>
__global__ void k_strength_sym(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                               const double* __restrict__ a, const double* __restrict__ rowmax, double theta,
                               int* __restrict__ scol, double* __restrict__ sw, int* __restrict__ cnt,
                               int* nonsym) {
    // G lanes per row, one entry per lane (coalesced loads).  With VS each
    // group takes 4 rows at once (rows g, g + groups, ...), whose loads are
    // all issued before the first is used (level 0 of the 128^3 heat
    // operator: 135 -> 112 us); the transpose lookups' registers leave room
    // for one row a group otherwise.  The transpose of entry (i, c) is
    // looked up in row c.  With nonsym, an entry without
    // its transpose raises the flag (the pattern was not known to be
    // symmetric: the result is then discarded).
    constexpr int R = VS ? 4 : 1;
    const int lane = int(threadIdx.x) % G;
    const int groups = int(gridDim.x) * (kBlock / G);
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[n] = 0;
    for (int i0 = int((int64_t(blockIdx.x) * kBlock + threadIdx.x) / G); i0 < n; i0 += groups * R) {
        int row[R], s[R], e[R], kept[R];
        double mi[R];
        int len = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            row[k] = i0 + k * groups;
            const bool valid = row[k] < n;
            const int ii = valid ? row[k] : i0;
            s[k] = rp[ii];
            e[k] = valid ? rp[ii + 1] : s[k];
            mi[k] = rowmax[ii];
            kept[k] = 0;
            len = max(len, e[k] - s[k]);
        }
        for (int t0 = 0; t0 < len; t0 += G) {
            int c[R], jj[R];
            double av[R];
            bool live[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const int j = s[k] + t0 + lane;
                live[k] = j < e[k];
                jj[k] = live[k] ? j : 0;
                c[k] = ci[jj[k]];
                av[k] = a[jj[k]];
            }
            double mc[R], at[R];
            bool has[R];
#pragma unroll
            for (int k = 0; k < R; ++k) mc[k] = rowmax[live[k] ? c[k] : 0];
            if constexpr (VS) {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    at[k] = av[k];
                    has[k] = true;
                }
            }
            else {
                int pos[R];
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int ck = live[k] ? c[k] : 0;
                    pos[k] = live[k] && c[k] != row[k] ? find_sorted_col(ci, rp[ck], rp[ck + 1], row[k]) : -1;
                }
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    has[k] = pos[k] >= 0;
                    at[k] = a[has[k] ? pos[k] : 0];
                    if (live[k] && c[k] != row[k] && !has[k] && nonsym) atomicOr(nonsym, 1);
                }
            }
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (!live[k]) continue;
                double w = 0.0;
                if (c[k] != row[k]) {
                    const double wi = mi[k] > 0 ? fabs(av[k]) / mi[k] : 0.0;
                    const double qi = (wi >= theta && wi > 0) ? wi : 0.0;
                    const double wc = (mc[k] > 0 && has[k]) ? fabs(at[k]) / mc[k] : 0.0;
                    const double qc = (wc >= theta && wc > 0) ? wc : 0.0;
                    w = qi > qc ? qi : qc;
                }
                scol[jj[k]] = w > 0 ? c[k] : -1;
                sw[jj[k]] = w;
                kept[k] += w > 0 ? 1 : 0;
            }
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int kt = group_sum<G>(kept[k]);
            if (lane == 0 && row[k] < n) cnt[row[k]] = kt;
        }
    }
}

// Learned weights in place of the strength measure: every off-diagonal
// entry with a positive learned weight is an edge of that weight (the
// weights are symmetric, so the graph is).
__global__ void k_learned_graph(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                const double* __restrict__ lw, int* __restrict__ scol, double* __restrict__ sw,
                                int* __restrict__ cnt) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i > n) return;
    if (i == n) {
        cnt[n] = 0;
        return;
    }
    int kept = 0;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
        const int c = ci[j];
        const double w = c != i && lw[j] > 0 ? lw[j] : 0.0;
        scol[j] = w > 0 ? c : -1;
        sw[j] = w;
        kept += w > 0 ? 1 : 0;
    }
    cnt[i] = kept;
}

// Packs the kept entries of each row in order: G lanes per row, positions
// from a ballot within the group.
template <int G>
__global__ void k_compact_rows(int n, const int* __restrict__ rp, const int* __restrict__ scol,
                               const double* __restrict__ sw, const int* __restrict__ orp, int* __restrict__ oci,
                               double* __restrict__ ow) {
    const int i = int((int64_t(blockIdx.x) * kBlock + threadIdx.x) / G);
    const int lane = int(threadIdx.x) % G;
    if (i >= n) return;
    const int shift = (int(threadIdx.x) & (kWave - 1)) / G * G;
    const u64 groupMask = G == 64 ? ~0ull : ((1ull << G) - 1);
    const int e = rp[i + 1];
    int o = orp[i];
    for (int j0 = rp[i]; j0 < e; j0 += G) {
        const int j = j0 + lane;
        const int c = j < e ? scol[j] : -1;
        const u64 m = (u64(__ballot(c >= 0)) >> shift) & groupMask;
        if (c >= 0) {
            const int at = o + __popcll(m & ((1ull << lane) - 1));
            oci[at] = c;
            ow[at] = sw[j];
        }
        o += __popcll(m);
    }
}

// Lanes per row for the row-parallel setup kernels: the smallest power of
// two at least the average row length, 4..32.
inline int lanes_per_row(int n, int64_t nnz) {
    const double avg = n > 0 ? double(nnz) / double(n) : 1.0;
    int g = 4;
    while (g < 32 && g < avg) g <<= 1;
    return g;
}

inline dim3 grid_rows(int64_t rows, int g) {
    const int64_t threads = rows * g;
    return dim3(unsigned((threads + kBlock - 1) / kBlock > 0 ? (threads + kBlock - 1) / kBlock : 1));
}
// The same for grid-stride kernels: at most kRowGroupBlocks blocks (two
// rounds of resident 256-thread blocks on 256 CUs), each group of lanes
// taking rows i, i + groups, ...
constexpr unsigned kRowGroupBlocks = 4096;
inline dim3 grid_rows_capped(int64_t rows, int g) {
    const dim3 d = grid_rows(rows, g);
    return dim3(d.x < kRowGroupBlocks ? d.x : kRowGroupBlocks);
}

// Output row I gathers the rows of its members (all rows of `offsets`
// order; offsets == nullptr means the identity aggregate {I}); columns go
// through map (nullptr: identity).  With `graph` set, intra-aggregate and
// non-positive entries are dropped (mapGraph, GpuAggregation.cu:534-553).
__device__ __forceinline__ void member_range(int I, const int* offsets, int& b, int& e) {
    if (offsets) {
        b = offsets[I];
        e = offsets[I + 1];
    }
    else {
        b = I;
        e = I + 1;
    }
}

__global__ void k_gather_count(int nc, const int* __restrict__ offsets, const int* __restrict__ members,
                               const int* __restrict__ map, const int* __restrict__ rp, const int* __restrict__ ci,
                               const double* __restrict__ v, int graph, int* cnt, int* rawLen) {
    const int I = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (I == nc) cnt[nc] = 0;
    if (I >= nc) return;
    int mb, me;
    member_range(I, offsets, mb, me);
    int total = 0, raw = 0;
    if (!graph) {
        // four member rows per step, loads clamped into range and issued together
        for (int k = mb; k < me; k += 4) {
            int rows[4], len[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = min(k + u, me - 1);
                rows[u] = offsets ? members[q] : q;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) len[u] = rp[rows[u] + 1] - rp[rows[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) raw += k + u < me ? len[u] : 0;
        }
    }
    else {
        for (int k = mb; k < me; ++k) {
            const int m = offsets ? members[k] : k;
            const int s = rp[m], e = rp[m + 1];
            raw += e - s;
            // 8 entries per step: column, weight and mapped column loads overlap
            for (int j0 = s; j0 < e; j0 += 8) {
                int cc[8];
                double vv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int j = min(j0 + u, e - 1);
                    cc[u] = ci[j];
                    vv[u] = v[j];
                }
                if (map) {
#pragma unroll
                    for (int u = 0; u < 8; ++u) cc[u] = map[cc[u]];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) total += (j0 + u < e && cc[u] != I && vv[u] > 0) ? 1 : 0;
            }
        }
    }
    cnt[I] = graph ? total : raw;
    if (rawLen) rawLen[I] = raw;  // LDS staging holds unfiltered entries
}

// Maximum of a[0..n): block maxima to partials, then one block combines them.
__global__ __launch_bounds__(kBlock) void k_max_partials(int n, const int* __restrict__ a, int* __restrict__ part) {
    __shared__ int wm[kBlock / kWave];
    int m = 0;
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock) m = max(m, a[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) wm[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        int r = 0;
        for (int i = 0; i < kBlock / kWave; ++i) r = max(r, wm[i]);
        part[blockIdx.x] = r;
    }
}
__global__ __launch_bounds__(kBlock) void k_max_final(int nparts, const int* __restrict__ part, int* out) {
    __shared__ int wm[kBlock / kWave];
    int m = 0;
    for (int i = int(threadIdx.x); i < nparts; i += kBlock) m = max(m, part[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) wm[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        int r = 0;
        for (int i = 0; i < kBlock / kWave; ++i) r = max(r, wm[i]);
        *out = r;
    }
}

// ---- per-output-row merges (canonical form, contraction, Galerkin) ----
// Output row I gathers its raw entries -- the rows of its members in member
// order, each in stored order; graph mode drops intra-aggregate and
// non-positive entries (mapGraph, GpuAggregation.cu:534-553) -- into LDS as
// keys (column << 32 | encounter index) and values, sorts the keys with a
// bitonic network, so that equal columns become adjacent in encounter order,
// and reduces each run of equal columns sequentially in encounter order (sum,
// or max for the graph: the order of the stable sort and reduce-by-key of
// GpuAggregation.cu:178-221, which the host oracle fixes).  A run's rank
// among the row's runs is its place in the sorted row.  Results go to
// staging at gstart[I]; ucnt[I] = distinct columns.  Rows of up to 16 or 32
// raw entries are merged by groups of as many lanes, one key a lane, sorted
// in registers; longer rows (up to kMergeCap) are queued for a wave each,
// whose 64..512 keys are sorted in registers too (1..8 a lane); longer still
// raise the overflow flag and the whole product takes the global sort path.
constexpr int kNoCol = 0x7fffffff;  // dropped entries and padding sort last
constexpr int kMergeSmall = 32;
constexpr int kMergeCap = 512;

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// raw[I] = entries output row I gathers (before filtering); raw[nc] = 0.
__global__ void k_raw_count(int nc, const int* __restrict__ offsets, const int* __restrict__ members,
                            const int* __restrict__ rp, int* __restrict__ raw) {
    const int I = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (I == nc) raw[nc] = 0;
    if (I >= nc) return;
    int mb, me;
    member_range(I, offsets, mb, me);
    int total = 0;
    // four member rows per step, loads clamped into range and issued together
    for (int k = mb; k < me; k += 4) {
        int rows[4], len[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = min(k + u, me - 1);
            rows[u] = offsets ? members[q] : q;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) len[u] = rp[rows[u] + 1] - rp[rows[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u) total += k + u < me ? len[u] : 0;
    }
    raw[I] = total;
}

// Gathers output row I's raw entries into K / V with the group's G lanes
// (up to G member rows at a time: lane k looks up member k's row range, a
// group prefix sum places the rows one after another in member order, then
// all lanes gather the entries of those rows together); returns the count.
template <int G>
__device__ int merge_gather(int I, int mb, int me, const int* __restrict__ offsets, const int* __restrict__ members,
                            const int* __restrict__ map, const int* __restrict__ rp, const int* __restrict__ ci,
                            const double* __restrict__ v, int graph, u64* K, double* V, int* mstart, int* moff,
                            int lane) {
    int m = 0;
    for (int k0 = mb; k0 < me; k0 += G) {
        const int k = k0 + lane;
        int s = 0, len = 0;
        if (k < me) {
            const int row = offsets ? members[k] : k;
            s = rp[row];
            len = rp[row + 1] - s;
        }
        int incl = len;
#pragma unroll
        for (int o = 1; o < G; o <<= 1) {
            const int t = __shfl_up(incl, o, G);
            if (lane >= o) incl += t;
        }
        const int total = __shfl(incl, G - 1, G);
        mstart[lane] = s;
        moff[lane] = incl - len;
        wave_lds_sync();
        const int nm = min(G, me - k0);
        for (int t = lane; t < total; t += G) {
            int q = 0;
            for (int b = 1; b < nm; ++b)
                if (moff[b] <= t) q = b;
            const int at = mstart[q] + (t - moff[q]);
            const int J = map ? map[ci[at]] : ci[at];
            const double x = v[at];
            const bool keep = !graph || (J != I && x > 0);
            K[m + t] = pack_key(keep ? J : kNoCol, m + t);
            V[m + t] = x;
        }
        m += total;
        wave_lds_sync();
    }
    return m;
}

template <int G>
__device__ __forceinline__ u64 shfl_xor_u64(u64 v, int j) {
    const int lo = __shfl_xor(int(unsigned(v)), j, G), hi = __shfl_xor(int(unsigned(v >> 32)), j, G);
    return (u64(unsigned(hi)) << 32) | unsigned(lo);
}

// Ascending bitonic sort of W = G R keys held by a group of G lanes in
// registers (key r of lane l is element G r + l): partners less than G
// apart are exchanged by shuffles, the others between the lane's own
// registers.
template <int G, int R>
__device__ __forceinline__ void group_bitonic(u64 (&key)[R], int lane) {
    constexpr int W = R * G;
#pragma unroll
    for (int k = 2; k <= W; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= G) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int p = r ^ (j / G);
                    if (p > r) {
                        const bool up = ((r * G + lane) & k) == 0;
                        const u64 a = key[r], b = key[p];
                        if ((a > b) == up) {
                            key[r] = b;
                            key[p] = a;
                        }
                    }
                }
            }
            else {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const u64 other = shfl_xor_u64<G>(key[r], j);
                    const bool low = (lane & j) == 0, up = ((r * G + lane) & k) == 0;
                    key[r] = (low == up) ? (other < key[r] ? other : key[r]) : (other > key[r] ? other : key[r]);
                }
            }
        }
}

// Sorts K[0, W) (W = G R) of one group through registers.
template <int G, int R>
__device__ __forceinline__ void group_sort_lds(u64* K, int lane) {
    u64 key[R];
#pragma unroll
    for (int r = 0; r < R; ++r) key[r] = K[r * G + lane];
    group_bitonic<G, R>(key, lane);
#pragma unroll
    for (int r = 0; r < R; ++r) K[r * G + lane] = key[r];
    wave_lds_sync();
}

// Reduces the runs of equal columns of the sorted K[0, W) (W >= G): lane l
// owns positions [l P, l P + P), P = W / G, and the runs that start there.
// Returns the row's distinct columns.
template <int G, bool Max>
__device__ int merge_runs(const u64* K, const double* V, int W, int lane, int base, int* __restrict__ scol,
                          double* __restrict__ sval, int* bad) {
    const int P = W / G, p0 = lane * P;
    int heads = 0;
    for (int q = 0; q < P; ++q) {
        const int p = p0 + q, J = key_col(K[p]);
        heads += (J != kNoCol && (p == 0 || key_col(K[p - 1]) != J)) ? 1 : 0;
    }
    int incl = heads;
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
        const int t = __shfl_up(incl, o, G);
        if (lane >= o) incl += t;
    }
    int rank = incl - heads;
    for (int q = 0; q < P; ++q) {
        const int p = p0 + q, J = key_col(K[p]);
        if (J == kNoCol || (p > 0 && key_col(K[p - 1]) == J)) continue;
        double acc = V[unsigned(K[p] & 0xffffffffu)];
        bool merged = false;
        for (int r = p + 1; r < W; ++r) {
            const u64 kr = K[r];
            if (key_col(kr) != J) break;
            const double x = V[unsigned(kr & 0xffffffffu)];
            acc = Max ? (acc > x ? acc : x) : acc + x;
            merged = true;
        }
        if (merged && !isfinite(acc)) atomicOr(bad, 1);
        scol[base + rank] = J;
        sval[base + rank] = acc;
        ++rank;
    }
    return __shfl(incl, G - 1, G);
}

// Rows of up to G (16 or 32) raw entries, G lanes each; longer rows are
// appended to the worklist wl (count in *wlCount) for k_merge_large.
template <bool Max, int G>
__global__ __launch_bounds__(kBlock) void k_merge_small(int nc, const int* __restrict__ offsets,
                                                        const int* __restrict__ members, const int* __restrict__ map,
                                                        const int* __restrict__ rp, const int* __restrict__ ci,
                                                        const double* __restrict__ v, int graph,
                                                        const int* __restrict__ raw, const int* __restrict__ gstart,
                                                        int* __restrict__ scol, double* __restrict__ sval,
                                                        int* __restrict__ ucnt, int* __restrict__ wl, int* wlCount,
                                                        int* bad) {
    constexpr int NG = kBlock / G;
    __shared__ u64 lk[NG][G];
    __shared__ double lv[NG][G];
    __shared__ int mstart[NG][G], moff[NG][G];
    const int g = int(threadIdx.x) / G, lane = int(threadIdx.x) % G;
    const int I = int(blockIdx.x) * NG + g;
    if (blockIdx.x == 0 && threadIdx.x == 0) ucnt[nc] = 0;
    if (I >= nc) return;
    // the row's length, member range and staging offset loaded together
    int mb, me;
    member_range(I, offsets, mb, me);
    const int r = raw[I], base = gstart[I];
    if (r > G) {
        if (lane == 0) wl[atomicAdd(wlCount, 1)] = I;
        return;
    }
    const int m = merge_gather<G>(I, mb, me, offsets, members, map, rp, ci, v, graph, lk[g], lv[g], mstart[g],
                                  moff[g], lane);
    // one key a lane, sorted in registers (compare-exchange with lane ^ j)
    u64 key = lane < m ? lk[g][lane] : pack_key(kNoCol, lane);
#pragma unroll
    for (int k = 2; k <= G; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const u64 other = __shfl_xor(key, j, G);
            const bool low = (lane & j) == 0, up = (lane & k) == 0;
            key = (low == up) ? (other < key ? other : key) : (other > key ? other : key);
        }
    lk[g][lane] = key;
    wave_lds_sync();
    const int u = merge_runs<G, Max>(lk[g], lv[g], G, lane, base, scol, sval, bad);
    if (lane == 0) ucnt[I] = u;
}

// The worklist's rows, a wave each, sorted at the next power of two >= 64
// of their raw length; rows past kMergeCap raise *overflow.  (Groups of 32
// lanes with two keys a lane measured slower: the LDS of eight rows a block
// halves the resident waves.)
template <bool Max>
__global__ __launch_bounds__(kBlock) void k_merge_large(int nc, const int* __restrict__ offsets,
                                                        const int* __restrict__ members, const int* __restrict__ map,
                                                        const int* __restrict__ rp, const int* __restrict__ ci,
                                                        const double* __restrict__ v, int graph,
                                                        const int* __restrict__ raw, const int* __restrict__ gstart,
                                                        int* __restrict__ scol, double* __restrict__ sval,
                                                        int* __restrict__ ucnt, const int* __restrict__ wl,
                                                        const int* wlCount, int* bad, int* overflow) {
    constexpr int G = kWave, NG = kBlock / G;
    __shared__ u64 lk[NG][kMergeCap];
    __shared__ double lv[NG][kMergeCap];
    __shared__ int mstart[NG][G], moff[NG][G];
    const int g = int(threadIdx.x) / G, lane = int(threadIdx.x) % G;
    const int count = wl ? *wlCount : nc;
    if (!wl && blockIdx.x == 0 && threadIdx.x == 0) ucnt[nc] = 0;
    for (int q = int(blockIdx.x) * NG + g; q < count; q += int(gridDim.x) * NG) {
        const int I = wl ? wl[q] : q;
        const int r = raw[I];
        if (r > kMergeCap) {
            if (lane == 0) {
                atomicOr(overflow, 1);
                ucnt[I] = 0;
            }
            continue;
        }
        int W = G;
        while (W < r) W <<= 1;
        int mb, me;
        member_range(I, offsets, mb, me);
        const int m = merge_gather<G>(I, mb, me, offsets, members, map, rp, ci, v, graph, lk[g], lv[g], mstart[g],
                                      moff[g], lane);
        for (int e = m + lane; e < W; e += G) lk[g][e] = pack_key(kNoCol, e);
        wave_lds_sync();
        switch (W / G) {
            case 1: group_sort_lds<G, 1>(lk[g], lane); break;
            case 2: group_sort_lds<G, 2>(lk[g], lane); break;
            case 4: group_sort_lds<G, 4>(lk[g], lane); break;
            default: group_sort_lds<G, 8>(lk[g], lane); break;
        }
        const int u = merge_runs<G, Max>(lk[g], lv[g], W, lane, gstart[I], scol, sval, bad);
        if (lane == 0) ucnt[I] = u;
        wave_lds_sync();
    }
}

// dst[0] = src[0] (a device scalar into a flag slot, without a copy command).
__global__ void k_copy_int(const int* __restrict__ src, int* __restrict__ dst) { *dst = *src; }

// Staging -> CSR, 16 lanes per row (coalesced row copies).
__global__ __launch_bounds__(kBlock) void k_dedup_copy(int nc, const int* __restrict__ gstart,
                                                       const int* __restrict__ scol, const double* __restrict__ sval,
                                                       const int* __restrict__ orp, int* __restrict__ oci,
                                                       double* __restrict__ ov) {
    const int I = int(blockIdx.x) * (kBlock / 16) + int(threadIdx.x) / 16;
    const int lane = int(threadIdx.x) % 16;
    if (I >= nc) return;
    const int b = gstart[I], o = orp[I], u = orp[I + 1] - o;
    for (int k = lane; k < u; k += 16) {
        oci[o + k] = scol[b + k];
        ov[o + k] = sval[b + k];
    }
}

__global__ void k_gather_fill(int nc, const int* __restrict__ offsets, const int* __restrict__ members,
                              const int* __restrict__ map, const int* __restrict__ rp, const int* __restrict__ ci,
                              const double* __restrict__ v, int graph, const int* __restrict__ start,
                              u64* __restrict__ keys, double* __restrict__ vals) {
    const int I = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (I >= nc) return;
    int mb, me;
    member_range(I, offsets, mb, me);
    int pos = 0;
    const int base = start[I];
    for (int k = mb; k < me; ++k) {
        const int m = offsets ? members[k] : k;
        for (int j = rp[m]; j < rp[m + 1]; ++j) {
            const int J = map ? map[ci[j]] : ci[j];
            if (graph && (J == I || !(v[j] > 0))) continue;
            keys[base + pos] = pack_key(J, pos);
            vals[base + pos] = v[j];
            ++pos;
        }
    }
}

constexpr int kSmallSegment = 32;
constexpr int kLdsSegment = 2048;

// Short segments: stable insertion sort by key in place; long ones are
// queued for the block-wide sort.
__global__ void k_segsort_small(int rows, const int* __restrict__ start, u64* keys, double* vals, int* large,
                                int* largeCount) {
    const int r = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (r >= rows) return;
    const int b = start[r], len = start[r + 1] - b;
    if (len > kSmallSegment) {
        large[atomicAdd(largeCount, 1)] = r;
        return;
    }
    u64* k = keys + b;
    double* x = vals + b;
    for (int i = 1; i < len; ++i) {
        const u64 kk = k[i];
        const double xx = x[i];
        int j = i;
        while (j > 0 && k[j - 1] > kk) {
            k[j] = k[j - 1];
            x[j] = x[j - 1];
            --j;
        }
        k[j] = kk;
        x[j] = xx;
    }
}

// Bitonic network whose compare-exchanges all move the minimum to the lower
// index (mirror step, then half-cleaners); indices >= len act as +inf and
// never move, so any len sorts without padding storage.
template <class KeyAt, class Swap>
__device__ void bitonic(int len, KeyAt keyAt, Swap swapAt) {
    int N = 1;
    while (N < len) N <<= 1;
    for (int k = 2; k <= N; k <<= 1) {
        const int half = k >> 1;
        for (int t = int(threadIdx.x); t < N / 2; t += int(blockDim.x)) {
            const int blk = t / half, off = t % half;
            const int i = blk * k + off, p = blk * k + k - 1 - off;
            if (p < len && keyAt(i) > keyAt(p)) swapAt(i, p);
        }
        __syncthreads();
        for (int j = half >> 1; j > 0; j >>= 1) {
            for (int t = int(threadIdx.x); t < N / 2; t += int(blockDim.x)) {
                const int i = (t / j) * 2 * j + (t % j), p = i + j;
                if (p < len && keyAt(i) > keyAt(p)) swapAt(i, p);
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_segsort_large(const int* __restrict__ start, u64* keys, double* vals,
                                                          const int* __restrict__ large,
                                                          const int* __restrict__ largeCount) {
    __shared__ u64 sk[kLdsSegment];
    __shared__ double sv[kLdsSegment];
    const int count = *largeCount;
    for (int q = int(blockIdx.x); q < count; q += int(gridDim.x)) {
        const int r = large[q];
        const int b = start[r], len = start[r + 1] - b;
        u64* gk = keys + b;
        double* gv = vals + b;
        if (len <= kLdsSegment) {
            for (int i = int(threadIdx.x); i < len; i += kBlock) {
                sk[i] = gk[i];
                sv[i] = gv[i];
            }
            __syncthreads();
            bitonic(
                len, [&](int i) { return sk[i]; },
                [&](int i, int p) {
                    const u64 tk = sk[i];
                    sk[i] = sk[p];
                    sk[p] = tk;
                    const double tv = sv[i];
                    sv[i] = sv[p];
                    sv[p] = tv;
                });
            for (int i = int(threadIdx.x); i < len; i += kBlock) {
                gk[i] = sk[i];
                gv[i] = sv[i];
            }
            __syncthreads();
        }
        else {
            bitonic(
                len, [&](int i) { return gk[i]; },
                [&](int i, int p) {
                    const u64 tk = gk[i];
                    gk[i] = gk[p];
                    gk[p] = tk;
                    const double tv = gv[i];
                    gv[i] = gv[p];
                    gv[p] = tv;
                });
        }
    }
}

__global__ void k_unique_count(int rows, const int* __restrict__ start, const u64* __restrict__ keys, int* ucnt) {
    const int r = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (r == rows) ucnt[rows] = 0;
    if (r >= rows) return;
    int u = 0, prev = -1;
    for (int j = start[r]; j < start[r + 1]; ++j) {
        const int c = key_col(keys[j]);
        if (u == 0 || c != prev) {
            ++u;
            prev = c;
        }
    }
    ucnt[r] = u;
}

// Sequential merge of equal columns: sum (a + b in order) or max
// (a > b ? a : b, GpuAggregation.cu:17-23).
template <bool Max>
__global__ void k_merge_write(int rows, const int* __restrict__ start, const u64* __restrict__ keys,
                              const double* __restrict__ vals, const int* __restrict__ orp, int* __restrict__ oci,
                              double* __restrict__ ov, int* bad) {
    const int r = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (r >= rows) return;
    int o = orp[r] - 1, prev = -1;
    bool first = true;
    for (int j = start[r]; j < start[r + 1]; ++j) {
        const int c = key_col(keys[j]);
        const double x = vals[j];
        if (first || c != prev) {
            ++o;
            oci[o] = c;
            ov[o] = x;
            prev = c;
            first = false;
        }
        else {
            const double acc = ov[o];
            const double m = Max ? (acc > x ? acc : x) : acc + x;
            if (!isfinite(m)) atomicOr(bad, 1);
            ov[o] = m;
        }
    }
}

// ---- matching (GpuAggregation.cu:402-505) ----
// A round after one that accepted nothing would reproduce the same
// proposals and accept nothing again, so it is skipped (prevChanged == 0).
// A vertex whose scan found no candidate is marked kNoCandidate (still an
// unmatched vertex to every reader of mate, which only tests mate < 0): the
// unmatched vertices only ever become fewer, so it can never find one later,
// and later rounds skip its scan (its proposal stays "none").
constexpr int kNoCandidate = -2;

// The proposal of vertex i (its best unmatched candidate, or -1).
__device__ __forceinline__ void propose_vertex(int i, const int* __restrict__ rp, const int* __restrict__ ci,
                                               const double* __restrict__ w, int* __restrict__ mate,
                                               int* __restrict__ proposal, int mark) {
    int best = -1;
    double score = 0;
    u64 priority = 0;
    const int mi = mate[i];
    if (mi < 0 && mi != kNoCandidate) {
        // The best candidate is the maximum of a total order (weight, then
        // priority, then the smaller column), so entries can be examined in
        // chunks: the loads of 8 entries and their mates' lookups go out
        // together (indices clamped into the row, extra slots ignored).
        const int e = rp[i + 1];
        for (int j0 = rp[i]; j0 < e; j0 += 8) {
            int cc[8], mm[8];
            double ss[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int j = min(j0 + u, e - 1);
                cc[u] = ci[j];
                ss[u] = w[j];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) mm[u] = mate[cc[u]];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int c = cc[u];
                const double s = ss[u];
                if (j0 + u >= e || c == i || mm[u] >= 0 || s <= 0 || s < score) continue;
                const u64 cand = edge_priority(i, c);
                if (s > score || (s == score && (best < 0 || cand > priority || (cand == priority && c < best)))) {
                    best = c;
                    score = s;
                    priority = cand;
                }
            }
        }
        if (best < 0 && mark) mate[i] = kNoCandidate;
    }
    proposal[i] = best;
}

// Sorts four ints ascending (a five-exchange network).
__device__ __forceinline__ void sort4(int& a, int& b, int& c, int& d) {
    auto cx = [](int& x, int& y) {
        const int lo = min(x, y), hi = max(x, y);
        x = lo;
        y = hi;
    };
    cx(a, b);
    cx(c, d);
    cx(a, c);
    cx(b, d);
    cx(b, c);
}

// Member-table entry (sizes[a], slots[4a..]) of the aggregate a that root i
// and its mate m (m < 0: unmatched) form, as the atomic slotting and
// per-aggregate sort of GpuAggregation.cu:586-632 would leave it: with no
// input table the members are {i} or {i, m} (i < m); with one (a pairing of
// aggregates, each of at most two members) the sorted union of the member
// lists of i and m.
__device__ __forceinline__ void write_members(int i, int m, int a, const int* __restrict__ inSizes,
                                              const int* __restrict__ inSlots, int* __restrict__ sizes,
                                              int* __restrict__ slots) {
    int* out = slots + size_t(a) * 4;
    if (!inSizes) {
        out[0] = i;
        if (m >= 0) out[1] = m;
        sizes[a] = m >= 0 ? 2 : 1;
        return;
    }
    constexpr int kNone = 0x7fffffff;
    const int na = inSizes[i], nb = m >= 0 ? inSizes[m] : 0;
    const int* xs = inSlots + size_t(i) * 4;
    const int* ys = inSlots + size_t(m >= 0 ? m : i) * 4;
    int x0 = xs[0], x1 = na > 1 ? xs[1] : kNone;
    int y0 = nb > 0 ? ys[0] : kNone, y1 = nb > 1 ? ys[1] : kNone;
    sort4(x0, x1, y0, y1);
    const int total = na + nb;
    out[0] = x0;
    if (total > 1) out[1] = x1;
    if (total > 2) out[2] = y0;
    if (total > 3) out[3] = y1;
    sizes[a] = total;
}

// Mutual proposals become pairs; true when vertex i was matched.
__device__ __forceinline__ bool accept_vertex(int i, const int* __restrict__ proposal, int* __restrict__ mate) {
    const int p = proposal[i];
    if (p >= 0 && proposal[p] == i) {
        mate[i] = p;
        return true;
    }
    return false;
}

__global__ void k_propose(int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ w,
                          int* __restrict__ mate, int* __restrict__ proposal, const int* prevChanged, int mark) {
    if (prevChanged && *prevChanged == 0) return;
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= n) return;
    propose_vertex(i, rp, ci, w, mate, proposal, mark);
}

__global__ void k_accept(int n, const int* __restrict__ proposal, int* __restrict__ mate, const int* prevChanged,
                         int* changed) {
    if (prevChanged && *prevChanged == 0) return;
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= n) return;
    if (accept_vertex(i, proposal, mate) && changed) *changed = 1;
}

// The whole matching of a small graph (n <= kPairBlockRows) in one
// workgroup: the rounds (a round without an acceptance ends them, as the
// per-round kernels skip every later round), then the roots' exclusive
// scan in LDS and the aggregate ids.  The same operations in the same
// order as the per-round kernels, so the same matching.
constexpr int kPairThreads = 1024;
constexpr int kPairBlockRows = 4096;
__global__ __launch_bounds__(kPairThreads) void k_pair_block(int n, const int* __restrict__ rp,
                                                             const int* __restrict__ ci,
                                                             const double* __restrict__ w, int rounds,
                                                             int* __restrict__ mate, int* __restrict__ proposal,
                                                             int* __restrict__ map, int* count,
                                                             const int* __restrict__ inSizes,
                                                             const int* __restrict__ inSlots, int* __restrict__ sizes,
                                                             int* __restrict__ slots) {
    __shared__ int prefix[kPairBlockRows + 1];
    __shared__ int wsum[kPairThreads / kWave];
    __shared__ int changed;
    const int t = int(threadIdx.x);
    for (int i = t; i < n; i += kPairThreads) mate[i] = -1;
    __syncthreads();
    for (int r = 0; r < rounds; ++r) {
        if (t == 0) changed = 0;
        for (int i = t; i < n; i += kPairThreads) propose_vertex(i, rp, ci, w, mate, proposal, 1);
        __syncthreads();
        bool any = false;
        for (int i = t; i < n; i += kPairThreads) any |= accept_vertex(i, proposal, mate);
        if (any) changed = 1;
        __syncthreads();
        if (!changed) break;
        __syncthreads();  // (changed is reset only after every thread has read it)
    }
    // roots: flag[i] = 1 for an unmatched vertex or the smaller end of a pair;
    // thread t scans items [4t, 4t + 4) of the n + 1 flags
    constexpr int kItems = (kPairBlockRows + 1 + kPairThreads - 1) / kPairThreads;
    int f[kItems], local = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = t * kItems + k;
        f[k] = (i < n && (mate[i] < 0 || i < mate[i])) ? 1 : 0;
        local += f[k];
    }
    int incl = local;
    for (int o = 1; o < kWave; o <<= 1) {
        const int v = __shfl_up(incl, o, kWave);
        if ((t & (kWave - 1)) >= o) incl += v;
    }
    if ((t & (kWave - 1)) == kWave - 1) wsum[t / kWave] = incl;
    __syncthreads();
    int before = 0;
    for (int q = 0; q < t / kWave; ++q) before += wsum[q];
    int run = before + incl - local;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = t * kItems + k;
        if (i <= n && i <= kPairBlockRows) prefix[i] = run;
        run += f[k];
    }
    __syncthreads();
    for (int i = t; i < n; i += kPairThreads) {
        const int m = mate[i];
        const int root = m < 0 ? i : min(i, m);
        map[i] = prefix[root];
        if (sizes && root == i) write_members(i, m, prefix[root], inSizes, inSlots, sizes, slots);
    }
    if (t == 0) {
        *count = prefix[n];
        if (sizes) sizes[prefix[n]] = 0;
    }
}

__global__ void k_roots(int n, const int* __restrict__ mate, int* __restrict__ flags) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i <= n) flags[i] = (i < n && (mate[i] < 0 || i < mate[i])) ? 1 : 0;
}

// Aggregate ids from the roots' prefix (launched over n + 1 threads); with a
// member table the roots also write their aggregates' entries and thread n
// the table's closing zero.
__global__ void k_assign(int n, const int* __restrict__ mate, const int* __restrict__ prefix, int* __restrict__ map,
                         const int* __restrict__ inSizes, const int* __restrict__ inSlots, int* __restrict__ sizes,
                         int* __restrict__ slots) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i > n) return;
    if (i == n) {
        if (sizes) sizes[prefix[n]] = 0;
        return;
    }
    const int m = mate[i];
    const int root = m < 0 ? i : min(i, m);
    const int a = prefix[root];
    map[i] = a;
    if (sizes && root == i) write_members(i, m, a, inSizes, inSlots, sizes, slots);
}

__global__ void k_compose(int n, int* __restrict__ map, const int* __restrict__ second) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i < n) map[i] = second[map[i]];
}

// ---- members (GpuAggregation.cu:586-632) ----
// sizes[0, count) = 0 and the overflow flag cleared, in one launch
__global__ void k_members_reset(int count, int* __restrict__ sizes, int* flag) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i < count) sizes[i] = 0;
    if (i == 0) *flag = 0;
}

__global__ void k_slot_members(int n, const int* __restrict__ map, int* sizes, int* __restrict__ slots) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= n) return;
    const int c = map[i];
    const int slot = atomicAdd(sizes + c, 1);
    if (slot < 4) slots[size_t(c) * 4 + slot] = i;
}

__global__ void k_sort_members(int nc, const int* __restrict__ sizes, int* __restrict__ slots, int* overflow) {
    const int c = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (c >= nc) return;
    const int sz = sizes[c];
    if (sz > 4) {
        atomicOr(overflow, 1);
        return;
    }
    int* m = slots + size_t(c) * 4;
    for (int i = 1; i < sz; ++i) {
        const int x = m[i];
        int j = i;
        while (j > 0 && m[j - 1] > x) {
            m[j] = m[j - 1];
            --j;
        }
        m[j] = x;
    }
}

// ---- absorption (GpuAggregation.cu:646-720) ----
// (also copies map into next, the map the acceptance then edits)
__global__ void k_propose_absorption(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                     const double* __restrict__ w, const int* __restrict__ map,
                                     const int* __restrict__ sizes, int* __restrict__ proposal,
                                     double* __restrict__ scores, int* __restrict__ next) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= n) return;
    next[i] = map[i];
    int best = -1;
    double score = 0;
    if (sizes[map[i]] == 1) {
        // maximum under (weight desc, aggregate asc): chunks of 8 entries
        const int e = rp[i + 1];
        for (int j0 = rp[i]; j0 < e; j0 += 8) {
            int cc[8], sz[8];
            double ww[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int j = min(j0 + u, e - 1);
                cc[u] = ci[j];
                ww[u] = w[j];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) cc[u] = map[cc[u]];
#pragma unroll
            for (int u = 0; u < 8; ++u) sz[u] = sizes[cc[u]];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int c = cc[u];
                const double s = ww[u];
                if (j0 + u >= e || sz[u] < 2 || sz[u] >= 4 || s <= 0) continue;
                if (s > score || (s == score && (best < 0 || c < best))) {
                    best = c;
                    score = s;
                }
            }
        }
    }
    proposal[i] = best;
    scores[i] = score;
}

// Aggregates of 2 or 3 members take their best singleton proposers (up to
// 4 members).  The member table (sizes, slots) is brought up to date in
// place, as rebuilding it from the new map would leave it: the aggregate's
// sorted list gains the absorbed rows and each absorbed row's old
// (singleton) aggregate, cur[row], becomes empty.  Only accepting aggregates
// (sizes 2..3) are read here and only they and emptied singletons written.
__global__ void k_accept_absorption(int nc, const int* __restrict__ rp, const int* __restrict__ ci,
                                    const int* __restrict__ cur, int* __restrict__ sizes, int* __restrict__ slots,
                                    const int* __restrict__ proposal, const double* __restrict__ scores,
                                    int* __restrict__ nextMap) {
    const int c = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (c >= nc) return;
    const int size = sizes[c];
    if (size < 2 || size >= 4) return;
    int selected[2] = {-1, -1};
    double weights[2] = {-1, -1};
    const int capacity = 4 - size;
    // The chosen rows are the top `capacity` proposers under (score desc,
    // row asc), a total order, so entries can be visited in chunks of 8
    // whose loads overlap (indices clamped into the row, extra slots ignored).
    for (int k = 0; k < size; ++k) {
        const int row = slots[size_t(c) * 4 + k];
        const int e = rp[row + 1];
        for (int j0 = rp[row]; j0 < e; j0 += 8) {
            int ii[8], pp[8];
            double ss[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) ii[u] = ci[min(j0 + u, e - 1)];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                pp[u] = proposal[ii[u]];
                ss[u] = scores[ii[u]];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
            const int i = ii[u];
            if (j0 + u >= e || pp[u] != c || i == selected[0] || i == selected[1]) continue;
            const double s = ss[u];
            for (int q = 0; q < capacity; ++q) {
                if (s > weights[q] || (s == weights[q] && i < selected[q])) {
                    for (int r = capacity - 1; r > q; --r) {
                        selected[r] = selected[r - 1];
                        weights[r] = weights[r - 1];
                    }
                    selected[q] = i;
                    weights[q] = s;
                    break;
                }
            }
            }
        }
    }
    constexpr int kNone = 0x7fffffff;
    int* m = slots + size_t(c) * 4;
    int m0 = m[0], m1 = m[1], m2 = size > 2 ? m[2] : kNone, m3 = kNone;
    int added = 0;
    for (int q = 0; q < capacity; ++q) {
        const int row = selected[q];
        if (row < 0) continue;
        nextMap[row] = c;
        sizes[cur[row]] = 0;
        if (size + added == 2)
            m2 = row;
        else
            m3 = row;
        ++added;
    }
    if (!added) return;
    sort4(m0, m1, m2, m3);
    m[0] = m0;
    m[1] = m1;
    m[2] = m2;
    if (size + added > 3) m[3] = m3;
    sizes[c] = size + added;
}

__global__ void k_occupied(int nc, const int* __restrict__ sizes, int* __restrict__ flags) {
    const int c = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (c <= nc) flags[c] = (c < nc && sizes[c] > 0) ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_pack_members(int nc, const int* __restrict__ sizes,
                                                         const int* __restrict__ slots,
                                                         const int* __restrict__ offsets, int* __restrict__ out,
                                                         int* histogram) {
    __shared__ int h[4];
    if (threadIdx.x < 4) h[threadIdx.x] = 0;
    __syncthreads();
    const int c = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (c < nc) {
        const int sz = sizes[c];
        for (int k = 0; k < sz; ++k) out[offsets[c] + k] = slots[size_t(c) * 4 + k];
        if (histogram && sz) atomicAdd(&h[sz - 1], 1);
    }
    __syncthreads();
    if (histogram && threadIdx.x < 4 && h[threadIdx.x]) atomicAdd(histogram + threadIdx.x, h[threadIdx.x]);
}

// The compacted member list after absorption: old aggregate c (of `bound`)
// that kept members becomes aggregate prefix[c]; its members go to
// out[oldscan[c]...] (oldscan: exclusive scan of the old sizes, which is
// also the compacted offsets), offsets[prefix[c]] = oldscan[c], and thread
// `bound` closes the offsets.  The size histogram as k_pack_members.
__global__ __launch_bounds__(kBlock) void k_pack_compact(int bound, const int* __restrict__ sizes,
                                                         const int* __restrict__ slots,
                                                         const int* __restrict__ prefix,
                                                         const int* __restrict__ oldscan, int* __restrict__ offsets,
                                                         int* __restrict__ out, int* histogram) {
    __shared__ int h[4];
    if (threadIdx.x < 4) h[threadIdx.x] = 0;
    __syncthreads();
    const int c = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (c == bound) {
        offsets[prefix[bound]] = oldscan[bound];
    }
    else if (c < bound) {
        const int sz = sizes[c];
        if (sz > 0) {
            const int at = oldscan[c];
            offsets[prefix[c]] = at;
            for (int k = 0; k < sz; ++k) out[at + k] = slots[size_t(c) * 4 + k];
            atomicAdd(&h[sz - 1], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x < 4 && h[threadIdx.x]) atomicAdd(histogram + threadIdx.x, h[threadIdx.x]);
}

// ---------------- host helpers ----------------

// mark: record vertices left without a candidate (kNoCandidate) so later
// rounds skip them; the single-round entry point leaves mate as the
// reference's matching would.
void match_round_flags(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* w, int* mate,
                       int* proposal, const int* prevChanged, int* changed, bool mark = true) {
    (void)nnz;  // one thread per row measured faster here than a group of lanes per row
    hipLaunchKernelGGL(k_propose, grid1(n), dim3(kBlock), 0, c.stream, n, rp, ci, w, mate, proposal, prevChanged,
                       int(mark));
    check(hipGetLastError(), "propose launch");
    hipLaunchKernelGGL(k_accept, grid1(n), dim3(kBlock), 0, c.stream, n, proposal, mate, prevChanged, changed);
    check(hipGetLastError(), "accept launch");
}

void fill_i32(tf_ctx& c, int n, int value, int* p) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_i32, grid1(n), dim3(kBlock), 0, c.stream, n, value, p);
    check(hipGetLastError(), "fill launch");
}

// Sort each row segment of (keys, vals) by key.
void segsort(tf_ctx& c, int rows, const int* start, u64* keys, double* vals) {
    if (rows <= 0) return;
    int* large = c.wst<int>(kLarge, size_t(rows) + 1);
    int* largeCount = large + rows;
    fill_i32(c, 1, 0, largeCount);
    hipLaunchKernelGGL(k_segsort_small, grid1(rows), dim3(kBlock), 0, c.stream, rows, start, keys, vals, large,
                       largeCount);
    check(hipGetLastError(), "segsort small launch");
    const int blocks = rows < 512 ? rows : 512;
    hipLaunchKernelGGL(k_segsort_large, dim3(blocks), dim3(kBlock), 0, c.stream, start, keys, vals, large, largeCount);
    check(hipGetLastError(), "segsort large launch");
}

// Merge equal columns of sorted row segments into a new CSR.
DevCsr merge_rows(tf_ctx& c, int rows, const int* start, const u64* keys, const double* vals, bool maximum) {
    DevCsr out;
    out.n = rows;
    int* ucnt = c.wst<int>(kUnique, size_t(rows) + 1);
    hipLaunchKernelGGL(k_unique_count, grid1(int64_t(rows) + 1), dim3(kBlock), 0, c.stream, rows, start, keys, ucnt);
    check(hipGetLastError(), "unique count launch");
    out.rp = dev_alloc<int>(c, size_t(rows) + 1);
    scan_excl(c, rows + 1, ucnt, out.rp);
    out.nnz = c.read_int(out.rp + rows);
    out.ci = dev_alloc<int>(c, size_t(out.nnz));
    out.v = dev_alloc<double>(c, size_t(out.nnz));
    c.clear_flags(1);
    if (maximum)
        hipLaunchKernelGGL(k_merge_write<true>, grid1(rows), dim3(kBlock), 0, c.stream, rows, start, keys, vals,
                           out.rp, out.ci, out.v, c.dflags);
    else
        hipLaunchKernelGGL(k_merge_write<false>, grid1(rows), dim3(kBlock), 0, c.stream, rows, start, keys, vals,
                           out.rp, out.ci, out.v, c.dflags);
    check(hipGetLastError(), "merge launch");
    c.read_flags(1);
    if (c.hflags[0]) {
        free_csr(c, out);
        throw Error(TF_E_NUMERIC, "Nonfinite aggregate reduction");
    }
    return out;
}

// Gather member rows into per-output-row segments and merge duplicate
// columns: LDS merges (k_merge_small / k_merge_large) when every row fits
// kMergeCap raw entries, otherwise staged keys + per-segment sort +
// sequential merge.  nnzIn: entries of the rows gathered (bounds the
// staging; < 0 reads it back).
DevCsr gather_merge(tf_ctx& c, int nc, const int* offsets, const int* members, const int* map, const int* rp,
                    const int* ci, const double* v, bool graph, bool maximum, int64_t nnzIn) {
    int* raw = c.wst<int>(kRawLen, size_t(nc) + 1);
    hipLaunchKernelGGL(k_raw_count, grid1(int64_t(nc) + 1), dim3(kBlock), 0, c.stream, nc, offsets, members, rp, raw);
    check(hipGetLastError(), "raw count launch");
    int* start = c.wst<int>(kStart, size_t(nc) + 1);
    scan_excl(c, nc + 1, raw, start);
    if (nnzIn < 0) nnzIn = c.read_int(start + nc);
    int* scol = c.wst<int>(kKeys, size_t(nnzIn) + 1);
    double* sval = c.wst<double>(kVals, size_t(nnzIn) + 1);
    int* ucnt = c.wst<int>(kUnique, size_t(nc) + 1);
    int* wl = c.wst<int>(kLarge, size_t(nc) + 1);
    // flags: 0 non-finite merge, 1 worklist length, 2 overflow, 3 output nnz
    c.clear_flags(3);
    {
        // Rows short on average: groups of 16 lanes (<= 12 entries on
        // average), 32 (<= 28) or a wave with one key a lane (<= 60), the
        // few longer rows queued for waves with up to 512 keys; otherwise
        // the latter throughout.  (The one-key-a-lane wave keeps its LDS
        // small, so many more rows are in flight than in the queue's
        // kernel: level-1 Galerkin products average ~41 raw entries.)
        const bool tiny = nnzIn <= int64_t(12) * nc;
        const bool small = nnzIn <= int64_t(28) * nc;
        const bool medium = nnzIn <= int64_t(60) * nc;
        const int ng = kBlock / (tiny ? 16 : small ? kMergeSmall : kWave);
        const dim3 grid((nc + ng - 1) / ng), large(std::max(1, std::min((nc + 3) / 4, 4096)));
        int* list = medium ? wl : nullptr;
#define TF_MERGE(MX)                                                                                           \
    if (tiny) {                                                                                                \
        hipLaunchKernelGGL((k_merge_small<MX, 16>), grid, dim3(kBlock), 0, c.stream, nc, offsets, members, map, rp, \
                           ci, v, int(graph), raw, start, scol, sval, ucnt, wl, c.dflags + 1, c.dflags);         \
    }                                                                                                          \
    else if (small) {                                                                                          \
        hipLaunchKernelGGL((k_merge_small<MX, kMergeSmall>), grid, dim3(kBlock), 0, c.stream, nc, offsets, members, \
                           map, rp, ci, v, int(graph), raw, start, scol, sval, ucnt, wl, c.dflags + 1, c.dflags); \
    }                                                                                                          \
    else if (medium) {                                                                                         \
        hipLaunchKernelGGL((k_merge_small<MX, kWave>), grid, dim3(kBlock), 0, c.stream, nc, offsets, members, map, \
                           rp, ci, v, int(graph), raw, start, scol, sval, ucnt, wl, c.dflags + 1, c.dflags);     \
    }                                                                                                          \
    hipLaunchKernelGGL(k_merge_large<MX>, large, dim3(kBlock), 0, c.stream, nc, offsets, members, map, rp, ci, v, \
                       int(graph), raw, start, scol, sval, ucnt, list, c.dflags + 1, c.dflags, c.dflags + 2)
        if (maximum) {
            TF_MERGE(true);
        }
        else {
            TF_MERGE(false);
        }
#undef TF_MERGE
        check(hipGetLastError(), "merge launch");
    }
    DevCsr out;
    out.n = nc;
    out.rp = dev_alloc<int>(c, size_t(nc) + 1);
    scan_excl(c, nc + 1, ucnt, out.rp);
    hipLaunchKernelGGL(k_copy_int, dim3(1), dim3(1), 0, c.stream, out.rp + nc, c.dflags + 3);
    c.read_flags(4);
    if (!c.hflags[2]) {
        if (c.hflags[0]) {
            dev_free(c, out.rp);
            throw Error(TF_E_NUMERIC, "Nonfinite aggregate reduction");
        }
        out.nnz = c.hflags[3];
        out.ci = dev_alloc<int>(c, size_t(out.nnz));
        out.v = dev_alloc<double>(c, size_t(out.nnz));
        hipLaunchKernelGGL(k_dedup_copy, dim3((nc + kBlock / 16 - 1) / (kBlock / 16)), dim3(kBlock), 0, c.stream, nc,
                           start, scol, sval, out.rp, out.ci, out.v);
        check(hipGetLastError(), "dedup copy launch");
        return out;
    }
    // a row longer than kMergeCap: every row through the global sort path
    dev_free(c, out.rp);
    int* cnt = c.wst<int>(kCount, size_t(nc) + 1);
    hipLaunchKernelGGL(k_gather_count, grid1(int64_t(nc) + 1), dim3(kBlock), 0, c.stream, nc, offsets, members, map,
                       rp, ci, v, int(graph), cnt, static_cast<int*>(nullptr));
    check(hipGetLastError(), "gather count launch");
    scan_excl(c, nc + 1, cnt, start);
    const int total = c.read_int(start + nc);
    u64* keys = c.wst<u64>(kKeys, size_t(total));
    double* vals = c.wst<double>(kVals, size_t(total));
    hipLaunchKernelGGL(k_gather_fill, grid1(nc), dim3(kBlock), 0, c.stream, nc, offsets, members, map, rp, ci, v,
                       int(graph), start, keys, vals);
    check(hipGetLastError(), "gather fill launch");
    segsort(c, nc, start, keys, vals);
    return merge_rows(c, nc, start, keys, vals, maximum);
}

// Member table: sizes (nc+1, last entry 0) and 4-slot sorted members.
void build_members(tf_ctx& c, int n, int nc, const int* map, int** sizesOut, int** slotsOut) {
    int* sizes = c.wst<int>(kSizes, size_t(nc) + 1);
    int* slots = c.wst<int>(kSlots, size_t(nc) * 4 + 1);
    hipLaunchKernelGGL(k_members_reset, grid1(int64_t(nc) + 1), dim3(kBlock), 0, c.stream, nc + 1, sizes, c.dflags);
    hipLaunchKernelGGL(k_slot_members, grid1(n), dim3(kBlock), 0, c.stream, n, map, sizes, slots);
    check(hipGetLastError(), "slot members launch");
    hipLaunchKernelGGL(k_sort_members, grid1(nc), dim3(kBlock), 0, c.stream, nc, sizes, slots, c.dflags);
    check(hipGetLastError(), "sort members launch");
    *sizesOut = sizes;
    *slotsOut = slots;
}

}  // namespace

void free_csr(tf_ctx& c, DevCsr& a) {
    dev_free(c, a.rp);
    dev_free(c, a.ci);
    dev_free(c, a.v);
    a.rp = a.ci = nullptr;
    a.v = nullptr;
}

void free_agg(tf_ctx& c, DevAgg& t) {
    dev_free(c, t.map);
    dev_free(c, t.offsets);
    dev_free(c, t.members);
    t.map = t.offsets = t.members = nullptr;
}

int csr_check(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v) {
    c.clear_flags(1);
    hipLaunchKernelGGL(k_csr_check, grid1(n), dim3(kBlock), 0, c.stream, n, int(nnz), rp, ci, v, c.dflags);
    check(hipGetLastError(), "csr check launch");
    c.read_flags(1);
    return c.hflags[0];
}

bool csr_structure_ok(tf_ctx& c, int rows, int cols, int64_t nnz, const int* rp, const int* ci) {
    c.clear_flags(1);
    hipLaunchKernelGGL(k_csr_structure, grid1(rows > 0 ? rows : 1), dim3(kBlock), 0, c.stream, rows, cols, nnz, rp,
                       ci, c.dflags);
    check(hipGetLastError(), "csr structure launch");
    c.read_flags(1);
    return c.hflags[0] == 0;
}

DevCsr copy_csr(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v) {
    DevCsr out;
    out.n = n;
    out.nnz = nnz;
    out.rp = dev_alloc<int>(c, size_t(n) + 1);
    out.ci = dev_alloc<int>(c, size_t(nnz));
    out.v = dev_alloc<double>(c, size_t(nnz));
    check(hipMemcpyAsync(out.rp, rp, (size_t(n) + 1) * sizeof(int), hipMemcpyDefault, c.stream), "copy rp");
    if (nnz) {
        check(hipMemcpyAsync(out.ci, ci, size_t(nnz) * sizeof(int), hipMemcpyDefault, c.stream), "copy ci");
        check(hipMemcpyAsync(out.v, v, size_t(nnz) * sizeof(double), hipMemcpyDefault, c.stream), "copy v");
    }
    return out;
}

DevCsr canonicalize(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v) {
    require(n >= 1 && n < 0x7fffffff && nnz <= 0x3fffffff,
            "AMG requires nonempty square int32-capacity CSR on workspace device");
    const int state = csr_check(c, n, nnz, rp, ci, v);
    if (state & 2) throw Error(TF_E_INVALID, "AMG invalid CSR or nonfinite coefficient");
    if (state == 0) return copy_csr(c, n, nnz, rp, ci, v);
    return gather_merge(c, n, nullptr, nullptr, nullptr, rp, ci, v, false, false, nnz);
}

DevCsr merge_csr_rows(tf_ctx& c, int rows, int64_t nnz, const int* rp, const int* ci, const double* v) {
    require(rows >= 1 && nnz >= 1 && nnz <= 0x3fffffff, "row merge needs rows >= 1 and 1..2^30 entries");
    return gather_merge(c, rows, nullptr, nullptr, nullptr, rp, ci, v, false, false, nnz);
}

void validate_rowmax_async(tf_ctx& c, int n, const int* rp, const int* ci, const double* v, double* rowmax,
                           int* flag) {
    hipLaunchKernelGGL(k_validate, grid1(n), dim3(kBlock), 0, c.stream, n, rp, ci, v, flag, rowmax);
    check(hipGetLastError(), "validate launch");
}

template <class T>
void validate_inverse_async(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v,
                            double* rowmax, T* inv, int* flag, int* flagInv) {
    if (n <= 0) return;
    switch (lanes_per_row(n, nnz)) {
#define TF_VALIDATE(G)                                                                                           \
    case G:                                                                                                      \
        hipLaunchKernelGGL((k_validate_inv<T, G>), grid_rows_capped(n, G), dim3(kBlock), 0, c.stream, n, rp, ci, v, \
                           flag, rowmax, inv, flagInv);                                                          \
        break;
        TF_VALIDATE(4) TF_VALIDATE(8) TF_VALIDATE(16) TF_VALIDATE(32)
#undef TF_VALIDATE
    }
    check(hipGetLastError(), "validate launch");
}
template void validate_inverse_async<float>(tf_ctx&, int, int64_t, const int*, const int*, const double*, double*, float*,
                                            int*, int*);
template void validate_inverse_async<double>(tf_ctx&, int, int64_t, const int*, const int*, const double*, double*, double*,
                                             int*, int*);

bool validate_rowmax(tf_ctx& c, int n, const int* rp, const int* ci, const double* v, double* rowmax) {
    c.clear_flags(1);
    hipLaunchKernelGGL(k_validate, grid1(n), dim3(kBlock), 0, c.stream, n, rp, ci, v, c.dflags, rowmax);
    check(hipGetLastError(), "validate launch");
    c.read_flags(1);
    return c.hflags[0] == 0;
}

double* rowmax_workspace(tf_ctx& c, int n) { return c.wst<double>(kRowmax, size_t(n)); }

DevCsr strength_graph(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v, double theta,
                      int symmetric, bool prevalidated, int* symmetricOut, const double* learned, int deferred,
                      bool valueSymmetric) {
    if (valueSymmetric) symmetric = 1;
    require(std::isfinite(theta) && theta >= 0 && theta <= 1, "Invalid strength threshold");
    if (nnz > int64_t(0x3fffffff)) throw Error(TF_E_CAPACITY, "Native AMG: strength graph exceeds 32-bit capacity");
    double* rowmax = rowmax_workspace(c, n);
    if (!prevalidated && !validate_rowmax(c, n, rp, ci, v, rowmax))
        throw Error(TF_E_NUMERIC, "Native AMG: missing/zero diagonal, non-finite value or reciprocal in hierarchy");
    // An unknown pattern takes the symmetric path with a transpose check in
    // the same pass (flag slot `deferred`, read back with the graph's size);
    // a matrix that fails it takes the general path.
    const bool probe = symmetric < 0;
    if (probe && learned) {
        c.clear_flags(1);
        hipLaunchKernelGGL(k_struct_sym, grid1(n), dim3(kBlock), 0, c.stream, n, rp, ci, c.dflags);
        check(hipGetLastError(), "symmetry check launch");
        c.read_flags(1);
        symmetric = c.hflags[0] == 0 ? 1 : 0;
    }
    require(!learned || symmetric == 1, "learned affinities need a matrix with a symmetric pattern");
    const int slot = std::max(deferred, 0);
    if (symmetric != 0) {
        int* nonsym = nullptr;
        if (symmetric < 0) {
            nonsym = c.dflags + slot;
            check(hipMemsetAsync(nonsym, 0, sizeof(int), c.stream), "symmetry flag clear");
        }
        int* cnt = c.wst<int>(kCount, size_t(n) + 1);
        int* scol = c.wst<int>(kKeys, size_t(nnz) + 1);
        double* sw = c.wst<double>(kVals, size_t(nnz) + 1);
        const int lanes = lanes_per_row(n, nnz);
        if (learned) {
            hipLaunchKernelGGL(k_learned_graph, grid1(n + 1), dim3(kBlock), 0, c.stream, n, rp, ci, learned, scol, sw,
                               cnt);
        }
        else {
#define TF_STRENGTH(G)                                                                                          \
    case G:                                                                                                     \
        if (valueSymmetric)                                                                                     \
            hipLaunchKernelGGL((k_strength_sym<G, true>), grid_rows_capped(n, G), dim3(kBlock), 0, c.stream,     \
                               n, rp, ci, v, rowmax, theta, scol, sw, cnt, nonsym);                             \
        else                                                                                                    \
            hipLaunchKernelGGL((k_strength_sym<G, false>), grid_rows(int64_t(n) + 1, G), dim3(kBlock), 0,         \
                               c.stream, n, rp, ci, v, rowmax, theta, scol, sw, cnt, nonsym);                   \
        break;
            switch (lanes) {
                TF_STRENGTH(4) TF_STRENGTH(8) TF_STRENGTH(16) TF_STRENGTH(32)
            }
#undef TF_STRENGTH
        }
        check(hipGetLastError(), "strength launch");
        DevCsr g;
        g.n = n;
        g.rp = dev_alloc<int>(c, size_t(n) + 1);
        scan_excl(c, n + 1, cnt, g.rp);
        const int flags = nonsym ? slot + 1 : deferred;
        g.nnz = flags > 0 ? c.read_int_with_flags(g.rp + n, flags) : c.read_int(g.rp + n);
        if (nonsym) symmetric = c.hflags[slot] == 0 ? 1 : 0;
        if (symmetricOut) *symmetricOut = symmetric;
        if (!symmetric) {
            dev_free(c, g.rp);
            return strength_graph(c, n, nnz, rp, ci, v, theta, 0, true, nullptr, nullptr, deferred);
        }
        g.ci = dev_alloc<int>(c, size_t(g.nnz));
        g.v = dev_alloc<double>(c, size_t(g.nnz));
#define TF_COMPACT(G)                                                                                            \
    case G:                                                                                                      \
        hipLaunchKernelGGL(k_compact_rows<G>, grid_rows(n, G), dim3(kBlock), 0, c.stream, n, rp, scol, sw, g.rp, g.ci, \
                           g.v);                                                                                 \
        break;
        switch (lanes) {
            TF_COMPACT(4) TF_COMPACT(8) TF_COMPACT(16) TF_COMPACT(32)
        }
#undef TF_COMPACT
        check(hipGetLastError(), "strength compaction launch");
        return g;
    }
    if (symmetricOut) *symmetricOut = 0;
    int* cnt = c.wst<int>(kCount, size_t(n) + 1);
    fill_i32(c, n + 1, 0, cnt);
    hipLaunchKernelGGL(k_strength_count, grid1(n), dim3(kBlock), 0, c.stream, n, rp, ci, v, rowmax, theta, cnt);
    check(hipGetLastError(), "strength count launch");
    int* start = c.wst<int>(kStart, size_t(n) + 1);
    scan_excl(c, n + 1, cnt, start);
    const int total = deferred > 0 ? c.read_int_with_flags(start + n, deferred) : c.read_int(start + n);
    if (total == 0) {
        DevCsr g;
        g.n = n;
        g.nnz = 0;
        g.rp = dev_alloc<int>(c, size_t(n) + 1);
        fill_i32(c, n + 1, 0, g.rp);
        g.ci = dev_alloc<int>(c, 0);
        g.v = dev_alloc<double>(c, 0);
        return g;
    }
    int* fill = c.wst<int>(kFill, size_t(n));
    fill_i32(c, n, 0, fill);
    u64* keys = c.wst<u64>(kKeys, size_t(total));
    double* vals = c.wst<double>(kVals, size_t(total));
    hipLaunchKernelGGL(k_strength_fill, grid1(n), dim3(kBlock), 0, c.stream, n, rp, ci, v, rowmax, theta, start, fill,
                       keys, vals);
    check(hipGetLastError(), "strength fill launch");
    segsort(c, n, start, keys, vals);
    return merge_rows(c, n, start, keys, vals, true);
}

void match_round(tf_ctx& c, int n, const int* rp, const int* ci, const double* w, int* mate, int* proposal) {
    match_round_flags(c, n, n > 0 ? c.read_int(rp + n) : 0, rp, ci, w, mate, proposal, nullptr, nullptr, false);
}

int pair(tf_ctx& c, int n, const int* rp, const int* ci, const double* w, int rounds, int* map, int64_t nnz,
         const int* inSizes, const int* inSlots, int* sizes, int* slots) {
    int* mate = c.wst<int>(kMate, size_t(n));
    int* proposal = c.wst<int>(kProposal, size_t(n));
    if (n >= 1 && n <= kPairBlockRows) {
        // one launch instead of two per round
        int* count = c.wst<int>(kChanged, 8);
        hipLaunchKernelGGL(k_pair_block, dim3(1), dim3(kPairThreads), 0, c.stream, n, rp, ci, w, rounds, mate,
                           proposal, map, count, inSizes, inSlots, sizes, slots);
        check(hipGetLastError(), "small matching launch");
        return c.read_int(count);
    }
    if (nnz < 0) nnz = n > 0 ? c.read_int(rp + n) : 0;
    fill_i32(c, n, -1, mate);
    int* changed = c.wst<int>(kChanged, 8);
    fill_i32(c, 8, 0, changed);
    for (int r = 0; r < rounds; ++r)
        match_round_flags(c, n, nnz, rp, ci, w, mate, proposal, r ? changed + r - 1 : nullptr, changed + r);
    int* flags = c.wst<int>(kFlags, size_t(n) + 1);
    hipLaunchKernelGGL(k_roots, grid1(int64_t(n) + 1), dim3(kBlock), 0, c.stream, n, mate, flags);
    check(hipGetLastError(), "roots launch");
    int* prefix = c.wst<int>(kPrefix, size_t(n) + 1);
    scan_excl(c, n + 1, flags, prefix);
    hipLaunchKernelGGL(k_assign, grid1(int64_t(n) + 1), dim3(kBlock), 0, c.stream, n, mate, prefix, map, inSizes,
                       inSlots, sizes, slots);
    check(hipGetLastError(), "assign launch");
    return c.read_int(prefix + n);
}

namespace {
// The contracted graph from a member table (sizes: nc + 1 entries, the last
// zero; slots: 4 sorted members an aggregate).
DevCsr contract_members(tf_ctx& c, int n, const int* rp, const int* ci, const double* w, const int* map, int nc,
                        int64_t nnz, const int* sizes, const int* slots) {
    int* moff = c.wst<int>(kMemOffsets, size_t(nc) + 1);
    scan_excl(c, nc + 1, sizes, moff);
    int* mem = c.wst<int>(kMembers, size_t(n));
    hipLaunchKernelGGL(k_pack_members, grid1(nc), dim3(kBlock), 0, c.stream, nc, sizes, slots, moff, mem,
                       static_cast<int*>(nullptr));
    check(hipGetLastError(), "pack members launch");
    return gather_merge(c, nc, moff, mem, map, rp, ci, w, true, true, nnz);
}
}  // namespace

DevCsr contract_graph(tf_ctx& c, int n, const int* rp, const int* ci, const double* w, const int* map, int nc,
                      int64_t nnz) {
    int *sizes, *slots;
    build_members(c, n, nc, map, &sizes, &slots);
    return contract_members(c, n, rp, ci, w, map, nc, nnz, sizes, slots);
}

DevAgg aggregate(tf_ctx& c, const DevCsr& g, int rounds, bool absorb, int* hostHist) {
    require(rounds >= 1 && rounds <= 8 && g.n >= 1, "Invalid matching configuration");
    const int n = g.n;
    DevAgg t;
    t.n = n;
    t.map = dev_alloc<int>(c, size_t(n));
    try {
        // Member tables come from the pairings themselves (a root writes its
        // aggregate's sorted members), and absorption updates them in place.
        int* sizes1 = c.wst<int>(kSizes, size_t(n) + 1);
        int* slots1 = c.wst<int>(kSlots, size_t(n) * 4 + 1);
        const int first = pair(c, n, g.rp, g.ci, g.v, rounds, t.map, g.nnz, nullptr, nullptr, sizes1, slots1);
        DevCsr contracted = contract_members(c, n, g.rp, g.ci, g.v, t.map, first, g.nnz, sizes1, slots1);
        int* second = c.wst<int>(kSecond, size_t(first));
        int* sizes = c.wst<int>(kSizes2, size_t(first) + 1);
        int* slots = c.wst<int>(kSlots2, size_t(first) * 4 + 1);
        int coarse = 0;
        try {
            coarse = pair(c, first, contracted.rp, contracted.ci, contracted.v, rounds, second, contracted.nnz, sizes1,
                          slots1, sizes, slots);
        }
        catch (...) {
            free_csr(c, contracted);
            throw;
        }
        free_csr(c, contracted);
        hipLaunchKernelGGL(k_compose, grid1(n), dim3(kBlock), 0, c.stream, n, t.map, second);
        check(hipGetLastError(), "compose launch");
        int* hist = c.wst<int>(kHist, 4);
        if (absorb) {
            int* proposal = c.wst<int>(kProposal, size_t(n));
            double* scores = c.wst<double>(kPropWeight, size_t(n));
            int* next = c.wst<int>(kNextMap, size_t(n));
            int* cur = t.map;
            for (int round = 0; round < 2; ++round) {
                hipLaunchKernelGGL(k_propose_absorption, grid1(n), dim3(kBlock), 0, c.stream, n, g.rp, g.ci, g.v, cur,
                                   sizes, proposal, scores, next);
                check(hipGetLastError(), "propose absorption launch");
                hipLaunchKernelGGL(k_accept_absorption, grid1(coarse), dim3(kBlock), 0, c.stream, coarse, g.rp, g.ci,
                                   cur, sizes, slots, proposal, scores, next);
                check(hipGetLastError(), "accept absorption launch");
                std::swap(cur, next);
            }
            if (cur != t.map)
                check(hipMemcpyAsync(t.map, cur, size_t(n) * sizeof(int), hipMemcpyDeviceToDevice, c.stream),
                      "absorbed map copy");
            // compaction: occupied aggregates keep their order
            const int bound = coarse;
            int* flags = c.wst<int>(kFlags, size_t(bound) + 1);
            hipLaunchKernelGGL(k_occupied, grid1(int64_t(bound) + 1), dim3(kBlock), 0, c.stream, bound, sizes, flags);
            check(hipGetLastError(), "occupied launch");
            int* prefix = c.wst<int>(kPrefix, size_t(bound) + 1);
            scan_excl(c, bound + 1, flags, prefix);
            hipLaunchKernelGGL(k_compose, grid1(n), dim3(kBlock), 0, c.stream, n, t.map, prefix);
            check(hipGetLastError(), "compaction launch");
            int* oldscan = c.wst<int>(kOldScan, size_t(bound) + 1);
            scan_excl(c, bound + 1, sizes, oldscan);
            fill_i32(c, 4, 0, hist);
            coarse = c.read_int(prefix + bound);
            t.nc = coarse;
            t.offsets = dev_alloc<int>(c, size_t(coarse) + 1);
            t.members = dev_alloc<int>(c, size_t(n));
            hipLaunchKernelGGL(k_pack_compact, grid1(int64_t(bound) + 1), dim3(kBlock), 0, c.stream, bound, sizes,
                               slots, prefix, oldscan, t.offsets, t.members, hist);
            check(hipGetLastError(), "pack members launch");
        }
        else {
            t.nc = coarse;
            t.offsets = dev_alloc<int>(c, size_t(coarse) + 1);
            t.members = dev_alloc<int>(c, size_t(n));
            scan_excl(c, coarse + 1, sizes, t.offsets);
            fill_i32(c, 4, 0, hist);
            hipLaunchKernelGGL(k_pack_members, grid1(coarse), dim3(kBlock), 0, c.stream, coarse, sizes, slots,
                               t.offsets, t.members, hist);
            check(hipGetLastError(), "pack members launch");
        }
        if (hostHist) {
            check(hipMemcpyAsync(hostHist, hist, 4 * sizeof(int), hipMemcpyDeviceToHost, c.stream),
                  "histogram download");
        }
        else {
            check(hipMemcpyAsync(c.hflags, hist, 4 * sizeof(int), hipMemcpyDeviceToHost, c.stream),
                  "histogram download");
            c.sync();
            for (int k = 0; k < 4; ++k) t.hist[k] = c.hflags[k];
        }
    }
    catch (...) {
        free_agg(c, t);
        throw;
    }
    return t;
}

DevCsr galerkin(tf_ctx& c, const DevCsr& a, const int* map, int nc, const int* offsets, const int* members) {
    require(nc >= 1, "AMG transfer dimension mismatch");
    return gather_merge(c, nc, offsets, members, map, a.rp, a.ci, a.v, false, false, a.nnz);
}

}  // namespace tf
