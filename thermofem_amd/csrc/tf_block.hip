// Multi-right-hand-side (block) operations on gfx950: SpMM with the fused
// epilogues of the single-vector kernels, the AMG cycle over row-major
// blocks, and CG with one control block per right-hand side.  A block is
// row-major n x k: row i holds entry i of each of the k right-hand sides
// (the DenseMatrix layout, DenseMatrix.cuh:9-19; the SpMM of
// linalg::multiply(CsrMatrix, DenseMatrix), Operations.cuh:425-442 and
// Backend.cu:256-281).
//
// Launch shape: G lanes per row (G = k rounded up to a power of two, at most
// 64; lane = column inside a chunk of G columns, blockIdx.y = chunk) and
// kBlock / G rows per block step.  At G = 64 a row is one wave: the row's
// column indices and values are staged by one coalesced vector load (64
// entries per load, lane l holding entry l) and handed to every lane with
// readlane, and each gathered block row is one coalesced 64-lane load.
// Row-parallel launches are XCD-major: XCD x takes the x-th contiguous slab
// of rows and its blocks walk the slab together, so the block rows that
// neighbouring rows gather stay in that XCD's L2.
//
// Every column is computed with the arithmetic of the single-vector kernels
// (row sums of rounded products added in stored order, unfused epilogues,
// sequential restriction sums, the coarse inverse's column-run partials
// added in run order), so column j of a block cycle equals the single-vector cycle
// of column j bit for bit.  Krylov dot products are per column but reduced
// in a different order than the single-vector solver's.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "tf_device.hpp"
#include "tf_engine.hpp"
#include "tf_krylov.hpp"

namespace tf {

namespace {

// Most blocks of a row-parallel launch with a fused column reduction
// (= partials per column): 8 blocks of 4 waves per CU fill the chip.
constexpr int kBParts = 2048;

// Row groups of this block: XCD-major slabs, strided by the blocks per XCD.
__device__ __forceinline__ void block_groups(int groups, int& first, int& step, int& end) {
    const int per = int(gridDim.x) / kXcds;
    const int x = int(blockIdx.x) % kXcds, local = int(blockIdx.x) / kXcds;
    const int slab = (groups + kXcds - 1) / kXcds;
    const int begin = min(groups, x * slab);
    first = begin + local;
    step = per;
    end = min(groups, begin + slab);
}

// Lanes per row = columns per chunk, at most a wave.  Measured at k = 64 on
// a 15-point P1 tet matrix (274,625 rows, 4.0M nonzeros), FP64 / FP32:
// whole-wave rows with staged indices 178 / 88 us, with wave-uniform scalar
// index loads 180 / 133 us, two rows per wave (32-column chunks) with
// per-lane index loads 314 / 124 us, with scalar loads 355 / 262 us, with
// staged loads 245 / 158 us -- all bit-identical.
int lanes_for(int k) {
    int g = 1;
    while (g < k && g < kWave) g <<= 1;
    return g;
}

// Blocks of a row-parallel launch over `groups` row groups: one group per
// block without a reduction, at most kBParts blocks with one.
int row_grid(int groups, bool reduce = true) {
    int per = std::max(1, (groups + kXcds - 1) / kXcds);
    if (reduce) per = std::min(per, kBParts / kXcds);
    return per * kXcds;
}

// Row sum of row [s, e) against column col of X (k columns): products
// rounded and added one by one in stored order; chunks of 8 issue their
// loads first, slots past the row's end re-read its last entry and are not
// added.
template <class T>
__device__ __forceinline__ T brow_sum(int s, int e, const int* __restrict__ ci, const T* __restrict__ v,
                                      const T* __restrict__ X, int k, int col) {
#pragma clang fp contract(off)
    T acc = 0;
    for (int j0 = s; j0 < e; j0 += 8) {
        T pv[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = min(j0 + u, e - 1);
            pv[u] = v[j];
            xv[u] = X[size_t(ci[j]) * size_t(k) + size_t(col)];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const T p = pv[u] * xv[u];
            const T sum = acc + p;
            acc = j0 + u < e ? sum : acc;
        }
    }
    return acc;
}

__device__ __forceinline__ int lane_value(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ float lane_value(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double lane_value(double x, int l) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(b), l));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(b >> 32), l));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Row sum of a wave-uniform row [s, e) against column col (lane = column):
// the row's indices and values are staged 64 at a time by one coalesced
// load (lane l holds entry l of the window) and broadcast with readlane;
// products rounded and added one by one in stored order as in brow_sum.
// Slots past the window re-read its last entry and are not added.
template <class T>
__device__ __forceinline__ T brow_sum_staged(int s, int e, const int* __restrict__ ci, const T* __restrict__ v,
                                             const T* __restrict__ X, int k, int col) {
#pragma clang fp contract(off)
    const int lane = int(threadIdx.x) & (kWave - 1);
    T acc = 0;
    for (int b = s; b < e; b += kWave) {
        int cl = 0;
        T vl = 0;
        if (b + lane < e) {
            cl = ci[b + lane];
            vl = v[b + lane];
        }
        const int cnt = min(kWave, e - b);
        for (int u0 = 0; u0 < cnt; u0 += 8) {
            T pv[8], xv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int q = min(u0 + u, cnt - 1);
                pv[u] = lane_value(vl, q);
                xv[u] = X[size_t(lane_value(cl, q)) * size_t(k) + size_t(col)];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const T p = pv[u] * xv[u];
                const T sum = acc + p;
                acc = u0 + u < cnt ? sum : acc;
            }
        }
    }
    return acc;
}

// Reductions fused into a row-parallel launch (one partial per column and
// block, written to partials[col * kBParts + block]).
enum BDot {
    kBNone = 0,
    kBDotX = 1,    // sum X[r] * y[r]   (CG curvature p.Ap)
    kBSquares = 2, // sum y[r]^2         (residual norms)
    kBTrue = 3,    // residual + squares only in columns whose control block asks for a refresh
};

__device__ __forceinline__ void block_partials(double contrib, int G, int col, int k, double* red,
                                               double* __restrict__ partials) {
    red[threadIdx.x] = contrib;
    __syncthreads();
    if (int(threadIdx.x) < G && col < k) {
        double s = 0;
        for (int q = 0; q < kBlock / G; ++q) s += red[q * G + int(threadIdx.x)];
        partials[size_t(col) * kBParts + blockIdx.x] = s;
    }
}

// Y = epilogue(A X) over a block; E as the single-vector kernels
// (kSpmv, kResidual, kJacobi), D a fused column reduction.
template <class T, int G, int E, int D>
__global__ __launch_bounds__(kBlock) void k_bspmm(int n, int k, const int* __restrict__ rp,
                                                  const int* __restrict__ ci, const T* __restrict__ v,
                                                  const T* __restrict__ X, T* __restrict__ Y,
                                                  const T* __restrict__ B, const T* __restrict__ inv, T w,
                                                  double* __restrict__ partials, const KCtl* __restrict__ ctl) {
#pragma clang fp contract(off)
    __shared__ double red[kBlock];
    constexpr int R = kBlock / G;
    const int sub = int(threadIdx.x) / G, lane = int(threadIdx.x) % G;
    const int col = int(blockIdx.y) * G + lane;
    const bool live = col < k;
    const int cc = live ? col : k - 1;
    bool active = live;
    if constexpr (D == kBTrue) {
        // launched every iteration; only columns due for a refresh compute
        active = live && !stopped(ctl + cc) && ctl[cc].check && !ctl[cc].early;
        if (!__syncthreads_or(active ? 1 : 0)) return;  // (their partials are never read)
    }
    double contrib = 0;
    int first, step, end;
    block_groups((n + R - 1) / R, first, step, end);
    for (int g = first; g < end; g += step) {
        T acc;
        int row;
        if constexpr (G == kWave) {
            // the wave's row (wave index made uniform): staged index loads
            row = g * R + __builtin_amdgcn_readfirstlane(int(threadIdx.x) / kWave);
            if (row >= n) continue;
            acc = brow_sum_staged<T>(__builtin_amdgcn_readfirstlane(rp[row]),
                                     __builtin_amdgcn_readfirstlane(rp[row + 1]), ci, v, X, k, cc);
        }
        else {
            row = g * R + sub;
            if (row >= n) continue;
            acc = brow_sum<T>(rp[row], rp[row + 1], ci, v, X, k, cc);
        }
        const size_t at = size_t(row) * size_t(k) + size_t(cc);
        T y;
        if constexpr (E == kSpmv) {
            y = acc;
        }
        else if constexpr (E == kResidual) {
            y = B[at] - acc;
        }
        else {
            const T wd = w * inv[row];
            const T d = B[at] - acc;
            const T upd = wd * d;
            y = X[at] + upd;
        }
        if (active) {
            Y[at] = y;
            if constexpr (D == kBDotX) contrib += double(X[at]) * double(y);
            if constexpr (D == kBSquares || D == kBTrue) contrib += double(y) * double(y);
        }
    }
    if constexpr (D != kBNone) block_partials(contrib, G, col, k, red, partials);
}

// The same product without a fused reduction, each lane computing V =
// 16 / sizeof(T) adjacent columns from one 16-byte load per gathered block
// row (k a multiple of V, the blocks 16-byte aligned): GL lanes per row
// (GL * V columns per pass), 64 / GL rows per wave.  A row's column indices
// and values are staged GL at a time by its lanes and handed out with
// shuffles.  Every column is computed with k_bspmm's arithmetic (products
// rounded and added in stored order, the same epilogues), so the results are
// bit-identical; the wider loads halve the gathers' load instructions (P1
// tets on 64^3 cells, k = 64, FP64: 171 -> 131 us).  FP64 launches may fuse
// a column reduction D as k_bspmm does (partials[col * kBParts + block],
// each block's rows added in wave, then row-in-wave order).
template <class T, int GL, int E, int D = kBNone>
__global__ __launch_bounds__(kBlock) void k_bspmm_vec(int n, int k, const int* __restrict__ rp,
                                                      const int* __restrict__ ci, const T* __restrict__ v,
                                                      const T* __restrict__ X, T* __restrict__ Y,
                                                      const T* __restrict__ B, const T* __restrict__ inv, T w,
                                                      double* __restrict__ partials, const KCtl* __restrict__ ctl) {
#pragma clang fp contract(off)
    constexpr int V = 16 / int(sizeof(T));
    static_assert(D == kBNone || sizeof(T) == 8, "fused column reductions are FP64");
    typedef T Vec __attribute__((ext_vector_type(V)));
    constexpr int RW = kWave / GL;  // rows per wave
    constexpr int R = kBlock / GL;  // rows per block step
    __shared__ double red[D != kBNone ? kBlock * V : 1];
    const int lane = int(threadIdx.x) & (kWave - 1);
    const int sub = lane / GL, l = lane % GL;
    const int wave = int(threadIdx.x) / kWave;
    const int col = int(blockIdx.y) * GL * V + l * V;
    const bool live = col < k;
    const int cc = live ? col : 0;
    // columns this lane writes (kBTrue: only those due for a refresh)
    bool act[V];
#pragma unroll
    for (int t = 0; t < V; ++t) act[t] = live;
    if constexpr (D == kBTrue) {
        bool any = false;
#pragma unroll
        for (int t = 0; t < V; ++t) {
            const KCtl* c = ctl + cc + t;
            act[t] = live && !stopped(c) && c->check && !c->early;
            any = any || act[t];
        }
        if (!__syncthreads_or(any ? 1 : 0)) return;  // (their partials are never read)
    }
    double contrib[V];
#pragma unroll
    for (int t = 0; t < V; ++t) contrib[t] = 0;
    int first, step, end;
    block_groups((n + R - 1) / R, first, step, end);
    for (int g = first; g < end; g += step) {
        const int row = g * R + wave * RW + sub;
        const bool rv = row < n;
        const int rr = rv ? row : n - 1;
        const int s = rp[rr], cnt = rv ? rp[rr + 1] - s : 0;
        // wave-uniform trip counts: the longest row of the wave
        int m = cnt;
#pragma unroll
        for (int o = kWave / 2; o >= GL; o >>= 1) m = max(m, __shfl_xor(m, o));
        Vec acc;
#pragma unroll
        for (int t = 0; t < V; ++t) acc[t] = T(0);
        for (int b = 0; b < m; b += GL) {
            int cl = 0;
            T vl = T(0);
            if (b + l < cnt) {
                cl = ci[s + b + l];
                vl = v[s + b + l];
            }
            const int wcnt = min(cnt - b, GL);  // entries of this row in the window (may be <= 0)
            const int span = min(GL, m - b);
            for (int u0 = 0; u0 < span; u0 += 8) {
                T pv[8];
                Vec xv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int q = max(min(u0 + u, wcnt - 1), 0);
                    pv[u] = __shfl(vl, sub * GL + q);
                    const int c = __shfl(cl, sub * GL + q);
                    xv[u] = *reinterpret_cast<const Vec*>(X + size_t(c) * size_t(k) + size_t(cc));
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
#pragma unroll
                    for (int t = 0; t < V; ++t) {
                        const T p = pv[u] * xv[u][t];
                        const T sum = acc[t] + p;
                        acc[t] = u0 + u < wcnt ? sum : acc[t];
                    }
                }
            }
        }
        if (!rv || !live) continue;
        const size_t at = size_t(row) * size_t(k) + size_t(cc);
        Vec y;
        if constexpr (E == kSpmv) {
            y = acc;
        }
        else if constexpr (E == kResidual) {
            const Vec bv = *reinterpret_cast<const Vec*>(B + at);
#pragma unroll
            for (int t = 0; t < V; ++t) y[t] = bv[t] - acc[t];
        }
        else {
            const Vec bv = *reinterpret_cast<const Vec*>(B + at);
            const Vec xo = *reinterpret_cast<const Vec*>(X + at);
            const T wd = w * inv[row];
#pragma unroll
            for (int t = 0; t < V; ++t) {
                const T d = bv[t] - acc[t];
                const T upd = wd * d;
                y[t] = xo[t] + upd;
            }
        }
        if constexpr (D == kBTrue) {
#pragma unroll
            for (int t = 0; t < V; ++t)
                if (act[t]) {
                    Y[at + size_t(t)] = y[t];
                    contrib[t] += double(y[t]) * double(y[t]);
                }
        }
        else {
            *reinterpret_cast<Vec*>(Y + at) = y;
            if constexpr (D == kBDotX) {
                const Vec xo = *reinterpret_cast<const Vec*>(X + at);
#pragma unroll
                for (int t = 0; t < V; ++t) contrib[t] += double(xo[t]) * double(y[t]);
            }
            if constexpr (D == kBSquares) {
#pragma unroll
                for (int t = 0; t < V; ++t) contrib[t] += double(y[t]) * double(y[t]);
            }
        }
    }
    if constexpr (D != kBNone) {
#pragma unroll
        for (int t = 0; t < V; ++t) red[t * kBlock + int(threadIdx.x)] = contrib[t];
        __syncthreads();
        const int q = int(threadIdx.x);  // column of this pass
        const int qcol = int(blockIdx.y) * GL * V + q;
        if (q < GL * V && qcol < k) {
            const int ql = q / V, qt = q % V;
            double s = 0;
            for (int wv = 0; wv < kBlock / kWave; ++wv)
#pragma unroll
                for (int sb = 0; sb < RW; ++sb) s += red[qt * kBlock + wv * kWave + sb * GL + ql];
            partials[size_t(qcol) * kBParts + blockIdx.x] = s;
        }
    }
}

// Column dot products sum X[r][j] * Y[r][j] (Y widened exactly to FP64;
// zero: Y = 0).
template <int G, class YT>
__global__ __launch_bounds__(kBlock) void k_bdot(int n, int k, const double* __restrict__ X,
                                                 const YT* __restrict__ Y, int zero, double* __restrict__ partials) {
    __shared__ double red[kBlock];
    constexpr int R = kBlock / G;
    const int sub = int(threadIdx.x) / G, lane = int(threadIdx.x) % G;
    const int col = int(blockIdx.y) * G + lane;
    double s = 0;
    int first, step, end;
    block_groups((n + R - 1) / R, first, step, end);
    if (col < k)
        for (int g = first; g < end; g += step) {
            const int row = g * R + sub;
            if (row < n) {
                const size_t at = size_t(row) * size_t(k) + size_t(col);
                s += X[at] * (zero ? 0.0 : double(Y[at]));
            }
        }
    block_partials(s, G, col, k, red, partials);
}

// Finaliser, one block per column: the column's partials added by the
// block (sum_partials), then the single-vector recurrence on that column's
// control block (mode >= 0), or the plain sum written to out (mode < 0).
// (One 1024-thread block walking all columns took ~32 us at k = 64 with
// 2048 partials a column; the column blocks run side by side.)
__global__ __launch_bounds__(kBlock) void k_bfinal(const double* __restrict__ partials, int nparts, KCtl* ctl,
                                                   int mode, double* __restrict__ out) {
    const int col = int(blockIdx.x);
    if (mode >= 0) {
        const KCtl* c = ctl + col;
        if (stopped(c)) return;
        if (mode == kFinTrue && (!c->check || c->early)) return;
    }
    const double s = sum_partials(partials + size_t(col) * kBParts, nparts);
    if (threadIdx.x == 0) {
        if (mode >= 0)
            finalize(mode, s, ctl + col);
        else
            out[col] = s;
    }
}

// Element-wise block kernels walk rows without dividing flat indices:
// 2^lg lanes a row (lanes_for(k)), lane l taking the row's columns l,
// l + 2^lg, ...; rows grid-strided.  `each` hands out one element at a
// time; the streaming kernels below take kWalkRows rows a step instead, so
// that several rows' loads are in flight (flat loops with one element a
// thread ran the 64-column prolongation and CG direction at 2.8-3.6 TB/s).
constexpr int kWalkRows = 4;
struct RowWalk {
    int n, k, lg;
    template <class F>
    __device__ __forceinline__ void each(F&& f) const {
        const int G = 1 << lg, R = kBlock >> lg;
        const int sub = int(threadIdx.x) >> lg, l = int(threadIdx.x) & (G - 1);
        for (int row = int(blockIdx.x) * R + sub; row < n; row += int(gridDim.x) * R) {
            const size_t base = size_t(row) * size_t(k);
            for (int j = l; j < k; j += G) f(row, j, base + size_t(j));
        }
    }
    // lane's first column, step between its columns, first row, row stride
    __device__ __forceinline__ int lane() const { return int(threadIdx.x) & ((1 << lg) - 1); }
    __device__ __forceinline__ int step() const { return 1 << lg; }
    __device__ __forceinline__ int row0() const { return int(blockIdx.x) * (kBlock >> lg) + (int(threadIdx.x) >> lg); }
    __device__ __forceinline__ int stride() const { return int(gridDim.x) * (kBlock >> lg); }
};

// CG direction per column: p = z (fresh) or p = z + beta p.
// (z in FP64 or the cycle's level-0 iterate in level precision, widened
// exactly; zero: z = 0)
template <class Z>
__global__ void k_bcg_p(RowWalk rw, const Z* __restrict__ z, int zero, double* __restrict__ p,
                        const KCtl* __restrict__ ctl) {
    const int n = rw.n, k = rw.k, stride = rw.stride();
    for (int j = rw.lane(); j < k; j += rw.step()) {
        const KCtl* c = ctl + j;
        if (stopped(c)) continue;
        const bool fresh = c->fresh != 0;
        const double beta = c->beta;
        for (int r0 = rw.row0(); r0 < n; r0 += stride * kWalkRows) {
            size_t at[kWalkRows];
            double zi[kWalkRows], pi[kWalkRows];
#pragma unroll
            for (int u = 0; u < kWalkRows; ++u) {
                at[u] = size_t(min(r0 + u * stride, n - 1)) * size_t(k) + size_t(j);
                zi[u] = zero ? 0.0 : double(z[at[u]]);
                pi[u] = fresh ? 0.0 : p[at[u]];
            }
#pragma unroll
            for (int u = 0; u < kWalkRows; ++u)
                if (r0 + u * stride < n) p[at[u]] = fresh ? zi[u] : zi[u] + beta * pi[u];
        }
    }
}

// x += alpha p; r -= alpha q; column sums of r^2.
template <int G>
__global__ __launch_bounds__(kBlock) void k_bcg_update(int n, int k, const double* __restrict__ p,
                                                       const double* __restrict__ q, double* __restrict__ x,
                                                       double* __restrict__ r, double* __restrict__ partials,
                                                       const KCtl* __restrict__ ctl) {
    __shared__ double red[kBlock];
    constexpr int R = kBlock / G;
    const int sub = int(threadIdx.x) / G, lane = int(threadIdx.x) % G;
    const int col = int(blockIdx.y) * G + lane;
    double s = 0;
    int first, step, end;
    block_groups((n + R - 1) / R, first, step, end);
    if (col < k && !stopped(ctl + col)) {
        const double alpha = ctl[col].alpha;
        for (int g = first; g < end; g += step) {
            const int row = g * R + sub;
            if (row < n) {
                const size_t at = size_t(row) * size_t(k) + size_t(col);
                x[at] = alpha * p[at] + x[at];
                const double ri = -alpha * q[at] + r[at];
                r[at] = ri;
                s += ri * ri;
            }
        }
    }
    block_partials(s, G, col, k, red, partials);
}

// First Jacobi sweep from a zero iterate: x = (w * inv) * b.
template <class T>
__global__ void k_bjacobi_zero(RowWalk rw, const T* __restrict__ inv, T w, const T* __restrict__ b,
                               T* __restrict__ x) {
#pragma clang fp contract(off)
    rw.each([&](int row, int, size_t i) {
        const T wd = w * inv[row];
        x[i] = wd * b[i];
    });
}

// coarse[c][j] = sum over the members m of aggregate c (ascending) of f[m][j];
// sweep: also the coarse level's first Jacobi sweep from zero,
// x[c][j] = (w * inv[c]) * coarse[c][j] (k_bjacobi_zero's arithmetic).
template <class T>
__global__ void k_brestrict(RowWalk rw, const int* __restrict__ off, const int* __restrict__ mem,
                            const T* __restrict__ f, T* __restrict__ coarse, const T* __restrict__ inv, T w,
                            T* __restrict__ x, int sweep) {
#pragma clang fp contract(off)
    const int k = rw.k;
    rw.each([&](int c, int j, size_t i) {
        const int q0 = off[c], q1 = off[c + 1];
        T s = 0;
        if (q1 > q0) {
            // the first four members' loads issued together, added in order
            T fv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) fv[u] = f[size_t(mem[min(q0 + u, q1 - 1)]) * size_t(k) + size_t(j)];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const T t = s + fv[u];
                s = q0 + u < q1 ? t : s;
            }
            for (int q = q0 + 4; q < q1; ++q) s += f[size_t(mem[q]) * size_t(k) + size_t(j)];
        }
        coarse[i] = s;
        if (sweep) {
            const T wd = w * inv[c];
            x[i] = wd * s;
        }
    });
}

// Cycle entry: b = R in level precision; sweep: also the first Jacobi sweep
// from zero, x = (w * inv) * b (k_bjacobi_zero's arithmetic).
template <class T>
__global__ void k_benter(RowWalk rw, const double* __restrict__ R, const T* __restrict__ inv, T w,
                         T* __restrict__ b, T* __restrict__ x, int sweep) {
#pragma clang fp contract(off)
    const int n = rw.n, k = rw.k, stride = rw.stride();
    for (int j = rw.lane(); j < k; j += rw.step())
        for (int r0 = rw.row0(); r0 < n; r0 += stride * kWalkRows) {
            size_t at[kWalkRows];
            T bi[kWalkRows], iv[kWalkRows];
#pragma unroll
            for (int u = 0; u < kWalkRows; ++u) {
                const int row = min(r0 + u * stride, n - 1);
                at[u] = size_t(row) * size_t(k) + size_t(j);
                bi[u] = T(R[at[u]]);
                iv[u] = sweep ? inv[row] : T(0);
            }
#pragma unroll
            for (int u = 0; u < kWalkRows; ++u)
                if (r0 + u * stride < n) {
                    b[at[u]] = bi[u];
                    if (sweep) {
                        const T wd = w * iv[u];
                        x[at[u]] = wd * bi[u];
                    }
                }
        }
}

// fine += P coarse, or fine = P coarse when the fine iterate is known zero.
template <class T>
__global__ void k_bprolong(RowWalk rw, const int* __restrict__ map, const T* __restrict__ coarse,
                           T* __restrict__ fine, int set) {
    const int n = rw.n, k = rw.k, stride = rw.stride();
    for (int j = rw.lane(); j < k; j += rw.step())
        for (int r0 = rw.row0(); r0 < n; r0 += stride * kWalkRows) {
            size_t at[kWalkRows];
            int m[kWalkRows];
            T cv[kWalkRows], fv[kWalkRows];
#pragma unroll
            for (int u = 0; u < kWalkRows; ++u) {
                const int row = min(r0 + u * stride, n - 1);
                at[u] = size_t(row) * size_t(k) + size_t(j);
                m[u] = map[row];
                fv[u] = set ? T(0) : fine[at[u]];
            }
#pragma unroll
            for (int u = 0; u < kWalkRows; ++u) cv[u] = coarse[size_t(m[u]) * size_t(k) + size_t(j)];
#pragma unroll
            for (int u = 0; u < kWalkRows; ++u)
                if (r0 + u * stride < n) fine[at[u]] = set ? cv[u] : fv[u] + cv[u];
        }
}

template <class S, class D>
__global__ void k_bconvert(size_t count, const S* __restrict__ s, D* __restrict__ d) {
    for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < count; i += size_t(gridDim.x) * kBlock)
        d[i] = D(s[i]);
}

// x = A^-1 b on the dense coarsest level from the column-major inverse, one
// block per (row r, chunk of 64 columns), lane = column: wave w forms the
// row's partials over the column runs s = w, w + 4, ... (fused
// multiply-adds, columns ascending, loads issued 8 at a time) and wave 0
// adds the partials in run order -- the operations coarse_inverse_apply
// performs for one vector.  (One wave per (row, column) with a lane per run
// took 23 us at 140 rows x 64 columns.)
template <class T>
__global__ __launch_bounds__(kBlock) void k_bcoarse(int n, int k, const double* __restrict__ it,
                                                    const T* __restrict__ b, T* __restrict__ x) {
    __shared__ double part[16 * kWave];  // coarse_segments(n) <= 16
    const int lane = int(threadIdx.x) & (kWave - 1), w = int(threadIdx.x) / kWave;
    const int chunks = (k + kWave - 1) / kWave;
    const int r = int(blockIdx.x) / chunks, j = (int(blockIdx.x) % chunks) * kWave + lane;
    const bool live = j < k;
    const int jj = live ? j : k - 1;
    const int segs = coarse_segments(n), run = coarse_run(n);
    for (int s = w; s < segs; s += kBlock / kWave) {
        const int c0 = s * run, c1 = min(n, c0 + run);
        double p = 0;
        for (int cb = c0; cb < c1; cb += 8) {
            double a[8], bv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int c = min(cb + u, c1 - 1);
                a[u] = it[size_t(c) * size_t(n) + size_t(r)];
                bv[u] = double(b[size_t(c) * size_t(k) + size_t(jj)]);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (cb + u < c1 && a[u] != 0.0) p = fma(a[u], bv[u], p);
        }
        part[s * kWave + lane] = p;
    }
    __syncthreads();
    if (w == 0 && live) {
        double acc = part[lane];
        for (int s = 1; s < segs; ++s) acc += part[s * kWave + lane];
        x[size_t(r) * size_t(k) + size_t(j)] = T(acc);
    }
}

unsigned flat_blocks(size_t count) {
    return unsigned(std::max<size_t>(1, std::min<size_t>((count + kBlock - 1) / kBlock, 8192)));
}

// Row walk over n rows of k columns and its grid (at most 8192 blocks).
RowWalk row_walk(int n, int k) {
    int lg = 0;
    while ((1 << lg) < lanes_for(k)) ++lg;
    return RowWalk{n, k, lg};
}
unsigned walk_blocks(int n, int k) {
    const int R = kBlock / lanes_for(k);
    return unsigned(std::max(1, std::min((n + R - 1) / R, 8192)));
}

// G: lanes per row = columns per pass (0: lanes_for(k)).
template <class T, int E, int D>
void bspmm(tf_ctx& c, int n, int k, const int* rp, const int* ci, const T* v, const T* X, T* Y, const T* B,
           const T* inv, T w, double* partials, const KCtl* ctl, int G = 0) {
    if (n <= 0 || k <= 0) return;
    if (G <= 0) G = lanes_for(k);
    constexpr int V = 16 / int(sizeof(T));
    auto aligned = [](const void* p) { return p == nullptr || reinterpret_cast<uintptr_t>(p) % 16 == 0; };
    const bool aligned16 = aligned(X) && aligned(Y) && aligned(B);
    if constexpr (D == kBNone || sizeof(T) == 8) {
        // whole-width passes: 16-byte column loads (k_bspmm_vec).  A reduced
        // launch always takes this form when it applies, so that the number
        // of partials it leaves is spmm_parts's; the solver's blocks are
        // 16-byte aligned whenever k is a multiple of V.
        if (G == lanes_for(k) && k % V == 0 && (aligned16 || D != kBNone)) {
            require(aligned16, "block SpMM with a reduction needs 16-byte aligned blocks", TF_E_INVALID);
            const int GL = lanes_for(k / V);
            const int RV = kBlock / GL;
            const dim3 gv(row_grid((n + RV - 1) / RV, D != kBNone), (k + GL * V - 1) / (GL * V));
#define TF_BV(GG)                                                                                                \
    case GG:                                                                                                     \
        hipLaunchKernelGGL((k_bspmm_vec<T, GG, E, D>), gv, dim3(kBlock), 0, c.stream, n, k, rp, ci, v, X, Y, B, \
                           inv, w, partials, ctl);                                                               \
        break;
            switch (GL) { TF_BV(1) TF_BV(2) TF_BV(4) TF_BV(8) TF_BV(16) TF_BV(32) TF_BV(64) }
#undef TF_BV
            check(hipGetLastError(), "block spmm launch");
            return;
        }
    }
    const int R = kBlock / G;
    const dim3 grid(row_grid((n + R - 1) / R, D != kBNone), (k + G - 1) / G);
#define TF_BS(GG)                                                                                              \
    case GG:                                                                                                   \
        hipLaunchKernelGGL((k_bspmm<T, GG, E, D>), grid, dim3(kBlock), 0, c.stream, n, k, rp, ci, v, X, Y, B, \
                           inv, w, partials, ctl);                                                             \
        break;
    switch (G) { TF_BS(1) TF_BS(2) TF_BS(4) TF_BS(8) TF_BS(16) TF_BS(32) TF_BS(64) }
#undef TF_BS
    check(hipGetLastError(), "block spmm launch");
}

// Number of partials a row-parallel launch over n rows writes per column
// (k_bdot, k_bcg_update: one column a lane).
int row_parts(int n, int k) {
    const int R = kBlock / lanes_for(k);
    return row_grid((n + R - 1) / R);
}

// Partials per column of bspmm's reduced launch over n rows with G lanes a
// row: the 16-byte column form where bspmm takes it, else one column a lane.
template <class T>
int spmm_parts(int n, int k, int G) {
    if (G <= 0) G = lanes_for(k);
    constexpr int V = 16 / int(sizeof(T));
    const int R = G == lanes_for(k) && k % V == 0 ? kBlock / lanes_for(k / V) : kBlock / G;
    return row_grid((n + R - 1) / R);
}

template <class YT>
void bdot(tf_ctx& c, int n, int k, const double* X, const YT* Y, int zero, double* partials) {
    const int G = lanes_for(k);
    const int R = kBlock / G;
    const dim3 grid(row_grid((n + R - 1) / R), (k + G - 1) / G);
#define TF_BD(GG)                                                                                              \
    case GG:                                                                                                   \
        hipLaunchKernelGGL((k_bdot<GG, YT>), grid, dim3(kBlock), 0, c.stream, n, k, X, Y, zero, partials);     \
        break;
    switch (G) { TF_BD(1) TF_BD(2) TF_BD(4) TF_BD(8) TF_BD(16) TF_BD(32) TF_BD(64) }
#undef TF_BD
    check(hipGetLastError(), "block dot launch");
}

void bcg_update(tf_ctx& c, int n, int k, const double* p, const double* q, double* x, double* r, double* partials,
                const KCtl* ctl) {
    const int G = lanes_for(k);
    const int R = kBlock / G;
    const dim3 grid(row_grid((n + R - 1) / R), (k + G - 1) / G);
#define TF_BU(GG)                                                                                              \
    case GG:                                                                                                   \
        hipLaunchKernelGGL((k_bcg_update<GG>), grid, dim3(kBlock), 0, c.stream, n, k, p, q, x, r, partials,    \
                           ctl);                                                                               \
        break;
    switch (G) { TF_BU(1) TF_BU(2) TF_BU(4) TF_BU(8) TF_BU(16) TF_BU(32) TF_BU(64) }
#undef TF_BU
    check(hipGetLastError(), "block CG update launch");
}

void bfinal(tf_ctx& c, int k, const double* partials, int nparts, KCtl* ctl, int mode, double* out) {
    if (k <= 0) return;
    hipLaunchKernelGGL(k_bfinal, dim3(unsigned(k)), dim3(kBlock), 0, c.stream, partials, nparts, ctl, mode, out);
    check(hipGetLastError(), "block finaliser launch");
}

// Autotuned right-hand-side tile width: the columns one pass of a block
// SpMM covers (= lanes per row).  Narrower tiles read the matrix once per
// tile but keep a smaller window of gathered block rows in the XCD's L2;
// which wins depends on the operator and k, so each candidate (the widest,
// lanes_for(k), then halves down to 8 columns) is timed on the operator
// itself -- best of 3 runs of two launches each -- and the fastest kept, per
// operator shape, for the process.  Every width gives the same bits.
struct LaneKey {
    int op, n;
    int64_t nnz;
    int k, bytes;
    bool operator<(const LaneKey& o) const {
        return std::tie(op, n, nnz, k, bytes) < std::tie(o.op, o.n, o.nnz, o.k, o.bytes);
    }
};
std::map<LaneKey, int> g_lanes;
std::mutex g_lanes_mutex;

template <class F>
int tuned_lanes(tf_ctx& c, const LaneKey& key, F&& run) {
    const int widest = lanes_for(key.k);
    if (key.k < 16 || key.nnz < 100000) return widest;
    {
        std::lock_guard<std::mutex> lock(g_lanes_mutex);
        auto it = g_lanes.find(key);
        if (it != g_lanes.end()) return it->second;
    }
    hipEvent_t e[2];
    for (auto& ev : e) check(hipEventCreate(&ev), "hipEventCreate");
    int best = widest;
    float bestT = 1e30f;
    for (int G = widest; G >= 8; G /= 2) {
        run(G);
        float t = 1e30f;
        for (int i = 0; i < 3; ++i) {
            check(hipEventRecord(e[0], c.stream), "event record");
            run(G);
            run(G);
            check(hipEventRecord(e[1], c.stream), "event record");
            check(hipEventSynchronize(e[1]), "event sync");
            float ms = 0;
            check(hipEventElapsedTime(&ms, e[0], e[1]), "event time");
            t = std::min(t, ms);
        }
        if (t < bestT) {
            bestT = t;
            best = G;
        }
    }
    for (auto& ev : e) (void)hipEventDestroy(ev);
    std::lock_guard<std::mutex> lock(g_lanes_mutex);
    g_lanes[key] = best;
    return best;
}

}  // namespace

// ------------------------------------------------------------ block SpMM

template <class T>
void block_spmm(tf_ctx& c, int n, const int* rp, const int* ci, const T* v, int k, const T* X, T* Y) {
    bspmm<T, kSpmv, kBNone>(c, n, k, rp, ci, v, X, Y, nullptr, nullptr, T(0), nullptr, nullptr);
}
template <class T>
void block_jacobi(tf_ctx& c, int n, const int* rp, const int* ci, const T* v, const T* inv, T w, int k, const T* B,
                  const T* X, T* Y) {
    bspmm<T, kJacobi, kBNone>(c, n, k, rp, ci, v, X, Y, B, inv, w, nullptr, nullptr);
}
template void block_jacobi<double>(tf_ctx&, int, const int*, const int*, const double*, const double*, double, int,
                                   const double*, const double*, double*);
template void block_spmm<double>(tf_ctx&, int, const int*, const int*, const double*, int, const double*, double*);
template void block_spmm<float>(tf_ctx&, int, const int*, const int*, const float*, int, const float*, float*);

// ------------------------------------------------------------- block cycle

BlockAmg::BlockAmg(tf_ctx* ctx, AmgEngine* amg) : c_(ctx), amg_(amg) {}

BlockAmg::~BlockAmg() { release(); }

void BlockAmg::release() {
    if (mem_) {
        (void)hipStreamSynchronize(c_->stream);
        dev_free(*c_, mem_);
        mem_ = nullptr;
    }
    lv_.clear();
    shape_.clear();
    k_ = 0;
}

void BlockAmg::reserve(int k) {
    const size_t es = amg_->opt_.precision == 32 ? sizeof(float) : sizeof(double);
    std::vector<int> shape;
    for (auto& l : amg_->levels_) shape.push_back(l.native.n);
    shape.push_back(int(es));
    if (k == k_ && shape == shape_ && mem_) return;
    release();
    size_t total = 0;
    auto round = [](size_t b) { return (b + 255) / 256 * 256; };
    for (auto& l : amg_->levels_) total += 4 * round(size_t(l.native.n) * size_t(k) * es);
    mem_ = dev_alloc<char>(*c_, total);
    char* p = static_cast<char*>(mem_);
    lv_.resize(amg_->levels_.size());
    for (size_t d = 0; d < lv_.size(); ++d) {
        const size_t bytes = round(size_t(amg_->levels_[d].native.n) * size_t(k) * es);
        lv_[d].b = p;
        lv_[d].x = p + bytes;
        lv_[d].x2 = p + 2 * bytes;
        lv_[d].r = p + 3 * bytes;
        p += 4 * bytes;
    }
    k_ = k;
    shape_ = shape;
    visits.assign(lv_.size(), 0);
    if (amg_->opt_.precision == 32)
        tune_lanes<float>();
    else
        tune_lanes<double>();
}

// Tile width of every level's sweeps and residuals (timed on the level's
// Jacobi sweep from zero blocks).
template <class L>
void BlockAmg::tune_lanes() {
    lanes_.assign(lv_.size(), 0);
    const L w = L(amg_->opt_.omega);
    for (size_t d = 0; d < lv_.size(); ++d) {
        const AmgLevel& a = amg_->levels_[d];
        const DevCsr& A = a.native;
        const size_t bytes = size_t(A.n) * size_t(k_) * sizeof(L);
        fill_zero(*c_, lv_[d].b, bytes);
        fill_zero(*c_, lv_[d].x, bytes);
        lanes_[d] = tuned_lanes(*c_, LaneKey{1, A.n, A.nnz, k_, int(sizeof(L))}, [&](int G) {
            bspmm<L, kJacobi, kBNone>(*c_, A.n, k_, A.rp, A.ci, static_cast<const L*>(a.vals),
                                      static_cast<const L*>(lv_[d].x), static_cast<L*>(lv_[d].x2),
                                      static_cast<const L*>(lv_[d].b), static_cast<const L*>(a.inv), w, nullptr,
                                      nullptr, G);
        });
    }
}

template <class L>
void BlockAmg::smooth(size_t d, int sweeps) {
    const AmgLevel& a = amg_->levels_[d];
    Lvl& l = lv_[d];
    const DevCsr& A = a.native;
    const L w = L(amg_->opt_.omega);
    for (int s = 0; s < sweeps; ++s) {
        if (l.zero) {
            hipLaunchKernelGGL(k_bjacobi_zero<L>, dim3(walk_blocks(A.n, k_)), dim3(kBlock), 0, c_->stream,
                               row_walk(A.n, k_), static_cast<const L*>(a.inv), w, static_cast<const L*>(l.b),
                               static_cast<L*>(l.x));
            check(hipGetLastError(), "block Jacobi launch");
            l.zero = false;
        }
        else {
            bspmm<L, kJacobi, kBNone>(*c_, A.n, k_, A.rp, A.ci, static_cast<const L*>(a.vals),
                                      static_cast<const L*>(l.x), static_cast<L*>(l.x2), static_cast<const L*>(l.b),
                                      static_cast<const L*>(a.inv), w, nullptr, nullptr, lanes_[d]);
            std::swap(l.x, l.x2);
        }
    }
}

template <class L>
void BlockAmg::restrict_to(size_t d, bool ofResidual, bool sweepNext) {
    const AmgLevel& a = amg_->levels_[d];
    Lvl& l = lv_[d];
    const DevAgg& t = a.transfer;
    const L* f = static_cast<const L*>(l.b);
    if (ofResidual && !l.zero) {
        bspmm<L, kResidual, kBNone>(*c_, a.native.n, k_, a.native.rp, a.native.ci, static_cast<const L*>(a.vals),
                                    static_cast<const L*>(l.x), static_cast<L*>(l.r), static_cast<const L*>(l.b),
                                    nullptr, L(0), nullptr, nullptr, lanes_[d]);
        f = static_cast<const L*>(l.r);
    }
    Lvl& lc = lv_[d + 1];
    hipLaunchKernelGGL(k_brestrict<L>, dim3(walk_blocks(t.nc, k_)), dim3(kBlock), 0, c_->stream,
                       row_walk(t.nc, k_), t.offsets, t.members, f, static_cast<L*>(lc.b), static_cast<const L*>(amg_->levels_[d + 1].inv),
                       L(amg_->opt_.omega), static_cast<L*>(lc.x), int(sweepNext));
    check(hipGetLastError(), "block restrict launch");
    lc.swept = sweepNext;
}

template <class L>
void BlockAmg::prolong(size_t d) {
    const AmgLevel& a = amg_->levels_[d];
    Lvl& l = lv_[d];
    hipLaunchKernelGGL(k_bprolong<L>, dim3(walk_blocks(a.native.n, k_)), dim3(kBlock), 0, c_->stream,
                       row_walk(a.native.n, k_), a.transfer.map, static_cast<const L*>(lv_[d + 1].x), static_cast<L*>(l.x), int(l.zero));
    check(hipGetLastError(), "block prolong launch");
    l.zero = false;
}

template <class L>
void BlockAmg::coarse(size_t d) {
    const AmgLevel& a = amg_->levels_[d];
    Lvl& l = lv_[d];
    const double* inverse = amg_->coarse_inverse();
    if (inverse) {
        const int blocks = a.native.n * ((k_ + kWave - 1) / kWave);
        hipLaunchKernelGGL(k_bcoarse<L>, dim3(blocks), dim3(kBlock), 0, c_->stream, a.native.n, k_, inverse,
                           static_cast<const L*>(l.b), static_cast<L*>(l.x));
        check(hipGetLastError(), "block coarse solve launch");
        l.zero = false;
    }
    else {
        smooth<L>(d, amg_->opt_.coarseSweeps);
    }
}

// executeCycle (Cycle.cuh:35-63), the operations of AmgEngine::cycle.
template <class L>
void BlockAmg::cycle(int kind, size_t d, bool zero) {
    ++visits[d];
    Lvl& l = lv_[d];
    if (zero || kind == 2) l.zero = true;
    // a zero-start first sweep already done by the kernel that wrote b
    const int done = l.swept ? 1 : 0;
    if (l.swept) {
        l.zero = false;
        l.swept = false;
    }
    if (d + 1 == lv_.size()) {
        coarse<L>(d);
        return;
    }
    const AmgOptionsC& o = amg_->opt_;
    if (kind == 2) {
        restrict_to<L>(d, false, false);
        cycle<L>(2, d + 1, true);
        prolong<L>(d);
        cycle<L>(0, d, false);
        return;
    }
    smooth<L>(d, o.pre - done);
    // the coarser level starts with a zero-start sweep unless it is the
    // coarsest one
    restrict_to<L>(d, true, o.pre >= 1 && d + 2 < lv_.size());
    cycle<L>(kind, d + 1, true);
    if (kind == 1) cycle<L>(kind, d + 1, false);
    prolong<L>(d);
    smooth<L>(d, o.post);
}

// The cycle on R, its result left in lv_[0].x (or lv_[0].zero).  The entry
// kernel also does the finest level's first (zero-start) sweep when the
// cycle begins with one.
template <class L>
void BlockAmg::run_typed(const double* R) {
    const AmgOptionsC& o = amg_->opt_;
    const int n = amg_->levels_[0].native.n;
    for (auto& l : lv_) {
        l.zero = false;
        l.swept = false;
    }
    const bool sweep = o.cycle != 2 && o.pre >= 1 && lv_.size() > 1;
    hipLaunchKernelGGL((k_benter<L>), dim3(walk_blocks(n, k_)), dim3(kBlock), 0, c_->stream, row_walk(n, k_), R,
                       static_cast<const L*>(amg_->levels_[0].inv), L(o.omega), static_cast<L*>(lv_[0].b),
                       static_cast<L*>(lv_[0].x), int(sweep));
    check(hipGetLastError(), "block enter launch");
    lv_[0].swept = sweep;
    std::fill(visits.begin(), visits.end(), 0);
    cycle<L>(o.cycle, 0, true);
}

void BlockAmg::apply(int k, const double* R, double* Z) {
    int bits = 0;
    bool zero = false;
    const void* x = apply_in_place(k, R, bits, zero);
    const size_t nk = size_t(amg_->levels_[0].native.n) * size_t(k);
    if (zero)
        fill_zero(*c_, Z, nk * sizeof(double));
    else if (bits == 32)
        hipLaunchKernelGGL((k_bconvert<float, double>), dim3(flat_blocks(nk)), dim3(kBlock), 0, c_->stream, nk,
                           static_cast<const float*>(x), Z);
    else
        hipLaunchKernelGGL((k_bconvert<double, double>), dim3(flat_blocks(nk)), dim3(kBlock), 0, c_->stream, nk,
                           static_cast<const double*>(x), Z);
    check(hipGetLastError(), "block exit launch");
}

const void* BlockAmg::apply_in_place(int k, const double* R, int& bits, bool& zero) {
    require(amg_->ready(), "AMG requires setup", TF_E_INVALID);
    require(k >= 1, "block apply needs at least one column");
    reserve(k);
    bits = amg_->opt_.precision == 32 ? 32 : 64;
    if (bits == 32)
        run_typed<float>(R);
    else
        run_typed<double>(R);
    zero = lv_[0].zero;
    return lv_[0].x;
}

// ---------------------------------------------------------- block Krylov

void SolverEngine::reserve_block(int k) {
    if (k == bk_ && bvec_) return;
    release_block();
    const size_t nk = size_t(n_) * size_t(k);
    bvec_ = dev_alloc<double>(*c_, 5 * nk);  // B, X, R, P, Q
    bctl_ = dev_alloc<KCtl>(*c_, size_t(k));
    bparts_ = dev_alloc<double>(*c_, size_t(kBParts) * size_t(k));
    bsums_ = dev_alloc<double>(*c_, size_t(k));
    check(hipHostMalloc(&hbctl_, sizeof(KCtl) * size_t(k), hipHostMallocDefault), "hipHostMalloc block control");
    check(hipHostMalloc(&hbsums_, sizeof(double) * size_t(k), hipHostMallocDefault), "hipHostMalloc block sums");
    bk_ = k;
}

void SolverEngine::release_block() {
    if (!bvec_) return;
    (void)hipStreamSynchronize(c_->stream);
    dev_free(*c_, bvec_);
    dev_free(*c_, bctl_);
    dev_free(*c_, bparts_);
    dev_free(*c_, bsums_);
    (void)hipHostFree(hbctl_);
    (void)hipHostFree(hbsums_);
    bvec_ = nullptr;
    bctl_ = nullptr;
    bparts_ = nullptr;
    bsums_ = nullptr;
    hbctl_ = nullptr;
    hbsums_ = nullptr;
    bk_ = 0;
    bLanes_ = 0;
    bgraph_.reset();
}

// Column sums of a row-parallel reduction -> host.
static void column_sums(tf_ctx& c, int k, const double* partials, int nparts, double* dsums, double* hsums) {
    bfinal(c, k, partials, nparts, nullptr, -1, dsums);
    check(hipMemcpyAsync(hsums, dsums, sizeof(double) * size_t(k), hipMemcpyDeviceToHost, c.stream),
          "block sums download");
    c.sync();
}

// One block CG iteration (Cg.cuh:138-187 per column) as a fixed launch
// sequence: preconditioner, rho, direction, A p with p.Ap, the update with
// |r|^2, and the true-residual refresh, which only columns due for one
// compute.  Stopped columns are left as they are.
void SolverEngine::launch_block_iteration(int k) {
    const size_t nk = size_t(n_) * size_t(k);
    double* bB = bvec_;
    double* bX = bvec_ + nk;
    double* bR = bvec_ + 2 * nk;
    double* bP = bvec_ + 3 * nk;
    double* bQ = bvec_ + 4 * nk;
    KCtl* ctl = static_cast<KCtl*>(bctl_);
    const int np = row_parts(n_, k), npA = spmm_parts<double>(n_, k, bLanes_);
    // z = M^-1 r read where the cycle left it (level precision; z = r
    // without a preconditioner)
    const void* z = bR;
    int bits = 64;
    bool zzero = false;
    if (amg_) z = amg_->block().apply_in_place(k, bR, bits, zzero);
    const RowWalk rw = row_walk(n_, k);
    const unsigned wb = walk_blocks(n_, k);
    if (bits == 32) {
        bdot(*c_, n_, k, bR, static_cast<const float*>(z), int(zzero), bparts_);
        bfinal(*c_, k, bparts_, np, ctl, kFinRhoCg, nullptr);
        hipLaunchKernelGGL(k_bcg_p<float>, dim3(wb), dim3(kBlock), 0, c_->stream, rw, static_cast<const float*>(z),
                           int(zzero), bP, ctl);
    }
    else {
        bdot(*c_, n_, k, bR, static_cast<const double*>(z), int(zzero), bparts_);
        bfinal(*c_, k, bparts_, np, ctl, kFinRhoCg, nullptr);
        hipLaunchKernelGGL(k_bcg_p<double>, dim3(wb), dim3(kBlock), 0, c_->stream, rw,
                           static_cast<const double*>(z), int(zzero), bP, ctl);
    }
    bspmm<double, kSpmv, kBDotX>(*c_, n_, k, A_.rp, A_.ci, A_.v, bP, bQ, nullptr, nullptr, 0.0, bparts_, ctl,
                                 bLanes_);
    bfinal(*c_, k, bparts_, npA, ctl, kFinCurvCg, nullptr);
    bcg_update(*c_, n_, k, bP, bQ, bX, bR, bparts_, ctl);
    bfinal(*c_, k, bparts_, np, ctl, kFinUpdCg, nullptr);
    bspmm<double, kResidual, kBTrue>(*c_, n_, k, A_.rp, A_.ci, A_.v, bX, bR, bB, nullptr, 0.0, bparts_, ctl,
                                     bLanes_);
    bfinal(*c_, k, bparts_, npA, ctl, kFinTrue, nullptr);
    check(hipGetLastError(), "block iteration launch");
}

// CG (Cg.cuh:123-189) on k right-hand sides at once: every column follows
// the single-vector recurrence on its own control block and stops on its
// own; the preconditioner and the products run on the whole block.  The
// iteration is captured once as a hipGraph and launched in batches between
// reads of the k control blocks (krylov_batch).
void SolverEngine::apply_block(int k, const double* B, double* X, tf_solve_result* out) {
    require(n_ > 0, "CG requires setup", TF_E_INVALID);
    require(kind_ == 0, "the multi-right-hand-side solve runs CG", TF_E_INVALID);
    require(k >= 1, "block solve needs at least one right-hand side");
    reserve_block(k);
    if (amg_) amg_->block().reserve(k);
    const size_t nk = size_t(n_) * size_t(k);
    double* bB = bvec_;
    double* bX = bvec_ + nk;
    double* bR = bvec_ + 2 * nk;
    double* bP = bvec_ + 3 * nk;
    double* bQ = bvec_ + 4 * nk;
    KCtl* ctl = static_cast<KCtl*>(bctl_);
    KCtl* h = static_cast<KCtl*>(hbctl_);
    double* hs = static_cast<double*>(hbsums_);
    if (bLanes_ == 0) {
        fill_zero(*c_, bP, nk * sizeof(double));
        bLanes_ = tuned_lanes(*c_, LaneKey{0, n_, A_.nnz, k, 8}, [&](int G) {
            bspmm<double, kSpmv, kBNone>(*c_, n_, k, A_.rp, A_.ci, A_.v, bP, bQ, nullptr, nullptr, 0.0, nullptr,
                                         nullptr, G);
        });
    }
    const int np = row_parts(n_, k), npA = spmm_parts<double>(n_, k, bLanes_);
    const unsigned fb = flat_blocks(nk);
    hipLaunchKernelGGL((k_bconvert<double, double>), dim3(fb), dim3(kBlock), 0, c_->stream, nk, B, bB);
    hipLaunchKernelGGL((k_bconvert<double, double>), dim3(fb), dim3(kBlock), 0, c_->stream, nk, X, bX);
    check(hipGetLastError(), "block copy launch");

    auto residual_norms = [&](std::vector<double>& res) {
        bspmm<double, kResidual, kBSquares>(*c_, n_, k, A_.rp, A_.ci, A_.v, bX, bR, bB, nullptr, 0.0, bparts_,
                                            nullptr, bLanes_);
        column_sums(*c_, k, bparts_, npA, bsums_, hs);
        res.resize(size_t(k));
        for (int j = 0; j < k; ++j) res[size_t(j)] = std::sqrt(hs[j]);
    };
    bdot(*c_, n_, k, bB, static_cast<const double*>(bB), 0, bparts_);
    column_sums(*c_, k, bparts_, np, bsums_, hs);
    std::vector<double> normb(static_cast<size_t>(k)), tol(static_cast<size_t>(k)), res;
    for (int j = 0; j < k; ++j) {
        normb[size_t(j)] = std::sqrt(hs[j]);
        tol[size_t(j)] = std::max(o_.atol, o_.rtol * normb[size_t(j)]);
    }
    residual_norms(res);

    int running = 0;
    for (int j = 0; j < k; ++j) {
        KCtl& c = h[j];
        std::memset(&c, 0, sizeof(KCtl));
        c.normb = normb[size_t(j)];
        c.tol = tol[size_t(j)];
        c.res = res[size_t(j)];
        c.rho_prev = 0.0;
        c.alpha = 1.0;
        c.omega = 1.0;
        c.fresh = 1;
        c.refresh = o_.refresh;
        if (!std::isfinite(normb[size_t(j)]) || !std::isfinite(res[size_t(j)]))
            c.status = kNonFinite;
        else if (res[size_t(j)] <= tol[size_t(j)])
            c.status = kConverged;
        else {
            c.status = kRunning;
            ++running;
        }
    }
    check(hipMemcpyAsync(ctl, h, sizeof(KCtl) * size_t(k), hipMemcpyHostToDevice, c_->stream), "block control upload");
    const std::vector<double> res0 = res;
    // the graph names the block work buffers (made for k columns) and the
    // hierarchy's block buffers (made for its shape): both follow the
    // preconditioner's generation and k
    const uint64_t key = (amg_ ? amg_->generation() : 0) * 65536 + uint64_t(k);
    int done = 0;
    while (done < o_.maxit && running > 0) {
        const int batch = std::min(krylov_batch(h, res0.data(), k, done), o_.maxit - done);
        for (int q = 0; q < batch; ++q)
            launch_captured(*c_, bgraph_, o_.graph, key, [&] { launch_block_iteration(k); });
        done += batch;
        check(hipMemcpyAsync(h, ctl, sizeof(KCtl) * size_t(k), hipMemcpyDeviceToHost, c_->stream),
              "block control download");
        c_->sync();
        running = 0;
        for (int j = 0; j < k; ++j) running += h[j].status == kRunning;
    }
    // Final true residuals decide each column's status (SolverEngine::finish).
    residual_norms(res);
    for (int j = 0; j < k; ++j) {
        int st = h[j].status == kRunning ? kMaxIt : h[j].status;
        const int its = h[j].status == kRunning ? o_.maxit : h[j].iter;
        const double r = res[size_t(j)];
        if (!std::isfinite(r) || !std::isfinite(normb[size_t(j)]))
            st = kNonFinite;
        else if (r <= tol[size_t(j)])
            st = kConverged;
        else if (st == kConverged)
            st = kMaxIt;
        out[j].status = st;
        out[j].iterations = its;
        out[j].residual_norm = r;
        out[j].relative_residual = normb[size_t(j)] > 0 ? r / normb[size_t(j)] : r;
    }
    hipLaunchKernelGGL((k_bconvert<double, double>), dim3(fb), dim3(kBlock), 0, c_->stream, nk, bX, X);
    check(hipGetLastError(), "block copy launch");
    c_->sync();
}

}  // namespace tf
