// Device-wide exclusive scan (reduce / scan-of-totals / apply, 2048 items
// per 256-thread block) and the coarse-level dense solver: partial-pivoting
// LU in FP64 in one workgroup, explicit inverse by column substitution, and
// the per-cycle coarse solve as an FP64 dense product.
#include "tf_device.hpp"
#include "tf_internal.hpp"

namespace tf {
namespace {

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = int(threadIdx.x) & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of the per-thread totals of one block; returns the
// exclusive prefix of this thread and the block total through *total.
__device__ int block_excl_scan(int v, int* total) {
    __shared__ int waveTotals[kBlock / kWave];
    const int incl = wave_incl_scan(v);
    const int w = int(threadIdx.x) / kWave;
    if ((int(threadIdx.x) & (kWave - 1)) == kWave - 1) waveTotals[w] = incl;
    __syncthreads();
    int offset = 0, all = 0;
    for (int i = 0; i < kBlock / kWave; ++i) {
        if (i < w) offset += waveTotals[i];
        all += waveTotals[i];
    }
    __syncthreads();
    *total = all;
    return offset + incl - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(int n, const int* __restrict__ in, int* __restrict__ aux) {
    const int base = int(blockIdx.x) * kScanTile + int(threadIdx.x) * kScanItems;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += in[base + k];
    s = block_sum(s);
    if (threadIdx.x == 0) aux[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(int n, const int* in, int* out, const int* __restrict__ aux) {
    const int base = int(blockIdx.x) * kScanTile + int(threadIdx.x) * kScanItems;
    int items[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        items[k] = base + k < n ? in[base + k] : 0;
        s += items[k];
    }
    int total;
    int run = block_excl_scan(s, &total) + (aux ? aux[blockIdx.x] : 0);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += items[k];
    }
}

void scan_rec(tf_ctx& c, int n, const int* in, int* out, int* scratch) {
    const int nb = (n + kScanTile - 1) / kScanTile;
    if (nb <= 1) {
        hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(kBlock), 0, c.stream, n, in, out, nullptr);
        check(hipGetLastError(), "scan launch");
        return;
    }
    int* aux = scratch;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, c.stream, n, in, aux);
    check(hipGetLastError(), "scan reduce launch");
    scan_rec(c, nb, aux, aux, scratch + nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kBlock), 0, c.stream, n, in, out, aux);
    check(hipGetLastError(), "scan apply launch");
}

// ---------------- coarse dense LU ----------------

constexpr int kLuThreads = 1024;

__global__ void k_dense_fill(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                             const double* __restrict__ v, double* __restrict__ m) {
    const int r = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (r >= n) return;
    for (int j = 0; j < n; ++j) m[size_t(r) * n + j] = 0.0;
    for (int j = rp[r]; j < rp[r + 1]; ++j) m[size_t(r) * n + ci[j]] = v[j];
}

// Partial-pivoting LU of GpuAggregation.cu:827-893 for n <= 1024 in one
// workgroup: the pivot is the first maximal |m[i][k]| over i >= k, rejected
// when non-finite or <= 64*eps*max|A|; rows are swapped whole; multipliers
// are stored below the diagonal.
// Orders whose matrix fits kLuLds doubles of LDS work there (the steps are
// a chain of dependent loads; LDS makes each one short) with the same
// operations, then write back.
constexpr int kLuLds = 16384;  // 128 KB
__global__ __launch_bounds__(kLuThreads) void k_lu_factor(int n, double* gm, int* piv, int* bad) {
    __shared__ double sv[kLuThreads / kWave];
    __shared__ int si[kLuThreads / kWave];
    __shared__ double scale;
    __shared__ int pivot, failed;
    __shared__ double lds[kLuLds];
    const int tid = int(threadIdx.x);
    const int lane = tid & (kWave - 1), w = tid / kWave;
    const bool inLds = n * n <= kLuLds;
    double* m = inLds ? lds : gm;
    if (inLds) {
        for (int i = tid; i < n * n; i += kLuThreads) lds[i] = gm[i];
        __syncthreads();
    }
    double mx = 0;
    for (int i = tid; i < n * n; i += kLuThreads) mx = fmax(mx, fabs(m[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, kWave));
    if (lane == 0) sv[w] = mx;
    __syncthreads();
    if (tid == 0) {
        double s = 0;
        for (int i = 0; i < kLuThreads / kWave; ++i) s = fmax(s, sv[i]);
        scale = s;
        failed = 0;
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        // First maximal |m[i][k]| over i >= k (strict comparisons skip NaN,
        // as the sequential scan does); a NaN diagonal keeps p = k.
        double bv = -1.0;
        int bi = -1;
        for (int i = k + tid; i < n; i += kLuThreads) {
            const double a = fabs(m[size_t(i) * n + k]);
            if (a > bv) {
                bv = a;
                bi = i;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(bv, o, kWave);
            const int oi = __shfl_xor(bi, o, kWave);
            if (oi >= 0 && (bi < 0 || ov > bv || (ov == bv && oi < bi))) {
                bv = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            sv[w] = bv;
            si[w] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            double cv = -1.0;
            int p = -1;
            for (int q = 0; q < kLuThreads / kWave; ++q) {
                if (si[q] >= 0 && (p < 0 || sv[q] > cv || (sv[q] == cv && si[q] < p))) {
                    cv = sv[q];
                    p = si[q];
                }
            }
            if (p < 0 || isnan(m[size_t(k) * n + k])) p = k;
            const double pv = m[size_t(p) * n + k];
            if (!isfinite(pv) || fabs(pv) <= 64 * 2.2204460492503131e-16 * scale) failed = 1;
            pivot = p;
            piv[k] = p;
        }
        __syncthreads();
        if (failed) {
            if (tid == 0) *bad = 1;
            if (inLds)
                for (int i = tid; i < n * n; i += kLuThreads) gm[i] = lds[i];
            return;
        }
        const int p = pivot;
        if (p != k) {
            for (int j = tid; j < n; j += kLuThreads) {
                const double t = m[size_t(p) * n + j];
                m[size_t(p) * n + j] = m[size_t(k) * n + j];
                m[size_t(k) * n + j] = t;
            }
        }
        __syncthreads();
        const double d = m[size_t(k) * n + k];
        for (int r = k + 1 + tid; r < n; r += kLuThreads) m[size_t(r) * n + k] /= d;
        __syncthreads();
        const int t = n - k - 1;
        for (int idx = tid; idx < t * t; idx += kLuThreads) {
            const int r = k + 1 + idx / t, j = k + 1 + idx % t;
            m[size_t(r) * n + j] -= m[size_t(r) * n + k] * m[size_t(k) * n + j];
        }
        __syncthreads();
    }
    int nonfinite = 0;
    for (int i = tid; i < n * n; i += kLuThreads) {
        if (!isfinite(m[i])) nonfinite = 1;
        if (inLds) gm[i] = lds[i];
    }
    if (nonfinite) atomicOr(bad, 1);
}

// A^-1 from the pivoted LU by substitution on all columns at once (one
// workgroup): V = P I (the row swaps of the factorisation in order), then
// for each k the unit-lower and upper updates run in parallel over rows and
// columns.  Per column this is the operation order of the triangular solve.
__global__ __launch_bounds__(kLuThreads) void k_lu_inverse(int n, const double* __restrict__ glu,
                                                           const int* __restrict__ piv, double* ginv) {
    __shared__ double lds[kLuLds];
    const int tid = int(threadIdx.x);
    const int nn = n * n;
    // both factors and the inverse in LDS when they fit
    const bool inLds = 2 * nn <= kLuLds;
    double* inv = inLds ? lds : ginv;
    const double* lu = glu;
    if (inLds) {
        for (int i = tid; i < nn; i += kLuThreads) lds[nn + i] = glu[i];
        lu = lds + nn;
    }
    for (int idx = tid; idx < nn; idx += kLuThreads) inv[idx] = (idx / n == idx % n) ? 1.0 : 0.0;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const int p = piv[k];
        if (p != k) {
            for (int j = tid; j < n; j += kLuThreads) {
                const double t = inv[size_t(k) * n + j];
                inv[size_t(k) * n + j] = inv[size_t(p) * n + j];
                inv[size_t(p) * n + j] = t;
            }
        }
        __syncthreads();
    }
    for (int k = 0; k < n; ++k) {
        const int rows = n - k - 1;
        for (int idx = tid; idx < rows * n; idx += kLuThreads) {
            const int r = k + 1 + idx / n, j = idx % n;
            inv[size_t(r) * n + j] -= lu[size_t(r) * n + k] * inv[size_t(k) * n + j];
        }
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
        const double d = lu[size_t(k) * n + k];
        for (int j = tid; j < n; j += kLuThreads) inv[size_t(k) * n + j] /= d;
        __syncthreads();
        for (int idx = tid; idx < k * n; idx += kLuThreads) {
            const int r = idx / n, j = idx % n;
            inv[size_t(r) * n + j] -= lu[size_t(r) * n + k] * inv[size_t(k) * n + j];
        }
        __syncthreads();
    }
    if (inLds)
        for (int i = tid; i < nn; i += kLuThreads) ginv[i] = lds[i];
}

// x = inv * b in FP64; b is staged in LDS first so rhs and x may alias.
template <class T>
__global__ __launch_bounds__(kLuThreads) void k_inv_apply(int n, const double* __restrict__ inv, const T* b, T* x) {
    __shared__ double bs[kCoarseMax];
    coarse_inverse_apply<T>(n, inv, b, x, bs);
}

// Triangular solve of GpuAggregation.cu:904-939 (pivots, forward, backward)
// with the vector in LDS; rhs and x may alias.
template <class T>
__global__ __launch_bounds__(kLuThreads) void k_lu_solve(int n, const double* __restrict__ lu,
                                                         const int* __restrict__ piv, const T* b, T* x) {
    __shared__ double v[1024];
    const int row = int(threadIdx.x);
    if (row < n) v[row] = double(b[row]);
    __syncthreads();
    if (row == 0) {
        for (int k = 0; k < n; ++k) {
            const double t = v[k];
            v[k] = v[piv[k]];
            v[piv[k]] = t;
        }
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (row > k && row < n) v[row] -= lu[size_t(row) * n + k] * v[k];
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
        if (row == k) v[row] /= lu[size_t(row) * n + row];
        __syncthreads();
        if (row < k) v[row] -= lu[size_t(row) * n + k] * v[k];
        __syncthreads();
    }
    if (row < n) x[row] = T(v[row]);
}

}  // namespace

void scan_excl(tf_ctx& c, int n, const int* in, int* out) {
    if (n <= 0) return;
    size_t need = 0;
    for (int m = n; m > kScanTile;) {
        m = (m + kScanTile - 1) / kScanTile;
        need += size_t(m);
    }
    int* scratch = c.scan_scratch(need + 16);
    scan_rec(c, n, in, out, scratch);
}

bool coarse_lu_factor(tf_ctx& c, int n, const int* rp, const int* ci, const double* v, double* lu, int* piv,
                      double* inverse) {
    require(n >= 1 && n <= 1024, "GPU LU requires order 1..1024");
    c.clear_flags(1);
    hipLaunchKernelGGL(k_dense_fill, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c.stream, n, rp, ci, v, lu);
    check(hipGetLastError(), "dense fill launch");
    hipLaunchKernelGGL(k_lu_factor, dim3(1), dim3(kLuThreads), 0, c.stream, n, lu, piv, c.dflags);
    check(hipGetLastError(), "lu factor launch");
    c.read_flags(1);
    const bool ok = c.hflags[0] == 0;
    if (ok && inverse) {
        hipLaunchKernelGGL(k_lu_inverse, dim3(1), dim3(kLuThreads), 0, c.stream, n, lu, piv, inverse);
        check(hipGetLastError(), "lu inverse launch");
    }
    return ok;
}

template <class T>
void coarse_inv_apply(tf_ctx& c, int n, const double* inverse, const T* b, T* x) {
    require(n >= 1 && n <= 1024, "coarse solve order must be 1..1024");
    hipLaunchKernelGGL(k_inv_apply<T>, dim3(1), dim3(kLuThreads), 0, c.stream, n, inverse, b, x);
    check(hipGetLastError(), "coarse inverse apply launch");
}

template <class T>
void coarse_lu_solve(tf_ctx& c, int n, const double* lu, const int* piv, const T* b, T* x) {
    require(n >= 1 && n <= 1024, "coarse solve order must be 1..1024");
    hipLaunchKernelGGL(k_lu_solve<T>, dim3(1), dim3(kLuThreads), 0, c.stream, n, lu, piv, b, x);
    check(hipGetLastError(), "coarse lu solve launch");
}

template void coarse_inv_apply<double>(tf_ctx&, int, const double*, const double*, double*);
template void coarse_inv_apply<float>(tf_ctx&, int, const double*, const float*, float*);
template void coarse_lu_solve<double>(tf_ctx&, int, const double*, const int*, const double*, double*);
template void coarse_lu_solve<float>(tf_ctx&, int, const double*, const int*, const float*, float*);

}  // namespace tf
