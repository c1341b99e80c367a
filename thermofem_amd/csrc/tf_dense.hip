// The remaining level-1 API of the reference's linalg namespace on gfx950
// (include/LinearAlgebra/Operations.cuh:160-525 over Backend.cu): the
// 1-norm, dense transpose / matrix-vector / matrix-matrix products over
// row-major DenseMatrix storage, and sparse add / transpose / product.  The
// reference hands these to cuBLAS and cuSPARSE; here every sparse operation
// is an expansion into per-row candidate lists followed by the library's
// per-row merge (sorted columns, duplicates summed in candidate order), so
// results are deterministic.
#include <algorithm>
#include <vector>

#include "tf_device.hpp"
#include "tf_internal.hpp"

namespace tf {

namespace {

unsigned grid_for(int64_t count) {
    return unsigned(std::max<int64_t>(1, std::min<int64_t>((count + kBlock - 1) / kBlock, 16384)));
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_asum_partial(int n, const T* __restrict__ x, double* __restrict__ part) {
    double s = 0;
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock) s += fabs(double(x[i]));
    s = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void k_sum_final(const double* __restrict__ part, int parts, double* out) {
    const double s = sum_partials(part, parts);
    if (threadIdx.x == 0) *out = s;
}

// At[c][r] = A[r][c] through 32 x 33 LDS tiles (the extra column avoids
// bank conflicts on the transposed reads).
template <class T>
__global__ __launch_bounds__(kBlock) void k_transpose(int rows, int cols, const T* __restrict__ A, T* __restrict__ At) {
    __shared__ T tile[32][33];
    const int c0 = int(blockIdx.x) * 32, r0 = int(blockIdx.y) * 32;
    const int tx = int(threadIdx.x) % 32, ty = int(threadIdx.x) / 32;  // 32 x 8 threads
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = A[size_t(r) * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) At[size_t(c) * rows + r] = tile[tx][k];
    }
}

// y = A x: one wave per row, lanes stride over the columns, wave reduction.
template <class T>
__global__ __launch_bounds__(kBlock) void k_gemv(int rows, int cols, const T* __restrict__ A, const T* __restrict__ x,
                                                 T* __restrict__ y) {
    const int row = int(blockIdx.x) * (kBlock / kWave) + int(threadIdx.x) / kWave;
    const int lane = int(threadIdx.x) % kWave;
    if (row >= rows) return;
    T s = 0;
    for (int c = lane; c < cols; c += kWave) s += A[size_t(row) * cols + c] * x[c];
    s = wave_sum(s);
    if (lane == 0) y[row] = s;
}

// C = A B (m x k times k x n, row-major): 64 x 64 tiles of C per block,
// 16-deep slabs of A and B in LDS, 4 x 4 outputs per thread.
template <class T>
__global__ __launch_bounds__(kBlock) void k_gemm(int m, int n, int k, const T* __restrict__ A, const T* __restrict__ B,
                                                 T* __restrict__ C) {
    constexpr int TM = 64, TN = 64, TK = 16;
    __shared__ T sa[TK][TM + 1];
    __shared__ T sb[TK][TN + 1];
    const int m0 = int(blockIdx.y) * TM, n0 = int(blockIdx.x) * TN;
    const int tr = int(threadIdx.x) / 16, tc = int(threadIdx.x) % 16;  // 16 x 16 threads, 4 x 4 each
    T acc[4][4] = {};
    for (int k0 = 0; k0 < k; k0 += TK) {
        for (int q = int(threadIdx.x); q < TM * TK; q += kBlock) {
            const int i = q / TK, kk = q % TK;
            const int r = m0 + i, c = k0 + kk;
            sa[kk][i] = (r < m && c < k) ? A[size_t(r) * k + c] : T(0);
        }
        for (int q = int(threadIdx.x); q < TK * TN; q += kBlock) {
            const int kk = q / TN, j = q % TN;
            const int r = k0 + kk, c = n0 + j;
            sb[kk][j] = (r < k && c < n) ? B[size_t(r) * n + c] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
            T av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                av[u] = sa[kk][tr * 4 + u];
                bv[u] = sb[kk][tc * 4 + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int w = 0; w < 4; ++w) acc[u][w] += av[u] * bv[w];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const int r = m0 + tr * 4 + u, c = n0 + tc * 4 + w;
            if (r < m && c < n) C[size_t(r) * n + c] = acc[u][w];
        }
}

// ---- sparse: candidate counts and fills, then the per-row merge

__global__ void k_add_count(int rows, const int* __restrict__ ra, const int* __restrict__ rb, int* cnt) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i > rows) return;
    cnt[i] = i == rows ? 0 : (ra[i + 1] - ra[i]) + (rb[i + 1] - rb[i]);
}

// row i's candidates: A's entries, then beta times B's entries
__global__ void k_add_fill(int rows, const int* __restrict__ ra, const int* __restrict__ ca, const double* __restrict__ va,
                           const int* __restrict__ rb, const int* __restrict__ cb, const double* __restrict__ vb,
                           double beta, const int* __restrict__ start, int* __restrict__ oc, double* __restrict__ ov) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= rows) return;
    int at = start[i];
    for (int j = ra[i]; j < ra[i + 1]; ++j) {
        oc[at] = ca[j];
        ov[at++] = va[j];
    }
    for (int j = rb[i]; j < rb[i + 1]; ++j) {
        oc[at] = cb[j];
        ov[at++] = beta * vb[j];
    }
}

__global__ void k_col_count(int64_t nnz, const int* __restrict__ ci, int* cnt) {
    const int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (j < nnz) atomicAdd(cnt + ci[j], 1);
}

// transposed candidates: entry (r, c) goes to row c (slot order is arbitrary;
// the merge sorts each row by column, and a transposed row has no duplicates)
__global__ void k_transpose_fill(int rows, const int* __restrict__ rp, const int* __restrict__ ci,
                                 const double* __restrict__ v, int* __restrict__ slot, int* __restrict__ oc,
                                 double* __restrict__ ov) {
    const int r = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (r >= rows) return;
    for (int j = rp[r]; j < rp[r + 1]; ++j) {
        const int at = atomicAdd(slot + ci[j], 1);
        oc[at] = r;
        ov[at] = v[j];
    }
}

__global__ void k_gemm_count(int rows, const int* __restrict__ ra, const int* __restrict__ ca,
                             const int* __restrict__ rb, int64_t* cnt64, int* cnt) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i > rows) return;
    if (i == rows) {
        cnt[rows] = 0;
        return;
    }
    int64_t s = 0;
    for (int j = ra[i]; j < ra[i + 1]; ++j) s += rb[ca[j] + 1] - rb[ca[j]];
    cnt64[i] = s;
    cnt[i] = int(s < 0x3fffffff ? s : 0x3fffffff);
}

// row i's candidates: for each a_ik in stored order, a_ik * b_kj over row k of B
__global__ void k_gemm_fill(int rows, const int* __restrict__ ra, const int* __restrict__ ca,
                            const double* __restrict__ va, const int* __restrict__ rb, const int* __restrict__ cb,
                            const double* __restrict__ vb, const int* __restrict__ start, int* __restrict__ oc,
                            double* __restrict__ ov) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= rows) return;
    int at = start[i];
    for (int j = ra[i]; j < ra[i + 1]; ++j) {
        const int k = ca[j];
        const double a = va[j];
        for (int q = rb[k]; q < rb[k + 1]; ++q) {
            oc[at] = cb[q];
            ov[at++] = a * vb[q];
        }
    }
}

}  // namespace

template <class T>
double asum(tf_ctx& c, int n, const T* x) {
    if (n <= 0) return 0.0;
    const unsigned parts = std::min<unsigned>(grid_for(n), 1024);
    double* part = dev_alloc<double>(c, parts);
    hipLaunchKernelGGL(k_asum_partial<T>, dim3(parts), dim3(kBlock), 0, c.stream, n, x, part);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(1024), 0, c.stream, part, int(parts), c.dscalars);
    check(hipGetLastError(), "asum launch");
    check(hipMemcpyAsync(c.hscalars, c.dscalars, sizeof(double), hipMemcpyDeviceToHost, c.stream), "asum download");
    c.sync();
    dev_free(c, part);
    return c.hscalars[0];
}

template <class T>
void dense_transpose(tf_ctx& c, int rows, int cols, const T* A, T* At) {
    if (rows <= 0 || cols <= 0) return;
    hipLaunchKernelGGL(k_transpose<T>, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(kBlock), 0, c.stream, rows, cols,
                       A, At);
    check(hipGetLastError(), "transpose launch");
}

template <class T>
void gemv(tf_ctx& c, int rows, int cols, const T* A, const T* x, T* y) {
    if (rows <= 0) return;
    if (cols <= 0) {
        fill_zero(c, y, size_t(rows) * sizeof(T));
        return;
    }
    const int per = kBlock / kWave;
    hipLaunchKernelGGL(k_gemv<T>, dim3((rows + per - 1) / per), dim3(kBlock), 0, c.stream, rows, cols, A, x, y);
    check(hipGetLastError(), "gemv launch");
}

template <class T>
void gemm(tf_ctx& c, int m, int n, int k, const T* A, const T* B, T* C) {
    if (m <= 0 || n <= 0) return;
    if (k <= 0) {
        fill_zero(c, C, size_t(m) * size_t(n) * sizeof(T));
        return;
    }
    hipLaunchKernelGGL(k_gemm<T>, dim3((n + 63) / 64, (m + 63) / 64), dim3(kBlock), 0, c.stream, m, n, k, A, B, C);
    check(hipGetLastError(), "gemm launch");
}

#define TF_DENSE(T)                                                                 \
    template double asum<T>(tf_ctx&, int, const T*);                                \
    template void dense_transpose<T>(tf_ctx&, int, int, const T*, T*);              \
    template void gemv<T>(tf_ctx&, int, int, const T*, const T*, T*);               \
    template void gemm<T>(tf_ctx&, int, int, int, const T*, const T*, T*);
TF_DENSE(float)
TF_DENSE(double)
#undef TF_DENSE

// Merge of per-row candidate lists `rows` x (start, oc, ov): sorted columns,
// duplicates summed in candidate order.
static DevCsr merge_candidates(tf_ctx& c, int rows, int64_t total, const int* start, const int* oc, const double* ov) {
    require(total <= 0x3fffffff, "sparse result exceeds 32-bit capacity", TF_E_CAPACITY);
    if (total == 0) {  // no candidates: every row empty
        DevCsr out;
        out.n = rows;
        out.nnz = 0;
        out.rp = dev_alloc<int>(c, size_t(rows) + 1);
        fill_zero(c, out.rp, (size_t(rows) + 1) * sizeof(int));
        out.ci = dev_alloc<int>(c, 1);
        out.v = dev_alloc<double>(c, 1);
        return out;
    }
    return merge_csr_rows(c, rows, total, start, oc, ov);
}

DevCsr csr_add(tf_ctx& c, int rows, const int* ra, const int* ca, const double* va, const int* rb, const int* cb,
               const double* vb, double beta) {
    int* cnt = dev_alloc<int>(c, size_t(rows) + 1);
    int* start = dev_alloc<int>(c, size_t(rows) + 1);
    hipLaunchKernelGGL(k_add_count, dim3(grid_for(int64_t(rows) + 1)), dim3(kBlock), 0, c.stream, rows, ra, rb, cnt);
    scan_excl(c, rows + 1, cnt, start);
    const int64_t total = c.read_int(start + rows);
    int* oc = dev_alloc<int>(c, size_t(std::max<int64_t>(total, 1)));
    double* ov = dev_alloc<double>(c, size_t(std::max<int64_t>(total, 1)));
    hipLaunchKernelGGL(k_add_fill, dim3(grid_for(rows)), dim3(kBlock), 0, c.stream, rows, ra, ca, va, rb, cb, vb, beta,
                       start, oc, ov);
    check(hipGetLastError(), "sparse add launch");
    DevCsr out = merge_candidates(c, rows, total, start, oc, ov);
    c.sync();
    for (void* p : {static_cast<void*>(cnt), static_cast<void*>(start), static_cast<void*>(oc), static_cast<void*>(ov)})
        dev_free(c, p);
    return out;
}

DevCsr csr_transpose(tf_ctx& c, int rows, int cols, int64_t nnz, const int* rp, const int* ci, const double* v) {
    // duplicates merged first, so a transposed row holds each column once and
    // the arbitrary slot order below cannot change a sum
    DevCsr in;
    if (nnz > 0) {
        in = merge_csr_rows(c, rows, nnz, rp, ci, v);
        nnz = in.nnz;
        rp = in.rp;
        ci = in.ci;
        v = in.v;
    }
    int* cnt = dev_alloc<int>(c, size_t(cols) + 1);
    int* start = dev_alloc<int>(c, size_t(cols) + 1);
    int* slot = dev_alloc<int>(c, size_t(cols) + 1);
    fill_zero(c, cnt, (size_t(cols) + 1) * sizeof(int));
    if (nnz > 0)
        hipLaunchKernelGGL(k_col_count, dim3(grid_for(nnz)), dim3(kBlock), 0, c.stream, nnz, ci, cnt);
    scan_excl(c, cols + 1, cnt, start);
    check(hipMemcpyAsync(slot, start, (size_t(cols) + 1) * sizeof(int), hipMemcpyDeviceToDevice, c.stream),
          "slot copy");
    int* oc = dev_alloc<int>(c, size_t(std::max<int64_t>(nnz, 1)));
    double* ov = dev_alloc<double>(c, size_t(std::max<int64_t>(nnz, 1)));
    if (rows > 0)
        hipLaunchKernelGGL(k_transpose_fill, dim3(grid_for(rows)), dim3(kBlock), 0, c.stream, rows, rp, ci, v, slot, oc,
                           ov);
    check(hipGetLastError(), "sparse transpose launch");
    DevCsr out = merge_candidates(c, cols, nnz, start, oc, ov);
    c.sync();
    for (void* p : {static_cast<void*>(cnt), static_cast<void*>(start), static_cast<void*>(slot), static_cast<void*>(oc),
                    static_cast<void*>(ov)})
        dev_free(c, p);
    if (in.rp) free_csr(c, in);
    return out;
}

DevCsr csr_multiply(tf_ctx& c, int rows, const int* ra, const int* ca, const double* va, const int* rb, const int* cb,
                    const double* vb) {
    int* cnt = dev_alloc<int>(c, size_t(rows) + 1);
    int64_t* cnt64 = dev_alloc<int64_t>(c, size_t(rows) + 1);
    int* start = dev_alloc<int>(c, size_t(rows) + 1);
    hipLaunchKernelGGL(k_gemm_count, dim3(grid_for(int64_t(rows) + 1)), dim3(kBlock), 0, c.stream, rows, ra, ca, rb,
                       cnt64, cnt);
    // the expanded size in 64 bits (the int counts saturate)
    std::vector<int64_t> host(static_cast<size_t>(rows));
    if (rows > 0)
        check(hipMemcpyAsync(host.data(), cnt64, size_t(rows) * sizeof(int64_t), hipMemcpyDeviceToHost, c.stream),
              "product size download");
    c.sync();
    int64_t total = 0;
    for (int64_t q : host) total += q;
    require(total <= 0x3fffffff, "sparse product exceeds 32-bit capacity", TF_E_CAPACITY);
    scan_excl(c, rows + 1, cnt, start);
    int* oc = dev_alloc<int>(c, size_t(std::max<int64_t>(total, 1)));
    double* ov = dev_alloc<double>(c, size_t(std::max<int64_t>(total, 1)));
    hipLaunchKernelGGL(k_gemm_fill, dim3(grid_for(rows)), dim3(kBlock), 0, c.stream, rows, ra, ca, va, rb, cb, vb, start,
                       oc, ov);
    check(hipGetLastError(), "sparse product launch");
    DevCsr out = merge_candidates(c, rows, total, start, oc, ov);
    c.sync();
    for (void* p : {static_cast<void*>(cnt), static_cast<void*>(cnt64), static_cast<void*>(start), static_cast<void*>(oc),
                    static_cast<void*>(ov)})
        dev_free(c, p);
    return out;
}

}  // namespace tf
