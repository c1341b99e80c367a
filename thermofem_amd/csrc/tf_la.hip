// Linear-algebra kernels for gfx950: CSR SpMV over LDS-staged row blocks
// (or sub-wave row groups) with fused epilogues (residual, weighted Jacobi),
// level-1 BLAS with deterministic reductions, aggregate
// restriction/prolongation and element conversions.
#include "tf_device.hpp"
#include "tf_internal.hpp"

namespace tf {
namespace {

template <class T, int E>
__global__ __launch_bounds__(kBlock) void k_csr_stream(int n, int nblocks, int rpb, const int* __restrict__ rp,
                                                       const int* __restrict__ ci, const T* __restrict__ v,
                                                       const T* __restrict__ x, T* __restrict__ y,
                                                       const T* __restrict__ b, const T* __restrict__ inv, T w) {
    __shared__ T prod[StreamCap<T>::value];
    __shared__ int srp[kMaxStreamRows + 1];
    const int blk = xcd_block(nblocks);
    if (blk < 0) return;
    const int row0 = blk * rpb;
    const int rows = min(rpb, n - row0);
    T acc[kMaxStreamRows / kBlock];
    csr_stream_rows<T>(row0, rows, rp, ci, v, [&](int col) { return x[col]; }, prod, srp, acc);
#pragma unroll
    for (int k = 0; k < kMaxStreamRows / kBlock; ++k) {
        const int r = int(threadIdx.x) + k * kBlock;
        if (r < rows) epilogue<T, E>(row0 + r, acc[k], x, y, b, inv, w);
    }
}

template <int G, class T, int E>
__global__ __launch_bounds__(kBlock) void k_csr_rows(int n, int nblocks, const int* __restrict__ rp,
                                                     const int* __restrict__ ci, const T* __restrict__ v,
                                                     const T* __restrict__ x, T* __restrict__ y,
                                                     const T* __restrict__ b, const T* __restrict__ inv, T w) {
    const int blk = xcd_block(nblocks);
    if (blk < 0) return;
    constexpr int R = kBlock / G;
    const int row = blk * R + int(threadIdx.x) / G;
    const int lane = int(threadIdx.x) % G;
    if (row >= n) return;  // whole groups leave together
    const int s = rp[row], e = rp[row + 1];
    T acc = 0;
    for (int j = s + lane; j < e; j += G) acc += ld_entry(v + j) * x[ld_entry(ci + j)];
    acc = group_sum<G>(acc);
    if (lane == 0) epilogue<T, E>(row, acc, x, y, b, inv, w);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_jacobi_zero(int n, const T* __restrict__ inv, T w,
                                                        const T* __restrict__ b, T* __restrict__ xout) {
#pragma clang fp contract(off)
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i < n) {
        const T wd = w * inv[i];
        xout[i] = wd * b[i];
    }
}

int group_width(int n, int64_t nnz) {
    const double avg = n > 0 ? double(nnz) / double(n) : 1.0;
    int g = 2;
    while (g < 64 && g < avg) g <<= 1;
    return g;
}

template <class T, int E>
void launch_rows(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const T* v, const T* x, T* y,
                 const T* b, const T* inv, T w) {
    if (n <= 0) return;
    const double avg = double(nnz) / double(n);
    if (c.spmv_layout == 0 && avg <= 256) {
        const int rpb = stream_rows_per_block<T>(n, nnz);
        const int nb = (n + rpb - 1) / rpb;
        hipLaunchKernelGGL((k_csr_stream<T, E>), dim3(round_xcd(nb)), dim3(kBlock), 0, c.stream, n, nb, rpb, rp, ci,
                           v, x, y, b, inv, w);
        check(hipGetLastError(), "csr stream kernel launch");
        return;
    }
    const int g = group_width(n, nnz);
    const int rows = kBlock / g;
    const int nb = (n + rows - 1) / rows;
    const dim3 grid(round_xcd(nb));
#define TF_ROWS(G)                                                                                          \
    case G:                                                                                                 \
        hipLaunchKernelGGL((k_csr_rows<G, T, E>), grid, dim3(kBlock), 0, c.stream, n, nb, rp, ci, v, x, y, b, \
                           inv, w);                                                                         \
        break;
    switch (g) {
        TF_ROWS(2) TF_ROWS(4) TF_ROWS(8) TF_ROWS(16) TF_ROWS(32) TF_ROWS(64)
    }
#undef TF_ROWS
    check(hipGetLastError(), "csr row kernel launch");
}

template <class T>
__global__ void k_jacobi_update(int n, T w, const T* __restrict__ inv, const T* __restrict__ b,
                                const T* __restrict__ p, T* __restrict__ x) {
#pragma clang fp contract(off)
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i < n) {
        const T wd = w * inv[i];
        const T d = b[i] - p[i];
        const T upd = wd * d;
        x[i] = x[i] + upd;
    }
}

template <class T>
__global__ void k_axpy(int n, T alpha, const T* __restrict__ x, T* __restrict__ y) {
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock)
        y[i] = alpha * x[i] + y[i];
}

template <class T>
__global__ void k_scal(int n, T alpha, T* __restrict__ x) {
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock) x[i] = alpha * x[i];
}

// Fixed grid and grid-stride order make the partial sums, and thus the
// result, independent of timing.
template <class T>
__global__ __launch_bounds__(kBlock) void k_dot(int n, const T* __restrict__ x, const T* __restrict__ y,
                                                double* partials) {
    double s = 0;
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock)
        s += double(x[i]) * double(y[i]);
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void k_dot_final(const double* __restrict__ partials, int nparts, double* dst) {
    const double s = sum_partials(partials, nparts);
    if (threadIdx.x == 0) *dst = s;
}

template <class A, class B>
__global__ void k_convert(size_t n, const A* __restrict__ in, B* __restrict__ out) {
    for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock)
        out[i] = static_cast<B>(in[i]);
}

// coarse[c] = sum of fine[members] in ascending member order, starting from 0.
// With xc set, also the coarse level's first Jacobi sweep from a zero
// iterate: xc[c] = (w*inv[c]) * coarse[c].
template <class T>
__global__ void k_restrict(int nc, const int* __restrict__ off, const int* __restrict__ mem, const T* __restrict__ f,
                           T* __restrict__ cvec, const T* __restrict__ inv, T w, T* __restrict__ xc) {
#pragma clang fp contract(off)
    const int c = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (c >= nc) return;
    const int k0 = off[c], k1 = off[c + 1];
    T s = 0;
    if (k1 > k0) {
        // the first four members' loads and gathers are issued together
        // (slots past the aggregate re-read its last member and are not
        // added), then added in member order
        int mm[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) mm[u] = mem[min(k0 + u, k1 - 1)];
        T fv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) fv[u] = f[mm[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const T t = s + fv[u];
            s = k0 + u < k1 ? t : s;
        }
        for (int k = k0 + 4; k < k1; ++k) s += f[mem[k]];
    }
    cvec[c] = s;
    if (xc) {
        const T wd = w * inv[c];
        xc[c] = wd * s;
    }
}

// Residual and restriction in one pass, four lanes per aggregate (one
// member row each, aggregates have at most four; more are taken four at a
// time): coarse[c] = sum over members m (ascending) of (b[m] - sum_j a_mj
// x[j]), each row sum sequential in stored order and the members' residuals
// added in member order from 0 -- the same roundings as the separate
// residual and restriction kernels.  Optional first coarse sweep from zero
// as in k_restrict.
template <class T>
__global__ __launch_bounds__(kBlock) void k_residual_restrict(int nc, const int* __restrict__ off,
                                                              const int* __restrict__ mem,
                                                              const int* __restrict__ rp, const int* __restrict__ ci,
                                                              const T* __restrict__ v, const T* __restrict__ b,
                                                              const T* __restrict__ x, T* __restrict__ cvec,
                                                              const T* __restrict__ inv, T w, T* __restrict__ xc) {
#pragma clang fp contract(off)
    constexpr int G = 4;
    const int c = int((int64_t(blockIdx.x) * kBlock + threadIdx.x) / G);
    const int q = int(threadIdx.x) % G;
    const bool live = c < nc;
    const int cc = live ? c : nc - 1;  // (every lane of a group takes part in the shuffles)
    const int k0 = off[cc], k1 = off[cc + 1];
    T s = 0;
    for (int kb = k0; kb < k1; kb += G) {
        const int k = kb + q;
        T r = 0;
        if (k < k1) {
            const int m = mem[k];
            const int e = rp[m + 1];
            T acc = 0;
            for (int j0 = rp[m]; j0 < e; j0 += 8) {
                T pv[8], xv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int j = min(j0 + u, e - 1);
                    pv[u] = v[j];
                    xv[u] = x[ci[j]];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const T p = pv[u] * xv[u];
                    const T t = acc + p;
                    acc = j0 + u < e ? t : acc;
                }
            }
            r = b[m] - acc;
        }
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const T ru = __shfl(r, u, G);
            const T t = s + ru;
            s = kb + u < k1 ? t : s;
        }
    }
    if (live && q == 0) {
        cvec[c] = s;
        if (xc) {
            const T wd = w * inv[c];
            xc[c] = wd * s;
        }
    }
}

// Prolongation fused into the first post-smoothing sweep: with
// x'_j = x_j + xc[map_j] (or xc[map_j] when x is zero), computes
// xout = x' + (w*inv)*(b - A x'), never storing x'.
template <class T>
__global__ __launch_bounds__(kBlock) void k_prolong_jacobi(int n, int nblocks, int rpb, const int* __restrict__ rp,
                                                           const int* __restrict__ ci, const T* __restrict__ v,
                                                           const int* __restrict__ map, const T* __restrict__ xcv,
                                                           const T* __restrict__ x, int zero, T* __restrict__ y,
                                                           const T* __restrict__ b, const T* __restrict__ inv, T w) {
#pragma clang fp contract(off)
    __shared__ T prod[StreamCap<T>::value];
    __shared__ int srp[kMaxStreamRows + 1];
    const int blk = xcd_block(nblocks);
    if (blk < 0) return;
    const int row0 = blk * rpb;
    const int rows = min(rpb, n - row0);
    auto xp = [&](int col) -> T { return zero ? xcv[map[col]] : x[col] + xcv[map[col]]; };
    T acc[kMaxStreamRows / kBlock];
    csr_stream_rows<T>(row0, rows, rp, ci, v, xp, prod, srp, acc);
#pragma unroll
    for (int k = 0; k < kMaxStreamRows / kBlock; ++k) {
        const int r = int(threadIdx.x) + k * kBlock;
        if (r < rows) {
            const int row = row0 + r;
            const T xr = xp(row);
            const T wd = w * inv[row];
            const T d = b[row] - acc[k];
            const T upd = wd * d;
            y[row] = xr + upd;
        }
    }
}

// Level entry: b = (T) rhs and, optionally, the first sweep from zero x = (w*inv)*b.
template <class T>
__global__ void k_enter(int n, const double* __restrict__ rhs, T* __restrict__ b, const T* __restrict__ inv, T w,
                        T* __restrict__ x) {
#pragma clang fp contract(off)
#pragma unroll 4
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock) {
        const T bi = T(rhs[i]);
        b[i] = bi;
        if (x) {
            const T wd = w * inv[i];
            x[i] = wd * bi;
        }
    }
}

// fine += P coarse, or fine = P coarse when the fine iterate is known zero.
template <class T>
__global__ void k_prolong(int n, const int* __restrict__ map, const T* __restrict__ cvec, T* __restrict__ f, int set) {
#pragma unroll 4
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock)
        f[i] = set ? cvec[map[i]] : f[i] + cvec[map[i]];
}

// Inverse diagonal: duplicate diagonal entries summed in stored order.
template <class T>
__global__ void k_inverse_diagonal(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                   const T* __restrict__ v, T* __restrict__ inv, int* bad) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= n) return;
    T d = 0;
    for (int j = rp[i]; j < rp[i + 1]; ++j)
        if (ci[j] == i) d += v[j];
    const T r = T(1) / d;
    if (!isfinite(d) || d == T(0) || !isfinite(r)) atomicOr(bad, 1);
    inv[i] = r;
}

__global__ void k_fill_zero(size_t words, uint32_t* __restrict__ p) {
    for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < words; i += size_t(gridDim.x) * kBlock) p[i] = 0u;
}
__global__ void k_fill_zero_bytes(size_t bytes, uint8_t* __restrict__ p) {
    for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < bytes; i += size_t(gridDim.x) * kBlock) p[i] = 0;
}

int blocks_for(size_t n, int cap = 4096) {
    size_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    return int(b > size_t(cap) ? size_t(cap) : b);
}

}  // namespace

template <class T>
void spmv(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const T* v, const T* x, T* y) {
    launch_rows<T, kSpmv>(c, n, nnz, rp, ci, v, x, y, nullptr, nullptr, T(0));
}
template <class T>
void residual(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const T* v, const T* b, const T* x, T* r) {
    launch_rows<T, kResidual>(c, n, nnz, rp, ci, v, x, r, b, nullptr, T(0));
}
template <class T>
void jacobi_fused(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const T* v, const T* inv, T w,
                  const T* b, const T* xin, T* xout) {
    if (n <= 0) return;
    if (!xin) {
        hipLaunchKernelGGL(k_jacobi_zero<T>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c.stream, n, inv, w, b,
                           xout);
        check(hipGetLastError(), "jacobi zero launch");
        return;
    }
    launch_rows<T, kJacobi>(c, n, nnz, rp, ci, v, xin, xout, b, inv, w);
}
template <class T>
void jacobi_update(tf_ctx& c, int n, T w, const T* inv, const T* b, const T* product, T* x) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_jacobi_update<T>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c.stream, n, w, inv, b,
                       product, x);
    check(hipGetLastError(), "jacobi update launch");
}
template <class T>
void axpy(tf_ctx& c, int n, T alpha, const T* x, T* y) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_axpy<T>, dim3(blocks_for(size_t(n))), dim3(kBlock), 0, c.stream, n, alpha, x, y);
    check(hipGetLastError(), "axpy launch");
}
template <class T>
void scal(tf_ctx& c, int n, T alpha, T* x) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scal<T>, dim3(blocks_for(size_t(n))), dim3(kBlock), 0, c.stream, n, alpha, x);
    check(hipGetLastError(), "scal launch");
}
template <class T>
void dot_device(tf_ctx& c, int n, const T* x, const T* y, double* dst, int slot) {
    require(slot >= 0 && slot < kReduceSlots, "reduction slot out of range");
    const int blocks = blocks_for(size_t(n > 0 ? n : 1), 8192);
    double* part = c.partials + size_t(slot) * kMaxReduceBlocks;
    hipLaunchKernelGGL(k_dot<T>, dim3(blocks), dim3(kBlock), 0, c.stream, n, x, y, part);
    check(hipGetLastError(), "dot launch");
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(1024), 0, c.stream, part, blocks, dst);
    check(hipGetLastError(), "dot final launch");
}
template <class A, class B>
void convert(tf_ctx& c, size_t n, const A* in, B* out) {
    if (!n) return;
    hipLaunchKernelGGL((k_convert<A, B>), dim3(blocks_for(n)), dim3(kBlock), 0, c.stream, n, in, out);
    check(hipGetLastError(), "convert launch");
}
template <class T>
void restrict_agg(tf_ctx& c, int nc, const int* offsets, const int* members, const T* fine, T* coarse, const T* inv,
                  T w, T* xc) {
    if (nc <= 0) return;
    hipLaunchKernelGGL(k_restrict<T>, dim3((nc + kBlock - 1) / kBlock), dim3(kBlock), 0, c.stream, nc, offsets,
                       members, fine, coarse, inv, w, xc);
    check(hipGetLastError(), "restrict launch");
}
template <class T>
void residual_restrict(tf_ctx& c, int nc, const int* offsets, const int* members, const int* rp, const int* ci,
                       const T* v, const T* b, const T* x, T* coarse, const T* inv, T w, T* xc) {
    if (nc <= 0) return;
    hipLaunchKernelGGL(k_residual_restrict<T>, dim3(unsigned((int64_t(nc) * 4 + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, c.stream, nc, offsets, members, rp, ci, v, b, x, coarse, inv, w, xc);
    check(hipGetLastError(), "residual restrict launch");
}
template <class T>
void prolong_jacobi(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const T* v, const int* map,
                    const T* xc, const T* x, bool zero, T* xout, const T* b, const T* inv, T w) {
    if (n <= 0) return;
    const int rpb = stream_rows_per_block<T>(n, nnz);
    const int nb = (n + rpb - 1) / rpb;
    hipLaunchKernelGGL(k_prolong_jacobi<T>, dim3(round_xcd(nb)), dim3(kBlock), 0, c.stream, n, nb, rpb, rp, ci, v, map,
                       xc, x, int(zero), xout, b, inv, w);
    check(hipGetLastError(), "prolong jacobi launch");
}
template <class T>
void enter_level(tf_ctx& c, int n, const double* rhs, T* b, const T* inv, T w, T* x) {
    if (n <= 0) return;
    // at most 8 resident blocks a CU, each thread walking several elements
    hipLaunchKernelGGL(k_enter<T>, dim3(blocks_for(size_t(n), 2048)), dim3(kBlock), 0, c.stream, n, rhs, b, inv, w,
                       x);
    check(hipGetLastError(), "level entry launch");
}
template <class T>
void prolong_agg(tf_ctx& c, int n, const int* map, const T* coarse, T* fine, bool set) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_prolong<T>, dim3(blocks_for(size_t(n), 2048)), dim3(kBlock), 0, c.stream, n, map, coarse,
                       fine, int(set));
    check(hipGetLastError(), "prolong launch");
}
template <class T>
void inverse_diagonal(tf_ctx& c, int n, const int* rp, const int* ci, const T* v, T* inv, int* dbad) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_inverse_diagonal<T>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c.stream, n, rp, ci,
                       v, inv, dbad);
    check(hipGetLastError(), "inverse diagonal launch");
}
void fill_zero(tf_ctx& c, void* dst, size_t bytes) {
    if (!bytes) return;
    if ((reinterpret_cast<uintptr_t>(dst) % 4) == 0 && bytes % 4 == 0) {
        hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(bytes / 4)), dim3(kBlock), 0, c.stream, bytes / 4,
                           static_cast<uint32_t*>(dst));
    }
    else {
        hipLaunchKernelGGL(k_fill_zero_bytes, dim3(blocks_for(bytes)), dim3(kBlock), 0, c.stream, bytes,
                           static_cast<uint8_t*>(dst));
    }
    check(hipGetLastError(), "fill launch");
}

#define TF_LA(T)                                                                                                   \
    template void spmv<T>(tf_ctx&, int, int64_t, const int*, const int*, const T*, const T*, T*);                \
    template void residual<T>(tf_ctx&, int, int64_t, const int*, const int*, const T*, const T*, const T*, T*);  \
    template void jacobi_fused<T>(tf_ctx&, int, int64_t, const int*, const int*, const T*, const T*, T, const T*, \
                                  const T*, T*);                                                                 \
    template void jacobi_update<T>(tf_ctx&, int, T, const T*, const T*, const T*, T*);                           \
    template void axpy<T>(tf_ctx&, int, T, const T*, T*);                                                        \
    template void scal<T>(tf_ctx&, int, T, T*);                                                                  \
    template void dot_device<T>(tf_ctx&, int, const T*, const T*, double*, int);                                 \
    template void restrict_agg<T>(tf_ctx&, int, const int*, const int*, const T*, T*, const T*, T, T*);         \
    template void residual_restrict<T>(tf_ctx&, int, const int*, const int*, const int*, const int*, const T*,   \
                                       const T*, const T*, T*, const T*, T, T*);                                 \
    template void prolong_jacobi<T>(tf_ctx&, int, int64_t, const int*, const int*, const T*, const int*, const T*, \
                                    const T*, bool, T*, const T*, const T*, T);                                  \
    template void enter_level<T>(tf_ctx&, int, const double*, T*, const T*, T, T*);                             \
    template void prolong_agg<T>(tf_ctx&, int, const int*, const T*, T*, bool);                                        \
    template void inverse_diagonal<T>(tf_ctx&, int, const int*, const int*, const T*, T*, int*);
TF_LA(float)
TF_LA(double)
#undef TF_LA
template void convert<double, float>(tf_ctx&, size_t, const double*, float*);
template void convert<float, double>(tf_ctx&, size_t, const float*, double*);
template void convert<int64_t, int>(tf_ctx&, size_t, const int64_t*, int*);
template void convert<int, int64_t>(tf_ctx&, size_t, const int*, int64_t*);
template void convert<double, double>(tf_ctx&, size_t, const double*, double*);
template void convert<float, float>(tf_ctx&, size_t, const float*, float*);

}  // namespace tf
