// Structured-stencil storage of an operator whose rows all take their columns
// from one short list of offsets from the row: finite-difference and
// structured-grid matrices (the 7-point heat operator on an nx*ny*nz grid
// has the offsets -nx*ny, -nx, -1, 0, 1, nx, nx*ny).  No column indices are
// stored.  Each row keeps a 16-bit mask of the terms it has, and the values
// sit in dense arrays, one per stored offset (the DIA layout), interleaved by
// tiles of 64 rows: a wave reads its rows' values as one contiguous run and
// gathers x at fixed offsets.  When the operator is symmetric bit for bit, the arrays of the
// negative offsets are not stored: a(i, i-d) is read as a(i-d, i) from the
// +d array, which holds the off-diagonal bytes to half.  For the 128^3 heat
// operator the FP64 product then moves 4 doubles and 2 mask bytes per row
// against 7 x (8 + 2) bytes in sliced ELL.  Every kernel adds a row's
// rounded products one by one in ascending column order (the CSR order),
// so the results equal the CSR and sliced-ELL kernels' bit for bit.
//
// The offset list comes from a sample of rows (512 around the middle, 512
// spread over the matrix); the fill pass then checks every entry against it
// and every mirrored value against its transpose bit for bit, so a matrix that
// is not a stencil, or not symmetric, is never stored wrongly: it keeps its
// sliced-ELL or CSR copy, or its full (non-mirrored) stencil arrays.
#include <algorithm>
#include <climits>
#include <vector>

#include "tf_device.hpp"
#include "tf_internal.hpp"

namespace tf {
namespace {

constexpr int kStencilWsSlot = 29;
constexpr int kProbeThreads = 1024;
constexpr int kProbeSlots = 32;
constexpr int kEmptyOffset = INT_MIN;

// Distinct column offsets of the sampled rows, ascending: out[0] = their
// count (-1 when more than kStencilTerms), out[1 + t] = offset t.
__global__ __launch_bounds__(kProbeThreads) void k_stencil_probe(int n, const int* __restrict__ rp,
                                                                 const int* __restrict__ ci, int* out) {
    __shared__ int set[kProbeSlots];
    __shared__ int overflow;
    const int t = int(threadIdx.x);
    if (t < kProbeSlots) set[t] = kEmptyOffset;
    if (t == 0) overflow = 0;
    __syncthreads();
    const int half = kProbeThreads / 2;
    const int64_t r = t < half ? int64_t(n) / 2 - half / 2 + t : int64_t(t - half) * n / half + n / (2 * half);
    const int row = int(r < 0 ? 0 : (r >= n ? n - 1 : r));
    for (int j = rp[row]; j < rp[row + 1] && !overflow; ++j) {
        const int off = ci[j] - row;
        int k = 0;
        // linear probing from slot 0: a filled slot never changes, so every
        // thread inserting an offset finds it or the same empty slot
        for (; k < kProbeSlots; ++k) {
            const int prev = atomicCAS(&set[k], kEmptyOffset, off);
            if (prev == kEmptyOffset || prev == off) break;
        }
        if (k == kProbeSlots) overflow = 1;
    }
    __syncthreads();
    if (t != 0) return;
    int v[kProbeSlots];
    int cnt = 0;
    for (int k = 0; k < kProbeSlots; ++k)
        if (set[k] != kEmptyOffset) v[cnt++] = set[k];
    for (int a = 1; a < cnt; ++a) {
        const int x = v[a];
        int b = a;
        while (b > 0 && v[b - 1] > x) {
            v[b] = v[b - 1];
            --b;
        }
        v[b] = x;
    }
    out[0] = (overflow || cnt > kStencilTerms) ? -1 : cnt;
    for (int k = 0; k < kStencilTerms; ++k) out[1 + k] = k < cnt ? v[k] : 0;
}

// The term list as the build kernels use it: term t has offset off[t];
// own[t] >= 0 is 64 * the index of the array that stores it, -1 when its
// values are read mirrored from the +|off| array at mir[t].
struct StencilTable {
    int terms, arrays, n, tile;
    int off[kStencilTerms];
    int own[kStencilTerms];
    int mir[kStencilTerms];
};

__device__ __forceinline__ int find_term(const StencilTable& tb, int off) {
    int t = -1;
#pragma unroll
    for (int k = 0; k < kStencilTerms; ++k)
        if (k < tb.terms && tb.off[k] == off) t = k;
    return t;
}

// Row masks and stored values; every stored array slot of the row is
// written (0 for a term the row lacks).  flags[0] = an entry whose offset
// is not in the list.  For a mirrored table, every entry of a negative term
// is compared (bit for bit) with its transpose, looked up in the row it
// mirrors: flags[1] = an entry whose transpose differs or is missing.
__global__ __launch_bounds__(kBlock) void k_stencil_fill(StencilTable tb, const int* __restrict__ rp,
                                                         const int* __restrict__ ci, const double* __restrict__ v,
                                                         unsigned short* __restrict__ mask,
                                                         double* __restrict__ vals, int* flags) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= tb.n) return;
    for (int a = 0; a < tb.arrays; ++a) vals[stencil_index(i, a * 64, tb.tile)] = 0.0;
    unsigned m = 0;
    bool unknown = false, differs = false;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
        const int c = ci[j];
        const int t = find_term(tb, c - i);
        if (t < 0) {
            unknown = true;
            continue;
        }
        m |= 1u << t;
        if (tb.own[t] >= 0) {
            vals[stencil_index(i, tb.own[t], tb.tile)] = v[j];
        }
        else {
            // the transpose (c, i) in row c < i
            const int q = find_sorted_col(ci, rp[c], rp[c + 1], i);
            differs |= q < 0 || __double_as_longlong(v[q]) != __double_as_longlong(v[j]);
        }
    }
    mask[i] = static_cast<unsigned short>(m);
    if (unknown) flags[0] = 1;
    if (differs) flags[1] = 1;
}

template <class T, int E, int NC>
__global__ __launch_bounds__(kBlock) void k_stencil(StencilView m, const T* __restrict__ vals,
                                                    const T* __restrict__ x, T* __restrict__ y,
                                                    const T* __restrict__ b, const T* __restrict__ inv, T w) {
    const int blk = xcd_block((m.n + kBlock - 1) / kBlock);
    if (blk < 0) return;
    const int row = blk * kBlock + int(threadIdx.x);
    if constexpr (NC == 0) {
        const T acc = stencil7_row_sum<T>(m, vals, x, row);  // (every lane of the wave takes part)
        if (row < m.n) epilogue<T, E>(row, acc, x, y, b, inv, w);
    }
    else {
        if (row >= m.n) return;
        const T acc = stencil_row_sum<T, NC>(m, vals, row, [&](int c) { return x[c]; });
        epilogue<T, E>(row, acc, x, y, b, inv, w);
    }
}

// The table for the offsets off[0..terms): mirrored stores the offsets >= 0
// only (ascending), the others all of them.
StencilTable make_table(int n, int terms, const int* off, bool mirrored) {
    StencilTable tb{};
    tb.terms = terms;
    tb.n = n;
    int arrays = 0;
    for (int t = 0; t < terms; ++t) {
        tb.off[t] = off[t];
        tb.own[t] = -1;
        tb.mir[t] = -1;
        if (!mirrored || off[t] >= 0) tb.own[t] = (arrays++) * 64;
    }
    tb.arrays = arrays;
    tb.tile = arrays * 64;
    if (mirrored)
        for (int t = 0; t < terms; ++t)
            if (off[t] < 0)
                for (int u = 0; u < terms; ++u)
                    if (off[u] == -off[t]) tb.mir[t] = tb.own[u];
    return tb;
}

// Runs the fill (with, mirrored, the transpose check) and reads the flags.
void fill_stencil(tf_ctx& c, const StencilTable& tb, const int* rp, const int* ci, const double* v,
                  unsigned short* mask, double* vals, int* flags, int* hostFlags) {
    check(hipMemsetAsync(flags, 0, 2 * sizeof(int), c.stream), "stencil flag clear");
    const dim3 grid((tb.n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_stencil_fill, grid, dim3(kBlock), 0, c.stream, tb, rp, ci, v, mask, vals, flags);
    check(hipGetLastError(), "stencil fill launch");
    check(hipMemcpyAsync(c.hflags, flags, 2 * sizeof(int), hipMemcpyDeviceToHost, c.stream), "stencil flags");
    c.sync();
    hostFlags[0] = c.hflags[0];
    hostFlags[1] = c.hflags[1];
}

template <class T, int E>
void stencil_rows(tf_ctx& c, const DevStencil& s, const T* v, const T* x, T* y, const T* b, const T* inv, T w) {
    if (s.n <= 0) return;
    const dim3 grid(round_xcd((s.n + kBlock - 1) / kBlock));
    // the shuffled 7-point row sum pays for FP64 rows (12 of 17 loads left);
    // FP32 sweeps measured ~1 us slower with it (11.2 -> 12.1 us at 128^3)
    if (s.chunks() == 0 && sizeof(T) == 8)
        hipLaunchKernelGGL((k_stencil<T, E, 0>), grid, dim3(kBlock), 0, c.stream, s.view(), v, x, y, b, inv, w);
    else if (s.chunks() <= 1)
        hipLaunchKernelGGL((k_stencil<T, E, 1>), grid, dim3(kBlock), 0, c.stream, s.view(), v, x, y, b, inv, w);
    else
        hipLaunchKernelGGL((k_stencil<T, E, 2>), grid, dim3(kBlock), 0, c.stream, s.view(), v, x, y, b, inv, w);
    check(hipGetLastError(), "stencil kernel launch");
}

}  // namespace

bool stencil_build(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v, DevStencil& s,
                   double** vals64) {
    s = DevStencil{};
    *vals64 = nullptr;
    if (n < 1 || nnz > int64_t(n) * kStencilTerms) return false;
    int* buf = c.wst<int>(kStencilWsSlot, 1 + kStencilTerms + 2);
    hipLaunchKernelGGL(k_stencil_probe, dim3(1), dim3(kProbeThreads), 0, c.stream, n, rp, ci, buf);
    check(hipGetLastError(), "stencil probe launch");
    check(hipMemcpyAsync(c.hflags, buf, (1 + kStencilTerms) * sizeof(int), hipMemcpyDeviceToHost, c.stream),
          "stencil probe download");
    c.sync();
    const int terms = c.hflags[0];
    const int64_t tiles = (int64_t(n) + 63) / 64;
    if (terms <= 0 || tiles * 64 * terms >= int64_t(INT_MAX)) return false;  // element indices are ints
    int off[kStencilTerms];
    for (int t = 0; t < terms; ++t) off[t] = c.hflags[1 + t];
    bool mirrored = true;
    for (int t = 0; t < terms; ++t)
        mirrored = mirrored && std::find(off, off + terms, -off[t]) != off + terms;
    int* flags = buf + 1 + kStencilTerms;
    unsigned short* mask = dev_alloc<unsigned short>(c, size_t(n));
    double* vals = nullptr;
    int hf[2] = {0, 0};
    for (int attempt = 0; attempt < 2; ++attempt) {
        const StencilTable tb = make_table(n, terms, off, mirrored);
        const size_t stored = size_t(tiles) * size_t(tb.tile);
        vals = dev_alloc<double>(c, stored);
        fill_stencil(c, tb, rp, ci, v, mask, vals, flags, hf);
        if (hf[0]) break;  // an offset outside the list: not a stencil
        if (!mirrored || !hf[1]) {
            s.n = n;
            s.terms = terms;
            s.arrays = tb.arrays;
            s.stored = int64_t(stored);
            s.mirrored = mirrored;
            s.mask = mask;
            for (int t = 0; t < kStencilTerms; ++t) {
                const int u = std::min(t, terms - 1);  // padding terms repeat the last one
                s.off[t] = off[u];
                s.shift[t] = tb.own[u] >= 0 ? 0 : off[u];
                s.src[t] = tb.own[u] >= 0 ? tb.own[u] : tb.mir[u];
            }
            *vals64 = vals;
            return true;
        }
        dev_free(c, vals);  // not symmetric: store every offset
        vals = nullptr;
        mirrored = false;
    }
    dev_free(c, vals);
    dev_free(c, mask);
    return false;
}

void stencil_copy(tf_ctx& c, const DevStencil& from, DevStencil& to) {
    to = from;
    to.mask = dev_alloc<unsigned short>(c, size_t(from.n));
    check(hipMemcpyAsync(to.mask, from.mask, size_t(from.n) * sizeof(unsigned short), hipMemcpyDeviceToDevice,
                         c.stream),
          "stencil mask copy");
}

void free_stencil(tf_ctx& c, DevStencil& s) {
    dev_free(c, s.mask);
    s = DevStencil{};
}

template <class T>
void stencil_spmv(tf_ctx& c, const DevStencil& s, const T* v, const T* x, T* y) {
    stencil_rows<T, kSpmv>(c, s, v, x, y, nullptr, nullptr, T(0));
}
template <class T>
void stencil_residual(tf_ctx& c, const DevStencil& s, const T* v, const T* b, const T* x, T* r) {
    stencil_rows<T, kResidual>(c, s, v, x, r, b, nullptr, T(0));
}
template <class T>
void stencil_jacobi(tf_ctx& c, const DevStencil& s, const T* v, const T* inv, T w, const T* b, const T* xin,
                    T* xout) {
    stencil_rows<T, kJacobi>(c, s, v, xin, xout, b, inv, w);
}

#define TF_STENCIL(T)                                                                                     \
    template void stencil_spmv<T>(tf_ctx&, const DevStencil&, const T*, const T*, T*);                    \
    template void stencil_residual<T>(tf_ctx&, const DevStencil&, const T*, const T*, const T*, T*);      \
    template void stencil_jacobi<T>(tf_ctx&, const DevStencil&, const T*, const T*, T, const T*, const T*, T*);
TF_STENCIL(float)
TF_STENCIL(double)
#undef TF_STENCIL

}  // namespace tf
