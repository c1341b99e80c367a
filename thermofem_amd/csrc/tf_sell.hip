// Sliced-ELL storage of a CSR matrix for the row-product kernels on gfx950.
//
// Rows are grouped in slices of 64 consecutive rows (one wave).  Slice s
// stores its entries slice-major: entry k of row 64s+l sits at
// base[s] + 64k + l (plain layout), for k < width(s) = the longest row of
// the slice, so the wave's k-th loads are 64 consecutive elements (one or
// two cache lines per array) and no LDS staging or block barrier is needed.
// Matrices with rows of 10 or more entries on average (aggregation-coarsened
// levels, FEM operators) are stored packed instead: entries in groups of 4
// per lane, entry k at base[s] + 256 (k / 4) + 4 l + k % 4 with widths
// rounded up to a multiple of 4, so a lane's 4 consecutive entries are one
// 16-byte value load (FP32; two for FP64) and one 8-byte column load (on
// the 598k-row level 1 of the 128^3 heat hierarchy an FP32 sweep took
// 15.4 us against 18.5 us with one load per entry); the slots of a row's
// last group past its end hold value 0 and the row's own column.  Short rows
// (5- and 7-point stencils) keep the plain layout, where rounding the widths
// up would add most of the padding.  Columns are kept as
// 16-bit offsets from the row whenever every offset of the matrix fits
// (stencil and aggregation-coarsened matrices), which cuts the index bytes
// per entry from 4 to 2; otherwise as 32-bit columns.  Row lengths are one
// byte per row (rows longer than 255 entries are not representable: the
// CSR kernels serve those matrices).  When rows of very different lengths
// share slices (aggregation-coarsened levels), rows are reordered by
// decreasing length inside windows of 256 rows (SELL-C-sigma with C = 64,
// sigma = 256) and a position -> row table is kept; results are written to
// the original rows, so the reordering is invisible.  Each lane adds its row's rounded
// products sequentially in stored (ascending column) order, so the row sums
// equal the CSR kernels' and the oracle's bit for bit.
#include <algorithm>
#include <climits>
#include <vector>

#include "tf_device.hpp"
#include "tf_internal.hpp"

namespace tf {
namespace {

constexpr int kSliceRows = kWave;
constexpr int kSlicesPerBlock = kBlock / kWave;
constexpr int kSortWindow = 256;

__global__ __launch_bounds__(kBlock) void k_sell_width(int n, int slices, const int* __restrict__ rp,
                                                       const int* __restrict__ ci, int* __restrict__ width64,
                                                       int* __restrict__ flags, int pack) {
    const int slice = int(blockIdx.x) * kSlicesPerBlock + int(threadIdx.x) / kWave;
    const int lane = int(threadIdx.x) & (kWave - 1);
    const int row = slice * kSliceRows + lane;
    int len = 0;
    if (slice < slices && row < n) {
        const int s = rp[row], e = rp[row + 1];
        len = e - s;
        if (len > 255) flags[0] = 1;
        bool wide = false;
        for (int j = s; j < e; ++j) {
            const int d = ci[j] - row;
            wide |= d < -32768 || d > 32767;
        }
        if (wide) flags[1] = 1;
    }
    int m = len;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, kWave));
    if (slice < slices && lane == 0) width64[slice] = (pack ? (m + 3) / 4 * 4 : m) * kSliceRows;
    if (slice == slices - 1 && lane == 0) width64[slices] = 0;
}

// Rows of each 256-row window ranked by (length descending, row ascending);
// writes the position -> row table and the width of every slice.
__global__ __launch_bounds__(kSortWindow) void k_sell_sort(int n, int slices, const int* __restrict__ rp,
                                                           int* __restrict__ perm, int* __restrict__ width64,
                                                           int pack) {
    __shared__ int lens[kSortWindow];
    const int w0 = int(blockIdx.x) * kSortWindow;
    const int i = int(threadIdx.x);
    const int row = w0 + i;
    const int len = row < n ? rp[row + 1] - rp[row] : -1;
    lens[i] = len;
    __syncthreads();
    if (row >= n) return;
    int rank = 0;
    for (int j = 0; j < kSortWindow; ++j) {
        const int lj = lens[j];
        rank += (lj > len) || (lj == len && j < i);
    }
    const int pos = w0 + rank;
    perm[pos] = row;
    if ((pos & (kSliceRows - 1)) == 0) width64[pos / kSliceRows] = (pack ? (len + 3) / 4 * 4 : len) * kSliceRows;
    if (pos / kSliceRows == slices - 1 && (pos & (kSliceRows - 1)) == 0) width64[slices] = 0;
}

template <class C>
__global__ __launch_bounds__(kBlock) void k_sell_fill(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                      const int* __restrict__ perm, const int* __restrict__ base,
                                                      unsigned char* __restrict__ len, C* __restrict__ col, int pack) {
    const int pos = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (pos >= n) return;
    const int row = perm ? perm[pos] : pos;
    const int s = rp[row], e = rp[row + 1];
    len[pos] = static_cast<unsigned char>(e - s);
    const int b = base[pos / kSliceRows], l = pos & (kSliceRows - 1);
    const int filled = pack ? (e - s + 3) / 4 * 4 : e - s;  // packed: the last group's spare slots too
    for (int k = 0; k < filled; ++k) {
        const int j = s + k;
        if constexpr (sizeof(C) == 2)
            col[b + sell_slot(k, l, pack)] = static_cast<C>(j < e ? ci[j] - row : 0);
        else
            col[b + sell_slot(k, l, pack)] = j < e ? ci[j] : row;
    }
}

template <class A, class T>
__global__ __launch_bounds__(kBlock) void k_sell_values(int n, const int* __restrict__ rp, const int* __restrict__ perm,
                                                        const int* __restrict__ base, const A* __restrict__ v,
                                                        T* __restrict__ out, int pack) {
    const int pos = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (pos >= n) return;
    const int row = perm ? perm[pos] : pos;
    const int s = rp[row], e = rp[row + 1];
    const int b = base[pos / kSliceRows], l = pos & (kSliceRows - 1);
    const int filled = pack ? (e - s + 3) / 4 * 4 : e - s;
    for (int k = 0; k < filled; ++k) out[b + sell_slot(k, l, pack)] = s + k < e ? static_cast<T>(v[s + k]) : T(0);
}

template <class T, class C, int E>
__global__ __launch_bounds__(kBlock) void k_sell(SellView<C> m, const T* __restrict__ v, const T* __restrict__ x,
                                                 T* __restrict__ y, const T* __restrict__ b,
                                                 const T* __restrict__ inv, T w) {
    const int blk = xcd_block((m.slices + kSlicesPerBlock - 1) / kSlicesPerBlock);
    if (blk < 0) return;
    const int lane = int(threadIdx.x) & (kWave - 1);
    const int slice = blk * kSlicesPerBlock + int(threadIdx.x) / kWave;
    const int pos = slice * kSliceRows + lane;
    if (pos >= m.n) return;
    const int row = m.perm ? m.perm[pos] : pos;
    auto xf = [&](int c) { return x[c]; };
    const T acc = m.packed
                      ? sell_row_sum4<T>(row, int(m.len[pos]), v + m.base[slice] + 4 * lane,
                                         m.col + m.base[slice] + 4 * lane, xf)
                      : sell_row_sum<T>(row, int(m.len[pos]), v + m.base[slice] + lane, m.col + m.base[slice] + lane, xf);
    epilogue<T, E>(row, acc, x, y, b, inv, w);
}

// The same row products with R lanes per row: a wave takes 64 / R rows of a
// slice and lane segment q of a row the slots [q K', (q + 1) K') of the
// slice's width K (K' = ceil(K / R)), so a slice is R waves and a lane loads
// (and gathers) a fraction of the row at once.  The row sum stays the
// sequential one: segment 0 adds its rounded products from 0, each later
// segment continues from the previous segment's partial (handed over by a
// shuffle) in slot order, and the last segment applies the epilogue -- the
// same operations in the same order as k_sell, so results are identical.
// Levels whose rows are long relative to their count (few slices, long
// per-lane chains) run faster this way; the tuner picks R per level.
constexpr int kSplitRegs = 16;  // products held per lane before the hand-over
template <class T, class C, int E, int R>
__global__ __launch_bounds__(kBlock) void k_sell_split(SellView<C> m, const T* __restrict__ v,
                                                       const T* __restrict__ x, T* __restrict__ y,
                                                       const T* __restrict__ b, const T* __restrict__ inv, T w) {
#pragma clang fp contract(off)
    constexpr int W = kWave / R;
    const int units = m.slices * R;
    const int blk = xcd_block((units + kSlicesPerBlock - 1) / kSlicesPerBlock);
    if (blk < 0) return;
    const int unit = blk * kSlicesPerBlock + int(threadIdx.x) / kWave;
    if (unit >= units) return;  // (uniform over the wave)
    const int lane = int(threadIdx.x) & (kWave - 1);
    const int slice = unit / R, q = lane / W;
    const int rl = (unit % R) * W + lane % W;  // the row's lane in its slice
    const int pos = slice * kSliceRows + rl;
    const bool valid = pos < m.n;
    const int len = valid ? int(m.len[pos]) : 0;
    const int base = m.base[slice];
    const int kseg = ((m.base[slice + 1] - base) / kSliceRows + R - 1) / R;
    const int k0 = q * kseg, k1 = min(len, k0 + kseg);
    const int cnt = k1 > k0 ? k1 - k0 : 0;
    const int row = valid ? (m.perm ? m.perm[pos] : pos) : 0;
    const T* vp = v + base;
    const C* cp = m.col + base;
    T prod[kSplitRegs];
    if (cnt > 0) {
        T pv[kSplitRegs], xv[kSplitRegs];
        int pc[kSplitRegs];
#pragma unroll
        for (int u = 0; u < kSplitRegs; ++u) {
            const int k = min(k0 + u, k1 - 1);
            pv[u] = ld_entry(vp + sell_slot(k, rl, m.packed));
            pc[u] = sell_col(ld_entry(cp + sell_slot(k, rl, m.packed)), row);
        }
#pragma unroll
        for (int u = 0; u < kSplitRegs; ++u) xv[u] = x[pc[u]];
#pragma unroll
        for (int u = 0; u < kSplitRegs; ++u) prod[u] = pv[u] * xv[u];
    }
    T acc = 0;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const T in = __shfl_up(acc, W);
        if (q == t) {
            acc = t == 0 ? T(0) : in;
#pragma unroll
            for (int u = 0; u < kSplitRegs; ++u) {
                const T sum = acc + prod[u];
                acc = u < cnt ? sum : acc;
            }
            for (int k = k0 + kSplitRegs; k < k1; ++k) {  // rows wider than R * kSplitRegs
                const T p = ld_entry(vp + sell_slot(k, rl, m.packed)) *
                            x[sell_col(ld_entry(cp + sell_slot(k, rl, m.packed)), row)];
                acc = acc + p;
            }
        }
    }
    if (q == R - 1 && valid) epilogue<T, E>(row, acc, x, y, b, inv, w);
}

}  // namespace

namespace {
// Slice offsets from the widths (exclusive scan); false if the total would
// pass the int range or the padding limit.  Sets s.base and s.stored.
bool sell_offsets(tf_ctx& c, int slices, int64_t nnz, const int* width64, double maxPadding, DevSell& s,
                  int flags = 0) {
    if (int64_t(slices) * kSliceRows * 256 > int64_t(INT32_MAX)) {
        // The offsets are scanned in int: check the exact total first when it could overflow.
        std::vector<int> w(static_cast<size_t>(slices));
        check(hipMemcpyAsync(w.data(), width64, size_t(slices) * sizeof(int), hipMemcpyDeviceToHost, c.stream),
              "sell widths");
        c.sync();
        int64_t total = 0;
        for (int x : w) total += x;
        if (total > int64_t(INT32_MAX)) {
            if (flags > 0) c.read_flags(flags);
            return false;
        }
    }
    int* base = dev_alloc<int>(c, size_t(slices) + 1);
    scan_excl(c, slices + 1, width64, base);
    const int64_t stored = flags > 0 ? c.read_int_with_flags(base + slices, flags) : c.read_int(base + slices);
    if (double(stored) > double(nnz) * (1.0 + maxPadding)) {
        dev_free(c, base);
        return false;
    }
    s.base = base;
    s.stored = stored;
    return true;
}
}  // namespace

bool sell_build(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, DevSell& s, double maxPadding,
                double maxPackedPadding) {
    // Packed, in row order, when rows hold 10 or more entries on average:
    // a lane reads only its own row's groups, so the slices' padding costs
    // memory, not bandwidth, and row order keeps the gathers and stores of a
    // wave on neighbouring rows (aggregation-coarsened levels of the 128^3
    // heat hierarchy: 1.42-1.56x stored entries in row order, 1.22-1.25x
    // when sorted by length); a packed build past maxPackedPadding is made plain
    // instead, sorted by length when that keeps it under maxPadding (the
    // library's levels allow the packed layout 0.6).
    if (nnz >= 10 * int64_t(n) && sell_build(c, n, nnz, rp, ci, s, maxPackedPadding < 0 ? maxPadding : maxPackedPadding,
                                             true))
        return true;
    return sell_build(c, n, nnz, rp, ci, s, maxPadding, false);
}

bool sell_build(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, DevSell& s, double maxPadding,
                bool pack) {
    s = DevSell{};
    if (n <= 0) return false;
    const int slices = (n + kSliceRows - 1) / kSliceRows;
    int* width64 = c.wst<int>(kSellWsSlot, size_t(slices) + 1);
    c.clear_flags(2);
    hipLaunchKernelGGL(k_sell_width, dim3((slices + kSlicesPerBlock - 1) / kSlicesPerBlock), dim3(kBlock), 0,
                       c.stream, n, slices, rp, ci, width64, c.dflags, pack ? 1 : 0);
    check(hipGetLastError(), "sell width launch");
    // The width flags (a row over 255 entries, a column offset past 16 bits)
    // come back with the first offsets scan's total.
    const bool fits = sell_offsets(c, slices, nnz, width64, maxPadding, s, 2);
    const bool tooLong = c.hflags[0] != 0, wide = c.hflags[1] != 0;
    if (tooLong) {
        free_sell(c, s);
        return false;
    }
    if (!fits && pack) {
        free_sell(c, s);
        return false;
    }
    if (!fits) {
        // Rows of one slice differ too much in length: order rows by length in 256-row windows.
        int* perm = dev_alloc<int>(c, size_t(n));
        hipLaunchKernelGGL(k_sell_sort, dim3((n + kSortWindow - 1) / kSortWindow), dim3(kSortWindow), 0, c.stream, n,
                           slices, rp, perm, width64, pack ? 1 : 0);
        check(hipGetLastError(), "sell sort launch");
        if (!sell_offsets(c, slices, nnz, width64, maxPadding, s)) {
            dev_free(c, perm);
            return false;
        }
        s.perm = perm;
    }
    s.n = n;
    s.slices = slices;
    s.narrow = !wide;
    s.packed = pack;
    s.len = dev_alloc<unsigned char>(c, size_t(n));
    s.col = s.narrow ? static_cast<void*>(dev_alloc<short>(c, size_t(s.stored)))
                     : static_cast<void*>(dev_alloc<int>(c, size_t(s.stored)));
    const dim3 grid((n + kBlock - 1) / kBlock);
    if (s.narrow)
        hipLaunchKernelGGL(k_sell_fill<short>, grid, dim3(kBlock), 0, c.stream, n, rp, ci, s.perm, s.base, s.len,
                           static_cast<short*>(s.col), pack ? 1 : 0);
    else
        hipLaunchKernelGGL(k_sell_fill<int>, grid, dim3(kBlock), 0, c.stream, n, rp, ci, s.perm, s.base, s.len,
                           static_cast<int*>(s.col), pack ? 1 : 0);
    check(hipGetLastError(), "sell fill launch");
    return true;
}

template <class A, class T>
void sell_values(tf_ctx& c, const DevSell& s, const int* rp, const A* v, T* out) {
    if (s.n <= 0) return;
    hipLaunchKernelGGL((k_sell_values<A, T>), dim3((s.n + kBlock - 1) / kBlock), dim3(kBlock), 0, c.stream, s.n, rp,
                       s.perm, s.base, v, out, s.packed ? 1 : 0);
    check(hipGetLastError(), "sell values launch");
}

void free_sell(tf_ctx& c, DevSell& s) {
    dev_free(c, s.perm);
    dev_free(c, s.base);
    dev_free(c, s.len);
    dev_free(c, s.col);
    s = DevSell{};
}

template <class T, int E>
void sell_rows(tf_ctx& c, const DevSell& s, const T* v, const T* x, T* y, const T* b, const T* inv, T w) {
    if (s.n <= 0) return;
    if (s.lanes == 2 || s.lanes == 4) {
        const dim3 grid(round_xcd((s.slices * s.lanes + kSlicesPerBlock - 1) / kSlicesPerBlock));
#define TF_SPLIT(C, R)                                                                                         \
    hipLaunchKernelGGL((k_sell_split<T, C, E, R>), grid, dim3(kBlock), 0, c.stream, s.view<C>(), v, x, y, b, inv, w)
        if (s.narrow) {
            if (s.lanes == 2)
                TF_SPLIT(short, 2);
            else
                TF_SPLIT(short, 4);
        }
        else {
            if (s.lanes == 2)
                TF_SPLIT(int, 2);
            else
                TF_SPLIT(int, 4);
        }
#undef TF_SPLIT
        check(hipGetLastError(), "sell kernel launch");
        return;
    }
    const dim3 grid(round_xcd((s.slices + kSlicesPerBlock - 1) / kSlicesPerBlock));
    if (s.narrow)
        hipLaunchKernelGGL((k_sell<T, short, E>), grid, dim3(kBlock), 0, c.stream, s.view<short>(), v, x, y, b, inv,
                           w);
    else
        hipLaunchKernelGGL((k_sell<T, int, E>), grid, dim3(kBlock), 0, c.stream, s.view<int>(), v, x, y, b, inv, w);
    check(hipGetLastError(), "sell kernel launch");
}

template <class T>
void sell_spmv(tf_ctx& c, const DevSell& s, const T* v, const T* x, T* y) {
    sell_rows<T, kSpmv>(c, s, v, x, y, nullptr, nullptr, T(0));
}
template <class T>
void sell_residual(tf_ctx& c, const DevSell& s, const T* v, const T* b, const T* x, T* r) {
    sell_rows<T, kResidual>(c, s, v, x, r, b, nullptr, T(0));
}
template <class T>
void sell_jacobi(tf_ctx& c, const DevSell& s, const T* v, const T* inv, T w, const T* b, const T* xin, T* xout) {
    sell_rows<T, kJacobi>(c, s, v, xin, xout, b, inv, w);
}

template void sell_values<double, double>(tf_ctx&, const DevSell&, const int*, const double*, double*);
template void sell_values<double, float>(tf_ctx&, const DevSell&, const int*, const double*, float*);
template void sell_values<float, float>(tf_ctx&, const DevSell&, const int*, const float*, float*);
#define TF_SELL(T)                                                                                        \
    template void sell_spmv<T>(tf_ctx&, const DevSell&, const T*, const T*, T*);                          \
    template void sell_residual<T>(tf_ctx&, const DevSell&, const T*, const T*, const T*, T*);            \
    template void sell_jacobi<T>(tf_ctx&, const DevSell&, const T*, const T*, T, const T*, const T*, T*);
TF_SELL(float)
TF_SELL(double)
#undef TF_SELL

}  // namespace tf
