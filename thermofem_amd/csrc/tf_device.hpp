// Device-side building blocks shared by every gfx950 kernel of the library:
// 64-lane wave reductions, deterministic block/grid reductions and the
// XCD-aware block index remap.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace tf {

constexpr int kBlock = 256;  // 4 waves of 64 lanes
constexpr int kWave = 64;
constexpr int kXcds = 8;     // MI355X: 8 XCDs, each with its own L2

// Sum over groups of G consecutive lanes (G power of two, <= 64).
template <int G, class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
    return group_sum<kWave>(v);
}

// Deterministic block sum (blocks of up to 1024 threads): waves reduce by
// shuffles, wave totals are added in wave order by thread 0.  The result is
// valid in thread 0 only.
template <class T>
__device__ __forceinline__ T block_sum(T v) {
    __shared__ T partial[1024 / kWave];
    v = wave_sum(v);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) partial[w] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0) {
#pragma unroll 16
        for (int i = 0; i < int(blockDim.x) / kWave; ++i) r += partial[i];
    }
    __syncthreads();
    return r;
}

// Grid-wide deterministic reductions are two launches: every block writes
// its block_sum to partials[blockIdx.x]; a one-block kernel then adds the
// partials in block order (sum_partials).  No device-scope atomics on a
// shared counter are used: across XCDs those serialise at the memory side.
__device__ __forceinline__ double sum_partials(const double* __restrict__ partials, int nparts) {
    double s = 0;
#pragma unroll 8
    for (int i = int(threadIdx.x); i < nparts; i += int(blockDim.x)) s += partials[i];
    return block_sum(s);
}

// Blocks are dispatched round-robin over the XCDs (block b runs on XCD b%8).
// Remapping gives each XCD one contiguous slab of logical blocks, so the
// neighbouring rows a stencil SpMV gathers stay in that XCD's L2.  The grid
// must be a multiple of kXcds; logical ids >= limit are idle.
__device__ __forceinline__ int xcd_block(int limit) {
    const int per = int(gridDim.x) / kXcds;
    const int logical = int(blockIdx.x % kXcds) * per + int(blockIdx.x / kXcds);
    return logical < limit ? logical : -1;
}

inline int round_xcd(int blocks) { return (blocks + kXcds - 1) / kXcds * kXcds; }

// Position of column `col` in the sorted row [b, e) (unique columns), or
// -1.  A row of up to 8 entries is read as one chunk whose loads are issued
// together (indices clamped into the row; a clamped slot repeats the last
// column, and the first matching slot is the real position), so it costs one
// round of loads instead of a chain of dependent probes; longer rows are
// binary-searched.
__device__ __forceinline__ int find_sorted_col(const int* __restrict__ ci, int b, int e, int col) {
    if (e - b <= 8) {
        if (e <= b) return -1;
        int cc[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) cc[u] = ci[min(b + u, e - 1)];
        int k = -1;
#pragma unroll
        for (int u = 7; u >= 0; --u) k = cc[u] == col ? b + u : k;
        return k;
    }
    while (b < e) {
        const int m = (b + e) >> 1;
        const int c = ci[m];
        if (c == col) return m;
        if (c < col)
            b = m + 1;
        else
            e = m;
    }
    return -1;
}

// Loads of matrix entries (values, column indices) that a launch reads
// once.  Plain loads: on gfx950 the nontemporal hint measured slower for the
// level-0 SpMV (38.5 against 28.4 us with the matrix coming from HBM, 35.8
// against 26.0 us when it is resident in the 256 MB Infinity Cache, which
// nontemporal loads do not fill).
template <class T>
__device__ __forceinline__ T ld_entry(const T* p) {
    return *p;
}

// LDS capacity (entries) of one CSR-stream chunk: 16 KB of products.
template <class T>
struct StreamCap {
    static constexpr int value = 16384 / int(sizeof(T));
};
constexpr int kMaxStreamRows = 256;  // rows per row block (one per thread)

// Row sums of `rows` (<= kMaxStreamRows) consecutive CSR rows starting at
// row0, for one 256-thread block.  Phase 1 streams the block's contiguous
// column/value range with coalesced, unrolled loads and stores each product
// v*x[c] to LDS; phase 2 lets thread t add up rows row0+t, row0+t+256, ...
// sequentially in stored order.  Long rows are handled in chunks of the LDS
// capacity.  Products are rounded when stored and added one by one, so each
// row sum is the sequential sum of rounded products (no fused multiply-add).
// acc[k] receives the sum of row row0 + t + 256k.  All threads must call it.
template <class T, class XF>
__device__ __forceinline__ void csr_stream_rows(int row0, int rows, const int* __restrict__ rp,
                                                const int* __restrict__ ci, const T* __restrict__ v, XF&& xval,
                                                T* prod, int* srp, T (&acc)[kMaxStreamRows / kBlock]) {
    constexpr int CAP = StreamCap<T>::value;
    constexpr int U = CAP / kBlock;
    constexpr int RPT = kMaxStreamRows / kBlock;
    const int tid = int(threadIdx.x);
    for (int t = tid; t <= rows; t += kBlock) srp[t] = rp[row0 + t];
    __syncthreads();
    const int s = srp[0], e = srp[rows];
#pragma unroll
    for (int k = 0; k < RPT; ++k) acc[k] = 0;
    for (int c0 = s; c0 < e; c0 += CAP) {
        const int c1 = min(e, c0 + CAP);
        int cc[U];
        T vv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = c0 + tid + u * kBlock;
            if (j < c1) {
                cc[u] = ld_entry(ci + j);
                vv[u] = ld_entry(v + j);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = c0 + tid + u * kBlock;
            if (j < c1) prod[j - c0] = vv[u] * xval(cc[u]);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = tid + k * kBlock;
            if (r < rows) {
                const int a = max(srp[r], c0), z = min(srp[r + 1], c1);
                T sum = acc[k];
                for (int q = a; q < z; ++q) sum += prod[q - c0];
                acc[k] = sum;
            }
        }
        __syncthreads();
    }
}

// Rows per block so that an average row block fits one LDS chunk.
template <class T>
inline int stream_rows_per_block(int n, int64_t nnz) {
    const double avg = n > 0 ? double(nnz) / double(n) : 1.0;
    int r = kMaxStreamRows;
    while (r > 1 && double(r) * avg > double(StreamCap<T>::value)) r >>= 1;
    return r;
}

enum Epilogue { kSpmv = 0, kResidual = 1, kJacobi = 2 };

// Fused epilogues of the row-product kernels, evaluated without
// contraction: y = sum, r = b - sum, x + (w*inv)*(b - sum).
template <class T, int E>
__device__ __forceinline__ void epilogue(int row, T acc, const T* __restrict__ x, T* __restrict__ y,
                                         const T* __restrict__ b, const T* __restrict__ inv, T w) {
#pragma clang fp contract(off)
    if constexpr (E == kSpmv) {
        y[row] = acc;
    }
    else if constexpr (E == kResidual) {
        y[row] = b[row] - acc;
    }
    else {
        const T wd = w * inv[row];
        const T d = b[row] - acc;
        const T upd = wd * d;
        y[row] = x[row] + upd;
    }
}

// Read-only view of a sliced-ELL matrix (slices of 64 row positions, entries
// stored slice-major in groups of 4 per lane, see sell_slot; C = short:
// column offsets from the row, int: columns).
// With perm, rows are ordered by decreasing length inside windows of 256
// rows so that the rows sharing a slice have similar lengths.
template <class C>
struct SellView {
    int n;
    int slices;
    int packed;                // layout (sell_slot)
    const int* base;           // entry offset of each slice (slices + 1)
    const unsigned char* len;  // row length per position
    const C* col;
    const int* perm;           // row at each position (nullptr: identity)
};

// Read-only view of a structured-stencil operator (tf_stencil.hip): every
// row's columns are row + off[t] for terms t of one ascending offset list;
// mask[row] bit t says whether the row has term t.  The stored value arrays
// are interleaved by tiles of 64 rows (one wave): value a of row r sits at
// stencil_index(r, a, tile) with tile = 64 * arrays, so a wave's loads of
// its rows' values are one contiguous run.  Term t of row `row` reads
// array src[t] / 64 at row + shift[t] (shift is 0, or -d for a term -d whose
// values are read from the +d array of a symmetric operator).  Terms past
// the list's end repeat its last term (their mask bits are 0).
constexpr int kStencilTerms = 16;
struct StencilView {
    int n;
    int tile;
    const unsigned short* mask;
    int off[kStencilTerms];
    int shift[kStencilTerms];
    int src[kStencilTerms];  // 64 * array index
};
__host__ __device__ __forceinline__ int stencil_index(int r, int slot64, int tile) {
    return (r >> 6) * tile + slot64 + (r & 63);
}

// Sum of row `row` of a stencil operator for the calling lane: NC chunks of
// 8 terms, each issuing all its value loads and gathers before the adds
// (row indices clamped into range for terms the row lacks, which are not
// added).  Products are rounded and added one by one in ascending column
// order, the CSR order.
template <class T, int NC, class XF>
__device__ __forceinline__ T stencil_row_sum(const StencilView& m, const T* __restrict__ vals, int row,
                                             XF&& xval) {
#pragma clang fp contract(off)
    const unsigned mask = m.mask[row];
    T acc = 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        T pv[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int t = c * 8 + u;
            pv[u] = vals[stencil_index(min(max(row + m.shift[t], 0), m.n - 1), m.src[t], m.tile)];
            xv[u] = xval(min(max(row + m.off[t], 0), m.n - 1));
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const T p = pv[u] * xv[u];
            const T s = acc + p;
            acc = (mask >> (c * 8 + u)) & 1u ? s : acc;
        }
    }
    return acc;
}

// Row sum of the mirrored 7-point stencil -- offsets -P, -N, -1, 0, 1, N, P
// (0 < 1 < N < P), arrays diag, +1, +N, +P, the negative terms read from
// the positive ones' arrays -- for a wave whose 64 lanes hold 64 consecutive
// rows (row0 + lane; rows past n are clamped, their sums unused).  The -1
// and +1 neighbours' x and the -1 term's value (the +1 array at row - 1) come
// from the neighbouring lanes by shuffles, the wave's edge lanes loading
// the ones beyond it, so a row issues 12 loads instead of 17 (level-0
// FP64 SpMV of the 128^3 heat operator: 32.7 -> 29.3 us with the data in
// HBM).  The same products added in the same order as
// stencil_row_sum<T, 1>.  Every lane of the wave must call it.
template <class T>
__device__ __forceinline__ T stencil7_row_sum(const StencilView& m, const T* __restrict__ vals, const T* x, int row) {
#pragma clang fp contract(off)
    const int lane = int(threadIdx.x) & (kWave - 1);
    const int n = m.n, r = min(row, n - 1);
    const int N = m.off[5], P = m.off[6];
    const unsigned mask = m.mask[r];
    const int tb = stencil_index(r, 0, 4 * kWave);
    const T a0 = vals[tb], a1 = vals[tb + kWave], aN = vals[tb + 2 * kWave], aP = vals[tb + 3 * kWave];
    const int rmN = max(r - N, 0), rmP = max(r - P, 0);
    const T bN = vals[stencil_index(rmN, 2 * kWave, 4 * kWave)];
    const T bP = vals[stencil_index(rmP, 3 * kWave, 4 * kWave)];
    const T xr = x[r];
    const T xmN = x[rmN], xpN = x[min(r + N, n - 1)], xmP = x[rmP], xpP = x[min(r + P, n - 1)];
    T b1 = __shfl_up(a1, 1), xm1 = __shfl_up(xr, 1), xp1 = __shfl_down(xr, 1);
    if (lane == 0) {
        const int rm1 = max(r - 1, 0);
        b1 = vals[stencil_index(rm1, kWave, 4 * kWave)];
        xm1 = x[rm1];
    }
    if (lane == kWave - 1) xp1 = x[min(r + 1, n - 1)];
    const T pv[7] = {bP, bN, b1, a0, a1, aN, aP};
    const T xv[7] = {xmP, xmN, xm1, xr, xp1, xpN, xpP};
    T acc = 0;
#pragma unroll
    for (int t = 0; t < 7; ++t) {
        const T p = pv[t] * xv[t];
        const T s = acc + p;
        acc = (mask >> t) & 1u ? s : acc;
    }
    return acc;
}

__device__ __forceinline__ int sell_col(short c, int row) { return row + int(c); }
__device__ __forceinline__ int sell_col(int c, int) { return c; }

// Sliced-ELL slot of entry k of the row at lane l of its slice, relative to
// the slice's base.  Plain layout: entry k of every lane, then entry k + 1
// (64 k + l).  Packed layout: groups of 4 per lane (group g of every lane's
// row, then group g + 1), so a lane's 4 consecutive entries are contiguous
// and one vector load brings them in.
__host__ __device__ __forceinline__ int sell_slot(int k, int l, bool packed) {
    return packed ? 256 * (k >> 2) + 4 * l + (k & 3) : kWave * k + l;
}

// Four consecutive entries (16-byte aligned for values and 32-bit columns,
// 8-byte aligned for 16-bit columns).
typedef float SellF4 __attribute__((ext_vector_type(4)));
typedef double SellD2 __attribute__((ext_vector_type(2)));
typedef short SellS4 __attribute__((ext_vector_type(4)));
typedef int SellI4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void sell_load4(const float* p, float (&o)[4]) {
    const SellF4 q = *reinterpret_cast<const SellF4*>(p);
    o[0] = q.x;
    o[1] = q.y;
    o[2] = q.z;
    o[3] = q.w;
}
__device__ __forceinline__ void sell_load4(const double* p, double (&o)[4]) {
    const SellD2 a = *reinterpret_cast<const SellD2*>(p);
    const SellD2 b = *reinterpret_cast<const SellD2*>(p + 2);
    o[0] = a.x;
    o[1] = a.y;
    o[2] = b.x;
    o[3] = b.y;
}
__device__ __forceinline__ void sell_load4(const short* p, short (&o)[4]) {
    const SellS4 q = *reinterpret_cast<const SellS4*>(p);
    o[0] = q.x;
    o[1] = q.y;
    o[2] = q.z;
    o[3] = q.w;
}
__device__ __forceinline__ void sell_load4(const int* p, int (&o)[4]) {
    const SellI4 q = *reinterpret_cast<const SellI4*>(p);
    o[0] = q.x;
    o[1] = q.y;
    o[2] = q.z;
    o[3] = q.w;
}

// Sum of one row of a sliced-ELL matrix for the calling lane (plain
// layout): vp and cp point at the row's first entry (base + lane, stride 64
// between its entries), len >= 0.  Chunks of 8 entries issue all their loads
// and gathers before the adds; slots past the row's end re-read its last
// entry (no branch, no padding read) and are not added.  Products are
// rounded and added one by one in stored order, as csr_stream_rows does.
template <class T, class C, class XF>
__device__ __forceinline__ T sell_row_sum(int row, int len, const T* __restrict__ vp, const C* __restrict__ cp,
                                          XF&& xval) {
#pragma clang fp contract(off)
    T acc = 0;
    for (int k0 = 0; k0 < len; k0 += 8) {
        T pv[8], xv[8];
        int pc[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = min(k0 + u, len - 1);
            pv[u] = ld_entry(vp + k * kWave);
            pc[u] = sell_col(ld_entry(cp + k * kWave), row);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = xval(pc[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const T p = pv[u] * xv[u];
            const T s = acc + p;
            acc = k0 + u < len ? s : acc;
        }
    }
    return acc;
}

// The same for the packed layout: vp and cp point at the lane's first group
// (base + 4 lane).  Two groups (8 entries) go per round, each one vector
// load of values and one of columns, all gathers of the round issued before
// the adds; a round past the row's last group re-reads that group and adds
// nothing.  The slots of the last group past the row's end hold value 0 and
// the row's own column, and are not added.  The same arithmetic.
template <class T, class C, class XF>
__device__ __forceinline__ T sell_row_sum4(int row, int len, const T* __restrict__ vp, const C* __restrict__ cp,
                                           XF&& xval) {
#pragma clang fp contract(off)
    T acc = 0;
    const int groups = (len + 3) >> 2;
    for (int g0 = 0; g0 < groups; g0 += 2) {
        T pv[8], xv[8];
        C cv[8];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int g = min(g0 + h, groups - 1);
            T pq[4];
            C cq[4];
            sell_load4(vp + 256 * g, pq);
            sell_load4(cp + 256 * g, cq);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                pv[4 * h + u] = pq[u];
                cv[4 * h + u] = cq[u];
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = xval(sell_col(cv[u], row));
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const T p = pv[u] * xv[u];
            const T s = acc + p;
            acc = 4 * g0 + u < len ? s : acc;
        }
    }
    return acc;
}

// x = A^-1 b for the dense coarsest level (order n <= kCoarseMax) from the
// column-major inverse (it[j * n + r] = (A^-1)_rj), one 1024-thread block.
// b is staged in LDS (bs) as FP64 first, so rhs and x may alias.  The
// columns are cut into coarse_segments(n) runs of coarse_run(n) consecutive
// columns; thread s * n + r forms row r's partial over run s (fused
// multiply-adds, columns ascending, starting from 0; the loads, coalesced
// over r, are issued 8 at a time before their first use), the partials meet
// in part (segments * n <= 1024 doubles) and thread r adds them in run
// order.  Zero entries of the inverse are skipped, so a non-finite value in
// b reaches only the rows the inverse couples to it (the coarse inverse of a
// block-diagonal operator is block-diagonal: independent systems stay
// independent).  All threads of the block must call it.
constexpr int kCoarseMax = 1024;
__host__ __device__ constexpr int coarse_segments(int n) {
    return n >= 1024 ? 1 : (1024 / n < 16 ? 1024 / n : 16);
}
__host__ __device__ constexpr int coarse_run(int n) {
    return (n + coarse_segments(n) - 1) / coarse_segments(n);
}
template <class T>
__device__ __forceinline__ void coarse_inverse_apply(int n, const double* __restrict__ it, const T* b, T* x,
                                                     double* bs, double* part) {
    const int tid = int(threadIdx.x);
    for (int i = tid; i < n; i += 1024) bs[i] = double(b[i]);
    __syncthreads();
    const int segs = coarse_segments(n), run = coarse_run(n);
    if (tid < segs * n) {
        const int s = tid / n, r = tid - s * n;
        const int j0 = s * run, j1 = min(n, j0 + run);
        double p = 0;
        for (int jb = j0; jb < j1; jb += 8) {
            double a[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = it[size_t(min(jb + u, j1 - 1)) * size_t(n) + size_t(r)];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (jb + u < j1 && a[u] != 0.0) p = fma(a[u], bs[jb + u], p);
        }
        part[tid] = p;
    }
    __syncthreads();
    for (int r = tid; r < n; r += 1024) {
        double acc = part[r];
        for (int s = 1; s < segs; ++s) acc += part[s * n + r];
        x[r] = T(acc);
    }
}

}  // namespace tf
