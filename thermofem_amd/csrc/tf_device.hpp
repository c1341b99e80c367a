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

template <class T>
__device__ __forceinline__ T ld_stream(const T* p) {
    return __builtin_nontemporal_load(p);
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
                cc[u] = ld_stream(ci + j);
                vv[u] = ld_stream(v + j);
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

}  // namespace tf
