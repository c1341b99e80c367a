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

// Deterministic block sum: waves reduce by shuffles, wave totals are added in
// wave order by thread 0.  The result is valid in thread 0 only.
template <class T>
__device__ __forceinline__ T block_sum(T v) {
    __shared__ T partial[kBlock / kWave];
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

// Grid-wide deterministic sum with last-block finalisation.  Every block
// passes its block total (thread 0); the block that arrives last sums the
// per-block partials in block order and returns true in all its threads with
// *total valid in thread 0.  The ticket is reset by the last block so the
// same slot can be reused by the next launch (also under graph replay).
template <class T>
__device__ bool grid_sum(T blockTotal, T* partials, unsigned* ticket, T* total) {
    __shared__ int isLast;
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = blockTotal;
        const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        isLast = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!isLast) return false;
    T s = 0;
    for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x)
        s += __hip_atomic_load(partials + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s = block_sum(s);
    if (threadIdx.x == 0) {
        *total = s;
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return true;
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

}  // namespace tf
