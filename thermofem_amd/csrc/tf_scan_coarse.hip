// Device-wide exclusive scan (reduce / scan-of-totals / apply, 2048 items
// per 256-thread block) and the per-cycle coarse solves: the FP64 dense
// product with the explicit inverse and the pivoted triangular solve (the
// factorisation and the inverse are built in tf_lu.hip).
#include <mutex>

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

// Up to this many tiles, each tile adds the totals of the tiles before it
// itself (at most 16 KB of reads a block), which saves the launch that would
// scan the totals.
constexpr int kDirectPrefixTiles = 4096;

__global__ __launch_bounds__(kBlock) void k_scan_apply_direct(int n, const int* in, int* out,
                                                              const int* __restrict__ totals) {
    __shared__ int sBase;
    int before = 0;
    for (int t = int(threadIdx.x); t < int(blockIdx.x); t += kBlock) before += totals[t];
    before = block_sum(before);
    if (threadIdx.x == 0) sBase = before;
    __syncthreads();
    const int base = int(blockIdx.x) * kScanTile + int(threadIdx.x) * kScanItems;
    int items[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        items[k] = base + k < n ? in[base + k] : 0;
        s += items[k];
    }
    int total;
    int run = block_excl_scan(s, &total) + sBase;
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
    if (nb <= kDirectPrefixTiles) {
        hipLaunchKernelGGL(k_scan_apply_direct, dim3(nb), dim3(kBlock), 0, c.stream, n, in, out, aux);
        check(hipGetLastError(), "scan apply launch");
        return;
    }
    scan_rec(c, nb, aux, aux, scratch + nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kBlock), 0, c.stream, n, in, out, aux);
    check(hipGetLastError(), "scan apply launch");
}

// ---------------- coarse dense LU ----------------

constexpr int kLuThreads = 1024;

// x = A^-1 b in FP64 from the column-major inverse; b is staged in LDS first
// so rhs and x may alias.
template <class T>
__global__ __launch_bounds__(kLuThreads) void k_inv_apply(int n, const double* __restrict__ inv, const T* b, T* x) {
    __shared__ double bs[kCoarseMax];
    __shared__ double part[kLuThreads];
    coarse_inverse_apply<T>(n, inv, b, x, bs, part);
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

// Block totals every level of scan_rec stores for n items.
size_t scan_scratch_ints(int n) {
    size_t need = 0;
    for (int m = n; m > kScanTile;) {
        m = (m + kScanTile - 1) / kScanTile;
        need += size_t(m);
    }
    return need;
}

void scan_excl(tf_ctx& c, int n, const int* in, int* out) {
    if (n <= 0) return;
    int* scratch = c.scan_scratch(scan_scratch_ints(n) + 16);
    scan_rec(c, n, in, out, scratch);
}

// ---- the scan without allocation ----
// Its block totals live in a fixed device array of kScanNoAllocInts ints
// (516 KB, held on every GPU the library runs on until the process ends),
// enough for n <= kScanNoAllocMaxN.  The array's host-side name is not a
// device address: each device's copy is looked up once, with that device
// current, and only that address is handed to kernels.  The array keeps the
// last call's totals; scan_rec writes every total before reading it, so
// nothing needs clearing.  Calls that overlap on one GPU (two streams) share
// the array and race.
constexpr int kScanNoAllocInts = 132096;
constexpr int kScanNoAllocMaxN = 1 << 28;
static __device__ __align__(256) int scratch[kScanNoAllocInts];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
int* g_scratch_ptr[kMaxDevices] = {};

int* device_scratch(int device) {
    require(device >= 0 && device < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        int* p = nullptr;
        check(hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch)), "scan scratch address");
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}
}  // namespace

void scan_excl_noalloc(tf_ctx& c, int n, const int* in, int* out) {
    require(n <= kScanNoAllocMaxN && scan_scratch_ints(n) <= size_t(kScanNoAllocInts),
            "tf_scan_excl_i32_noalloc: n exceeds 268435456, the fixed scratch's limit");
    if (n <= 0) return;
    scan_rec(c, n, in, out, device_scratch(c.device));
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
