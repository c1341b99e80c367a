// The coarse end of a V/W/F cycle as one program run by one 1024-thread
// workgroup.  Below about a thousand rows every per-level kernel is
// dominated by its launch and its chain of dependent memory round trips, not
// by its bytes; here the whole sub-cycle (sweeps, residuals, restrictions,
// prolongations, the dense coarse solve) runs between workgroup barriers in
// a single launch, with every vector and index array of those levels held in
// LDS, so an operation costs one round of independent matrix loads.  Each operation repeats the arithmetic of the
// corresponding per-level kernel exactly: sequential row sums of rounded
// products, the same epilogue order, restriction over members in ascending
// order and the same dense inverse apply, so results do not depend on where
// the tail starts.  The program (operation list) is generated on the host
// from the same cycle recursion the per-level path follows.
#include "tf_device.hpp"
#include "tf_engine.hpp"

namespace tf {
namespace {

constexpr int kTailThreads = 1024;

template <class T>
__device__ __forceinline__ T* lds_at(char* lds, int off) {
    return reinterpret_cast<T*>(lds + off);
}

// Row sums of A x (x in LDS) for a tail level.  All of the level's column
// indices and values (up to kTailPrefetch * 1024 entries per pass) are
// loaded first with coalesced, independent loads -- one memory round trip --
// then, chunk by chunk, the rounded products go to LDS and thread t adds up
// rows t, t+1024, ... sequentially in stored order, carrying partial sums
// across chunks (the per-level kernels' arithmetic).  f(row, sum) runs once
// per row.  Rows per thread are bounded by kTailRowSlots.
template <class T>
struct TailPrefetch {
    static constexpr int value = sizeof(T) == 4 ? 24 : 16;  // entries per thread loaded ahead
};
constexpr int kTailChunk = 4;     // of which this many per LDS chunk
constexpr int kTailRowSlots = 4;  // rows per thread (levels up to 4096 rows)
template <class T, class F>
__device__ __forceinline__ void tail_rows(int n, const int* rp, const int* __restrict__ ci, const T* __restrict__ v,
                                          const T* x, T* prod, F&& f) {
    constexpr int kTailPrefetch = TailPrefetch<T>::value;
    constexpr int PASS = kTailPrefetch * kTailThreads;
    constexpr int CHUNK = kTailChunk * kTailThreads;
    const int tid = int(threadIdx.x);
    T acc[kTailRowSlots];
    int rs[kTailRowSlots], re[kTailRowSlots];
#pragma unroll
    for (int k = 0; k < kTailRowSlots; ++k) {
        const int i = tid + k * kTailThreads;
        rs[k] = i < n ? rp[i] : 0;
        re[k] = i < n ? rp[i + 1] : 0;
        acc[k] = 0;
    }
    const int e = rp[n];
    for (int p0 = 0; p0 < e; p0 += PASS) {
        int cc[kTailPrefetch];
        T vv[kTailPrefetch];
        // unconditional loads (indices clamped into the range): the compiler
        // then issues all of them before the first wait
#pragma unroll
        for (int u = 0; u < kTailPrefetch; ++u) {
            const int j = min(p0 + tid + u * kTailThreads, e - 1);
            cc[u] = ci[j];
            vv[u] = v[j];
        }
#pragma unroll
        for (int c = 0; c < kTailPrefetch / kTailChunk; ++c) {
            const int c0 = p0 + c * CHUNK;
            if (c0 >= e) break;
            const int c1 = min(e, c0 + CHUNK);
#pragma unroll
            for (int u = 0; u < kTailChunk; ++u) {
                const int q = c * kTailChunk + u;
                const int j = p0 + tid + q * kTailThreads;
                if (j < c1) prod[j - c0] = vv[q] * x[cc[q]];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kTailRowSlots; ++k) {
                // the row's products in this chunk, read 16 at a time (all
                // reads issued before the adds), added in order
                const int a = max(rs[k], c0), z = min(re[k], c1);
                T sum = acc[k];
                for (int q0 = a; q0 < z; q0 += 16) {
                    T pv[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u) pv[u] = prod[min(q0 + u, z - 1) - c0];
#pragma unroll
                    for (int u = 0; u < 16; ++u) {
                        const T t = sum + pv[u];
                        sum = q0 + u < z ? t : sum;
                    }
                }
                acc[k] = sum;
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int k = 0; k < kTailRowSlots; ++k) {
        const int i = tid + k * kTailThreads;
        if (i < n) f(i, acc[k]);
    }
}

template <class T>
__global__ __launch_bounds__(kTailThreads) void k_tail(const TailLevel* __restrict__ levels, int top, int count,
                                                       const void* __restrict__ image, int imageBytes,
                                                       const TailOp* __restrict__ ops, int nops,
                                                       const double* __restrict__ inverse, T w) {
#pragma clang fp contract(off)
    __shared__ __align__(16) char lds[kTailLds];
    __shared__ double bs[kCoarseMax];
    __shared__ T prod[kTailChunk * kTailThreads];
    __shared__ TailLevel lv[kMaxTailLevels];
    __shared__ TailOp prog[kMaxTailOps];
    const int tid = int(threadIdx.x);
    // The program and the level table go to LDS first: no operation then
    // waits on a global load of its own description.
    for (int i = tid; i < nops; i += kTailThreads) prog[i] = ops[i];
    for (int i = tid; i < count; i += kTailThreads) lv[i] = levels[top + i];
    __syncthreads();
    // Stage the LDS image (every level's inverse diagonal, row offsets and
    // transfer, laid out as in LDS) with 16-byte loads issued 4 at a time,
    // then the entry level's rhs and iterate.
    {
        const int4* src = static_cast<const int4*>(image);
        int4* dst = reinterpret_cast<int4*>(lds);
        const int words = imageBytes / 16;
        for (int i0 = 0; i0 < words; i0 += 4 * kTailThreads) {
            int4 t[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) t[u] = src[min(i0 + tid + u * kTailThreads, words - 1)];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = i0 + tid + u * kTailThreads;
                if (i < words) dst[i] = t[u];
            }
        }
        const TailLevel& L = lv[0];
        T* lb = lds_at<T>(lds, L.lb);
        T* lx = lds_at<T>(lds, L.lx0);
        const T* gb = static_cast<const T*>(L.b);
        const T* gx = static_cast<const T*>(L.x);
        for (int i = tid; i < L.n; i += kTailThreads) {
            const T bv = gb[i], xv = gx[i];
            lb[i] = bv;
            lx[i] = xv;
        }
    }
    __syncthreads();
    for (int k = 0; k < nops; ++k) {
        const TailOp op = prog[k];
        const TailLevel& L = lv[op.level - top];
        const int n = L.n;
        T* b = lds_at<T>(lds, L.lb);
        const T* inv = lds_at<T>(lds, L.linv);
        T* xin = lds_at<T>(lds, op.in ? L.lx1 : L.lx0);
        T* xout = lds_at<T>(lds, op.out ? L.lx1 : L.lx0);
        switch (op.kind) {
        case kTSweep:
            tail_rows<T>(n, lds_at<int>(lds, L.lrp), L.ci, static_cast<const T*>(L.v), xin, prod, [&](int i, T acc) {
                const T wd = w * inv[i];
                const T d = b[i] - acc;
                const T upd = wd * d;
                xout[i] = xin[i] + upd;
            });
            break;
        case kTResidual: {
            T* r = lds_at<T>(lds, L.lr);
            tail_rows<T>(n, lds_at<int>(lds, L.lrp), L.ci, static_cast<const T*>(L.v), xin, prod,
                         [&](int i, T acc) { r[i] = b[i] - acc; });
            break;
        }
        case kTSweepZero:
            for (int i = tid; i < n; i += kTailThreads) {
                const T wd = w * inv[i];
                xout[i] = wd * b[i];
            }
            break;
        case kTRestrict: {
            const T* src = op.in ? lds_at<T>(lds, L.lr) : b;
            const TailLevel& N = lv[op.level + 1 - top];
            T* nb = lds_at<T>(lds, N.lb);
            const T* ninv = lds_at<T>(lds, N.linv);
            T* nx = lds_at<T>(lds, op.out ? N.lx1 : N.lx0);
            const int* off = lds_at<int>(lds, L.loff);
            const int* mem = lds_at<int>(lds, L.lmem);
            for (int c = tid; c < L.nc; c += kTailThreads) {
                T s = 0;
                for (int m = off[c]; m < off[c + 1]; ++m) s += src[mem[m]];
                nb[c] = s;
                if (op.flag) {
                    const T wd = w * ninv[c];
                    nx[c] = wd * s;
                }
            }
            break;
        }
        case kTProlong: {
            const TailLevel& N = lv[op.level + 1 - top];
            const T* xc = lds_at<T>(lds, op.in ? N.lx1 : N.lx0);
            const int* map = lds_at<int>(lds, L.lmap);
            for (int i = tid; i < n; i += kTailThreads) xout[i] = op.flag ? xc[map[i]] : xout[i] + xc[map[i]];
            break;
        }
        case kTCoarse:
            coarse_inverse_apply<T>(n, inverse, b, xout, bs);
            break;
        case kTCopy:
            for (int i = tid; i < n; i += kTailThreads) xout[i] = xin[i];
            break;
        default:
            break;
        }
        __syncthreads();
    }
    const TailLevel& L = lv[0];
    const T* lx = lds_at<T>(lds, L.lx0);
    T* gx = static_cast<T*>(L.x);
    for (int i = tid; i < L.n; i += kTailThreads) gx[i] = lx[i];
}

}  // namespace

template <class L>
void tail_run(tf_ctx& c, const TailLevel* levels, int top, int count, const void* image, int imageBytes,
              const TailOp* ops, int nops, const double* inverse, L w) {
    hipLaunchKernelGGL(k_tail<L>, dim3(1), dim3(kTailThreads), 0, c.stream, levels, top, count, image, imageBytes,
                       ops, nops, inverse, w);
    check(hipGetLastError(), "tail launch");
}

template void tail_run<float>(tf_ctx&, const TailLevel*, int, int, const void*, int, const TailOp*, int,
                              const double*, float);
template void tail_run<double>(tf_ctx&, const TailLevel*, int, int, const void*, int, const TailOp*, int,
                               const double*, double);

}  // namespace tf
