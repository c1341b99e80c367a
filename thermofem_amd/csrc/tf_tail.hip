// The coarse end of a V/W/F cycle as one program run by one 1024-thread
// workgroup.  Below about a thousand rows every per-level kernel is
// dominated by its launch and its chain of dependent memory round trips, not
// by its bytes; here the whole sub-cycle (sweeps, residuals, restrictions,
// prolongations, the dense coarse solve) runs between workgroup barriers in
// a single launch, with every vector and index array of those levels held in
// LDS and each level matrix read from a packed copy (each lane's row in
// contiguous groups of 4 entries), so an operation costs about one round of
// independent matrix loads.  Each operation repeats the arithmetic of the
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

// The level table lives in LDS, so the compiler cannot tell that the
// pointers in it address global memory and would issue flat loads, which
// also count against the LDS wait counter and so serialise the LDS gathers
// behind outstanding matrix loads.  Matrix pointers go through this cast.
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> global_of(const T* p) {
    return reinterpret_cast<gptr<T>>(reinterpret_cast<uintptr_t>(p));
}

// Adds p[0], ..., p[cnt - 1] to sum one at a time, in order.  Whole groups
// of 8 are read before they are added; the last partial group reads up to 7
// entries past p[cnt - 1] (the caller's buffer is padded for it) and keeps
// them out of the sum by selection.
template <class T>
__device__ __forceinline__ T add_in_order(T sum, const T* p, int cnt) {
#pragma clang fp contract(off)
    for (; cnt >= 8; cnt -= 8, p += 8) {
        T pv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) pv[u] = p[u];
#pragma unroll
        for (int u = 0; u < 8; ++u) sum = sum + pv[u];
    }
    if (cnt > 0) {
        T pv[7];
#pragma unroll
        for (int u = 0; u < 7; ++u) pv[u] = p[u];
#pragma unroll
        for (int u = 0; u < 7; ++u) {
            const T t = sum + pv[u];
            sum = u < cnt ? t : sum;
        }
    }
    return sum;
}

// Row sums of A x (x in LDS) for a tail level.  A pass loads P * 1024 of the
// level's column indices and values with coalesced, independent loads (one
// memory round trip), writes the rounded products to the LDS product area
// (cap entries; a pass that does not fit is split into chunks), and after one
// barrier thread t adds up rows t, t+1024, ... sequentially in stored order,
// carrying partial sums across chunks and passes (the per-level kernels'
// arithmetic).  f(row, sum) runs once per row.  Rows per thread are bounded
// by kTailRowSlots; prod has room for cap + 7 entries.
constexpr int kTailRowSlots = 4;  // rows per thread (levels up to 4096 rows)
constexpr int kTailPad = 8;       // product-area entries kept for add_in_order's over-read
template <class T, int P, class F>
__device__ __forceinline__ void tail_rows_p(int n, int e, const int* rp, gptr<int> ci, gptr<T> v, const T* x, T* prod,
                                            int cap, F&& f) {
#pragma clang fp contract(off)
    constexpr int PASS = P * kTailThreads;
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
    for (int p0 = 0; p0 < e; p0 += PASS) {
        const int p1 = min(e, p0 + PASS);
        int cc[P];
        T vv[P];
        // unconditional loads (indices clamped into the range): the compiler
        // then issues all of them before the first wait
#pragma unroll
        for (int u = 0; u < P; ++u) {
            const int j = min(p0 + tid + u * kTailThreads, e - 1);
            cc[u] = ci[j];
            vv[u] = v[j];
        }
        for (int c0 = p0; c0 < p1; c0 += cap) {
            const int c1 = min(p1, c0 + cap);
#pragma unroll
            for (int u = 0; u < P; ++u) {
                const int j = p0 + tid + u * kTailThreads;
                if (j >= c0 && j < c1) prod[j - c0] = vv[u] * x[cc[u]];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kTailRowSlots; ++k) {
                const int a = max(rs[k], c0), z = min(re[k], c1);
                if (a < z) acc[k] = add_in_order<T>(acc[k], prod + (a - c0), z - a);
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

// Entries loaded ahead per thread: 8 when that covers the level in one pass
// (small levels issue no surplus loads), else 16 (deeper spills registers).
template <class T, class F>
__device__ __forceinline__ void tail_rows(int n, const int* rp, const int* __restrict__ ci, const T* __restrict__ v,
                                          const T* x, T* prod, int cap, F&& f) {
    const int e = rp[n];
    if (e <= 8 * kTailThreads)
        tail_rows_p<T, 8>(n, e, rp, global_of(ci), global_of(v), x, prod, cap, f);
    else
        tail_rows_p<T, 16>(n, e, rp, global_of(ci), global_of(v), x, prod, cap, f);
}

// Four consecutive entries of a packed level (16-byte aligned).
template <class T>
struct Quad {
    T e[4];
};
typedef float F32x4 __attribute__((ext_vector_type(4)));
typedef double F64x2 __attribute__((ext_vector_type(2)));
typedef unsigned short U16x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ Quad<float> load_quad(gptr<float> p) {
    const F32x4 q = *reinterpret_cast<gptr<F32x4>>(p);
    return Quad<float>{{q.x, q.y, q.z, q.w}};
}
__device__ __forceinline__ Quad<double> load_quad(gptr<double> p) {
    const F64x2 a = *reinterpret_cast<gptr<F64x2>>(p);
    const F64x2 b = *reinterpret_cast<gptr<F64x2>>(p + 2);
    return Quad<double>{{a.x, a.y, b.x, b.y}};
}

// Row sums of a tail level from its packed copy: wave w takes slices w,
// w+16, ..., lane l row l of the slice; each group of 4 of the row's entries
// is one 16-byte (FP32) or two (FP64) value loads and one 8-byte column load,
// up to GB groups in flight before the first wait, and only as many groups
// as the slice is wide.  The rounded products are added in stored order;
// slots past the row's end are not added (the per-level kernels'
// arithmetic).  No barrier.  Slice offsets and row offsets come from LDS.
template <class T, class F>
__device__ __forceinline__ void tail_packed_rows(const TailLevel& L, char* lds, const T* x, F&& f) {
#pragma clang fp contract(off)
    constexpr int GB = sizeof(T) == 4 ? 6 : 4;
    const int lane = int(threadIdx.x) & (kWave - 1);
    const int* base = lds_at<int>(lds, L.lsbase);
    const int* rp = lds_at<int>(lds, L.lrp);
    const gptr<T> pv = global_of(static_cast<const T*>(L.pv));
    const gptr<unsigned short> pcol = global_of(L.pcol);
    for (int sl = int(threadIdx.x) / kWave; sl < L.slices; sl += kTailThreads / kWave) {
        const int pos = sl * kWave + lane;
        const int len = pos < L.n ? rp[pos + 1] - rp[pos] : 0;
        const int b0 = __builtin_amdgcn_readfirstlane(base[sl]);
        const int groups = (__builtin_amdgcn_readfirstlane(base[sl + 1]) - b0) / (4 * kWave);
        const gptr<T> vp = pv + b0 + 4 * lane;
        const gptr<unsigned short> cp = pcol + b0 + 4 * lane;
        T acc = 0;
        for (int g0 = 0; g0 < groups; g0 += GB) {
            Quad<T> q[GB];
            U16x4 c[GB];
#pragma unroll
            for (int g = 0; g < GB; ++g) {
                if (g0 + g < groups) {
                    q[g] = load_quad(vp + (g0 + g) * 4 * kWave);
                    c[g] = *reinterpret_cast<gptr<U16x4>>(cp + (g0 + g) * 4 * kWave);
                }
            }
#pragma unroll
            for (int g = 0; g < GB; ++g) {
                if (g0 + g < groups) {
                    const int cols[4] = {int(c[g].x), int(c[g].y), int(c[g].z), int(c[g].w)};
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const T p = q[g].e[u] * x[cols[u]];
                        const T t = acc + p;
                        acc = 4 * (g0 + g) + u < len ? t : acc;
                    }
                }
            }
        }
        if (pos < L.n) f(pos, acc);
    }
}

// Row sums of a tail level, from its packed copy when it has one.
template <class T, class F>
__device__ __forceinline__ void level_rows(const TailLevel& L, char* lds, const T* x, T* prod, int cap, F&& f) {
    if (L.slices > 0)
        tail_packed_rows<T>(L, lds, x, f);
    else
        tail_rows<T>(L.n, lds_at<int>(lds, L.lrp), L.ci, static_cast<const T*>(L.v), x, prod, cap, f);
}

template <class T>
__global__ __launch_bounds__(kTailThreads) void k_tail(const TailLevel* __restrict__ levels, int top, int count,
                                                       const void* __restrict__ image, int imageBytes,
                                                       int ldsBytes, const TailOp* __restrict__ ops, int nops,
                                                       const double* __restrict__ inverse, T w) {
#pragma clang fp contract(off)
    __shared__ __align__(16) char lds[kTailLdsAll];
    __shared__ double bs[kCoarseMax];
    __shared__ TailLevel lv[kMaxTailLevels];
    __shared__ TailOp prog[kMaxTailOps];
    const int tid = int(threadIdx.x);
    // products of the CSR row sums go after the levels' arrays
    T* prod = lds_at<T>(lds, ldsBytes);
    const int cap = (kTailLdsAll - ldsBytes) / int(sizeof(T)) - kTailPad;
    // The program and the level table go to LDS first: no operation then
    // waits on a global load of its own description.
    for (int i = tid; i < nops; i += kTailThreads) prog[i] = ops[i];
    for (int i = tid; i < count; i += kTailThreads) lv[i] = levels[top + i];
    __syncthreads();
    // Stage the LDS image (every level's inverse diagonal, row offsets,
    // transfer and packed slices' offsets, laid out as in LDS; the vectors
    // follow it) with 16-byte loads issued 4 at a time, and the entry level's
    // rhs and iterate.  The stores are unconditional at clamped indices (a
    // clamped slot rewrites a word with its own value), so no load waits
    // behind a branch.
    {
        const TailLevel& L = lv[0];
        const gptr<T> gb = global_of(static_cast<const T*>(L.b));
        const gptr<T> gx = global_of(static_cast<const T*>(L.x));
        T bv[kTailRowSlots], xv[kTailRowSlots];
#pragma unroll
        for (int k = 0; k < kTailRowSlots; ++k) {
            const int i = min(tid + k * kTailThreads, L.n - 1);
            bv[k] = gb[i];
            xv[k] = gx[i];
        }
        const int4* src = static_cast<const int4*>(image);
        int4* dst = reinterpret_cast<int4*>(lds);
        const int words = imageBytes / 16;
        for (int i0 = 0; i0 < words; i0 += 4 * kTailThreads) {
            int4 t[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) t[u] = src[min(i0 + tid + u * kTailThreads, words - 1)];
#pragma unroll
            for (int u = 0; u < 4; ++u) dst[min(i0 + tid + u * kTailThreads, words - 1)] = t[u];
        }
        T* lb = lds_at<T>(lds, L.lb);
        T* lx = lds_at<T>(lds, L.lx0);
#pragma unroll
        for (int k = 0; k < kTailRowSlots; ++k) {
            const int i = min(tid + k * kTailThreads, L.n - 1);
            lb[i] = bv[k];
            lx[i] = xv[k];
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
            level_rows<T>(L, lds, xin, prod, cap, [&](int i, T acc) {
                const T wd = w * inv[i];
                const T d = b[i] - acc;
                const T upd = wd * d;
                xout[i] = xin[i] + upd;
            });
            break;
        case kTResidual: {
            T* r = lds_at<T>(lds, L.lr);
            level_rows<T>(L, lds, xin, prod, cap, [&](int i, T acc) { r[i] = b[i] - acc; });
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
            coarse_inverse_apply<T>(n, inverse, b, xout, bs, reinterpret_cast<double*>(prod));
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

template <class T>
__global__ __launch_bounds__(kBlock) void k_tail_pack(int n, int slices, const int* __restrict__ rp,
                                                      const int* __restrict__ ci, const T* __restrict__ v,
                                                      const int* __restrict__ base, T* __restrict__ pv,
                                                      unsigned short* __restrict__ pcol) {
    const int pos = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (pos >= slices * kWave) return;
    const int sl = pos / kWave, lane = pos % kWave;
    const int width = (base[sl + 1] - base[sl]) / kWave;
    const int s = pos < n ? rp[pos] : 0;
    const int len = pos < n ? rp[pos + 1] - s : 0;
    for (int k = 0; k < width; ++k) {
        const int o = base[sl] + (k / 4) * 4 * kWave + 4 * lane + k % 4;
        pv[o] = k < len ? v[s + k] : T(0);
        pcol[o] = static_cast<unsigned short>(k < len ? ci[s + k] : 0);
    }
}

}  // namespace

template <class L>
void tail_pack(tf_ctx& c, int n, int slices, const int* rp, const int* ci, const L* v, const int* base, L* pv,
               unsigned short* pcol) {
    if (slices <= 0) return;
    hipLaunchKernelGGL(k_tail_pack<L>, dim3((slices * kWave + kBlock - 1) / kBlock), dim3(kBlock), 0, c.stream, n,
                       slices, rp, ci, v, base, pv, pcol);
    check(hipGetLastError(), "tail pack launch");
}

template void tail_pack<float>(tf_ctx&, int, int, const int*, const int*, const float*, const int*, float*,
                               unsigned short*);
template void tail_pack<double>(tf_ctx&, int, int, const int*, const int*, const double*, const int*, double*,
                                unsigned short*);

template <class L>
void tail_run(tf_ctx& c, const TailLevel* levels, int top, int count, const void* image, int imageBytes,
              int ldsBytes, const TailOp* ops, int nops, const double* inverse, L w) {
    hipLaunchKernelGGL(k_tail<L>, dim3(1), dim3(kTailThreads), 0, c.stream, levels, top, count, image, imageBytes,
                       ldsBytes, ops, nops, inverse, w);
    check(hipGetLastError(), "tail launch");
}

template void tail_run<float>(tf_ctx&, const TailLevel*, int, int, const void*, int, int, const TailOp*, int,
                              const double*, float);
template void tail_run<double>(tf_ctx&, const TailLevel*, int, int, const void*, int, int, const TailOp*, int,
                               const double*, double);

}  // namespace tf
