// AMG hierarchy + cycle (Amg.cuh:81-161, Cycle.cuh:35-63) and the CG /
// BiCGStab iterations (Cg.cuh:123-189, BiCgStab.cuh:144-247) on gfx950.
//
// The Krylov scalars (rho, alpha, beta, omega, norms) never leave the
// device: each reduction kernel's last block evaluates the reference's
// scalar recurrences and breakdown/non-finite tests on a control block, and
// later kernels of the iteration read it.  One iteration is therefore a
// fixed launch sequence that is captured once as a hipGraph and replayed;
// the host reads the control block once per iteration to stop.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "tf_device.hpp"
#include "tf_engine.hpp"
#include "tf_krylov.hpp"

namespace tf {

namespace {

constexpr int kSlotA = 40, kSlotB = 41, kSlotC = 42, kSlotD = 43, kSlotE = 44;
// Levels up to this many rows compute residual and restriction in one
// aggregate-parallel kernel (four lanes per aggregate).  It saves a launch
// but measured slower than the streaming residual kernel + restriction on
// the latency-bound levels of the 128^3 hierarchy (20.58 -> 20.70 ms per FP32
// solve with it on levels up to 16,384 rows, 20.80 ms up to 65,536): a
// member row per lane is a long dependent chain.  Off by default.
constexpr int kFusedResidualRows = 0;
// Levels up to this many rows fuse prolongation into the first post-sweep
// (saves a launch); on larger levels the extra map gather per nonzero costs
// more than the separate prolongation pass.
constexpr int kFusedProlongRows = 65536;
// Operators are also stored in sliced ELL when that costs at most this much
// padding (relative to nnz) and the level is not the LU-solved coarsest one:
// one lane per row, 16-bit column offsets, no LDS staging (tf_sell.hip).
constexpr double kSellMaxPadding = 0.3;
// the packed layout (10+ entries a row) in row order: padding costs memory only
constexpr double kSellMaxPackedPadding = 0.6;
constexpr int kSellMinRows = 256;
constexpr int kStencilMinRows = 4096;  // below this the probe and fill cost more than they save
// Levels with at least this many nonzeros are tuned: sliced ELL and the CSR
// kernel are both timed on the level's own operator and the faster one is
// kept (sliced ELL wins on stencil-like levels, not always on coarsened ones).
constexpr int64_t kTuneMinNnz = 200000;

// Process-wide cache of tuning decisions, keyed by the operator's shape and
// the element type, so repeated setups of one problem time each level once.
struct TuneKey {
    int op;  // 0: Jacobi sweep of an AMG level, 1: the Krylov operator's product
    int n;
    int64_t nnz, stored;
    int bytes;
    bool operator<(const TuneKey& o) const {
        return std::tie(op, n, nnz, stored, bytes) < std::tie(o.op, o.n, o.nnz, o.stored, o.bytes);
    }
};
std::map<TuneKey, int> g_tuned;  // 0: CSR kernels; 1, 2, 4: sliced ELL with that many lanes per row
std::map<TuneKey, int> g_tail_start;  // chosen first level of the coarse tail (-1: none)
std::mutex g_tune_mutex;

// Times the two candidates alternately (one untimed launch each first), each
// sample a run of `reps` back-to-back launches (a single ~10 us launch is
// within the events' own noise), and compares their fastest samples, which
// is robust to the odd slow run; true unless the second is faster by more
// than the margin (a fraction of the first's time).
template <class A, class B>
bool first_is_faster(tf_ctx& c, A&& first, B&& second, double margin = 0.0, int samples = 5, int reps = 4) {
    hipEvent_t e[4];
    for (auto& ev : e) check(hipEventCreate(&ev), "hipEventCreate");
    first();
    second();
    float best[2] = {1e30f, 1e30f};
    for (int i = 0; i < samples; ++i) {
        check(hipEventRecord(e[0], c.stream), "event record");
        for (int r = 0; r < reps; ++r) first();
        check(hipEventRecord(e[1], c.stream), "event record");
        for (int r = 0; r < reps; ++r) second();
        check(hipEventRecord(e[2], c.stream), "event record");
        check(hipEventSynchronize(e[2]), "event sync");
        float t0 = 0, t1 = 0;
        check(hipEventElapsedTime(&t0, e[0], e[1]), "event time");
        check(hipEventElapsedTime(&t1, e[1], e[2]), "event time");
        best[0] = std::min(best[0], t0);
        best[1] = std::min(best[1], t1);
    }
    for (auto& ev : e) (void)hipEventDestroy(ev);
    return best[1] >= best[0] * float(1.0 - margin);
}

// CSR replaces a sliced-ELL copy only when it is this much faster in
// isolation: timed back to back the level's data stays in the caches, while
// in the cycle the sliced-ELL sweep moves fewer bytes (measured at 128^3:
// a level both timed as equals swept in 8.8 us as sliced ELL, 10.9 us as CSR).
constexpr double kSellPreference = 0.05;

// Least time of f over `samples` runs (after one untimed run), in ms.
template <class F>
float best_time(tf_ctx& c, F&& f, int samples = 3) {
    hipEvent_t e[2];
    for (auto& ev : e) check(hipEventCreate(&ev), "hipEventCreate");
    f();
    float best = 1e30f;
    for (int i = 0; i < samples; ++i) {
        check(hipEventRecord(e[0], c.stream), "event record");
        f();
        check(hipEventRecord(e[1], c.stream), "event record");
        check(hipEventSynchronize(e[1]), "event sync");
        float t = 0;
        check(hipEventElapsedTime(&t, e[0], e[1]), "event time");
        best = std::min(best, t);
    }
    for (auto& ev : e) (void)hipEventDestroy(ev);
    return best;
}

// Keeps the sliced-ELL copy of a level only where it beats the CSR sweep,
// with the lanes per row (1, 2 or 4) that sweep it fastest; the three give
// identical results.
template <class L>
bool sell_pays(tf_ctx& c, AmgLevel& l, double w) {
    const DevCsr& A = l.native;
    const TuneKey key{0, A.n, A.nnz, l.sell.stored, int(sizeof(L))};
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            l.sell.lanes = std::max(1, it->second);
            return it->second != 0;
        }
    }
    fill_zero(c, l.b, size_t(A.n) * sizeof(L));
    fill_zero(c, l.x, size_t(A.n) * sizeof(L));
    const L* vals = static_cast<const L*>(l.vals);
    const L* svals = static_cast<const L*>(l.svals);
    const L* inv = static_cast<const L*>(l.inv);
    const L* b = static_cast<const L*>(l.b);
    const L* x = static_cast<const L*>(l.x);
    L* x2 = static_cast<L*>(l.x2);
    auto sweep = [&] { sell_jacobi<L>(c, l.sell, svals, inv, L(w), b, x, x2); };
    int lanes = 1;
    float best = 1e30f;
    for (int r : {1, 2, 4}) {
        l.sell.lanes = r;
        const float t = best_time(c, [&] {
            for (int q = 0; q < 4; ++q) sweep();
        });
        if (t < best) {
            best = t;
            lanes = r;
        }
    }
    l.sell.lanes = lanes;
    const bool sell = A.nnz < kTuneMinNnz ||
                      first_is_faster(c, sweep, [&] { jacobi_fused<L>(c, A.n, A.nnz, A.rp, A.ci, vals, inv, L(w), b, x, x2); },
                                      kSellPreference);
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tuned[key] = sell ? lanes : 0;
    return sell;
}


// Loop over LDS-staged row blocks for the fused reduction kernels (XCD
// contiguous; with a grid below the block count each XCD walks its slab).
// f(row, sum) runs in the thread owning the row.
template <class XF, class F>
__device__ __forceinline__ void for_stream_blocks(int n, int nblocks, int rpb, const int* __restrict__ rp,
                                                  const int* __restrict__ ci, const double* __restrict__ v,
                                                  XF&& xval, F&& f) {
    __shared__ double prod[StreamCap<double>::value];
    __shared__ int srp[kMaxStreamRows + 1];
    const int per = int(gridDim.x) / kXcds;
    const int xcd = int(blockIdx.x % kXcds), idx = int(blockIdx.x / kXcds);
    const int chunk = (nblocks + kXcds - 1) / kXcds;
    const int end = min(nblocks, (xcd + 1) * chunk);
    for (int lb = xcd * chunk + idx; lb < end; lb += per) {
        const int row0 = lb * rpb;
        const int rows = min(rpb, n - row0);
        double acc[kMaxStreamRows / kBlock];
        csr_stream_rows<double>(row0, rows, rp, ci, v, xval, prod, srp, acc);
#pragma unroll
        for (int k = 0; k < kMaxStreamRows / kBlock; ++k) {
            const int r = int(threadIdx.x) + k * kBlock;
            if (r < rows) f(row0 + r, acc[k]);
        }
    }
}

// The same over a sliced-ELL operator: one lane per row, slices of 64 rows,
// 4 slices per block, blocks XCD-contiguous.
template <class C, class XF, class F>
__device__ __forceinline__ void for_sell_rows(const SellView<C>& m, const double* __restrict__ v, XF&& xval, F&& f) {
    constexpr int spb = kBlock / kWave;
    const int nblocks = (m.slices + spb - 1) / spb;
    const int per = int(gridDim.x) / kXcds;
    const int xcd = int(blockIdx.x % kXcds), idx = int(blockIdx.x / kXcds);
    const int chunk = (nblocks + kXcds - 1) / kXcds;
    const int end = min(nblocks, (xcd + 1) * chunk);
    const int lane = int(threadIdx.x) & (kWave - 1);
    for (int lb = xcd * chunk + idx; lb < end; lb += per) {
        const int slice = lb * spb + int(threadIdx.x) / kWave;
        const int pos = slice * kWave + lane;
        if (pos < m.n) {
            const int row = m.perm ? m.perm[pos] : pos;
            const int b = m.base[slice];
            f(row, m.packed ? sell_row_sum4<double>(row, int(m.len[pos]), v + b + 4 * lane, m.col + b + 4 * lane, xval)
                            : sell_row_sum<double>(row, int(m.len[pos]), v + b + lane, m.col + b + lane, xval));
        }
    }
}

// The same over a structured-stencil operator: one lane per row, blocks of
// 256 consecutive rows, XCD-contiguous.
template <int NC, class XF, class F>
__device__ __forceinline__ void for_stencil_rows(const StencilView& m, const double* __restrict__ v, XF&& xval,
                                                 F&& f) {
    const int nblocks = (m.n + kBlock - 1) / kBlock;
    const int per = int(gridDim.x) / kXcds;
    const int xcd = int(blockIdx.x % kXcds), idx = int(blockIdx.x / kXcds);
    const int chunk = (nblocks + kXcds - 1) / kXcds;
    const int end = min(nblocks, (xcd + 1) * chunk);
    for (int lb = xcd * chunk + idx; lb < end; lb += per) {
        const int row = lb * kBlock + int(threadIdx.x);
        if (row < m.n) f(row, stencil_row_sum<double, NC == 0 ? 1 : NC>(m, v, row, xval));
    }
}
// The same with x an array: the mirrored 7-point stencil (NC == 0) takes its
// -1 / +1 neighbours from the wave's lanes.
template <int NC, class F>
__device__ __forceinline__ void for_stencil_rows_x(const StencilView& m, const double* __restrict__ v,
                                                   const double* x, F&& f) {
    if constexpr (NC != 0) {
        for_stencil_rows<NC>(m, v, [&](int c) { return x[c]; }, f);
    }
    else {
        const int nblocks = (m.n + kBlock - 1) / kBlock;
        const int per = int(gridDim.x) / kXcds;
        const int xcd = int(blockIdx.x % kXcds), idx = int(blockIdx.x / kXcds);
        const int chunk = (nblocks + kXcds - 1) / kXcds;
        const int end = min(nblocks, (xcd + 1) * chunk);
        for (int lb = xcd * chunk + idx; lb < end; lb += per) {
            const int row = lb * kBlock + int(threadIdx.x);
            const double acc = stencil7_row_sum<double>(m, v, x, row);
            if (row < m.n) f(row, acc);
        }
    }
}

// Row sources of the fused kernels below.
struct CsrRows {
    int n, nblocks, rpb;
    const int* rp;
    const int* ci;
    const double* v;
    template <class F>
    __device__ __forceinline__ void each(const double* x, F&& f) const {
        for_stream_blocks(n, nblocks, rpb, rp, ci, v, [&](int c) { return x[c]; }, f);
    }
    // the same with x[c] given by xval(c)
    template <class XF, class F>
    __device__ __forceinline__ void each_fn(XF&& xval, F&& f) const {
        for_stream_blocks(n, nblocks, rpb, rp, ci, v, xval, f);
    }
};
template <class C>
struct SellRows {
    SellView<C> m;
    const double* v;
    template <class F>
    __device__ __forceinline__ void each(const double* x, F&& f) const {
        for_sell_rows(m, v, [&](int c) { return x[c]; }, f);
    }
    template <class XF, class F>
    __device__ __forceinline__ void each_fn(XF&& xval, F&& f) const {
        for_sell_rows(m, v, xval, f);
    }
};

template <int NC>
struct StencilRows {
    StencilView m;
    const double* v;
    template <class F>
    __device__ __forceinline__ void each(const double* x, F&& f) const {
        for_stencil_rows_x<NC>(m, v, x, f);
    }
    template <class XF, class F>
    __device__ __forceinline__ void each_fn(XF&& xval, F&& f) const {
        for_stencil_rows<NC>(m, v, xval, f);
    }
};

// Block partial of a reduction whose finaliser runs in k_kfinal.
__device__ __forceinline__ void block_partial(double contrib, double* partials) {
    contrib = block_sum(contrib);
    if (threadIdx.x == 0) partials[blockIdx.x] = contrib;
}

// dot(x, y) -> finalizer.
__global__ __launch_bounds__(kBlock) void k_kdot(int n, const double* __restrict__ x, const double* __restrict__ y,
                                                 double* partials, const KCtl* ctl) {
    if (stopped(ctl)) return;
    double s = 0;
#pragma unroll 4
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock) s += x[i] * y[i];
    block_partial(s, partials);
}

// y = A x fused with sum(w .* y) -> finalizer.
template <class R>
__global__ __launch_bounds__(kBlock) void k_kspmv_dot(R rows, const double* __restrict__ x, double* __restrict__ y,
                                                      const double* __restrict__ w, double* partials,
                                                      const KCtl* ctl) {
    if (stopped(ctl)) return;
    double contrib = 0;
    rows.each(x, [&](int row, double acc) {
        y[row] = acc;
        contrib += w[row] * acc;
    });
    block_partial(contrib, partials);
}

// t = A x fused with (t.t, t.s): BiCgStab.cuh:214-226 (omega).
template <class R>
__global__ __launch_bounds__(kBlock) void k_kspmv_omega(R rows, const double* __restrict__ x, double* __restrict__ t,
                                                        const double* __restrict__ s, double* partials,
                                                        const KCtl* ctl) {
    if (stopped(ctl) || ctl->early) return;
    double tt = 0, ts = 0;
    rows.each(x, [&](int row, double acc) {
        t[row] = acc;
        tt += acc * acc;
        ts += acc * s[row];
    });
    tt = block_sum(tt);
    ts = block_sum(ts);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = tt;
        partials[kMaxReduceBlocks + blockIdx.x] = ts;
    }
}

// r = b - A x and sum(r.^2) when the control block asks for it; with `shadow`
// the new residual is also copied into the shadow vector (BiCGStab restart).
template <class R>
__global__ __launch_bounds__(kBlock) void k_ktrue(R rows, const double* __restrict__ b, const double* __restrict__ x,
                                                  double* __restrict__ r, double* __restrict__ shadow,
                                                  double* partials, const KCtl* ctl, int onEarly) {
    if (stopped(ctl)) return;
    if (onEarly ? !ctl->early : (!ctl->check || ctl->early)) return;
    double contrib = 0;
    rows.each(x, [&](int row, double acc) {
        const double res = b[row] - acc;
        r[row] = res;
        if (shadow) shadow[row] = res;
        contrib += res * res;
    });
    block_partial(contrib, partials);
}

// Finalisation of one reduction: adds the producer's block partials in
// block order and applies the scalar recurrence.  It skips exactly when its
// producer skipped (same control-block state), so partials are never stale.
// mode kFinOmega reads two partial arrays (t.t, t.s).
constexpr int kFinOmega = 100;
__global__ __launch_bounds__(1024) void k_kfinal(const double* __restrict__ partials, int nparts, KCtl* ctl, int mode,
                                                 int onEarly) {
    if (stopped(ctl)) return;
    if (mode == kFinTrue && (onEarly ? !ctl->early : (!ctl->check || ctl->early))) return;
    if ((mode == kFinOmega || mode == kFinUpdBi) && ctl->early) return;
    const double a = sum_partials(partials, nparts);
    if (mode == kFinOmega) {
        const double b = sum_partials(partials + kMaxReduceBlocks, nparts);
        if (threadIdx.x == 0) {
            if (!isfinite(a) || !isfinite(b)) ctl->status = kNonFinite;
            else if (a == 0) ctl->status = kBreakdown;
            else {
                const double omega = b / a;
                ctl->omega = omega;
                if (!isfinite(omega)) ctl->status = kNonFinite;
                else if (omega == 0) ctl->status = kBreakdown;
            }
        }
        return;
    }
    if (threadIdx.x == 0) finalize(mode, a, ctl);
}

// The preconditioner's output step fused with the dot product that follows
// it: z = FP64(level-0 iterate) (or 0), sum(r .* z) per block with the
// thread order and block reduction of k_kdot, so the partials are the ones
// k_kdot would form from z.  With z == nullptr only the dot is formed (the
// consumer reads the iterate itself).
template <class L>
__global__ __launch_bounds__(kBlock) void k_out_dot(int n, const L* __restrict__ xl, int zero,
                                                    const double* __restrict__ r, double* __restrict__ z,
                                                    double* partials, const KCtl* ctl) {
    if (stopped(ctl)) return;
    double s = 0;
#pragma unroll 4
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock) {
        const double zi = zero ? 0.0 : double(xl[i]);
        if (z) z[i] = zi;
        s += r[i] * zi;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// BiCGStab direction: p = r (fresh) or p = r + beta (p - omega v).
__global__ void k_bi_p(int n, const double* __restrict__ r, const double* __restrict__ v, double* __restrict__ p,
                       const KCtl* ctl) {
    if (stopped(ctl)) return;
    const int fresh = ctl->fresh;
    const double beta = ctl->beta, omega = ctl->omega;
#pragma unroll 4
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock)
        p[i] = fresh ? r[i] : 1.0 * r[i] + beta * (-omega * v[i] + p[i]);
}

// s = r - alpha v; x += alpha ph; sum(s.^2).
__global__ __launch_bounds__(kBlock) void k_bi_s(int n, const double* __restrict__ r, const double* __restrict__ v,
                                                 const double* __restrict__ ph, double* __restrict__ s,
                                                 double* __restrict__ x, double* partials,
                                                 KCtl* ctl) {
    if (stopped(ctl)) return;
    const double alpha = ctl->alpha;
    double acc = 0;
#pragma unroll 4
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock) {
        const double si = -alpha * v[i] + r[i];
        s[i] = si;
        x[i] = alpha * ph[i] + x[i];
        acc += si * si;
    }
    block_partial(acc, partials);
}

// x += omega shat; r = s - omega t; sum(r.^2).
__global__ __launch_bounds__(kBlock) void k_bi_update(int n, const double* __restrict__ shat,
                                                      const double* __restrict__ s, const double* __restrict__ t,
                                                      double* __restrict__ x, double* __restrict__ r,
                                                      double* partials, const KCtl* ctl) {
    if (stopped(ctl) || ctl->early) return;
    const double omega = ctl->omega;
    double acc = 0;
#pragma unroll 4
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock) {
        x[i] = omega * shat[i] + x[i];
        const double ri = -omega * t[i] + s[i];
        r[i] = ri;
        acc += ri * ri;
    }
    block_partial(acc, partials);
}

// ---- CG iteration with the scalar recurrences folded into the consumers ----
// Each kernel below starts by adding its producer's block partials (every
// block in the same order, so every block reaches the same value), applies
// the recurrence step to a copy of the control block, and lets block 0 write
// the fields that step sets.  The fields a kernel reads are never ones it
// writes, except the status, whose new value every block derives itself,
// so blocks that start after block 0's write behave identically.  Between
// two iterations the control block holds the same values as after separate
// finalisers; "fresh = 0", the curvature step's, is written by the residual
// step that follows, where no block still reads it.

// Every thread's share of the producer's partials -> thread 0 (block order).
__device__ __forceinline__ double fold_partials(const double* __restrict__ parts, int n) {
    return sum_partials(parts, n);
}

// rho = r.z (finaliser of the preconditioner's output step), then the CG
// direction p = z (fresh) or z + beta p.  A true-residual refresh left
// pending by the previous iteration's check (its |r|^2 partials in pend) is
// finalised first, as its own finaliser would have done at the end of that
// iteration: every block redoes the sum, block 0 writes the result back.
// (z in FP64, or the preconditioner's level-0 iterate in level precision
// Z, widened exactly; zero: z = 0)
template <class Z>
__global__ __launch_bounds__(kBlock) void k_cg_dir(int n, const Z* __restrict__ z, int zero, double* __restrict__ p,
                                                   const double* __restrict__ partsIn, int nIn,
                                                   const double* __restrict__ pend, int nPend, KCtl* ctl) {
    __shared__ double sBeta;
    __shared__ int sRun, sFresh, sPend;
    if (stopped(ctl)) return;
    if (threadIdx.x == 0) sPend = pend && ctl->check && !ctl->early;
    __syncthreads();
    const double trueSq = sPend ? fold_partials(pend, nPend) : 0.0;
    const double rho = fold_partials(partsIn, nIn);
    if (threadIdx.x == 0) {
        KCtl c = *ctl;
        if (sPend) finalize(kFinTrue, trueSq, &c);  // (idempotent if block 0 already wrote it back)
        if (c.status == kRunning) finalize(kFinRhoCg, rho, &c);
        if (blockIdx.x == 0) {
            if (sPend) {
                ctl->res = c.res;
                ctl->check = c.check;
                ctl->fresh = c.fresh;
            }
            ctl->iter = c.iter;
            ctl->rho = c.rho;
            ctl->beta = c.beta;
            ctl->status = c.status;
        }
        sBeta = c.beta;
        sFresh = c.fresh;
        sRun = c.status == kRunning;
    }
    __syncthreads();
    if (!sRun) return;
    const double beta = sBeta;
    const bool fresh = sFresh != 0;
#pragma unroll 4
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock) {
        const double zi = zero ? 0.0 : double(z[i]);
        p[i] = fresh ? zi : zi + beta * p[i];
    }
}

// alpha = rho / p.q (finaliser of the product A p), then x += alpha p,
// r -= alpha q, and |r|^2 per block.
__global__ __launch_bounds__(kBlock) void k_cg_step(int n, const double* __restrict__ p, const double* __restrict__ q,
                                                    double* __restrict__ x, double* __restrict__ r,
                                                    const double* __restrict__ partsIn, int nIn,
                                                    double* __restrict__ partsOut, KCtl* ctl) {
    __shared__ double sAlpha;
    __shared__ int sRun;
    if (stopped(ctl)) return;
    const double curv = fold_partials(partsIn, nIn);
    if (threadIdx.x == 0) {
        KCtl c = *ctl;
        finalize(kFinCurvCg, curv, &c);
        if (blockIdx.x == 0) {
            ctl->alpha = c.alpha;
            ctl->status = c.status;
        }
        sAlpha = c.alpha;
        sRun = c.status == kRunning;
    }
    __syncthreads();
    if (!sRun) return;
    const double alpha = sAlpha;
    double s = 0;
#pragma unroll 4
    for (int i = int(blockIdx.x) * kBlock + int(threadIdx.x); i < n; i += int(gridDim.x) * kBlock) {
        x[i] = alpha * p[i] + x[i];
        const double ri = -alpha * q[i] + r[i];
        r[i] = ri;
        s += ri * ri;
    }
    block_partial(s, partsOut);
}

// |r| and the refresh test (Cg.cuh:173-187), then, where a refresh is due,
// r = b - A x and |r|^2 per block for the true-residual finaliser.
template <class R>
__global__ __launch_bounds__(kBlock) void k_cg_check(R rows, const double* __restrict__ b,
                                                     const double* __restrict__ x, double* __restrict__ r,
                                                     const double* __restrict__ partsIn, int nIn,
                                                     double* __restrict__ partsOut, KCtl* ctl) {
    __shared__ int sTrue;
    if (stopped(ctl)) return;
    const double rr = fold_partials(partsIn, nIn);
    if (threadIdx.x == 0) {
        KCtl c = *ctl;
        finalize(kFinUpdCg, rr, &c);
        c.fresh = 0;
        if (blockIdx.x == 0) {
            ctl->res = c.res;
            ctl->check = c.check;
            ctl->rho_prev = c.rho_prev;
            ctl->status = c.status;
            ctl->fresh = c.fresh;
        }
        sTrue = c.status == kRunning && c.check && !c.early;
    }
    __syncthreads();
    if (!sTrue) return;
    double contrib = 0;
    rows.each(x, [&](int row, double acc) {
        const double res = b[row] - acc;
        r[row] = res;
        contrib += res * res;
    });
    block_partial(contrib, partsOut);
}

// Element-wise Krylov kernels: at most `cap` blocks, each thread walking its
// elements with the loop unrolled by 4 so that several elements' loads are
// in flight.  2048 blocks (8 per CU, all resident at once) suit the kernels
// with up to three streams; the solution/residual updates (five or six
// streams) ran faster with one element a thread (at 128^3: 20.3 us at 8192
// blocks against 22.8 us at 2048 for the CG update).
constexpr int kVecBlocks = 2048, kUpdateBlocks = 8192;
// Most partials a reduction whose recurrence step is folded into its
// consumer may leave: every consumer block adds them all.
constexpr int kFoldParts = 2048;
int vec_blocks(int n, int cap = kVecBlocks) {
    int b = (n + kBlock - 1) / kBlock;
    if (b > cap) b = cap;
    return b < 1 ? 1 : b;
}

}  // namespace

// ------------------------------------------------------------------ AMG

AmgOptionsC amg_options_from(const tf_amg_options* o) {
    AmgOptionsC a;
    if (o) {
        a.maxLevels = o->max_levels;
        a.coarseRows = o->coarse_rows;
        a.rounds = o->matching_rounds;
        a.pre = o->pre_sweeps;
        a.post = o->post_sweeps;
        a.coarseSweeps = o->coarse_sweeps;
        a.omega = o->jacobi_weight;
        a.theta = o->strength_threshold;
        a.absorb = o->absorb_singletons != 0;
        a.cycle = o->cycle;
        a.precision = o->level_precision;
        a.tail = o->coarse_tail;
        a.learned = o->learned;
    }
    // Amg.cuh:66-75 (coarse_rows widened to 1..1024).
    require(a.maxLevels >= 1 && a.maxLevels <= 20 && a.coarseRows >= 1 && a.coarseRows <= 1024 && a.rounds >= 1 &&
                a.rounds <= 8 && a.pre >= 0 && a.post >= 0 && a.coarseSweeps >= 1 && std::isfinite(a.theta) &&
                a.theta >= 0 && a.theta <= 1 && std::isfinite(a.omega) && a.omega > 0 && a.omega < 2 &&
                a.cycle >= 0 && a.cycle <= 2 && (a.precision == 32 || a.precision == 64) && a.tail >= 0 &&
                a.tail <= 2 && a.learned >= 0 && a.learned <= 2,
            "Invalid AMG options");
    return a;
}

AmgEngine::AmgEngine(tf_ctx* ctx, const AmgOptionsC& o) : opt_(o), c_(ctx) {}
AmgEngine::~AmgEngine() {
    delete block_;
    release();
    if (histHost_) (void)hipHostFree(histHost_);
}

BlockAmg& AmgEngine::block() {
    if (!block_) block_ = new BlockAmg(c_, this);
    return *block_;
}

namespace {
void free_level(tf_ctx& c, AmgLevel& l, bool f32) {
    if (f32) dev_free(c, l.vals);
    l.vals = nullptr;
    free_sell(c, l.sell);
    dev_free(c, l.svals);
    l.svals = nullptr;
    free_stencil(c, l.st);
    dev_free(c, l.stvals);
    l.stvals = nullptr;
    for (void** p : {&l.inv, &l.b, &l.x, &l.x2, &l.r}) {
        dev_free(c, *p);
        *p = nullptr;
    }
    free_csr(c, l.native);
    free_agg(c, l.transfer);
}
}  // namespace

void AmgEngine::release() {
    ++generation_;
    if (c_) (void)hipStreamSynchronize(c_->stream);
    for (auto& l : levels_) free_level(*c_, l, opt_.precision == 32);
    levels_.clear();
    for (void* p : {static_cast<void*>(lu_), static_cast<void*>(piv_), static_cast<void*>(inverse_)})
        dev_free(*c_, p);
    free_tail();
    lu_ = nullptr;
    piv_ = nullptr;
    inverse_ = nullptr;
    factored_ = singular_ = false;
}

void AmgEngine::free_tail() {
    ++generation_;
    if (tailLevels_ || tailOps_ || tailImage_ || tailPack_) {
        if (c_) (void)hipStreamSynchronize(c_->stream);
        for (void* p : {static_cast<void*>(tailLevels_), static_cast<void*>(tailOps_), tailImage_, tailPack_})
            dev_free(*c_, p);
    }
    tailImage_ = nullptr;
    tailPack_ = nullptr;
    tailImageBytes_ = 0;
    tailLdsBytes_ = 0;
    tailLevels_ = nullptr;
    tailOps_ = nullptr;
    tail_ = -1;
}

namespace {
// Deterministic pseudo-random start vector in [-1, 1) (SplitMix64 of the index).
__global__ void k_hash_fill(int n, uint64_t seed, double* __restrict__ x) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= n) return;
    uint64_t z = seed + uint64_t(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x[i] = 2.0 * (double(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}
}  // namespace

// Hierarchy choice with learned affinities: 1 builds on the network's
// weights, 2 builds both hierarchies and keeps the learned one only when its
// cycle makes PCG converge faster (convergence_estimate; the plain hierarchy
// is the fallback).
void AmgEngine::setup(int n, int64_t nnz, const int* rp, const int* ci, const double* v, bool canonicalInput) {
    learnedUsed_ = false;
    rhoLearned_ = rhoPlain_ = -1.0;
    if (opt_.learned == 0 || !gnn_) {
        setup_hierarchy(n, nnz, rp, ci, v, canonicalInput, false);
        return;
    }
    setup_hierarchy(n, nnz, rp, ci, v, canonicalInput, true);
    if (opt_.learned == 1) {
        learnedUsed_ = true;
        return;
    }
    rhoLearned_ = convergence_estimate();
    AmgOptionsC po = opt_;
    po.learned = 0;
    AmgEngine plain(c_, po);
    plain.setup(n, nnz, rp, ci, v, canonicalInput);
    rhoPlain_ = plain.convergence_estimate();
    if (rhoLearned_ < rhoPlain_)
        learnedUsed_ = true;
    else
        swap_hierarchy(plain);
}

void AmgEngine::swap_hierarchy(AmgEngine& o) {
    ++generation_;
    ++o.generation_;
    (void)hipStreamSynchronize(c_->stream);
    std::swap(levels_, o.levels_);
    std::swap(visits_, o.visits_);
    std::swap(stop_, o.stop_);
    std::swap(coarseSolver_, o.coarseSolver_);
    std::swap(tail_, o.tail_);
    std::swap(tailLevels_, o.tailLevels_);
    std::swap(tailOps_, o.tailOps_);
    std::swap(tailImage_, o.tailImage_);
    std::swap(tailPack_, o.tailPack_);
    std::swap(tailImageBytes_, o.tailImageBytes_);
    std::swap(tailLdsBytes_, o.tailLdsBytes_);
    std::swap(tailPrograms_, o.tailPrograms_);
    std::swap(factored_, o.factored_);
    std::swap(singular_, o.singular_);
    std::swap(lu_, o.lu_);
    std::swap(piv_, o.piv_);
    std::swap(inverse_, o.inverse_);
}

// Convergence of the cycle as a CG preconditioner: 10 PCG iterations on
// A x = b from x = 0 with a fixed pseudo-random b, and the mean residual
// reduction per iteration (|r_10| / |r_0|)^(1/10) (smaller is better; the
// cycle alone need not contract as a stationary iteration).
double AmgEngine::convergence_estimate() {
    const DevCsr& A = levels_[0].native;
    const int n = A.n;
    double* r = dev_alloc<double>(*c_, size_t(n));
    double* z = dev_alloc<double>(*c_, size_t(n));
    double* p = dev_alloc<double>(*c_, size_t(n));
    double* q = dev_alloc<double>(*c_, size_t(n));
    hipLaunchKernelGGL(k_hash_fill, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c_->stream, n,
                       uint64_t(0x5EEDull), r);
    check(hipGetLastError(), "start vector launch");
    auto dot = [&](const double* x, const double* y) {
        dot_device<double>(*c_, n, x, y, c_->dscalars, 0);
        check(hipMemcpyAsync(c_->hscalars, c_->dscalars, sizeof(double), hipMemcpyDeviceToHost, c_->stream),
              "dot download");
        c_->sync();
        return c_->hscalars[0];
    };
    const double r0 = std::sqrt(dot(r, r));
    apply(r, z);
    convert<double, double>(*c_, size_t(n), z, p);
    double rz = dot(r, z);
    constexpr int kIts = 10;
    double rk = r0;
    for (int k = 0; k < kIts && rz > 0 && std::isfinite(rz); ++k) {
        spmv<double>(*c_, n, A.nnz, A.rp, A.ci, A.v, p, q);
        const double pq = dot(p, q);
        if (!(pq > 0) || !std::isfinite(pq)) break;
        const double alpha = rz / pq;
        axpy<double>(*c_, n, -alpha, q, r);
        rk = std::sqrt(dot(r, r));
        apply(r, z);
        const double rzNew = dot(r, z);
        scal<double>(*c_, n, rzNew / rz, p);
        axpy<double>(*c_, n, 1.0, z, p);
        rz = rzNew;
    }
    c_->sync();
    for (double* v : {r, z, p, q}) dev_free(*c_, v);
    const double rho = r0 > 0 ? std::pow(rk / r0, 1.0 / kIts) : 0.0;
    return std::isfinite(rho) ? rho : 1e300;
}

void AmgEngine::setup_hierarchy(int n, int64_t nnz, const int* rp, const int* ci, const double* v,
                                bool canonicalInput, bool learned) {
    const bool f32 = opt_.precision == 32;
    const size_t es = f32 ? sizeof(float) : sizeof(double);
    std::vector<AmgLevel> next;
    std::string stop;
    double *lu = nullptr, *inv = nullptr;
    int* piv = nullptr;
    bool factored = false, singular = false;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(c_->stream);
        if (c_->side) (void)hipStreamSynchronize(c_->side);  // (a coarse LU may still run there)
        for (auto& l : next) free_level(*c_, l, f32);
        dev_free(*c_, lu);
        dev_free(*c_, piv);
        dev_free(*c_, inv);
    };
    if (!histHost_)
        check(hipHostMalloc(&histHost_, sizeof(int) * 4 * kMaxHistLevels, hipHostMallocDefault), "hipHostMalloc");
    try {
        DevCsr native = canonicalInput ? copy_csr(*c_, n, nnz, rp, ci, v) : canonicalize(*c_, n, nnz, rp, ci, v);
        int symmetric = -1;
        for (int depth = 0;; ++depth) {
            next.emplace_back();
            AmgLevel& L = next.back();
            L.native = native;
            const int ln = native.n;
            // The level's checks raise flag slots 1 (row validity) and 2
            // (diagonal); they come back with the strength graph's size, or
            // in a read of their own on the coarsest level.
            c_->clear_flags(3);
            auto level_checks = [&]() {
                if (c_->hflags[1])
                    throw Error(TF_E_NUMERIC,
                                "Native AMG: missing/zero diagonal, non-finite value or reciprocal in hierarchy");
                if (c_->hflags[2]) throw Error(TF_E_NUMERIC, "Missing or invalid diagonal");
            };
            if (f32) {
                L.vals = dev_alloc<float>(*c_, size_t(native.nnz));
                convert<double, float>(*c_, size_t(native.nnz), native.v, static_cast<float*>(L.vals));
            }
            else {
                L.vals = native.v;
            }
            L.inv = dev_alloc<char>(*c_, size_t(ln) * es);
            L.b = dev_alloc<char>(*c_, size_t(ln) * es);
            L.x = dev_alloc<char>(*c_, size_t(ln) * es);
            L.x2 = dev_alloc<char>(*c_, size_t(ln) * es);
            L.r = dev_alloc<char>(*c_, size_t(ln) * es);
            // row checks, row maxima and the inverse diagonal in one pass
            // (the level's rows are canonical: one diagonal entry each)
            if (f32)
                validate_inverse_async<float>(*c_, ln, native.nnz, native.rp, native.ci, native.v, rowmax_workspace(*c_, ln),
                                              static_cast<float*>(L.inv), c_->dflags + 1, c_->dflags + 2);
            else
                validate_inverse_async<double>(*c_, ln, native.nnz, native.rp, native.ci, native.v, rowmax_workspace(*c_, ln),
                                               static_cast<double*>(L.inv), c_->dflags + 1, c_->dflags + 2);
            if (ln <= opt_.coarseRows || depth + 1 == opt_.maxLevels) {
                c_->read_flags(3);
                level_checks();
                stop = ln <= opt_.coarseRows ? "coarse size" : "level limit";
                break;
            }
            // Galerkin products of a pattern-symmetric matrix are pattern-symmetric, so
            // the check runs on the finest level only.
            double* lw = nullptr;
            if (learned) {
                // the network's weights of this level's entries (its own stream)
                lw = dev_alloc<double>(*c_, size_t(native.nnz));
                c_->sync();
                gnn_->infer(ln, native.nnz, native.rp, native.ci, native.v, lw);
                gnn_->context()->sync();
            }
            DevCsr g;
            try {
                // a finest level whose stencil storage was built mirrored has
                // bit-symmetric values (its fill compared every transpose)
                const bool valueSym = depth == 0 && fineStencil_ && fineStencil_->n == ln && fineStencil_->mirrored;
                g = strength_graph(*c_, ln, native.nnz, native.rp, native.ci, native.v, opt_.theta, symmetric, true,
                                   &symmetric, lw, 3, valueSym);
                level_checks();
            }
            catch (...) {
                free_csr(*c_, g);
                dev_free(*c_, lw);
                throw;
            }
            if (lw) {
                c_->sync();
                dev_free(*c_, lw);
            }
            if (g.nnz == 0) {
                free_csr(*c_, g);
                stop = "empty strength graph";
                break;
            }
            DevAgg t;
            try {
                t = aggregate(*c_, g, opt_.rounds, opt_.absorb,
                              depth < kMaxHistLevels ? histHost_ + 4 * depth : nullptr);
            }
            catch (...) {
                free_csr(*c_, g);
                throw;
            }
            free_csr(*c_, g);
            if (t.nc >= ln) {
                free_agg(*c_, t);
                stop = "no contraction";
                break;
            }
            L.transfer = t;
            for (int k = 0; k < 4; ++k) L.hist[k] = t.hist[k];  // (filled below when copied asynchronously)
            native = galerkin(*c_, L.native, t.map, t.nc, t.offsets, t.members);
        }
        // The coarse LU runs on the side stream while the levels' sweep
        // storage is built here (it reads only the coarsest matrix).
        const DevCsr& last = next.back().native;
        const bool luPending = last.n <= std::max(64, opt_.coarseRows);
        if (luPending) {
            lu = dev_alloc<double>(*c_, size_t(last.n) * size_t(last.n));
            inv = dev_alloc<double>(*c_, size_t(last.n) * size_t(last.n));
            piv = dev_alloc<int>(*c_, size_t(last.n));
            coarse_lu_launch(*c_, last.n, last.rp, last.ci, last.v, lu, piv, inv);
        }
        for (size_t d = 0; d < next.size(); ++d) {
            AmgLevel& L = next[d];
            if (L.native.n < kSellMinRows || (luPending && d + 1 == next.size())) continue;
            // The finest operator of a structured grid sweeps in the stencil
            // storage (coarsened levels never have one offset list).
            double* st64 = nullptr;
            const bool shared = d == 0 && fineStencil_ && fineStencil_->n == L.native.n;
            if (shared) {
                stencil_copy(*c_, *fineStencil_, L.st);
                st64 = const_cast<double*>(fineStencil64_);
            }
            if (shared || (d == 0 && L.native.n >= kStencilMinRows &&
                           stencil_build(*c_, L.native.n, L.native.nnz, L.native.rp, L.native.ci, L.native.v, L.st,
                                         &st64))) {
                if (shared && !f32) {
                    L.stvals = dev_alloc<double>(*c_, size_t(L.st.stored));
                    convert<double, double>(*c_, size_t(L.st.stored), st64, static_cast<double*>(L.stvals));
                }
                else if (f32) {
                    const size_t count = size_t(L.st.stored);
                    L.stvals = dev_alloc<float>(*c_, count);
                    convert<double, float>(*c_, count, st64, static_cast<float*>(L.stvals));
                    if (!shared) dev_free(*c_, st64);  // (stream-ordered reuse)
                }
                else {
                    L.stvals = st64;
                }
                continue;
            }
            if (!sell_build(*c_, L.native.n, L.native.nnz, L.native.rp, L.native.ci, L.sell, kSellMaxPadding,
                            kSellMaxPackedPadding))
                continue;
            bool keep;
            if (f32) {
                L.svals = dev_alloc<float>(*c_, size_t(L.sell.stored));
                sell_values<double, float>(*c_, L.sell, L.native.rp, L.native.v, static_cast<float*>(L.svals));
                keep = sell_pays<float>(*c_, L, opt_.omega);
            }
            else {
                L.svals = dev_alloc<double>(*c_, size_t(L.sell.stored));
                sell_values<double, double>(*c_, L.sell, L.native.rp, L.native.v, static_cast<double*>(L.svals));
                keep = sell_pays<double>(*c_, L, opt_.omega);
            }
            if (!keep) {
                c_->sync();
                free_sell(*c_, L.sell);
                dev_free(*c_, L.svals);
                L.svals = nullptr;
            }
        }
        if (luPending) {
            factored = coarse_lu_join(*c_);
            singular = !factored;
        }
        c_->sync();
        for (size_t d = 0; d + 1 < next.size() && d < size_t(kMaxHistLevels); ++d)
            for (int k = 0; k < 4; ++k) next[d].hist[k] = histHost_[4 * d + k];
    }
    catch (...) {
        cleanup();
        throw;
    }
    release();
    levels_ = std::move(next);
    lu_ = lu;
    piv_ = piv;
    inverse_ = inv;
    factored_ = factored;
    singular_ = singular;
    stop_ = stop;
    coarseSolver_ = factored ? "pivoted GPU LU" : (singular ? "Jacobi (singular LU)" : "Jacobi (coarsening stopped)");
    visits_.assign(levels_.size(), 0);
    build_tail();
    tune_tail();
}

// The one-workgroup tail is kept only where it beats launching the same
// operations level by level (timed once per hierarchy shape and cycle).
void AmgEngine::tune_tail() {
    if (tail_ < 0 || opt_.tail == 1) return;
    if (opt_.tail == 2) {
        free_tail();
        return;
    }
    int64_t sig = 0;
    for (const auto& l : levels_) sig = sig * 1000003 + l.native.nnz;
    const TuneKey key{2, levels_[0].native.n, sig, int64_t(opt_.cycle * 100 + opt_.pre * 10 + opt_.post),
                      opt_.precision};
    const int first = tail_;
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tail_start.find(key);
        if (it != g_tail_start.end()) {
            if (it->second < 0)
                free_tail();
            else if (it->second != first)
                build_tail(it->second);
            return;
        }
    }
    // Candidates: no tail, and the tail from the first level that fits its
    // LDS budget or from one of the next two levels (the levels above the
    // start then run as per-level kernels); the fastest cycle wins.
    const size_t n = size_t(levels_[0].native.n);
    double* tmp = dev_alloc<double>(*c_, 2 * n);
    fill_zero(*c_, tmp, 2 * n * sizeof(double));
    auto time_cycle = [&] { return best_time(*c_, [&] { apply(tmp, tmp + n); }); };
    tail_ = -1;
    float best = time_cycle();
    int chosen = -1;
    tail_ = first;
    for (int s = first; s < int(levels_.size()) && s <= first + 2; ++s) {
        if (s != tail_) build_tail(s);
        if (tail_ != s) break;
        const float t = time_cycle();
        if (t < best) {
            best = t;
            chosen = s;
        }
    }
    c_->sync();
    dev_free(*c_, tmp);
    if (chosen < 0)
        free_tail();
    else if (chosen != tail_)
        build_tail(chosen);
    std::fill(visits_.begin(), visits_.end(), 0);
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tail_start[key] = chosen;
}

namespace {
// Host mirror of AmgEngine::cycle that records operations instead of
// launching kernels (same flag logic: lazily cleared iterates, first sweeps
// fused into the restriction), with explicit iterate buffers per level.
struct TailGen {
    const std::vector<AmgLevel>& levels;
    const AmgOptionsC& opt;
    bool factored;
    std::vector<char> zero, pre;
    std::vector<int> cur;
    std::vector<int64_t> visits;
    std::vector<TailOp> ops;

    void emit(int kind, int level, int in, int out, int flag) { ops.push_back(TailOp{kind, level, in, out, flag}); }
    void smooth(size_t d, int sweeps, int first) {
        for (int s = first; s < sweeps; ++s) {
            if (zero[d]) {
                emit(kTSweepZero, int(d), 0, cur[d], 0);
                zero[d] = 0;
            }
            else {
                emit(kTSweep, int(d), cur[d], 1 - cur[d], 0);
                cur[d] = 1 - cur[d];
            }
        }
    }
    void restrict_to(size_t d, bool ofResidual, bool fuse) {
        if (!ofResidual || zero[d]) {
            emit(kTRestrict, int(d), 0, cur[d + 1], fuse);
        }
        else {
            emit(kTResidual, int(d), cur[d], 0, 0);
            emit(kTRestrict, int(d), 1, cur[d + 1], fuse);
        }
        pre[d + 1] = fuse;
    }
    void prolong(size_t d) {
        emit(kTProlong, int(d), cur[d + 1], cur[d], zero[d]);
        zero[d] = 0;
    }
    void cycle(int kind, size_t d, bool z) {
        ++visits[d];
        int done = 0;
        if (pre[d]) {
            zero[d] = 0;
            pre[d] = 0;
            done = 1;
        }
        else if (z || kind == 2) {
            zero[d] = 1;
        }
        if (d + 1 == levels.size()) {
            if (factored) {
                emit(kTCoarse, int(d), 0, cur[d], 0);
                zero[d] = 0;
            }
            else {
                smooth(d, opt.coarseSweeps, 0);
            }
            return;
        }
        const bool nxSmooths = d + 2 < levels.size() && opt.pre >= 1;
        if (kind == 2) {
            restrict_to(d, false, false);
            cycle(2, d + 1, true);
            prolong(d);
            cycle(0, d, false);
            return;
        }
        smooth(d, opt.pre, done);
        restrict_to(d, true, nxSmooths);
        cycle(kind, d + 1, true);
        if (kind == 1) cycle(kind, d + 1, false);
        prolong(d);
        smooth(d, opt.post, 0);
    }
};
}  // namespace

// The whole cycle of apply() as one operation list, from level 0 with its
// first sweep applied at entry when apply() fuses it there; curOut / zeroOut
// receive the level-0 iterate's buffer and whether it is still zero.
std::vector<TailOp> AmgEngine::cycle_program(bool& firstAtEntry, int& curOut, bool& zeroOut) const {
    TailGen g{levels_, opt_, factored_, std::vector<char>(levels_.size(), 0), std::vector<char>(levels_.size(), 0),
              std::vector<int>(levels_.size(), 0), std::vector<int64_t>(levels_.size(), 0), {}};
    firstAtEntry = opt_.cycle != 2 && levels_.size() > 1 && opt_.pre >= 1;
    g.pre[0] = firstAtEntry;
    g.cycle(opt_.cycle, 0, true);
    curOut = g.cur[0];
    zeroOut = g.zero[0] != 0;
    return g.ops;
}

// Chooses the tail's first level and records its programs for every entry
// state (cycle kind, zero start, first sweep already applied by the parent's
// restriction).
void AmgEngine::build_tail(int from) {
    free_tail();
    if (levels_.size() < 2 || (factored_ && levels_.back().native.n > kCoarseMax)) return;
    const size_t es = opt_.precision == 32 ? sizeof(float) : sizeof(double);
    auto al = [](size_t b) { return (b + 15) / 16 * 16; };
    // LDS bytes of levels d..coarsest: five vectors, row offsets, the transfer
    // and the packed slices' offsets.
    auto slices_of = [](int n) { return (n + kWave - 1) / kWave; };
    std::vector<size_t> need(levels_.size() + 1, 0);
    for (size_t d = levels_.size(); d-- > 0;) {
        const AmgLevel& l = levels_[d];
        const size_t n = size_t(l.native.n);
        size_t b = 5 * al(n * es) + al((n + 1) * 4) + al((size_t(slices_of(l.native.n)) + 1) * 4);
        if (l.transfer.nc > 0) b += 2 * al(n * 4) + al((size_t(l.transfer.nc) + 1) * 4);
        need[d] = need[d + 1] + b;
    }
    int t = -1;
    for (size_t d = size_t(std::max(1, from)); d < levels_.size(); ++d)
        if (need[d] <= size_t(kTailLds) && levels_[d].native.n <= kTailMaxRows) {
            t = int(d);
            break;
        }
    if (t < 0) return;
    std::vector<TailLevel> table(levels_.size());
    size_t at = 0;
    auto place = [&](size_t bytes) {
        const int off = int(at);
        at += al(bytes);
        return off;
    };
    for (size_t d = size_t(t); d < levels_.size(); ++d) {
        const AmgLevel& l = levels_[d];
        TailLevel& e = table[d];
        const size_t n = size_t(l.native.n);
        e.n = l.native.n;
        e.rp = l.native.rp;
        e.ci = l.native.ci;
        e.v = l.vals;
        e.inv = l.inv;
        e.b = l.b;
        e.x = l.x;
        e.nc = l.transfer.nc;
        e.map = l.transfer.map;
        e.off = l.transfer.offsets;
        e.mem = l.transfer.members;
        e.linv = place(n * es);
        e.lrp = place((n + 1) * 4);
        if (e.nc > 0) {
            e.lmap = place(n * 4);
            e.lmem = place(n * 4);
            e.loff = place((size_t(e.nc) + 1) * 4);
        }
        e.pv = nullptr;
        e.pcol = nullptr;
        e.slices = 0;
        e.lsbase = place((size_t(slices_of(e.n)) + 1) * 4);
    }
    // The constant arrays above form the image staged at each launch; the
    // vectors go after them.
    const size_t imageBytes = al(at);
    for (size_t d = size_t(t); d < levels_.size(); ++d) {
        TailLevel& e = table[d];
        const size_t n = size_t(e.n);
        e.lb = place(n * es);
        e.lx0 = place(n * es);
        e.lx1 = place(n * es);
        e.lr = place(n * es);
    }
    std::vector<TailOp> all;
    for (int key = 0; key < 12; ++key) {
        const int kind = key / 4;
        const bool zero = (key & 2) != 0, pre = (key & 1) != 0;
        TailGen g{levels_, opt_, factored_, std::vector<char>(levels_.size(), 0),
                  std::vector<char>(levels_.size(), 0), std::vector<int>(levels_.size(), 0),
                  std::vector<int64_t>(levels_.size(), 0), {}};
        g.pre[size_t(t)] = pre;
        g.cycle(kind, size_t(t), zero);
        if (g.cur[size_t(t)] != 0) g.emit(kTCopy, t, 1, 0, 0);
        if (g.ops.size() > size_t(kMaxTailOps)) return;  // (W-cycles over very deep tails)
        tailPrograms_[key].offset = int(all.size());
        tailPrograms_[key].count = int(g.ops.size());
        tailPrograms_[key].visits = g.visits;
        all.insert(all.end(), g.ops.begin(), g.ops.end());
    }
    // Packed copies of the levels whose rows the program sums (all but a
    // factored coarsest): slice widths from the row lengths, rounded up to
    // groups of 4; a level whose padding would pass 2x its entries (a few
    // very long rows) keeps CSR row sums.
    const size_t last = factored_ ? levels_.size() - 1 : levels_.size();
    std::vector<std::vector<int>> rows(levels_.size()), bases(levels_.size());
    for (size_t d = size_t(t); d < last; ++d) {
        rows[d].resize(size_t(levels_[d].native.n) + 1);
        check(hipMemcpyAsync(rows[d].data(), levels_[d].native.rp, rows[d].size() * sizeof(int),
                             hipMemcpyDeviceToHost, c_->stream),
              "tail row offsets");
    }
    c_->sync();
    size_t packBytes = 0;
    std::vector<size_t> packAt(levels_.size(), 0);
    for (size_t d = size_t(t); d < last; ++d) {
        const int n = levels_[d].native.n, slices = slices_of(n);
        std::vector<int>& base = bases[d];
        base.assign(size_t(slices) + 1, 0);
        int64_t total = 0;
        for (int s = 0; s < slices; ++s) {
            int w = 0;
            for (int i = s * kWave; i < std::min(n, (s + 1) * kWave); ++i) w = std::max(w, rows[d][i + 1] - rows[d][i]);
            total += int64_t((w + 3) / 4) * 4 * kWave;
            base[size_t(s) + 1] = int(std::min<int64_t>(total, INT32_MAX));
        }
        if (total > 2 * int64_t(levels_[d].native.nnz) + 4 * kWave * int64_t(slices)) {
            base.clear();
            continue;
        }
        packAt[d] = packBytes;
        packBytes += al(size_t(total) * es) + al(size_t(total) * 2);
    }
    if (packBytes > 0) tailPack_ = dev_alloc<char>(*c_, packBytes);
    // LDS image: each level's inverse diagonal, row offsets, transfer and
    // packed slices' offsets at their LDS offsets.
    tailImageBytes_ = int(imageBytes);
    tailLdsBytes_ = int(al(at));
    tailImage_ = dev_alloc<char>(*c_, imageBytes);
    char* img = static_cast<char*>(tailImage_);
    auto put = [&](int off, const void* src, size_t bytes) {
        if (bytes) check(hipMemcpyAsync(img + off, src, bytes, hipMemcpyDeviceToDevice, c_->stream), "tail image");
    };
    for (size_t d = size_t(t); d < levels_.size(); ++d) {
        const TailLevel& e = table[d];
        const size_t n = size_t(e.n);
        put(e.linv, e.inv, n * es);
        put(e.lrp, e.rp, (n + 1) * 4);
        if (e.nc > 0) {
            put(e.lmap, e.map, n * 4);
            put(e.lmem, e.mem, n * 4);
            put(e.loff, e.off, (size_t(e.nc) + 1) * 4);
        }
    }
    // packed levels: slice offsets into the image, matrices into tailPack_
    for (size_t d = size_t(t); d < last; ++d) {
        const std::vector<int>& base = bases[d];
        if (base.empty()) continue;
        TailLevel& e = table[d];
        const AmgLevel& l = levels_[d];
        const int slices = int(base.size()) - 1;
        const size_t total = size_t(base.back());
        char* at0 = static_cast<char*>(tailPack_) + packAt[d];
        unsigned short* pcol = reinterpret_cast<unsigned short*>(at0 + al(total * es));
        check(hipMemcpyAsync(img + e.lsbase, base.data(), base.size() * sizeof(int), hipMemcpyHostToDevice,
                             c_->stream),
              "tail slice offsets");
        const int* dbase = reinterpret_cast<const int*>(img + e.lsbase);
        if (opt_.precision == 32)
            tail_pack<float>(*c_, e.n, slices, l.native.rp, l.native.ci, static_cast<const float*>(l.vals), dbase,
                             reinterpret_cast<float*>(at0), pcol);
        else
            tail_pack<double>(*c_, e.n, slices, l.native.rp, l.native.ci, static_cast<const double*>(l.vals), dbase,
                              reinterpret_cast<double*>(at0), pcol);
        e.pv = at0;
        e.pcol = pcol;
        e.slices = slices;
    }
    tailLevels_ = dev_alloc<TailLevel>(*c_, table.size());
    tailOps_ = dev_alloc<TailOp>(*c_, all.size());
    check(hipMemcpyAsync(tailLevels_, table.data(), table.size() * sizeof(TailLevel), hipMemcpyHostToDevice,
                         c_->stream),
          "tail table upload");
    check(hipMemcpyAsync(tailOps_, all.data(), all.size() * sizeof(TailOp), hipMemcpyHostToDevice, c_->stream),
          "tail program upload");
    c_->sync();
    tail_ = t;
    ++generation_;
}

template <class L>
void AmgEngine::run_tail(int kind, bool zero) {
    AmgLevel& l = levels_[size_t(tail_)];
    const TailProgram& p = tailPrograms_[kind * 4 + (zero ? 2 : 0) + (l.preapplied ? 1 : 0)];
    tail_run<L>(*c_, tailLevels_, tail_, int(levels_.size()) - tail_, tailImage_, tailImageBytes_, tailLdsBytes_,
                tailOps_ + p.offset, p.count, inverse_, L(opt_.omega));
    for (size_t d = 0; d < visits_.size(); ++d) visits_[d] += p.visits[d];
    l.preapplied = false;
    l.zero = false;
}

template <class L>
void AmgEngine::smooth(size_t d, int sweeps, int first) {
    AmgLevel& l = levels_[d];
    const DevCsr& A = l.native;
    for (int s = first; s < sweeps; ++s) {
        if (l.zero) {
            jacobi_fused<L>(*c_, A.n, A.nnz, A.rp, A.ci, static_cast<const L*>(l.vals), static_cast<const L*>(l.inv),
                            L(opt_.omega), static_cast<const L*>(l.b), nullptr, static_cast<L*>(l.x));
            l.zero = false;
        }
        else {
            if (l.st)
                stencil_jacobi<L>(*c_, l.st, static_cast<const L*>(l.stvals), static_cast<const L*>(l.inv),
                                  L(opt_.omega), static_cast<const L*>(l.b), static_cast<const L*>(l.x),
                                  static_cast<L*>(l.x2));
            else if (l.sell)
                sell_jacobi<L>(*c_, l.sell, static_cast<const L*>(l.svals), static_cast<const L*>(l.inv),
                               L(opt_.omega), static_cast<const L*>(l.b), static_cast<const L*>(l.x),
                               static_cast<L*>(l.x2));
            else
                jacobi_fused<L>(*c_, A.n, A.nnz, A.rp, A.ci, static_cast<const L*>(l.vals),
                                static_cast<const L*>(l.inv), L(opt_.omega), static_cast<const L*>(l.b),
                                static_cast<const L*>(l.x), static_cast<L*>(l.x2));
            std::swap(l.x, l.x2);
        }
    }
}

// Restrict the rhs (ofResidual == false) or the residual of level d into the
// rhs of level d+1.  With fuseFirstSweep the coarse level's first Jacobi
// sweep from zero is computed in the same kernel.
template <class L>
void AmgEngine::restrict_to(size_t d, bool ofResidual, bool fuseFirstSweep) {
    AmgLevel& l = levels_[d];
    AmgLevel& nx = levels_[d + 1];
    const DevAgg& t = l.transfer;
    const L* inv = fuseFirstSweep ? static_cast<const L*>(nx.inv) : nullptr;
    L* xc = fuseFirstSweep ? static_cast<L*>(nx.x) : nullptr;
    const L w = L(opt_.omega);
    if (!ofResidual || l.zero) {
        restrict_agg<L>(*c_, t.nc, t.offsets, t.members, static_cast<const L*>(l.b), static_cast<L*>(nx.b), inv, w,
                        xc);
    }
    else if (l.native.n <= kFusedResidualRows) {
        residual_restrict<L>(*c_, t.nc, t.offsets, t.members, l.native.rp, l.native.ci,
                             static_cast<const L*>(l.vals), static_cast<const L*>(l.b), static_cast<const L*>(l.x),
                             static_cast<L*>(nx.b), inv, w, xc);
    }
    else {
        if (l.st)
            stencil_residual<L>(*c_, l.st, static_cast<const L*>(l.stvals), static_cast<const L*>(l.b),
                                static_cast<const L*>(l.x), static_cast<L*>(l.r));
        else if (l.sell)
            sell_residual<L>(*c_, l.sell, static_cast<const L*>(l.svals), static_cast<const L*>(l.b),
                             static_cast<const L*>(l.x), static_cast<L*>(l.r));
        else
            residual<L>(*c_, l.native.n, l.native.nnz, l.native.rp, l.native.ci, static_cast<const L*>(l.vals),
                        static_cast<const L*>(l.b), static_cast<const L*>(l.x), static_cast<L*>(l.r));
        restrict_agg<L>(*c_, t.nc, t.offsets, t.members, static_cast<const L*>(l.r), static_cast<L*>(nx.b), inv, w,
                        xc);
    }
    nx.preapplied = fuseFirstSweep;
}

template <class L>
void AmgEngine::coarse(size_t d) {
    AmgLevel& l = levels_[d];
    if (factored_) {
        coarse_inv_apply<L>(*c_, l.native.n, inverse_, static_cast<const L*>(l.b), static_cast<L*>(l.x));
        l.zero = false;
    }
    else {
        smooth<L>(d, opt_.coarseSweeps);
    }
}

// executeCycle (Cycle.cuh:35-63) with the reference's operations mapped to
// fused kernels; the clear of a zero initial guess is tracked lazily.
template <class L>
void AmgEngine::cycle(int kind, size_t d, bool zero) {
    if (int(d) == tail_) {
        run_tail<L>(kind, zero);
        return;
    }
    ++visits_[d];
    AmgLevel& l = levels_[d];
    int done = 0;
    if (l.preapplied) {
        l.zero = false;
        l.preapplied = false;
        done = 1;
    }
    else if (zero || kind == 2) {
        l.zero = true;
    }
    if (d + 1 == levels_.size()) {
        coarse<L>(d);
        return;
    }
    AmgLevel& nx = levels_[d + 1];
    const DevAgg& t = l.transfer;
    const bool nxSmooths = d + 2 < levels_.size() && opt_.pre >= 1;
    if (kind == 2) {
        restrict_to<L>(d, false, false);
        cycle<L>(2, d + 1, true);
        prolong_agg<L>(*c_, l.native.n, t.map, static_cast<const L*>(nx.x), static_cast<L*>(l.x), l.zero);
        l.zero = false;
        cycle<L>(0, d, false);
        return;
    }
    smooth<L>(d, opt_.pre, done);
    restrict_to<L>(d, true, nxSmooths);
    cycle<L>(kind, d + 1, true);
    if (kind == 1) cycle<L>(kind, d + 1, false);
    if (opt_.post >= 1 && l.native.n <= kFusedProlongRows) {
        const DevCsr& A = l.native;
        prolong_jacobi<L>(*c_, A.n, A.nnz, A.rp, A.ci, static_cast<const L*>(l.vals), t.map,
                          static_cast<const L*>(nx.x), static_cast<const L*>(l.x), l.zero, static_cast<L*>(l.x2),
                          static_cast<const L*>(l.b), static_cast<const L*>(l.inv), L(opt_.omega));
        std::swap(l.x, l.x2);
        l.zero = false;
        smooth<L>(d, opt_.post, 1);
    }
    else {
        prolong_agg<L>(*c_, l.native.n, t.map, static_cast<const L*>(nx.x), static_cast<L*>(l.x), l.zero);
        l.zero = false;
        smooth<L>(d, opt_.post);
    }
}

template <class L>
void AmgEngine::apply_typed(const double* rhs, double* x, const double* with, double* partials, int blocks,
                            const KCtl* ctl, Output* out) {
    AmgLevel& l0 = levels_[0];
    const int n = l0.native.n;
    const bool first = opt_.cycle != 2 && levels_.size() > 1 && opt_.pre >= 1;
    enter_level<L>(*c_, n, rhs, static_cast<L*>(l0.b), static_cast<const L*>(l0.inv), L(opt_.omega),
                   first ? static_cast<L*>(l0.x) : nullptr);
    for (auto& l : levels_) l.preapplied = false;
    l0.preapplied = first;
    std::fill(visits_.begin(), visits_.end(), 0);
    cycle<L>(opt_.cycle, 0, true);
    if (out) {
        out->x = l0.x;
        out->f32 = sizeof(L) == sizeof(float);
        out->zero = l0.zero;
    }
    if (with) {
        hipLaunchKernelGGL(k_out_dot<L>, dim3(blocks), dim3(kBlock), 0, c_->stream, n,
                           static_cast<const L*>(l0.x), l0.zero ? 1 : 0, with, x, partials, ctl);
        check(hipGetLastError(), "output dot launch");
    }
    else if (l0.zero) {
        fill_zero(*c_, x, size_t(n) * sizeof(double));
    }
    else {
        convert<L, double>(*c_, size_t(n), static_cast<const L*>(l0.x), x);
    }
}

void AmgEngine::apply(const double* rhs, double* x) {
    require(ready(), "AMG requires setup", TF_E_INVALID);
    if (opt_.precision == 32)
        apply_typed<float>(rhs, x);
    else
        apply_typed<double>(rhs, x);
}

void AmgEngine::apply_dot(const double* rhs, double* x, const double* with, double* partials, int blocks,
                          const KCtl* ctl, Output* out) {
    require(ready(), "AMG requires setup", TF_E_INVALID);
    require(x || out, "AMG output needs a vector or a descriptor", TF_E_INVALID);
    if (opt_.precision == 32)
        apply_typed<float>(rhs, x, with, partials, blocks, ctl, out);
    else
        apply_typed<double>(rhs, x, with, partials, blocks, ctl, out);
}

// --------------------------------------------------------------- Krylov

SolverEngine::SolverEngine(tf_ctx* ctx, int kind, const KrylovOptionsC& o, AmgEngine* amg)
    : c_(ctx), kind_(kind), o_(o), amg_(amg) {
    require(kind == 0 || kind == 1, "unknown Krylov kind");
    require(o.maxit >= 0 && o.refresh >= 1 && std::isfinite(o.rtol) && o.rtol >= 0 && o.rtol < 1 &&
                std::isfinite(o.atol) && o.atol >= 0,
            kind == 0 ? "Invalid CG options" : "Invalid BiCGStab options");
    check(hipMalloc(&ctl_, sizeof(KCtl)), "hipMalloc control");
    check(hipHostMalloc(&hctl_, sizeof(KCtl), hipHostMallocDefault), "hipHostMalloc control");
}

void SolverEngine::release_graphs() {
    graph_.reset();
    bgraph_.reset();
    sgraph_.reset();
}

SolverEngine::~SolverEngine() {
    if (c_) (void)hipStreamSynchronize(c_->stream);
    release_graphs();
    release_block();
    release_segments();
    free_csr(*c_, A_);
    free_sell(*c_, As_);
    dev_free(*c_, Asv_);
    free_stencil(*c_, Ast_);
    dev_free(*c_, Astv_);
    dev_free(*c_, vec_);
    if (ctl_) (void)hipFree(ctl_);
    if (hctl_) (void)hipHostFree(hctl_);
}

void SolverEngine::setup(int n, int64_t nnz, const int* rp, const int* ci, const double* v) {
    require(n >= 1, "Solver requires a nonempty square matrix");
    release_graphs();
    release_block();
    release_segments();
    DevCsr A = canonicalize(*c_, n, nnz, rp, ci, v);
    if (vec_) {
        dev_free(*c_, vec_);
        vec_ = nullptr;
    }
    free_csr(*c_, A_);
    free_sell(*c_, As_);
    dev_free(*c_, Asv_);
    Asv_ = nullptr;
    free_stencil(*c_, Ast_);
    dev_free(*c_, Astv_);
    Astv_ = nullptr;
    A_ = A;
    n_ = n;
    const int nv = kind_ == 0 ? 6 : 11;
    vec_ = dev_alloc<double>(*c_, size_t(n) * nv);
    double* p = vec_;
    auto take = [&]() {
        double* q = p;
        p += n;
        return q;
    };
    rhs_ = take();
    x_ = take();
    r_ = take();
    z_ = take();  // CG: M^-1 r
    p_ = take();
    q_ = take();  // CG: A p; BiCGStab: v
    if (kind_ == 1) {
        sh_ = take();
        s_ = take();
        t_ = take();
        ph_ = take();
        shat_ = take();
    }
    if (n >= kStencilMinRows && stencil_build(*c_, A_.n, A_.nnz, A_.rp, A_.ci, A_.v, Ast_, &Astv_)) {
        // structured-grid operator: the iteration's products read the stencil arrays
    }
    else if (n >= kSellMinRows &&
             sell_build(*c_, A_.n, A_.nnz, A_.rp, A_.ci, As_, kSellMaxPadding, kSellMaxPackedPadding)) {
        Asv_ = dev_alloc<double>(*c_, size_t(As_.stored));
        sell_values<double, double>(*c_, As_, A_.rp, A_.v, Asv_);
        if (!sell_spmv_pays()) {
            c_->sync();
            free_sell(*c_, As_);
            dev_free(*c_, Asv_);
            Asv_ = nullptr;
        }
    }
    if (amg_) {
        amg_->share_fine_stencil(Ast_ ? &Ast_ : nullptr, Astv_);
        try {
            amg_->setup(A_.n, A_.nnz, A_.rp, A_.ci, A_.v, true);
        }
        catch (...) {
            amg_->share_fine_stencil(nullptr, nullptr);
            throw;
        }
        amg_->share_fine_stencil(nullptr, nullptr);
    }
    c_->sync();
}

// The operator's products run in sliced ELL only where that beats the CSR
// kernel (timed once per operator shape, as for the AMG levels).
bool SolverEngine::sell_spmv_pays() {
    if (A_.nnz < kTuneMinNnz) return true;
    const TuneKey key{1, A_.n, A_.nnz, As_.stored, 8};
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) return it->second != 0;
    }
    fill_zero(*c_, p_, size_t(n_) * sizeof(double));
    const bool sell = first_is_faster(
        *c_, [&] { sell_spmv<double>(*c_, As_, Asv_, p_, q_); },
        [&] { spmv<double>(*c_, A_.n, A_.nnz, A_.rp, A_.ci, A_.v, p_, q_); }, kSellPreference);
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tuned[key] = sell;
    return sell;
}

void SolverEngine::precondition(const double* in, double* out) {
    if (amg_)
        amg_->apply(in, out);
    else
        convert<double, double>(*c_, size_t(n_), in, out);
}

double SolverEngine::norm_host(const double* x) {
    dot_device<double>(*c_, n_, x, x, c_->dscalars, 0);
    check(hipMemcpyAsync(c_->hscalars, c_->dscalars, sizeof(double), hipMemcpyDeviceToHost, c_->stream),
          "norm download");
    c_->sync();
    return std::sqrt(c_->hscalars[0]);
}

// One iteration of the solver as a fixed launch sequence.
void SolverEngine::launch_part() {
    if (Ast_) {
        const int nb = (Ast_.n + kBlock - 1) / kBlock;
        if (Ast_.chunks() == 0)
            launch_iteration(StencilRows<0>{Ast_.view(), Astv_}, nb);
        else if (Ast_.chunks() == 1)
            launch_iteration(StencilRows<1>{Ast_.view(), Astv_}, nb);
        else
            launch_iteration(StencilRows<2>{Ast_.view(), Astv_}, nb);
    }
    else if (!As_) {
        const int rpb = stream_rows_per_block<double>(A_.n, A_.nnz);
        const int nb = (A_.n + rpb - 1) / rpb;
        launch_iteration(CsrRows{A_.n, nb, rpb, A_.rp, A_.ci, A_.v}, nb);
    }
    else {
        const int nb = (As_.slices + kBlock / kWave - 1) / (kBlock / kWave);
        if (As_.narrow)
            launch_iteration(SellRows<short>{As_.view<short>(), Asv_}, nb);
        else
            launch_iteration(SellRows<int>{As_.view<int>(), Asv_}, nb);
    }
}

template <class R>
void SolverEngine::launch_iteration(const R& rows, int nb) {
    KCtl* ctl = static_cast<KCtl*>(ctl_);
    const dim3 grid(std::min(round_xcd(nb), kMaxReduceBlocks)), vb(vec_blocks(n_)),
        ub(vec_blocks(n_, kUpdateBlocks));
    double* P = c_->partials;
    auto pslot = [&](int s) { return P + size_t(s) * kMaxReduceBlocks; };
#define TF_G(NAME, ...) hipLaunchKernelGGL(NAME, grid, dim3(kBlock), 0, c_->stream, __VA_ARGS__)
    const int ng = int(grid.x), nv = int(vb.x), nu = int(ub.x);
    auto fin = [&](int slot, int parts, int mode, int onEarly) {
        hipLaunchKernelGGL(k_kfinal, dim3(1), dim3(1024), 0, c_->stream, pslot(slot), parts, ctl, mode, onEarly);
    };
    if (kind_ == 0) {
        // z = M r with rho = r.z formed in the same pass; each reduction's
        // recurrence step runs at the start of the kernel that consumes it
        // (producers keep to at most kFoldParts partials)
        const dim3 gf(std::min(int(grid.x), kFoldParts));
        const int nf = int(gf.x);
        pendParts_ = nf;
        // the direction reads z where the preconditioner leaves it (its
        // level-0 iterate, or r itself without one): z is not written out
        if (amg_) {
            AmgEngine::Output z;
            amg_->apply_dot(r_, nullptr, r_, pslot(kSlotA), nv, ctl, &z);
            if (z.f32)
                hipLaunchKernelGGL(k_cg_dir<float>, vb, dim3(kBlock), 0, c_->stream, n_,
                                   static_cast<const float*>(z.x), int(z.zero), p_, pslot(kSlotA), nv, pslot(kSlotD),
                                   nf, ctl);
            else
                hipLaunchKernelGGL(k_cg_dir<double>, vb, dim3(kBlock), 0, c_->stream, n_,
                                   static_cast<const double*>(z.x), int(z.zero), p_, pslot(kSlotA), nv,
                                   pslot(kSlotD), nf, ctl);
        }
        else {
            hipLaunchKernelGGL(k_out_dot<double>, vb, dim3(kBlock), 0, c_->stream, n_, r_, 0, r_,
                               static_cast<double*>(nullptr), pslot(kSlotA), ctl);
            hipLaunchKernelGGL(k_cg_dir<double>, vb, dim3(kBlock), 0, c_->stream, n_, r_, 0, p_, pslot(kSlotA), nv,
                               pslot(kSlotD), nf, ctl);
        }
        hipLaunchKernelGGL(k_kspmv_dot<R>, gf, dim3(kBlock), 0, c_->stream, rows, p_, q_, p_, pslot(kSlotB), ctl);
        hipLaunchKernelGGL(k_cg_step, vb, dim3(kBlock), 0, c_->stream, n_, p_, q_, x_, r_, pslot(kSlotB), nf,
                           pslot(kSlotC), ctl);
        hipLaunchKernelGGL(k_cg_check<R>, gf, dim3(kBlock), 0, c_->stream, rows, rhs_, x_, r_, pslot(kSlotC), nv,
                           pslot(kSlotD), ctl);
        // (a refresh's finaliser runs in the next iteration's direction
        // kernel, or in settle() before the host reads the control block)
        (void)nu;
    }
    else {
        double* v = q_;
        hipLaunchKernelGGL(k_kdot, vb, dim3(kBlock), 0, c_->stream, n_, sh_, r_, pslot(kSlotA), ctl);
        fin(kSlotA, nv, kFinRhoBi, 0);
        hipLaunchKernelGGL(k_bi_p, vb, dim3(kBlock), 0, c_->stream, n_, r_, v, p_, ctl);
        precondition(p_, ph_);
        TF_G(k_kspmv_dot<R>, rows, ph_, v, sh_, pslot(kSlotB), ctl);
        fin(kSlotB, ng, kFinDenBi, 0);
        hipLaunchKernelGGL(k_bi_s, ub, dim3(kBlock), 0, c_->stream, n_, r_, v, ph_, s_, x_, pslot(kSlotC),
                           ctl);
        fin(kSlotC, nu, kFinSBi, 0);
        TF_G(k_ktrue<R>, rows, rhs_, x_, r_, sh_, pslot(kSlotD), ctl, 1);
        fin(kSlotD, ng, kFinTrue, 1);
        precondition(s_, shat_);
        TF_G(k_kspmv_omega<R>, rows, shat_, t_, s_, pslot(kSlotE), ctl);
        fin(kSlotE, ng, kFinOmega, 0);
        hipLaunchKernelGGL(k_bi_update, ub, dim3(kBlock), 0, c_->stream, n_, shat_, s_, t_, x_, r_, pslot(kSlotA),
                           ctl);
        fin(kSlotA, nu, kFinUpdBi, 0);
        TF_G(k_ktrue<R>, rows, rhs_, x_, r_, sh_, pslot(kSlotD), ctl, 0);
        fin(kSlotD, ng, kFinTrue, 0);
    }
#undef TF_G
    check(hipGetLastError(), "Krylov iteration launch");
}

void SolverEngine::settle() {
    if (kind_ != 0 || pendParts_ <= 0) return;
    hipLaunchKernelGGL(k_kfinal, dim3(1), dim3(1024), 0, c_->stream, c_->partials + size_t(kSlotD) * kMaxReduceBlocks,
                       pendParts_, static_cast<KCtl*>(ctl_), int(kFinTrue), 0);
    check(hipGetLastError(), "refresh finaliser launch");
}

void SolverEngine::true_residual() {
    if (Ast_)
        stencil_residual<double>(*c_, Ast_, Astv_, rhs_, x_, r_);
    else if (As_)
        sell_residual<double>(*c_, As_, Asv_, rhs_, x_, r_);
    else
        residual<double>(*c_, A_.n, A_.nnz, A_.rp, A_.ci, A_.v, rhs_, x_, r_);
}

tf_solve_result SolverEngine::finish(int status, int iterations) {
    true_residual();
    const double res = norm_host(r_);
    if (!std::isfinite(res) || !std::isfinite(normb_))
        status = kNonFinite;
    else if (res <= tol_)
        status = kConverged;
    else if (status == kConverged)
        status = kMaxIt;
    tf_solve_result out;
    out.status = status;
    out.iterations = iterations;
    out.residual_norm = res;
    out.relative_residual = normb_ > 0 ? res / normb_ : res;
    return out;
}

tf_solve_result SolverEngine::apply(const double* b, double* x) {
    require(n_ > 0, kind_ == 0 ? "CG requires setup" : "BiCGStab requires setup", TF_E_INVALID);
    convert<double, double>(*c_, size_t(n_), b, rhs_);
    convert<double, double>(*c_, size_t(n_), x, x_);
    tf_solve_result out{};
    normb_ = norm_host(rhs_);
    tol_ = std::max(o_.atol, o_.rtol * normb_);
    true_residual();
    const double res = norm_host(r_);
    if (!std::isfinite(normb_) || !std::isfinite(res)) {
        out = finish(kNonFinite, 0);
    }
    else if (res <= tol_) {
        out = finish(kConverged, 0);
    }
    else {
        KCtl init{};
        init.normb = normb_;
        init.tol = tol_;
        init.res = res;
        init.rho_prev = kind_ == 0 ? 0.0 : 1.0;
        init.alpha = 1.0;
        init.omega = 1.0;
        init.status = kRunning;
        init.iter = 0;
        init.fresh = 1;
        init.refresh = o_.refresh;
        std::memcpy(hctl_, &init, sizeof(KCtl));
        check(hipMemcpyAsync(ctl_, hctl_, sizeof(KCtl), hipMemcpyHostToDevice, c_->stream), "control upload");
        if (kind_ == 1) convert<double, double>(*c_, size_t(n_), r_, sh_);
        bool done = false;
        KCtl* h = static_cast<KCtl*>(hctl_);
        // Iterations are launched in batches between reads of the control
        // block: the batch is one less than the iterations the mean residual
        // reduction so far says remain (at most 8, one near the tolerance),
        // so the read rarely comes late.  An iteration launched after the
        // solve has stopped changes nothing: its Krylov kernels return on the
        // stopped status, only the preconditioner runs once more.
        const double res0 = res;
        int k = 0;
        while (k < o_.maxit) {
            const int batch = std::min(krylov_batch(h, &res0, 1, k), o_.maxit - k);
            // the captured launches name the preconditioner's level buffers:
            // a shared AMG set up again since the capture is re-captured
            const uint64_t key = amg_ ? amg_->generation() : 0;
            for (int q = 0; q < batch; ++q) launch_captured(*c_, graph_, o_.graph, key, [&] { launch_part(); });
            k += batch;
            settle();
            check(hipMemcpyAsync(hctl_, ctl_, sizeof(KCtl), hipMemcpyDeviceToHost, c_->stream), "control download");
            c_->sync();
            if (h->status != kRunning) {
                out = finish(h->status, h->iter);
                done = true;
                break;
            }
        }
        if (!done) out = finish(kMaxIt, o_.maxit);
    }
    convert<double, double>(*c_, size_t(n_), x_, x);
    c_->sync();
    return out;
}

}  // namespace tf
