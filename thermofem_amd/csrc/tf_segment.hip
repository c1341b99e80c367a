// Segmented batch solve on gfx950: many independent systems stored as one
// block-diagonal operator (segment s = rows [off[s], off[s+1]), no entry
// couples two segments) and solved together by CG (Cg.cuh:123-189) with one
// control block per segment.  Every kernel of an iteration is one persistent
// launch over all segments: a fixed grid of at most 8 blocks per CU walks a
// table of row tiles (256 rows, never crossing a segment), so the cost of a
// launch is shared by all systems instead of paid per system.  Dot products
// are per segment and deterministic: each tile's block sum goes to its own
// slot, and a segment's slots are added in tile order.  The AMG hierarchy of
// the block-diagonal operator is block-diagonal too (aggregates never span
// two segments), so one cycle preconditions every system at once.
// apply_segmented_persistent runs the same solve -- initial residuals, every
// iteration with its whole cycle, each segment's stopping test, the final
// residuals -- as one cooperative launch whose phases are separated by a
// grid-wide barrier (k_seg_solve, below).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "tf_device.hpp"
#include "tf_engine.hpp"
#include "tf_krylov.hpp"

namespace tf {

namespace {

constexpr int kTileRows = kMaxStreamRows;  // 256 rows, one per thread
constexpr int kSegGrid = 2048;             // persistent blocks: 8 per CU

// A tile: `rows` (<= 256) consecutive rows from row0, all in segment seg.
struct STile {
    int row0, rows, seg, pad;
};

__global__ __launch_bounds__(kBlock) void k_seg_check(const STile* __restrict__ tiles, int ntiles,
                                                      const int* __restrict__ rp, const int* __restrict__ ci,
                                                      const int* __restrict__ off, int* bad) {
    for (int t = int(blockIdx.x); t < ntiles; t += int(gridDim.x)) {
        const STile tl = tiles[t];
        const int lo = off[tl.seg], hi = off[tl.seg + 1];
        const int row = tl.row0 + int(threadIdx.x);
        if (int(threadIdx.x) < tl.rows)
            for (int j = rp[row]; j < rp[row + 1]; ++j)
                if (ci[j] < lo || ci[j] >= hi) atomicOr(bad, 1);
    }
}

// sum x.y per tile (ctl: skip stopped segments).
__global__ __launch_bounds__(kBlock) void k_sdot(const STile* __restrict__ tiles, int ntiles,
                                                 const double* __restrict__ x, const double* __restrict__ y,
                                                 double* __restrict__ partials, const KCtl* __restrict__ ctl) {
    for (int t = int(blockIdx.x); t < ntiles; t += int(gridDim.x)) {
        const STile tl = tiles[t];
        if (ctl && stopped(ctl + tl.seg)) continue;
        double s = 0;
        if (int(threadIdx.x) < tl.rows) {
            const int row = tl.row0 + int(threadIdx.x);
            s = x[row] * y[row];
        }
        s = block_sum(s);
        if (threadIdx.x == 0) partials[t] = s;
    }
}

// q = A p and sum p.q per tile.
__global__ __launch_bounds__(kBlock) void k_sspmv_dot(const STile* __restrict__ tiles, int ntiles,
                                                      const int* __restrict__ rp, const int* __restrict__ ci,
                                                      const double* __restrict__ v, const double* __restrict__ p,
                                                      double* __restrict__ q, double* __restrict__ partials,
                                                      const KCtl* __restrict__ ctl) {
    __shared__ double prod[StreamCap<double>::value];
    __shared__ int srp[kMaxStreamRows + 1];
    for (int t = int(blockIdx.x); t < ntiles; t += int(gridDim.x)) {
        const STile tl = tiles[t];
        if (stopped(ctl + tl.seg)) continue;
        double acc[1];
        csr_stream_rows<double>(tl.row0, tl.rows, rp, ci, v, [&](int c) { return p[c]; }, prod, srp, acc);
        double s = 0;
        if (int(threadIdx.x) < tl.rows) {
            const int row = tl.row0 + int(threadIdx.x);
            q[row] = acc[0];
            s = p[row] * acc[0];
        }
        s = block_sum(s);
        if (threadIdx.x == 0) partials[t] = s;
    }
}

// r = b - A x and sum r^2 per tile: all segments (all != 0) or the running
// ones whose control block asks for a true-residual refresh.
__global__ __launch_bounds__(kBlock) void k_strue(const STile* __restrict__ tiles, int ntiles,
                                                  const int* __restrict__ rp, const int* __restrict__ ci,
                                                  const double* __restrict__ v, const double* __restrict__ b,
                                                  const double* __restrict__ x, double* __restrict__ r,
                                                  double* __restrict__ partials, const KCtl* __restrict__ ctl,
                                                  int all) {
    __shared__ double prod[StreamCap<double>::value];
    __shared__ int srp[kMaxStreamRows + 1];
    for (int t = int(blockIdx.x); t < ntiles; t += int(gridDim.x)) {
        const STile tl = tiles[t];
        if (!all) {
            const KCtl* c = ctl + tl.seg;
            if (stopped(c) || !c->check || c->early) continue;
        }
        double acc[1];
        csr_stream_rows<double>(tl.row0, tl.rows, rp, ci, v, [&](int c) { return x[c]; }, prod, srp, acc);
        double s = 0;
        if (int(threadIdx.x) < tl.rows) {
            const int row = tl.row0 + int(threadIdx.x);
            const double res = b[row] - acc[0];
            r[row] = res;
            s = res * res;
        }
        s = block_sum(s);
        if (threadIdx.x == 0) partials[t] = s;
    }
}

// p = z (fresh) or z + beta p, per segment.
__global__ __launch_bounds__(kBlock) void k_scg_p(const STile* __restrict__ tiles, int ntiles,
                                                  const double* __restrict__ z, double* __restrict__ p,
                                                  const KCtl* __restrict__ ctl) {
    for (int t = int(blockIdx.x); t < ntiles; t += int(gridDim.x)) {
        const STile tl = tiles[t];
        const KCtl* c = ctl + tl.seg;
        if (stopped(c) || int(threadIdx.x) >= tl.rows) continue;
        const int row = tl.row0 + int(threadIdx.x);
        p[row] = c->fresh ? z[row] : z[row] + c->beta * p[row];
    }
}

// x += alpha p; r -= alpha q; sum r^2 per tile.
__global__ __launch_bounds__(kBlock) void k_scg_update(const STile* __restrict__ tiles, int ntiles,
                                                       const double* __restrict__ p, const double* __restrict__ q,
                                                       double* __restrict__ x, double* __restrict__ r,
                                                       double* __restrict__ partials, const KCtl* __restrict__ ctl) {
    for (int t = int(blockIdx.x); t < ntiles; t += int(gridDim.x)) {
        const STile tl = tiles[t];
        const KCtl* c = ctl + tl.seg;
        if (stopped(c)) continue;
        const double alpha = c->alpha;
        double s = 0;
        if (int(threadIdx.x) < tl.rows) {
            const int row = tl.row0 + int(threadIdx.x);
            x[row] = alpha * p[row] + x[row];
            const double ri = -alpha * q[row] + r[row];
            r[row] = ri;
            s = ri * ri;
        }
        s = block_sum(s);
        if (threadIdx.x == 0) partials[t] = s;
    }
}

// A segment's tile partials [lo, hi), lane-strided: lane L adds entries
// lo + L, lo + L + 64, ... in order (the caller's wave_sum then combines the
// lanes); the loads go out 8 at a time before their adds.
__device__ __forceinline__ double lane_strided_sum(const double* __restrict__ p, int lo, int hi) {
    const int lane = int(threadIdx.x) & (kWave - 1);
    double sum = 0;
    for (int i0 = lo + lane; i0 < hi; i0 += 8 * kWave) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[min(i0 + u * kWave, hi - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + u * kWave < hi) sum += v[u];
    }
    return sum;
}

// Per segment, one single-wave block: its tiles' partials added in tile
// order (lane_strided_sum, then the wave reduction), then the recurrence
// (mode >= 0) or the plain sum written to out (mode < 0).  (One 1024-thread
// block walking the segments took ~22 us a launch with 32 segments.)
__global__ __launch_bounds__(kWave) void k_sfinal(const int* __restrict__ tileOff,
                                                  const double* __restrict__ partials, KCtl* ctl, int mode,
                                                  double* __restrict__ out) {
    const int lane = int(threadIdx.x) & (kWave - 1);
    const int s = int(blockIdx.x);
    if (mode >= 0) {
        const KCtl* c = ctl + s;
        if (stopped(c)) return;
        if (mode == kFinTrue && (!c->check || c->early)) return;
    }
    double sum = lane_strided_sum(partials, tileOff[s], tileOff[s + 1]);
    sum = wave_sum(sum);
    if (lane == 0) {
        if (mode >= 0)
            finalize(mode, sum, ctl + s);
        else
            out[s] = sum;
    }
}

// Preconditioner input: w = r on running segments, 0 on stopped ones, so a
// segment that stopped on a non-finite value feeds nothing into the shared
// cycle.
__global__ __launch_bounds__(kBlock) void k_smask(const STile* __restrict__ tiles, int ntiles,
                                                  const double* __restrict__ r, double* __restrict__ w,
                                                  const KCtl* __restrict__ ctl) {
    for (int t = int(blockIdx.x); t < ntiles; t += int(gridDim.x)) {
        const STile tl = tiles[t];
        if (int(threadIdx.x) >= tl.rows) continue;
        const int row = tl.row0 + int(threadIdx.x);
        w[row] = stopped(ctl + tl.seg) ? 0.0 : r[row];
    }
}

}  // namespace

// ------------------------------------------------------------------------
// The whole segmented solve in one cooperative launch: initial residuals,
// every CG iteration (the cycle included) and the final residuals run as
// phases of one persistent grid, separated by grid-wide barriers.  Each
// phase repeats the arithmetic of the per-kernel path above (the same tile
// partials and block sums, the same per-segment reductions and recurrences,
// the cycle's operations in the order AmgEngine::cycle issues them), so both
// paths give the same iterates.
namespace {

// A level of the hierarchy as the persistent solve reads it: CSR in level
// precision; x[0] and x[1] are the iterate and its Jacobi partner.
struct PLevel {
    int n, tiles;  // rows, 256-row tiles
    const int* rp;
    const int* ci;
    const void* v;
    const void* inv;
    void* b;
    void* x[2];
    void* r;
    int nc;  // rows of the next level (0 on the coarsest)
    const int* map;
    const int* off;
    const int* mem;
};

// A cycle operation (the TailOp kinds, plus a prolongation fused with the
// sweep that follows it: in = coarse buffer, out = the fine iterate it
// corrects, the sweep writing the other buffer) and whether block 0 runs it
// alone (small levels: no grid barrier between consecutive solo operations).
constexpr int kPProlongSweep = 16;
constexpr int kSoloRows = 512;
struct POp {
    int kind, level, in, out, flag, solo;
};

struct PSolve {
    const STile* tiles;
    const int* tileOff;
    int ntiles, segments;
    const int* rp;  // the FP64 operator
    const int* ci;
    const double* v;
    const double* b;
    double* x;
    double* r;
    double* z;
    double* p;
    double* q;
    double* parts;   // one slot per tile
    double* parts2;  // second slot per tile (initial b.b)
    KCtl* ctl;
    double* sums;    // final r.r per segment
    int* counters;   // [0] running segments; [1], [2] refresh requests by iteration parity
    unsigned* barrier;  // kBarWords zeroed words of the grid barrier
    const PLevel* levels;
    const POp* ops;
    int nops, cur0, zero0, first, amg;
    const double* inverse;  // column-major coarse inverse (factored coarsest)
    double rtol, atol;
    int refresh, maxit;
};

// Buffer k (0 or 1) of a level, without indexing a private array.
__device__ __forceinline__ void* buf(const PLevel& l, int k) { return k ? l.x[1] : l.x[0]; }

// Block order with one contiguous run of workers per XCD (grid: a multiple
// of kXcds), so neighbouring tiles of a phase share an L2.
__device__ __forceinline__ int xcd_order() {
    const int per = int(gridDim.x) / kXcds;
    return int(blockIdx.x % kXcds) * per + int(blockIdx.x / kXcds);
}

// Grid-wide barrier of the cooperative grid (all blocks resident): blocks
// arrive on the counter of their group (blockIdx % kBarGroups); the last
// block of a group arrives on the top counter, and the last group releases
// every group's word.  Counters only grow (barrier k waits for k arrivals a
// member), so nothing is reset between barriers.  Thread 0 of each block
// makes the block's writes visible device-wide before it arrives (agent-scope
// release: the L2 write-back that other XCDs need) and waits with relaxed
// loads, then acquires once (one L2/L1 invalidate instead of one per poll):
// measured on MI355X at 768 blocks, ~12 us against ~60 us for the
// cooperative-groups grid sync, whose polls each invalidate the caches.
constexpr int kBarGroups = 32;
constexpr int kBarStride = 64;  // counters 256 bytes apart
constexpr int kBarWords = kBarStride * (1 + 2 * kBarGroups);
struct GridBarrier {
    unsigned* words;  // [0] top, [1 + g] group arrivals, [1 + G + g] group release
    unsigned k = 0;
    __device__ void sync() {
        ++k;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned g = blockIdx.x % kBarGroups;
            const unsigned members = (gridDim.x - g + kBarGroups - 1) / kBarGroups;
            const unsigned groups = gridDim.x < unsigned(kBarGroups) ? gridDim.x : unsigned(kBarGroups);
            unsigned* rel = words + kBarStride * (1 + kBarGroups);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            const unsigned old =
                __hip_atomic_fetch_add(words + kBarStride * (1 + g), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (old + 1 == k * members) {
                // the group's last block passes on what its members released
                // (acquire their arrivals, release them with its own)
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                const unsigned t = __hip_atomic_fetch_add(words, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (t + 1 == k * groups) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                    for (unsigned q = 0; q < groups; ++q)
                        __hip_atomic_store(rel + kBarStride * q, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            while (__hip_atomic_load(rel + kBarStride * g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < k)
                __builtin_amdgcn_s_sleep(1);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
    }
};

__device__ __forceinline__ int load_counter(int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Initial b.b (parts2) and r = b - A x with r.r (parts) per tile (k_sdot and
// k_strue over all segments).
__device__ void p_init_tiles(const PSolve& P, double* prod, int* srp) {
    for (int t = xcd_order(); t < P.ntiles; t += int(gridDim.x)) {
        const STile tl = P.tiles[t];
        const bool own = int(threadIdx.x) < tl.rows;
        const int row = tl.row0 + int(threadIdx.x);
        double s = 0;
        if (own) s = P.b[row] * P.b[row];
        s = block_sum(s);
        if (threadIdx.x == 0) P.parts2[t] = s;
        double acc[1];
        csr_stream_rows<double>(tl.row0, tl.rows, P.rp, P.ci, P.v, [&](int c) { return P.x[c]; }, prod, srp, acc);
        double s2 = 0;
        if (own) {
            const double res = P.b[row] - acc[0];
            P.r[row] = res;
            s2 = res * res;
        }
        s2 = block_sum(s2);
        if (threadIdx.x == 0) P.parts[t] = s2;
    }
}

// Per segment, one wave: norms, tolerance and status (the host start of
// apply_segmented); running segments are counted.
__device__ void p_init_segments(const PSolve& P) {
    const int lane = int(threadIdx.x) & (kWave - 1);
    const int waves = int(gridDim.x) * (kBlock / kWave);
    for (int s = int(blockIdx.x) * (kBlock / kWave) + int(threadIdx.x) / kWave; s < P.segments; s += waves) {
        double bb = lane_strided_sum(P.parts2, P.tileOff[s], P.tileOff[s + 1]);
        double rr = lane_strided_sum(P.parts, P.tileOff[s], P.tileOff[s + 1]);
        bb = wave_sum(bb);
        rr = wave_sum(rr);
        if (lane == 0) {
            const double normb = sqrt(bb), res = sqrt(rr);
            const double rel = P.rtol * normb;
            KCtl c;
            c.normb = normb;
            c.tol = P.atol < rel ? rel : P.atol;
            c.res = res;
            c.rho = c.rho_prev = c.beta = c.den = c.ns = 0;
            c.alpha = 1.0;
            c.omega = 1.0;
            c.iter = 0;
            c.fresh = 1;
            c.check = 0;
            c.early = 0;
            c.refresh = P.refresh;
            if (!isfinite(normb) || !isfinite(res))
                c.status = kNonFinite;
            else if (res <= c.tol)
                c.status = kConverged;
            else {
                c.status = kRunning;
                atomicAdd(P.counters, 1);
            }
            P.ctl[s] = c;
        }
    }
}

// Per segment, one wave: its tiles' partials in tile order, then the
// recurrence (k_sfinal); mode < 0 writes the plain sum to sums.  A segment
// that stops is taken off the running count; with `checks`, a segment that
// asks for a true-residual refresh is counted there.
__device__ void p_finalize(const PSolve& P, int mode, int* checks) {
    const int lane = int(threadIdx.x) & (kWave - 1);
    const int waves = int(gridDim.x) * (kBlock / kWave);
    for (int s = int(blockIdx.x) * (kBlock / kWave) + int(threadIdx.x) / kWave; s < P.segments; s += waves) {
        KCtl* c = P.ctl + s;
        if (mode >= 0) {
            if (stopped(c)) continue;
            if (mode == kFinTrue && (!c->check || c->early)) continue;
        }
        double sum = lane_strided_sum(P.parts, P.tileOff[s], P.tileOff[s + 1]);
        sum = wave_sum(sum);
        if (lane == 0) {
            if (mode >= 0) {
                finalize(mode, sum, c);
                if (stopped(c))
                    atomicSub(P.counters, 1);
                else if (checks && c->check && !c->early)
                    atomicAdd(checks, 1);
            }
            else {
                P.sums[s] = sum;
            }
        }
    }
}

// Level entry of the preconditioner input (k_smask, then k_enter): b0 = r on
// running segments, 0 on stopped ones; the first sweep from zero when fused.
template <class L>
__device__ void p_enter(const PSolve& P, L w) {
#pragma clang fp contract(off)
    const PLevel l0 = P.levels[0];
    L* b0 = static_cast<L*>(l0.b);
    L* x0 = static_cast<L*>(l0.x[0]);
    const L* inv = static_cast<const L*>(l0.inv);
    for (int t = xcd_order(); t < P.ntiles; t += int(gridDim.x)) {
        const STile tl = P.tiles[t];
        if (int(threadIdx.x) >= tl.rows) continue;
        const int row = tl.row0 + int(threadIdx.x);
        const double wv = stopped(P.ctl + tl.seg) ? 0.0 : P.r[row];
        const L bi = L(wv);
        b0[row] = bi;
        if (P.first) {
            const L wd = w * inv[row];
            x0[row] = wd * bi;
        }
    }
}

// z = the cycle's level-0 iterate (or r without a preconditioner) and r.z
// per tile of a running segment.
template <class L>
__device__ void p_out(const PSolve& P) {
    const L* xl = P.amg ? static_cast<const L*>(buf(P.levels[0], P.cur0)) : nullptr;
    for (int t = xcd_order(); t < P.ntiles; t += int(gridDim.x)) {
        const STile tl = P.tiles[t];
        if (stopped(P.ctl + tl.seg)) continue;
        double s = 0;
        if (int(threadIdx.x) < tl.rows) {
            const int row = tl.row0 + int(threadIdx.x);
            const double zi = !P.amg ? P.r[row] : (P.zero0 ? 0.0 : double(xl[row]));
            P.z[row] = zi;
            s = P.r[row] * zi;
        }
        s = block_sum(s);
        if (threadIdx.x == 0) P.parts[t] = s;
    }
}

// p = z (fresh) or z + beta p (k_scg_p).
__device__ void p_dir(const PSolve& P) {
    for (int t = xcd_order(); t < P.ntiles; t += int(gridDim.x)) {
        const STile tl = P.tiles[t];
        const KCtl* c = P.ctl + tl.seg;
        if (stopped(c) || int(threadIdx.x) >= tl.rows) continue;
        const int row = tl.row0 + int(threadIdx.x);
        P.p[row] = c->fresh ? P.z[row] : P.z[row] + c->beta * P.p[row];
    }
}

// q = A p and p.q per tile (k_sspmv_dot).
__device__ void p_spmv_dot(const PSolve& P, double* prod, int* srp) {
    for (int t = xcd_order(); t < P.ntiles; t += int(gridDim.x)) {
        const STile tl = P.tiles[t];
        if (stopped(P.ctl + tl.seg)) continue;
        double acc[1];
        csr_stream_rows<double>(tl.row0, tl.rows, P.rp, P.ci, P.v, [&](int c) { return P.p[c]; }, prod, srp, acc);
        double s = 0;
        if (int(threadIdx.x) < tl.rows) {
            const int row = tl.row0 + int(threadIdx.x);
            P.q[row] = acc[0];
            s = P.p[row] * acc[0];
        }
        s = block_sum(s);
        if (threadIdx.x == 0) P.parts[t] = s;
    }
}

// x += alpha p; r -= alpha q; r.r per tile (k_scg_update).
__device__ void p_update(const PSolve& P) {
    for (int t = xcd_order(); t < P.ntiles; t += int(gridDim.x)) {
        const STile tl = P.tiles[t];
        const KCtl* c = P.ctl + tl.seg;
        if (stopped(c)) continue;
        const double alpha = c->alpha;
        double s = 0;
        if (int(threadIdx.x) < tl.rows) {
            const int row = tl.row0 + int(threadIdx.x);
            P.x[row] = alpha * P.p[row] + P.x[row];
            const double ri = -alpha * P.q[row] + P.r[row];
            P.r[row] = ri;
            s = ri * ri;
        }
        s = block_sum(s);
        if (threadIdx.x == 0) P.parts[t] = s;
    }
}

// r = b - A x and r.r per tile: every segment (all) or the running ones that
// ask for a true-residual refresh (k_strue).
__device__ void p_true(const PSolve& P, double* prod, int* srp, int all) {
    for (int t = xcd_order(); t < P.ntiles; t += int(gridDim.x)) {
        const STile tl = P.tiles[t];
        if (!all) {
            const KCtl* c = P.ctl + tl.seg;
            if (stopped(c) || !c->check || c->early) continue;
        }
        double acc[1];
        csr_stream_rows<double>(tl.row0, tl.rows, P.rp, P.ci, P.v, [&](int c) { return P.x[c]; }, prod, srp, acc);
        double s = 0;
        if (int(threadIdx.x) < tl.rows) {
            const int row = tl.row0 + int(threadIdx.x);
            const double res = P.b[row] - acc[0];
            P.r[row] = res;
            s = res * res;
        }
        s = block_sum(s);
        if (threadIdx.x == 0) P.parts[t] = s;
    }
}

// x = A^-1 b on the coarsest level by one 256-thread block: the arithmetic
// of coarse_inverse_apply (the same column runs, partials and run order),
// its 1024 threads' work taken four at a time.
template <class T>
__device__ void p_coarse(int n, const double* it, const T* b, T* x, double* bs, double* part) {
    const int tid = int(threadIdx.x);
    for (int i = tid; i < n; i += kBlock) bs[i] = double(b[i]);
    __syncthreads();
    const int segs = coarse_segments(n), run = coarse_run(n);
    for (int vt = tid; vt < segs * n; vt += kBlock) {
        const int s = vt / n, r = vt - s * n;
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
        part[vt] = p;
    }
    __syncthreads();
    for (int r = tid; r < n; r += kBlock) {
        double acc = part[r];
        for (int s = 1; s < segs; ++s) acc += part[s * n + r];
        x[r] = T(acc);
    }
}

// One cycle operation over workers wid, wid + wcount, ... (blocks of the
// grid, or block 0 alone with wid 0, wcount 1), in the arithmetic of the
// per-level kernels (k_sell / k_csr_stream epilogues, k_jacobi_zero,
// k_restrict, k_prolong, k_prolong_jacobi, the coarse inverse apply).
template <class L>
__device__ void p_op(const PSolve& P, const POp op, L w, int wid, int wcount, L* prod, int* srp, double* dbuf) {
#pragma clang fp contract(off)
    const PLevel lv = P.levels[op.level];
    const int n = lv.n, tid = int(threadIdx.x);
    const L* b = static_cast<const L*>(lv.b);
    const L* inv = static_cast<const L*>(lv.inv);
    auto rows = [&](auto&& xval, auto&& f) {
        for (int t = wid; t < lv.tiles; t += wcount) {
            const int row0 = t * kMaxStreamRows, cnt = min(kMaxStreamRows, n - row0);
            L acc[1];
            csr_stream_rows<L>(row0, cnt, lv.rp, lv.ci, static_cast<const L*>(lv.v), xval, prod, srp, acc);
            if (tid < cnt) f(row0 + tid, acc[0]);
        }
    };
    switch (op.kind) {
    case kTSweep: {
        const L* xi = static_cast<const L*>(buf(lv, op.in));
        L* xo = static_cast<L*>(buf(lv, op.out));
        rows([&](int c) { return xi[c]; },
             [&](int i, L acc) {
                 const L wd = w * inv[i];
                 const L d = b[i] - acc;
                 const L upd = wd * d;
                 xo[i] = xi[i] + upd;
             });
        break;
    }
    case kTResidual: {
        const L* xi = static_cast<const L*>(buf(lv, op.in));
        L* r = static_cast<L*>(lv.r);
        rows([&](int c) { return xi[c]; }, [&](int i, L acc) { r[i] = b[i] - acc; });
        break;
    }
    case kTSweepZero: {
        L* xo = static_cast<L*>(buf(lv, op.out));
        for (int i = wid * kBlock + tid; i < n; i += wcount * kBlock) {
            const L wd = w * inv[i];
            xo[i] = wd * b[i];
        }
        break;
    }
    case kTRestrict: {
        const L* src = op.in ? static_cast<const L*>(lv.r) : b;
        const PLevel nx = P.levels[op.level + 1];
        L* nb = static_cast<L*>(nx.b);
        const L* ninv = static_cast<const L*>(nx.inv);
        L* nxo = static_cast<L*>(buf(nx, op.out));
        for (int c = wid * kBlock + tid; c < lv.nc; c += wcount * kBlock) {
            L s = 0;
            for (int m = lv.off[c]; m < lv.off[c + 1]; ++m) s += src[lv.mem[m]];
            nb[c] = s;
            if (op.flag) {
                const L wd = w * ninv[c];
                nxo[c] = wd * s;
            }
        }
        break;
    }
    case kTProlong: {
        const L* xc = static_cast<const L*>(buf(P.levels[op.level + 1], op.in));
        L* xo = static_cast<L*>(buf(lv, op.out));
        for (int i = wid * kBlock + tid; i < n; i += wcount * kBlock)
            xo[i] = op.flag ? xc[lv.map[i]] : xo[i] + xc[lv.map[i]];
        break;
    }
    case kPProlongSweep: {
        const L* xc = static_cast<const L*>(buf(P.levels[op.level + 1], op.in));
        const L* xf = static_cast<const L*>(buf(lv, op.out));
        L* xo = static_cast<L*>(buf(lv, 1 - op.out));
        const int* map = lv.map;
        const bool set = op.flag != 0;
        auto xp = [&](int c) -> L { return set ? xc[map[c]] : xf[c] + xc[map[c]]; };
        rows(xp, [&](int i, L acc) {
            const L xr = xp(i);
            const L wd = w * inv[i];
            const L d = b[i] - acc;
            const L upd = wd * d;
            xo[i] = xr + upd;
        });
        break;
    }
    case kTCoarse:
        p_coarse<L>(n, P.inverse, b, static_cast<L*>(buf(lv, op.out)), dbuf, dbuf + kCoarseMax);
        break;
    case kTCopy: {
        const L* xi = static_cast<const L*>(buf(lv, op.in));
        L* xo = static_cast<L*>(buf(lv, op.out));
        for (int i = wid * kBlock + tid; i < n; i += wcount * kBlock) xo[i] = xi[i];
        break;
    }
    default:
        break;
    }
}

template <class L>
__global__ __launch_bounds__(kBlock) void k_seg_solve(PSolve P, L w) {
    // 16 KB: the CSR-stream products (FP64 or level precision) or the
    // coarse apply's staged rhs and partials
    __shared__ __align__(16) double dbuf[2 * kCoarseMax];
    __shared__ int srp[kMaxStreamRows + 1];
    L* lprod = reinterpret_cast<L*>(dbuf);
    const int wid = xcd_order(), wcount = int(gridDim.x);
    GridBarrier bar{P.barrier};
    p_init_tiles(P, dbuf, srp);
    bar.sync();
    p_init_segments(P);
    bar.sync();
    for (int it = 0; it < P.maxit; ++it) {
        if (load_counter(P.counters) == 0) break;
        if (P.amg) {
            p_enter<L>(P, w);
            bar.sync();
            for (int k = 0; k < P.nops;) {
                const POp op = P.ops[k];
                if (op.solo) {
                    int e = k + 1;
                    while (e < P.nops && P.ops[e].solo) ++e;
                    if (blockIdx.x == 0) {
                        for (int j = k; j < e; ++j) {
                            p_op<L>(P, P.ops[j], w, 0, 1, lprod, srp, dbuf);
                            __syncthreads();
                        }
                    }
                    k = e;
                }
                else {
                    p_op<L>(P, op, w, wid, wcount, lprod, srp, dbuf);
                    ++k;
                }
                bar.sync();
            }
        }
        p_out<L>(P);
        bar.sync();
        p_finalize(P, kFinRhoCg, nullptr);
        // the refresh counter of the next iteration: its last read (two
        // iterations back) is behind at least one barrier
        if (blockIdx.x == 0 && threadIdx.x == 0) P.counters[1 + ((it + 1) & 1)] = 0;
        bar.sync();
        p_dir(P);
        bar.sync();
        p_spmv_dot(P, dbuf, srp);
        bar.sync();
        p_finalize(P, kFinCurvCg, nullptr);
        bar.sync();
        p_update(P);
        bar.sync();
        int* checks = P.counters + 1 + (it & 1);
        p_finalize(P, kFinUpdCg, checks);
        bar.sync();
        if (load_counter(checks) > 0) {
            p_true(P, dbuf, srp, 0);
            bar.sync();
            p_finalize(P, kFinTrue, nullptr);
            bar.sync();
        }
    }
    p_true(P, dbuf, srp, 1);
    bar.sync();
    p_finalize(P, -1, nullptr);
}

}  // namespace

// Device and pinned work of a segment layout, kept while the offsets stay
// the same: tiles, tile offsets per segment, segment offsets, control
// blocks, sums, per-tile partials and the vectors b, x, r, z, p, q, w.
struct SolverEngine::Segments {
    std::vector<int> off;
    int segments = 0, tiles = 0;
    void* dev = nullptr;
    void* host = nullptr;
    STile* dTiles = nullptr;
    int* dTileOff = nullptr;
    int* dOff = nullptr;
    KCtl* ctl = nullptr;
    double* sums = nullptr;
    double* parts = nullptr;
    double* parts2 = nullptr;  // second per-tile slot (the persistent solve's b.b)
    int* counters = nullptr;   // the persistent solve's running / refresh counters
    double* v[7] = {};
    KCtl* hctl = nullptr;
    double* hsums = nullptr;
    // the persistent solve's level table and cycle program, made for one
    // preconditioner generation
    void* plan = nullptr;
    uint64_t planKey = ~uint64_t(0);
    int planLevels = 0, planOps = 0, planCur = 0, planZero = 0, planFirst = 0;
};

void SolverEngine::release_segments() {
    sgraph_.reset();
    if (!seg_) return;
    (void)hipStreamSynchronize(c_->stream);
    dev_free(*c_, seg_->dev);
    dev_free(*c_, seg_->plan);
    if (seg_->host) (void)hipHostFree(seg_->host);
    delete seg_;
    seg_ = nullptr;
}

// One segmented CG iteration (Cg.cuh:138-187 per segment) as a fixed launch
// sequence over every segment's tiles; stopped segments' tiles are skipped.
void SolverEngine::launch_segment_iteration() {
    Segments& S = *seg_;
    const int T = S.tiles, segments = S.segments;
    const dim3 grid(unsigned(std::max(1, std::min(T, kSegGrid))));
    double *sB = S.v[0], *sX = S.v[1], *sR = S.v[2], *sZ = S.v[3], *sP = S.v[4], *sQ = S.v[5], *sW = S.v[6];
    const KCtl* cctl = S.ctl;
    auto fin = [&](int mode) {
        hipLaunchKernelGGL(k_sfinal, dim3(unsigned(segments)), dim3(kWave), 0, c_->stream, S.dTileOff, S.parts,
                           S.ctl, mode, static_cast<double*>(nullptr));
    };
    hipLaunchKernelGGL(k_smask, grid, dim3(kBlock), 0, c_->stream, S.dTiles, T, sR, sW, cctl);
    precondition(sW, sZ);
    hipLaunchKernelGGL(k_sdot, grid, dim3(kBlock), 0, c_->stream, S.dTiles, T, sR, sZ, S.parts, cctl);
    fin(kFinRhoCg);
    hipLaunchKernelGGL(k_scg_p, grid, dim3(kBlock), 0, c_->stream, S.dTiles, T, sZ, sP, cctl);
    hipLaunchKernelGGL(k_sspmv_dot, grid, dim3(kBlock), 0, c_->stream, S.dTiles, T, A_.rp, A_.ci, A_.v, sP, sQ,
                       S.parts, cctl);
    fin(kFinCurvCg);
    hipLaunchKernelGGL(k_scg_update, grid, dim3(kBlock), 0, c_->stream, S.dTiles, T, sP, sQ, sX, sR, S.parts, cctl);
    fin(kFinUpdCg);
    hipLaunchKernelGGL(k_strue, grid, dim3(kBlock), 0, c_->stream, S.dTiles, T, A_.rp, A_.ci, A_.v, sB, sX, sR,
                       S.parts, cctl, 0);
    fin(kFinTrue);
    check(hipGetLastError(), "segmented iteration launch");
}

void SolverEngine::prepare_segments(int segments, const int* off) {
    require(n_ > 0, "CG requires setup", TF_E_INVALID);
    require(kind_ == 0, "the segmented batch solve runs CG", TF_E_INVALID);
    require(segments >= 1 && off != nullptr, "segmented solve needs at least one segment");
    require(off[0] == 0 && off[segments] == n_, "segment offsets must start at 0 and end at the row count");
    for (int s = 0; s < segments; ++s) require(off[s + 1] >= off[s], "segment offsets must be nondecreasing");
    if (seg_ && seg_->segments == segments && std::equal(off, off + segments + 1, seg_->off.begin())) return;
    release_segments();
    std::vector<STile> tiles;
    std::vector<int> tileOff(size_t(segments) + 1, 0);
    for (int s = 0; s < segments; ++s) {
        tileOff[size_t(s)] = int(tiles.size());
        for (int r = off[s]; r < off[s + 1]; r += kTileRows) tiles.push_back({r, std::min(kTileRows, off[s + 1] - r), s, 0});
    }
    tileOff[size_t(segments)] = int(tiles.size());
    const int T = int(tiles.size());
    auto rnd = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t bTiles = rnd(sizeof(STile) * size_t(std::max(T, 1)));
    const size_t bTOff = rnd(sizeof(int) * (size_t(segments) + 1));
    const size_t bCtl = rnd(sizeof(KCtl) * size_t(segments));
    const size_t bSums = rnd(sizeof(double) * size_t(segments));
    const size_t bParts = rnd(sizeof(double) * size_t(std::max(T, 1)));
    const size_t bCount = 256 + sizeof(unsigned) * size_t(kBarWords);
    const size_t bVec = rnd(sizeof(double) * size_t(n_));
    seg_ = new Segments();
    Segments& S = *seg_;
    S.off.assign(off, off + segments + 1);
    S.segments = segments;
    S.tiles = T;
    S.dev = dev_alloc<char>(*c_, bTiles + 2 * bTOff + bCtl + bSums + 2 * bParts + bCount + 7 * bVec);
    check(hipHostMalloc(&S.host, bCtl + bSums, hipHostMallocDefault), "hipHostMalloc segment control");
    char* p = static_cast<char*>(S.dev);
    auto take = [&](size_t bytes) {
        char* q = p;
        p += bytes;
        return q;
    };
    S.dTiles = reinterpret_cast<STile*>(take(bTiles));
    S.dTileOff = reinterpret_cast<int*>(take(bTOff));
    S.dOff = reinterpret_cast<int*>(take(bTOff));
    S.ctl = reinterpret_cast<KCtl*>(take(bCtl));
    S.sums = reinterpret_cast<double*>(take(bSums));
    S.parts = reinterpret_cast<double*>(take(bParts));
    S.parts2 = reinterpret_cast<double*>(take(bParts));
    S.counters = reinterpret_cast<int*>(take(bCount));
    for (auto& q : S.v) q = reinterpret_cast<double*>(take(bVec));
    S.hctl = static_cast<KCtl*>(S.host);
    S.hsums = reinterpret_cast<double*>(static_cast<char*>(S.host) + bCtl);
    check(hipMemcpyAsync(S.dTiles, tiles.data(), sizeof(STile) * size_t(T), hipMemcpyHostToDevice, c_->stream),
          "tile upload");
    check(hipMemcpyAsync(S.dTileOff, tileOff.data(), sizeof(int) * tileOff.size(), hipMemcpyHostToDevice,
                         c_->stream),
          "tile offset upload");
    check(hipMemcpyAsync(S.dOff, off, sizeof(int) * (size_t(segments) + 1), hipMemcpyHostToDevice, c_->stream),
          "segment offset upload");
    // no entry may couple two segments
    c_->clear_flags(1);
    hipLaunchKernelGGL(k_seg_check, dim3(unsigned(std::max(1, std::min(T, kSegGrid)))), dim3(kBlock), 0, c_->stream,
                       S.dTiles, T, A_.rp, A_.ci, S.dOff, c_->dflags);
    check(hipGetLastError(), "segment check launch");
    c_->read_flags(1);
    if (c_->hflags[0] != 0) {
        release_segments();
        throw Error(TF_E_INVALID, "the operator couples two segments (it must be block diagonal)");
    }
}

// Statuses as finish() decides them (Cg.cuh:105-121) from the control blocks
// (S.hctl) and the final residuals' squares.
void SolverEngine::segment_results(const std::vector<double>& normb, const std::vector<double>& tol,
                                   const std::vector<double>& rr, tf_solve_result* out) {
    const KCtl* h = seg_->hctl;
    for (int s = 0; s < seg_->segments; ++s) {
        const double r = std::sqrt(rr[size_t(s)]);
        int st = h[s].status == kRunning ? kMaxIt : h[s].status;
        const int its = h[s].status == kRunning ? o_.maxit : h[s].iter;
        if (!std::isfinite(r) || !std::isfinite(normb[size_t(s)]))
            st = kNonFinite;
        else if (r <= tol[size_t(s)])
            st = kConverged;
        else if (st == kConverged)
            st = kMaxIt;
        out[s].status = st;
        out[s].iterations = its;
        out[s].residual_norm = r;
        out[s].relative_residual = normb[size_t(s)] > 0 ? r / normb[size_t(s)] : r;
    }
}

// CG on every segment of the block-diagonal operator at once.  off: host
// row offsets of the segments (segments + 1 entries, off[0] = 0,
// off[segments] = n, nondecreasing).  The iteration is captured once per
// segment layout as a hipGraph and launched in batches between reads of
// the control blocks (krylov_batch); every segment stops on its own test.
void SolverEngine::apply_segmented(int segments, const int* off, const double* b, double* x,
                                   tf_solve_result* out) {
    prepare_segments(segments, off);
    Segments& S = *seg_;
    const int T = S.tiles;
    const dim3 grid(unsigned(std::max(1, std::min(T, kSegGrid))));
    double *sB = S.v[0], *sX = S.v[1], *sR = S.v[2];
    KCtl* h = S.hctl;
    convert<double, double>(*c_, size_t(n_), b, sB);
    convert<double, double>(*c_, size_t(n_), x, sX);

    auto sums = [&](double* dst) {
        hipLaunchKernelGGL(k_sfinal, dim3(unsigned(segments)), dim3(kWave), 0, c_->stream, S.dTileOff, S.parts,
                           S.ctl, -1, S.sums);
        check(hipMemcpyAsync(S.hsums, S.sums, sizeof(double) * size_t(segments), hipMemcpyDeviceToHost, c_->stream),
              "segment sums download");
        c_->sync();
        for (int s = 0; s < segments; ++s) dst[s] = S.hsums[s];
    };
    std::vector<double> normb(static_cast<size_t>(segments)), res(static_cast<size_t>(segments));
    hipLaunchKernelGGL(k_sdot, grid, dim3(kBlock), 0, c_->stream, S.dTiles, T, sB, sB, S.parts,
                       static_cast<const KCtl*>(nullptr));
    sums(normb.data());
    hipLaunchKernelGGL(k_strue, grid, dim3(kBlock), 0, c_->stream, S.dTiles, T, A_.rp, A_.ci, A_.v, sB, sX, sR,
                       S.parts, static_cast<const KCtl*>(nullptr), 1);
    sums(res.data());
    std::vector<double> tol(static_cast<size_t>(segments));
    int running = 0;
    for (int s = 0; s < segments; ++s) {
        normb[size_t(s)] = std::sqrt(normb[size_t(s)]);
        res[size_t(s)] = std::sqrt(res[size_t(s)]);
        tol[size_t(s)] = std::max(o_.atol, o_.rtol * normb[size_t(s)]);
        KCtl& c = h[s];
        std::memset(&c, 0, sizeof(KCtl));
        c.normb = normb[size_t(s)];
        c.tol = tol[size_t(s)];
        c.res = res[size_t(s)];
        c.alpha = 1.0;
        c.omega = 1.0;
        c.fresh = 1;
        c.refresh = o_.refresh;
        if (!std::isfinite(normb[size_t(s)]) || !std::isfinite(res[size_t(s)]))
            c.status = kNonFinite;
        else if (res[size_t(s)] <= tol[size_t(s)])
            c.status = kConverged;
        else {
            c.status = kRunning;
            ++running;
        }
    }
    check(hipMemcpyAsync(S.ctl, h, sizeof(KCtl) * size_t(segments), hipMemcpyHostToDevice, c_->stream),
          "segment control upload");
    const std::vector<double> res0 = res;
    // the graph names the layout's buffers (released with it) and the
    // preconditioner's level buffers
    const uint64_t key = amg_ ? amg_->generation() : 0;
    int done = 0;
    while (done < o_.maxit && running > 0) {
        const int batch = std::min(krylov_batch(h, res0.data(), segments, done), o_.maxit - done);
        for (int q = 0; q < batch; ++q)
            launch_captured(*c_, sgraph_, o_.graph, key, [&] { launch_segment_iteration(); });
        done += batch;
        check(hipMemcpyAsync(h, S.ctl, sizeof(KCtl) * size_t(segments), hipMemcpyDeviceToHost, c_->stream),
              "segment control download");
        c_->sync();
        running = 0;
        for (int s = 0; s < segments; ++s) running += h[s].status == kRunning;
    }
    hipLaunchKernelGGL(k_strue, grid, dim3(kBlock), 0, c_->stream, S.dTiles, T, A_.rp, A_.ci, A_.v, sB, sX, sR,
                       S.parts, static_cast<const KCtl*>(nullptr), 1);
    sums(res.data());
    segment_results(normb, tol, res, out);
    convert<double, double>(*c_, size_t(n_), sX, x);
    c_->sync();
}

// The same solve in one cooperative launch (k_seg_solve).  The level table
// and the cycle program are made once per preconditioner generation; the
// grid is as many 256-thread blocks as can be resident together, at most 4
// per CU, a multiple of 8 (one run of blocks per XCD).
void SolverEngine::apply_segmented_persistent(int segments, const int* off, const double* b, double* x,
                                              tf_solve_result* out) {
    prepare_segments(segments, off);
    Segments& S = *seg_;
    const bool f32 = amg_ && amg_->opt_.precision == 32;
    if (amg_) {
        require(amg_->ready(), "AMG requires setup", TF_E_INVALID);
        require(amg_->n() == n_, "the preconditioner was set up on a matrix of another size", TF_E_INVALID);
        const uint64_t key = amg_->generation();
        if (!S.plan || S.planKey != key) {
            dev_free(*c_, S.plan);
            S.plan = nullptr;
            const std::vector<AmgLevel>& L = amg_->levels_;
            require(!amg_->factored() || L.back().native.n <= kCoarseMax,
                    "the persistent segmented solve needs a coarsest level of at most 1024 rows", TF_E_INVALID);
            bool first = false, zero = false;
            int cur = 0;
            const std::vector<TailOp> prog = amg_->cycle_program(first, cur, zero);
            std::vector<PLevel> lv(L.size());
            for (size_t d = 0; d < L.size(); ++d) {
                const AmgLevel& l = L[d];
                lv[d] = PLevel{l.native.n, (l.native.n + kMaxStreamRows - 1) / kMaxStreamRows, l.native.rp,
                               l.native.ci, l.vals, l.inv, l.b, {l.x, l.x2}, l.r, l.transfer.nc, l.transfer.map,
                               l.transfer.offsets, l.transfer.members};
            }
            std::vector<POp> ops;
            for (size_t k = 0; k < prog.size(); ++k) {
                const TailOp& o = prog[k];
                if (o.kind == kTProlong && k + 1 < prog.size() && prog[k + 1].kind == kTSweep &&
                    prog[k + 1].level == o.level && prog[k + 1].in == o.out && prog[k + 1].out == 1 - o.out) {
                    ops.push_back(POp{kPProlongSweep, o.level, o.in, o.out, o.flag, 0});
                    ++k;
                }
                else {
                    ops.push_back(POp{o.kind, o.level, o.in, o.out, o.flag, 0});
                }
            }
            for (POp& o : ops) o.solo = o.kind == kTCoarse || lv[size_t(o.level)].n <= kSoloRows;
            const size_t bl = (sizeof(PLevel) * lv.size() + 255) / 256 * 256;
            S.plan = dev_alloc<char>(*c_, bl + sizeof(POp) * std::max<size_t>(ops.size(), 1));
            check(hipMemcpyAsync(S.plan, lv.data(), sizeof(PLevel) * lv.size(), hipMemcpyHostToDevice, c_->stream),
                  "persistent level table upload");
            if (!ops.empty())
                check(hipMemcpyAsync(static_cast<char*>(S.plan) + bl, ops.data(), sizeof(POp) * ops.size(),
                                     hipMemcpyHostToDevice, c_->stream),
                      "persistent program upload");
            c_->sync();
            S.planKey = key;
            S.planLevels = int(lv.size());
            S.planOps = int(ops.size());
            S.planCur = cur;
            S.planZero = zero ? 1 : 0;
            S.planFirst = first ? 1 : 0;
        }
    }
    convert<double, double>(*c_, size_t(n_), b, S.v[0]);
    convert<double, double>(*c_, size_t(n_), x, S.v[1]);
    fill_zero(*c_, S.counters, 256 + sizeof(unsigned) * size_t(kBarWords));
    PSolve P{};
    P.tiles = S.dTiles;
    P.tileOff = S.dTileOff;
    P.ntiles = S.tiles;
    P.segments = segments;
    P.rp = A_.rp;
    P.ci = A_.ci;
    P.v = A_.v;
    P.b = S.v[0];
    P.x = S.v[1];
    P.r = S.v[2];
    P.z = S.v[3];
    P.p = S.v[4];
    P.q = S.v[5];
    P.parts = S.parts;
    P.parts2 = S.parts2;
    P.ctl = S.ctl;
    P.sums = S.sums;
    P.counters = S.counters;
    P.barrier = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(S.counters) + 256);
    if (amg_) {
        const size_t bl = (sizeof(PLevel) * size_t(S.planLevels) + 255) / 256 * 256;
        P.levels = static_cast<const PLevel*>(S.plan);
        P.ops = reinterpret_cast<const POp*>(static_cast<const char*>(S.plan) + bl);
        P.nops = S.planOps;
        P.cur0 = S.planCur;
        P.zero0 = S.planZero;
        P.first = S.planFirst;
        P.amg = 1;
        P.inverse = amg_->coarse_inverse();
    }
    P.rtol = o_.rtol;
    P.atol = o_.atol;
    P.refresh = o_.refresh;
    P.maxit = o_.maxit;
    const void* fn = f32 ? reinterpret_cast<const void*>(&k_seg_solve<float>)
                         : reinterpret_cast<const void*>(&k_seg_solve<double>);
    int perCu = 0, cus = 0, coop = 0;
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, fn, kBlock, 0), "persistent occupancy");
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c_->device), "CU count");
    check(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c_->device), "cooperative attribute");
    require(coop != 0 && perCu >= 1 && cus >= 1, "the device cannot run the persistent segmented solve",
            TF_E_INVALID);
    const int grid = std::max(kXcds, (std::min(perCu, 4) * cus) / kXcds * kXcds);
    float wf = float(amg_ ? amg_->opt_.omega : 1.0);
    double wd = amg_ ? amg_->opt_.omega : 1.0;
    void* args[2] = {&P, f32 ? static_cast<void*>(&wf) : static_cast<void*>(&wd)};
    check(hipLaunchCooperativeKernel(fn, dim3(unsigned(grid)), dim3(kBlock), args, 0, c_->stream),
          "persistent segmented launch");
    check(hipMemcpyAsync(S.hctl, S.ctl, sizeof(KCtl) * size_t(segments), hipMemcpyDeviceToHost, c_->stream),
          "segment control download");
    check(hipMemcpyAsync(S.hsums, S.sums, sizeof(double) * size_t(segments), hipMemcpyDeviceToHost, c_->stream),
          "segment sums download");
    c_->sync();
    const size_t ns = static_cast<size_t>(segments);
    std::vector<double> normb(ns), tol(ns), rr(ns);
    for (int s = 0; s < segments; ++s) {
        normb[size_t(s)] = S.hctl[s].normb;
        tol[size_t(s)] = S.hctl[s].tol;
        rr[size_t(s)] = S.hsums[s];
    }
    segment_results(normb, tol, rr, out);
    convert<double, double>(*c_, size_t(n_), S.v[1], x);
    c_->sync();
}

}  // namespace tf
