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

// Per segment: its tiles' partials added in tile order (lane-strided, then
// the wave reduction), then the recurrence (mode >= 0) or the plain sum
// written to out (mode < 0).
__global__ __launch_bounds__(1024) void k_sfinal(int segments, const int* __restrict__ tileOff,
                                                 const double* __restrict__ partials, KCtl* ctl, int mode,
                                                 double* __restrict__ out) {
    const int lane = int(threadIdx.x) & (kWave - 1), w = int(threadIdx.x) / kWave;
    for (int s = w; s < segments; s += 1024 / kWave) {
        if (mode >= 0) {
            const KCtl* c = ctl + s;
            if (stopped(c)) continue;
            if (mode == kFinTrue && (!c->check || c->early)) continue;
        }
        double sum = 0;
        for (int i = tileOff[s] + lane; i < tileOff[s + 1]; i += kWave) sum += partials[i];
        sum = wave_sum(sum);
        if (lane == 0) {
            if (mode >= 0)
                finalize(mode, sum, ctl + s);
            else
                out[s] = sum;
        }
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
    double* v[7] = {};
    KCtl* hctl = nullptr;
    double* hsums = nullptr;
};

void SolverEngine::release_segments() {
    sgraph_.reset();
    if (!seg_) return;
    (void)hipStreamSynchronize(c_->stream);
    dev_free(*c_, seg_->dev);
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
        hipLaunchKernelGGL(k_sfinal, dim3(1), dim3(1024), 0, c_->stream, segments, S.dTileOff, S.parts, S.ctl, mode,
                           static_cast<double*>(nullptr));
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

// CG on every segment of the block-diagonal operator at once.  off: host
// row offsets of the segments (segments + 1 entries, off[0] = 0,
// off[segments] = n, nondecreasing).  The iteration is captured once per
// segment layout as a hipGraph and launched in batches between reads of
// the control blocks (krylov_batch); every segment stops on its own test.
void SolverEngine::apply_segmented(int segments, const int* off, const double* b, double* x,
                                   tf_solve_result* out) {
    require(n_ > 0, "CG requires setup", TF_E_INVALID);
    require(kind_ == 0, "the segmented batch solve runs CG", TF_E_INVALID);
    require(segments >= 1 && off != nullptr, "segmented solve needs at least one segment");
    require(off[0] == 0 && off[segments] == n_, "segment offsets must start at 0 and end at the row count");
    for (int s = 0; s < segments; ++s) require(off[s + 1] >= off[s], "segment offsets must be nondecreasing");
    if (!seg_ || seg_->segments != segments || !std::equal(off, off + segments + 1, seg_->off.begin())) {
        release_segments();
        std::vector<STile> tiles;
        std::vector<int> tileOff(size_t(segments) + 1, 0);
        for (int s = 0; s < segments; ++s) {
            tileOff[size_t(s)] = int(tiles.size());
            for (int r = off[s]; r < off[s + 1]; r += kTileRows)
                tiles.push_back({r, std::min(kTileRows, off[s + 1] - r), s, 0});
        }
        tileOff[size_t(segments)] = int(tiles.size());
        const int T = int(tiles.size());
        auto rnd = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t bTiles = rnd(sizeof(STile) * size_t(std::max(T, 1)));
        const size_t bTOff = rnd(sizeof(int) * (size_t(segments) + 1));
        const size_t bCtl = rnd(sizeof(KCtl) * size_t(segments));
        const size_t bSums = rnd(sizeof(double) * size_t(segments));
        const size_t bParts = rnd(sizeof(double) * size_t(std::max(T, 1)));
        const size_t bVec = rnd(sizeof(double) * size_t(n_));
        seg_ = new Segments();
        Segments& S = *seg_;
        S.off.assign(off, off + segments + 1);
        S.segments = segments;
        S.tiles = T;
        S.dev = dev_alloc<char>(*c_, bTiles + 2 * bTOff + bCtl + bSums + bParts + 7 * bVec);
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
        hipLaunchKernelGGL(k_seg_check, dim3(unsigned(std::max(1, std::min(T, kSegGrid)))), dim3(kBlock), 0,
                           c_->stream, S.dTiles, T, A_.rp, A_.ci, S.dOff, c_->dflags);
        check(hipGetLastError(), "segment check launch");
        c_->read_flags(1);
        if (c_->hflags[0] != 0) {
            release_segments();
            throw Error(TF_E_INVALID, "the operator couples two segments (it must be block diagonal)");
        }
    }
    Segments& S = *seg_;
    const int T = S.tiles;
    const dim3 grid(unsigned(std::max(1, std::min(T, kSegGrid))));
    double *sB = S.v[0], *sX = S.v[1], *sR = S.v[2];
    KCtl* h = S.hctl;
    convert<double, double>(*c_, size_t(n_), b, sB);
    convert<double, double>(*c_, size_t(n_), x, sX);

    auto sums = [&](double* dst) {
        hipLaunchKernelGGL(k_sfinal, dim3(1), dim3(1024), 0, c_->stream, segments, S.dTileOff, S.parts, S.ctl, -1,
                           S.sums);
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
    for (int s = 0; s < segments; ++s) {
        const double r = std::sqrt(res[size_t(s)]);
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
    convert<double, double>(*c_, size_t(n_), sX, x);
    c_->sync();
}

}  // namespace tf
