// Preconditioner and Krylov engines behind the C ABI: the AMG hierarchy and
// V/W/F cycle (Amg.cuh, Cycle.cuh) and CG / BiCGStab (Cg.cuh, BiCgStab.cuh)
// driven by device-resident scalars.
#pragma once
#include <string>
#include <vector>

#include "tf_internal.hpp"

namespace tf {

struct AmgOptionsC {
    int maxLevels = 20, coarseRows = 64, rounds = 8, pre = 1, post = 1, coarseSweeps = 8;
    double omega = 0.75, theta = 0.25;
    bool absorb = true;
    int cycle = 0;
    int precision = 64;
    int tail = 0;     // 0 timed, 1 always, 2 never
    int learned = 0;  // 0 strength weights, 1 learned affinities, 2 learned where PCG converges faster with it
};
AmgOptionsC amg_options_from(const tf_amg_options* o);

struct AmgLevel {
    DevCsr native;          // canonical FP64 operator of this level
    void* vals = nullptr;   // values in level precision (aliases native.v for FP64)
    void* inv = nullptr;    // inverse diagonal, level precision
    void* b = nullptr;      // right-hand side
    void* x = nullptr;      // iterate
    void* x2 = nullptr;     // Jacobi ping-pong partner
    void* r = nullptr;      // residual
    DevAgg transfer;        // to the next coarser level (empty on the coarsest)
    DevSell sell;           // sliced-ELL copy for the sweeps and residuals (empty: CSR kernels)
    void* svals = nullptr;  // its values, level precision
    DevStencil st;          // structured-stencil copy (level 0 of a stencil operator; replaces sell)
    void* stvals = nullptr; // its value arrays, level precision
    bool zero = false;      // iterate is known to be zero (cleared lazily)
    bool preapplied = false; // x already holds the first sweep from zero (fused upstream)
    int64_t hist[4] = {0, 0, 0, 0};
};

// The cycle below a small level runs as one program of one 1024-thread
// workgroup (tf_tail.hip): the same operations as the per-level kernels, in
// the same order and arithmetic, without a launch per operation.
enum TailKind { kTSweep = 0, kTSweepZero = 1, kTResidual = 2, kTRestrict = 3, kTProlong = 4, kTCoarse = 5, kTCopy = 6 };
struct TailOp {
    int kind, level;
    int in, out;  // iterate buffers (0: x, 1: x2); kTRestrict: in 0 restricts b, 1 restricts r
    int flag;     // kTRestrict: fuse the next level's first sweep; kTProlong: overwrite (zero iterate)
};
struct TailLevel {
    int n;
    const int* rp;
    const int* ci;
    const void* v;  // values, level precision
    const void* inv;
    void* b;        // global rhs and iterate: the entry level's are read at start,
    void* x;        // its iterate written back at the end
    int nc;         // rows of the next level (0 on the coarsest)
    const int* map;
    const int* off;
    const int* mem;
    // the level's packed copy for the program's row sums (slices of 64 rows,
    // each lane's entries in groups of 4 contiguous ones), when it has one
    const void* pv;
    const unsigned short* pcol;
    int slices;     // 0: row sums from the CSR arrays
    // byte offsets of this level's arrays in the workgroup's LDS (lsbase:
    // the packed slices' entry offsets)
    int lb, lx0, lx1, lr, linv, lrp, lmap, loff, lmem, lsbase;
};
// LDS bytes for the tail's vectors and index arrays, and for those plus the
// product area of the CSR row sums, which gets what they leave (at least
// 40 KB).
constexpr int kTailLds = 104 * 1024;
constexpr int kTailLdsAll = 143 * 1024;
constexpr int kTailMaxRows = 4096;  // rows of a tail level (4 per thread)
constexpr int kMaxTailLevels = 20;
constexpr int kMaxTailOps = 256;    // operations of one program
template <class L>
void tail_run(tf_ctx& c, const TailLevel* levels, int top, int count, const void* image, int imageBytes,
              int ldsBytes, const TailOp* ops, int nops, const double* inverse, L w);
// Packed copy of a tail level (n <= kTailMaxRows, so columns fit 16 bits):
// entry k of row 64 s + l at base[s] + 256 (k / 4) + 4 l + k % 4; slots past
// a row's end hold value 0 and column 0.  base (device, slices + 1 entry
// offsets, multiples of 256) gives each slice's width.
template <class L>
void tail_pack(tf_ctx& c, int n, int slices, const int* rp, const int* ci, const L* v, const int* base, L* pv,
               unsigned short* pcol);

class BlockAmg;
struct KCtl;

// A launch sequence captured once as a hipGraph and replayed.  It names
// device buffers by address, so it is valid only while they live: `key`
// identifies the buffers it was captured for (owners change it when they
// free or replace one), and a different key re-captures.
struct Captured {
    hipGraphExec_t exec = nullptr;
    uint64_t key = 0;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
};
// Runs launch() through `g` (captured on first use or when key changed), or
// directly when use is false.
template <class F>
void launch_captured(tf_ctx& c, Captured& g, bool use, uint64_t key, F&& launch) {
    if (!use) {
        launch();
        return;
    }
    if (g.exec && g.key != key) g.reset();
    if (!g.exec) {
        hipGraph_t graph = nullptr;
        check(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            launch();
        }
        catch (...) {
            (void)hipStreamEndCapture(c.stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        check(hipStreamEndCapture(c.stream, &graph), "end capture");
        const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        check(e, "graph instantiate");
        g.key = key;
    }
    check(hipGraphLaunch(g.exec, c.stream), "graph launch");
}

// Levels whose aggregate histograms setup downloads asynchronously (the
// option's level limit).
constexpr int kMaxHistLevels = 20;

class AmgEngine {
public:
    AmgEngine(tf_ctx* ctx, const AmgOptionsC& o);
    ~AmgEngine();
    // Device arrays; canonicalInput skips the canonical-form check/copy (the
    // caller guarantees sorted unique columns and finite values).
    void setup(int n, int64_t nnz, const int* rp, const int* ci, const double* v, bool canonicalInput = false);
    // The next setups' finest level takes its stencil storage from this one
    // (same matrix, FP64 values) instead of building its own; nullptr clears.
    void share_fine_stencil(const DevStencil* s, const double* vals64) {
        fineStencil_ = s;
        fineStencil64_ = vals64;
    }
    // z = M^-1 r (FP64 device vectors; may alias).  Enqueued on the context stream.
    void apply(const double* rhs, double* x);
    // apply(rhs, x) whose last step (the level-0 iterate written out as
    // FP64 x) also forms sum(with .* x) as block partials partials[b],
    // b < blocks, in the Krylov reductions' layout; that step is skipped once
    // the control block has stopped.  With x == nullptr nothing is written
    // and *out describes where the result stays: the level-0 iterate in
    // level precision (valid until the next apply), or zero.
    struct Output {
        const void* x = nullptr;
        bool f32 = false;
        bool zero = false;
    };
    void apply_dot(const double* rhs, double* x, const double* with, double* partials, int blocks, const KCtl* ctl,
                   Output* out = nullptr);
    bool ready() const { return !levels_.empty(); }
    // Explicit inverse of the coarsest operator (nullptr: coarse Jacobi).
    const double* coarse_inverse() const { return factored_ ? inverse_ : nullptr; }
    // The multi-right-hand-side cycle over this hierarchy (made on first use).
    BlockAmg& block();
    // Network whose edge weights replace the strength weights (not owned;
    // used when options.learned != 0).
    void set_affinity_model(GnnModel* m) { gnn_ = m; }
    bool learned_used() const { return learnedUsed_; }
    double rho_learned() const { return rhoLearned_; }
    double rho_plain() const { return rhoPlain_; }
    double convergence_estimate();
    int n() const { return levels_.empty() ? 0 : levels_[0].native.n; }
    // Changes whenever a device buffer the cycle's launches name may have been
    // freed or replaced (setup, release, tail rebuild, hierarchy swap): a
    // captured graph of the cycle is valid only for the generation it was
    // captured at.
    uint64_t generation() const { return generation_; }
    // apply()'s whole cycle as one operation list (tail-program form, from
    // level 0): firstAtEntry says whether level 0's first sweep from zero is
    // applied when the rhs enters; curOut / zeroOut give the buffer that holds
    // the level-0 iterate at the end and whether it is still zero.
    std::vector<TailOp> cycle_program(bool& firstAtEntry, int& curOut, bool& zeroOut) const;
    bool factored() const { return factored_; }

    std::vector<AmgLevel> levels_;
    std::vector<int64_t> visits_;
    std::string stop_, coarseSolver_;
    AmgOptionsC opt_;

private:
    const DevStencil* fineStencil_ = nullptr;
    const double* fineStencil64_ = nullptr;
    template <class L>
    void cycle(int kind, size_t d, bool zero);
    template <class L>
    void smooth(size_t d, int sweeps, int first = 0);
    template <class L>
    void restrict_to(size_t d, bool ofResidual, bool fuseFirstSweep);
    template <class L>
    void coarse(size_t d);
    template <class L>
    void apply_typed(const double* rhs, double* x, const double* with = nullptr, double* partials = nullptr,
                     int blocks = 0, const KCtl* ctl = nullptr, Output* out = nullptr);
    template <class L>
    void run_tail(int kind, bool zero);
    // Builds the tail from the first level at or below `from` that fits its
    // LDS budget (tail_ = that level, or -1 when none fits).
    void build_tail(int from = 1);
    void tune_tail();
    void free_tail();
    void release();
    void setup_hierarchy(int n, int64_t nnz, const int* rp, const int* ci, const double* v, bool canonicalInput,
                         bool learned);
    void swap_hierarchy(AmgEngine& o);

    // One-workgroup tail: first level it covers (-1: none), level table,
    // and the programs for every (cycle kind, zero start, pre-applied sweep).
    struct TailProgram {
        int offset = 0, count = 0;
        std::vector<int64_t> visits;
    };
    int tail_ = -1;
    TailLevel* tailLevels_ = nullptr;
    TailOp* tailOps_ = nullptr;
    void* tailImage_ = nullptr;  // the levels' constant arrays in their LDS layout
    void* tailPack_ = nullptr;   // the levels' packed matrices
    int tailImageBytes_ = 0;
    int tailLdsBytes_ = 0;      // image plus the levels' vectors
    TailProgram tailPrograms_[12];

    tf_ctx* c_;
    uint64_t generation_ = 0;
    int* histHost_ = nullptr;  // pinned: every level's aggregate-size histogram (4 ints a level)
    bool factored_ = false, singular_ = false;
    double* lu_ = nullptr;
    int* piv_ = nullptr;
    double* inverse_ = nullptr;
    BlockAmg* block_ = nullptr;
    GnnModel* gnn_ = nullptr;
    bool learnedUsed_ = false;
    double rhoLearned_ = -1.0, rhoPlain_ = -1.0;
};

// The cycle of an AmgEngine's hierarchy on k right-hand sides at once
// (row-major n x k blocks, tf_block.hip): AmgEngine::cycle's operations,
// column by column in the same arithmetic, without the coarse tail.
class BlockAmg {
public:
    BlockAmg(tf_ctx* ctx, AmgEngine* amg);
    ~BlockAmg();
    void apply(int k, const double* R, double* Z);  // Z = M^-1 R per column; R, Z distinct
    // The same cycle with the result left in the level-0 iterate: returns it
    // (float or double as `bits` says, valid until the next cycle) and
    // whether it is known zero.
    const void* apply_in_place(int k, const double* R, int& bits, bool& zero);
    // Makes the buffers (and tunes the launch widths) for k columns, so that
    // apply(k, ...) allocates nothing and can be captured.
    void reserve(int k);
    std::vector<int64_t> visits;

private:
    struct Lvl {
        void *b = nullptr, *x = nullptr, *x2 = nullptr, *r = nullptr;
        bool zero = false;
        bool swept = false;  // the zero-start sweep was fused into the kernel that wrote b
    };
    template <class L>
    void run_typed(const double* R);
    template <class L>
    void cycle(int kind, size_t d, bool zero);
    template <class L>
    void smooth(size_t d, int sweeps);
    template <class L>
    void restrict_to(size_t d, bool ofResidual, bool sweepNext);
    template <class L>
    void prolong(size_t d);
    template <class L>
    void coarse(size_t d);
    void release();
    template <class L>
    void tune_lanes();

    tf_ctx* c_;
    AmgEngine* amg_;
    int k_ = 0;
    std::vector<int> lanes_;  // columns per pass of each level's SpMM (tuned)
    std::vector<int> shape_;  // level sizes and element size the buffers were made for
    std::vector<Lvl> lv_;
    void* mem_ = nullptr;
};

struct KrylovOptionsC {
    int maxit = 1000;
    double rtol = 1e-8, atol = 0;
    int refresh = 50;
    bool graph = true;
};

class SolverEngine {
public:
    SolverEngine(tf_ctx* ctx, int kind, const KrylovOptionsC& o, AmgEngine* amg);
    ~SolverEngine();
    void setup(int n, int64_t nnz, const int* rp, const int* ci, const double* v);
    tf_solve_result apply(const double* b, double* x);  // device vectors
    // k right-hand sides at once (CG only): B, X row-major n x k device blocks,
    // out[k] host results, one per column.
    void apply_block(int k, const double* B, double* X, tf_solve_result* out);
    // Independent systems on the diagonal blocks of the operator (CG only):
    // segment s = rows [off[s], off[s+1]) (host offsets); out[segments].
    void apply_segmented(int segments, const int* off, const double* b, double* x, tf_solve_result* out);
    // The same solve as apply_segmented, every iteration and the whole cycle
    // in one cooperative launch with grid-wide barriers between phases.
    void apply_segmented_persistent(int segments, const int* off, const double* b, double* x,
                                    tf_solve_result* out);
    AmgEngine* amg() { return amg_; }
    int n() const { return n_; }
    double setupSeconds = 0, applySeconds = 0;

private:
    void precondition(const double* in, double* out);
    void launch_part();
    // finalises a true-residual refresh the last CG iteration left pending
    // (its finaliser otherwise runs in the next iteration's direction kernel)
    void settle();
    int pendParts_ = 0;
    void launch_block_iteration(int k);
    struct Segments;
    void launch_segment_iteration();
    // Checks the offsets and (re)makes the tile table and work buffers of a
    // segment layout.
    void prepare_segments(int segments, const int* off);
    // Results from the control blocks and the final residuals' squares.
    void segment_results(const std::vector<double>& normb, const std::vector<double>& tol,
                         const std::vector<double>& rr, tf_solve_result* out);
    template <class R>
    void launch_iteration(const R& rows, int nb);
    tf_solve_result finish(int status, int iterations);
    void true_residual();  // r = rhs - A x
    bool sell_spmv_pays();
    double norm_host(const double* x);
    void release_graphs();
    void reserve_block(int k);
    void release_block();
    void release_segments();

    tf_ctx* c_;
    int kind_;
    KrylovOptionsC o_;
    AmgEngine* amg_;
    DevCsr A_;
    DevSell As_;             // sliced-ELL copy of A_ for the iteration's products (empty: CSR)
    double* Asv_ = nullptr;
    DevStencil Ast_;         // structured-stencil copy of A_ (used instead of As_ when built)
    double* Astv_ = nullptr;
    int n_ = 0;
    double* vec_ = nullptr;  // all work vectors, contiguous
    double *rhs_ = nullptr, *x_ = nullptr, *r_ = nullptr, *z_ = nullptr, *p_ = nullptr, *q_ = nullptr;
    double *sh_ = nullptr, *s_ = nullptr, *t_ = nullptr, *ph_ = nullptr, *shat_ = nullptr;
    void* ctl_ = nullptr;      // device control block
    void* hctl_ = nullptr;     // pinned host mirror
    Captured graph_;  // one iteration (key: the preconditioner's generation)
    double normb_ = 0, tol_ = 0;
    // multi-right-hand-side work (allocated by the first block solve)
    int bk_ = 0;
    double* bvec_ = nullptr;  // B, X, R, Z, P, Q blocks
    void* bctl_ = nullptr;    // one control block per column
    void* hbctl_ = nullptr;
    double* bparts_ = nullptr;
    double* bsums_ = nullptr;
    void* hbsums_ = nullptr;
    int bLanes_ = 0;          // columns per pass of the operator's SpMM (tuned)
    Captured bgraph_;         // one block iteration
    // segmented batch work (tile table, control blocks, partials, vectors)
    Segments* seg_ = nullptr;
    Captured sgraph_;         // one segmented iteration
};

}  // namespace tf
