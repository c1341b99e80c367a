/* thermofem_hip.h - C ABI of the MI355X (gfx950) aggregation-AMG Krylov library.
 *
 * The reference (Mgepahmge/ThermoFEM) exposes its hot path only as a C++20
 * API; this header is the thin C layer underneath the C++ drop-in headers in
 * thermofem_amd/include/thermofem/ and the surface a ctypes / cgo / JNI
 * binding would load.  Each entry cites the reference interface it replaces.
 *
 * Conventions
 *  - Every function returns TF_OK (0) or a negative TF_E* code; the message of
 *    the last failure on the calling thread is tf_last_error().
 *  - Array arguments are device pointers unless the name says host_; indices
 *    are int32 (the reference's native AMG index type,
 *    src/Multigrid/GpuAggregation.cu:993-995,1113-1114).
 *  - Work is enqueued on the context's stream; entry points that return
 *    host values (counts, flags, dot products) synchronise that stream.
 *  - Arrays the library allocates for the caller are released with tf_free.
 */
#ifndef THERMOFEM_HIP_H
#define THERMOFEM_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TF_OK = 0,
    TF_E_INVALID = -1,  /* contract violation (std::invalid_argument / logic_error in the reference) */
    TF_E_CAPACITY = -2, /* std::length_error / overflow_error */
    TF_E_DEVICE = -3,   /* HIP runtime failure (CudaError in the reference, Memory.cuh:52-66) */
    TF_E_NUMERIC = -4   /* invalid values: non-finite, zero diagonal (std::invalid_argument in the reference) */
};

/* Memory strategies: include/DataStructure/Memory.cuh:17-42. */
enum { TF_MEM_GLOBAL = 0, TF_MEM_PINNED = 1, TF_MEM_MAPPED = 2, TF_MEM_MANAGED = 3 };

typedef struct tf_ctx tf_ctx;

const char* tf_last_error(void);
int tf_device_count(int* count);
/* Per-device execution context: one stream, reduction scratch, pinned
 * staging.  Replaces the per-call cuBLAS/cuSPARSE handles of
 * src/LinearAlgebra/Backend.cu:128-415 and IterationContext.cu:38-124. */
/* Device name (NUL-terminated, truncated to name_len) and HIP runtime version
 * (the cudaGetDeviceProperties / cudaRuntimeGetVersion report of
 * example/solve-matrix.cu:70-85); either output may be NULL. */
int tf_device_info(int device, char* name, size_t name_len, int* runtime_version);
int tf_get_device(int* device); /* current HIP device of the calling thread */
int tf_set_device(int device);
int tf_ctx_create(int device, tf_ctx** ctx);
int tf_ctx_destroy(tf_ctx* ctx);
int tf_ctx_synchronize(tf_ctx* ctx);
void* tf_ctx_stream(tf_ctx* ctx);
int tf_ctx_device(tf_ctx* ctx);

/* HIP events on the context stream (timing of enqueued work). */
typedef struct tf_event tf_event;
int tf_event_create(tf_event** ev);
int tf_event_destroy(tf_event* ev);
int tf_event_record(tf_ctx* ctx, tf_event* ev);
int tf_event_elapsed_ms(tf_event* start, tf_event* stop, float* ms); /* waits for stop */

/* ---- memory: src/DataStructure/Memory.cu:53-105,179-196,246-260 ---- */
int tf_alloc(int device, int strategy, size_t bytes, void** ptr);
int tf_free(int strategy, void* ptr);
/* Device view: Global/Managed -> ptr, Mapped -> device alias, Pinned -> TF_E_INVALID (Memory.cu:250-252). */
int tf_device_view(int strategy, void* ptr, void** device_ptr);
int tf_memcpy(void* dst, const void* src, size_t bytes); /* synchronous, any direction */
int tf_memcpy_async(tf_ctx* ctx, void* dst, const void* src, size_t bytes);
int tf_fill_zero(tf_ctx* ctx, void* dst, size_t bytes); /* kernel fill (graph-capture safe) */

/* ---- SpMV / SpMM: cuSPARSE SpMV in IterationContext.cu:139-159, SpMM in Backend.cu:256-281 ---- */
int tf_spmv_csr_f64(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v, const double* x,
                    double* y);
int tf_spmv_csr_f32(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const float* v, const float* x,
                    float* y);
/* r = b - A x  (IterationContext::residual, IterationContext.cu:166-175) */
int tf_residual_csr_f64(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v,
                        const double* b, const double* x, double* r);
int tf_residual_csr_f32(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const float* v,
                        const float* b, const float* x, float* r);
/* Y = A X with X, Y row-major n x k (DenseMatrix.cuh:9-19 layout); X != Y.
 * Column j of Y equals tf_spmv_csr_* of column j bit for bit. */
int tf_spmm_csr_f64(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v, int k,
                    const double* X, double* Y);
int tf_spmm_csr_f32(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const float* v, int k,
                    const float* X, float* Y);

/* ---- Remaining linalg functions of Operations.cuh:160-525 (cuBLAS / cuSPARSE in Backend.cu) ---- */
/* sum |x_i| (norm1, Operations.cuh:251-261) */
int tf_asum_f64(tf_ctx* ctx, int n, const double* x, double* host_result);
int tf_asum_f32(tf_ctx* ctx, int n, const float* x, float* host_result);
/* At (cols x rows) = A^T for row-major A (rows x cols) (transpose, Operations.cuh:309-321) */
int tf_dense_transpose_f64(tf_ctx* ctx, int rows, int cols, const double* A, double* At);
int tf_dense_transpose_f32(tf_ctx* ctx, int rows, int cols, const float* A, float* At);
/* y = A x, A row-major rows x cols (Operations.cuh:325-344) */
int tf_gemv_f64(tf_ctx* ctx, int rows, int cols, const double* A, const double* x, double* y);
int tf_gemv_f32(tf_ctx* ctx, int rows, int cols, const float* A, const float* x, float* y);
/* C = A B, row-major m x k and k x n (Operations.cuh:347-367) */
int tf_gemm_f64(tf_ctx* ctx, int m, int n, int k, const double* A, const double* B, double* C);
int tf_gemm_f32(tf_ctx* ctx, int m, int n, int k, const float* A, const float* B, float* C);
/* Sparse add A + beta B (Operations.cuh:449-467), transpose (:477-491) and
 * product A B (:494-513) of CSR operands (any column order, duplicates
 * allowed): results are canonical (columns ascending, duplicates summed in
 * candidate order: A's entries then B's for the sum, a_ik b_kj by k in A's
 * stored order for the product).  out_rp has result rows + 1 entries in
 * caller memory; out_ci / out_v are library allocations (tf_free). */
int tf_csr_add_f64(tf_ctx* ctx, int rows, int cols, int64_t nnz_a, const int* rp_a, const int* ci_a, const double* v_a,
                   int64_t nnz_b, const int* rp_b, const int* ci_b, const double* v_b, double beta, int* out_rp,
                   int** out_ci, double** out_v, int64_t* out_nnz);
int tf_csr_transpose_f64(tf_ctx* ctx, int rows, int cols, int64_t nnz, const int* rp, const int* ci, const double* v,
                         int* out_rp, int** out_ci, double** out_v);
int tf_csr_multiply_f64(tf_ctx* ctx, int rows_a, int cols_a, int cols_b, int64_t nnz_a, const int* rp_a,
                        const int* ci_a, const double* v_a, int64_t nnz_b, const int* rp_b, const int* ci_b,
                        const double* v_b, int* out_rp, int** out_ci, double** out_v, int64_t* out_nnz);

/* ---- Jacobi: update kernel IterationContext.cu:26-32 and one sweep of Jacobi.cuh:60-74 ---- */
/* x += w * inv * (b - product) */
int tf_jacobi_update_f64(tf_ctx* ctx, int n, double w, const double* inv, const double* b, const double* product,
                         double* x);
int tf_jacobi_update_f32(tf_ctx* ctx, int n, float w, const float* inv, const float* b, const float* product,
                         float* x);
/* x_out = x_in + w * inv * (b - A x_in) fused; x_in == NULL means x_in = 0 (no SpMV). x_out != x_in. */
int tf_jacobi_fused_f64(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v,
                        const double* inv, double w, const double* b, const double* x_in, double* x_out);
int tf_jacobi_fused_f32(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const float* v,
                        const float* inv, float w, const float* b, const float* x_in, float* x_out);
/* inv[i] = 1 / sum of diagonal entries of row i (Diagonal.cuh:13-36); *host_bad = 1 on a
 * missing, zero or non-finite diagonal or reciprocal. */
int tf_inverse_diagonal_f64(tf_ctx* ctx, int n, const int* rp, const int* ci, const double* v, double* inv,
                            int* host_bad);
int tf_inverse_diagonal_f32(tf_ctx* ctx, int n, const int* rp, const int* ci, const float* v, float* inv,
                            int* host_bad);

/* ---- level-1: cuBLAS axpy/scal/dot/nrm2 of IterationContext.cu:176-249 ---- */
int tf_axpy_f64(tf_ctx* ctx, int n, double alpha, const double* x, double* y);
int tf_axpy_f32(tf_ctx* ctx, int n, float alpha, const float* x, float* y);
int tf_scal_f64(tf_ctx* ctx, int n, double alpha, double* x);
int tf_scal_f32(tf_ctx* ctx, int n, float alpha, float* x);
/* Deterministic two-level reductions; results returned to the host. */
int tf_dot_f64(tf_ctx* ctx, int n, const double* x, const double* y, double* host_result);
int tf_dot_f32(tf_ctx* ctx, int n, const float* x, const float* y, float* host_result);
int tf_nrm2_f64(tf_ctx* ctx, int n, const double* x, double* host_result);
int tf_nrm2_f32(tf_ctx* ctx, int n, const float* x, float* host_result);

/* ---- conversions: convert<A,B> of GpuAggregation.cu:1101-1108 ---- */
int tf_convert_f64_f32(tf_ctx* ctx, size_t n, const double* in, float* out);
int tf_convert_f32_f64(tf_ctx* ctx, size_t n, const float* in, double* out);
int tf_convert_i64_i32(tf_ctx* ctx, size_t n, const int64_t* in, int* out);
int tf_convert_i32_i64(tf_ctx* ctx, size_t n, const int* in, int64_t* out);
/* dst = src as a plain streaming kernel (Vector::copyDataFrom device path, Memory.cu:179-196). */
int tf_copy_f64(tf_ctx* ctx, size_t n, const double* src, double* dst);

/* ---- scan: cub::DeviceScan::ExclusiveSum of GpuAggregation.cu:116-121 ---- */
int tf_scan_excl_i32(tf_ctx* ctx, int n, const int* in, int* out); /* out has n entries */
/* The same exclusive scan, allocating nothing: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops).  Its block totals live in a fixed
 * 516 KB device array per GPU, so n <= 268435456 (2^28; larger n returns
 * TF_E_INVALID); calls that overlap on one GPU share that array and race. */
int tf_scan_excl_i32_noalloc(tf_ctx* ctx, int n, const int* in, int* out);

/* ---- aggregation setup: GpuAggregation.cu (multigrid::GpuAggregation, GpuAggregation.cuh:34-69) ---- */
/* checkCanonical (:283-304): *host_flags bit0 = unsorted or duplicate columns, bit1 = invalid. */
int tf_csr_check(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v, int* host_flags);
/* Device CSR structure validation of CsrMatrix::fromBuffers (src/DataStructure/CsrMatrix.cu:12-58). */
int tf_csr_validate_structure(tf_ctx* ctx, int rows, int cols, int64_t nnz, const int* rp, const int* ci,
                              int* host_bad);
/* canonical (:991-1007): rows sorted, duplicates summed in stored order, explicit zeros kept.
 * Allocates *out_ci / *out_v (device) and writes out_rp (n+1, caller-owned). */
int tf_csr_canonicalize(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v, int* out_rp,
                        int** out_ci, double** out_v, int64_t* out_nnz);
/* validate (:247-271): *host_bad on non-finite values or an invalid diagonal; rowmax may be NULL. */
int tf_agg_validate_rowmax(tf_ctx* ctx, int n, const int* rp, const int* ci, const double* v, double* rowmax,
                           int* host_bad);
/* strength graph (:339-390): symmetric, max-merged, no self loops.  Allocates the graph arrays. */
int tf_agg_strength(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v, double theta,
                    int** g_rp, int** g_ci, double** g_w, int64_t* g_nnz);
/* One synchronous matching round, propose + accept (:402-449). mate initialised to -1 by the caller. */
int tf_agg_match_round(tf_ctx* ctx, int n, const int* g_rp, const int* g_ci, const double* g_w, int* mate,
                       int* proposal);
/* Full pairing (:485-505): rounds of matching, roots, scan, assign.  Returns the aggregate count. */
int tf_agg_pair(tf_ctx* ctx, int n, const int* g_rp, const int* g_ci, const double* g_w, int rounds, int* map,
                int* host_count);
/* Graph contraction (:534-576) through the first-phase map; allocates the coarse graph. */
int tf_agg_contract(tf_ctx* ctx, int n, const int* g_rp, const int* g_ci, const double* g_w, const int* map, int nc,
                    int** c_rp, int** c_ci, double** c_w, int64_t* c_nnz);
/* aggregate (:766-818,1020-1031): two phases, optional two absorption rounds, compaction, member CSR.
 * Allocates *map (n), *offsets (nc+1), *members (n). */
int tf_agg_aggregate(tf_ctx* ctx, int n, const int* g_rp, const int* g_ci, const double* g_w, int rounds,
                     int absorb, int** map, int** offsets, int** members, int* host_coarse_rows,
                     int64_t* host_histogram);
/* Galerkin Ac = P^T A P for piecewise-constant P (:136-146,227-234,1032-1038); duplicates summed
 * in original entry order.  Allocates Ac arrays. */
int tf_galerkin(tf_ctx* ctx, int n, const int* rp, const int* ci, const double* v, const int* map, int nc,
                const int* offsets, const int* members, int** c_rp, int** c_ci, double** c_v, int64_t* c_nnz);
/* restrictKernel (:950-960) and prolongKernel (:970-976). */
int tf_restrict_f64(tf_ctx* ctx, int nc, const int* offsets, const int* members, const double* fine, double* coarse);
int tf_restrict_f32(tf_ctx* ctx, int nc, const int* offsets, const int* members, const float* fine, float* coarse);
int tf_prolong_f64(tf_ctx* ctx, int n, const int* map, const double* coarse, double* fine);
int tf_prolong_f32(tf_ctx* ctx, int n, const int* map, const float* coarse, float* fine);

/* ---- coarse dense LU: factor (:827-893) and triangular (:904-939), widened to n <= 1024 ----
 * lu (n*n row-major, device), pivots (n, device), inverse (n*n column-major: inverse[j*n + i] = (A^-1)_ij;
 * device, may be NULL).
 * *host_singular = 1 when a pivot is <= 64*eps*max|A| or a factor is non-finite. */
int tf_coarse_lu_factor_f64(tf_ctx* ctx, int n, const int* rp, const int* ci, const double* v, double* lu,
                            int* pivots, double* inverse, int* host_singular);
/* x = A^-1 b by permuted forward/backward substitution in FP64 (rhs and x may alias). */
int tf_coarse_lu_solve_f64(tf_ctx* ctx, int n, const double* lu, const int* pivots, const double* b, double* x);
int tf_coarse_lu_solve_f32(tf_ctx* ctx, int n, const double* lu, const int* pivots, const float* b, float* x);
/* x = A^-1 b in FP64 from the column-major inverse above (the per-cycle coarse solve; rhs and x may alias). */
int tf_coarse_inv_apply_f64(tf_ctx* ctx, int n, const double* inverse, const double* b, double* x);
int tf_coarse_inv_apply_f32(tf_ctx* ctx, int n, const double* inverse, const float* b, float* x);

/* ---- preconditioner and solvers: Amg.cuh, Cg.cuh, BiCgStab.cuh, Jacobi.cuh ---- */
typedef struct tf_amg_options {
    int max_levels;         /* 1..20 (Amg.cuh:15) */
    int coarse_rows;        /* 1..1024 (reference 1..64, Amg.cuh:16) */
    int matching_rounds;    /* 1..8 */
    int pre_sweeps, post_sweeps, coarse_sweeps;
    double jacobi_weight;   /* (0,2) */
    double strength_threshold; /* [0,1] */
    int absorb_singletons;
    int cycle;              /* 0 V, 1 W, 2 F (Cycle.cuh:10) */
    int level_precision;    /* 64: cycle in FP64 (reference); 32: FP32 cycle under an FP64 caller */
    int coarse_tail;        /* coarse levels as one one-workgroup program: 0 timed at setup, 1 always, 2 never */
    int learned;            /* matching weights: 0 strength (reference), 1 learned affinities (needs
                               tf_amg_set_affinity_model), 2 learned only where PCG converges faster with it */
} tf_amg_options;

typedef struct tf_krylov_options {
    int max_iterations;     /* >= 0 */
    double relative_tolerance, absolute_tolerance;
    int residual_refresh_interval; /* >= 1 */
    int use_graph;          /* capture one iteration as a hipGraph and replay it */
} tf_krylov_options;

typedef struct tf_solve_result {
    int status; /* 0 Converged, 1 MaxIterations, 2 Breakdown, 3 NonFinite (Interfaces.cuh:10) */
    int iterations;
    double residual_norm, relative_residual;
} tf_solve_result;

typedef struct tf_amg tf_amg;
/* Amg<double,int> (level_precision 64) or the FP32-cycle adapter (32). */
int tf_amg_create(int device, const tf_amg_options* options, tf_amg** amg);
int tf_amg_destroy(tf_amg* amg);
/* setup (Amg.cuh:81-143); arrays are host (host_memory != 0) or device pointers. */
int tf_amg_setup(tf_amg* amg, int n, int64_t nnz, const int* rp, const int* ci, const double* v, int host_memory);
/* apply (Amg.cuh:149-161): x = cycle(rhs), rhs and x may alias. */
int tf_amg_apply(tf_amg* amg, const double* rhs, double* x, int host_memory);
/* k right-hand sides at once: R, Z distinct row-major n x k device blocks,
 * Z = cycle(R) column by column.  Replaces k calls of tf_amg_apply (Amg.cuh:
 * 149-161; the reference has no multi-vector cycle, its dense type is
 * DenseMatrix.cuh:9-19); column j equals tf_amg_apply of column j bit for bit. */
int tf_amg_apply_block(tf_amg* amg, int k, const double* R, double* Z);
/* Per-level visit counts of the last block cycle (level_count entries). */
int tf_amg_block_visits(tf_amg* amg, int64_t* visits);
int tf_amg_level_count(tf_amg* amg);
int tf_amg_level_info(tf_amg* amg, int level, int64_t* rows, int64_t* nnz, int64_t* histogram);
/* Host copies of one level's native FP64 matrix and transfer (for parity checks). */
int tf_amg_level_matrix(tf_amg* amg, int level, int* host_rp, int* host_ci, double* host_v);
int tf_amg_level_transfer(tf_amg* amg, int level, int* host_map, int* host_offsets, int* host_members);
const char* tf_amg_stop_reason(tf_amg* amg);
const char* tf_amg_coarse_solver(tf_amg* amg);
int tf_amg_visits(tf_amg* amg, int64_t* visits);

typedef struct tf_solver tf_solver;
/* kind 0: Cg (Cg.cuh), 1: BiCgStab (BiCgStab.cuh).  amg_options NULL: Identity preconditioner. */
int tf_solver_create(int device, int kind, const tf_krylov_options* options, const tf_amg_options* amg_options,
                     tf_solver** solver);
int tf_solver_destroy(tf_solver* solver);
int tf_solver_setup(tf_solver* solver, int n, int64_t nnz, const int* rp, const int* ci, const double* v,
                    int host_memory);
/* apply: x holds the initial guess and receives the solution. */
int tf_solver_apply(tf_solver* solver, const double* b, double* x, int host_memory, tf_solve_result* result);
/* A solver whose preconditioner is an existing AMG handle (not owned; it must
 * outlive the solver).  The solver runs on the AMG's stream; setup re-runs the
 * AMG setup on the solver's canonical matrix, as Cg::setup / BiCgStab::setup
 * call preconditioner->setup (Cg.cuh:60-67, BiCgStab.cuh:70-77). */
int tf_solver_create_with_amg(int kind, const tf_krylov_options* options, tf_amg* amg, tf_solver** solver);
tf_amg* tf_solver_amg(tf_solver* solver);
/* CG on k right-hand sides at once (Cg.cuh:123-189 per column, each column
 * with its own scalars and stopping test): B, X distinct row-major n x k
 * device blocks (X holds the initial guesses), results[k] host array.  The
 * preconditioner and the products run on the whole block (SpMM,
 * Backend.cu:256-281).  BiCGStab solvers return TF_E_INVALID. */
int tf_solver_apply_block(tf_solver* solver, int k, const double* B, double* X, tf_solve_result* results);
/* Independent systems in one solve: the solver's matrix is block diagonal
 * with segment s = rows [host_offsets[s], host_offsets[s+1]) (offsets from 0
 * to n, nondecreasing; an entry coupling two segments is TF_E_INVALID).  CG
 * (Cg.cuh:123-189) runs on every segment with its own scalars and stopping
 * test, each kernel of an iteration one persistent launch over all segments;
 * the AMG hierarchy of a block-diagonal matrix is block diagonal, so one
 * cycle preconditions all of them.  b, x distinct device vectors of n;
 * results[segments] host array.  No reference counterpart (one system per
 * Cg::apply, Cg.cuh:123-189).  BiCGStab solvers return TF_E_INVALID. */
/* ---- Learned edge affinities (no reference counterpart: the reference's
 * matching reads the strength weights of GpuAggregation.cu:339-364 at
 * :402-431; the network's weights replace them, the matching is unchanged).
 * A two-layer message-passing network (hidden 32, bfloat16 weights, FP32
 * accumulation) scores every off-diagonal entry of a matrix with a symmetric
 * pattern; forward, backward and Adam run on the GPU.  All matrix and vector
 * pointers are device pointers. ---- */
typedef struct tf_gnn tf_gnn;
int tf_gnn_param_count(void);
int tf_gnn_create(int device, tf_gnn** model); /* zero weights; set them with tf_gnn_set_params */
int tf_gnn_destroy(tf_gnn* model);
int tf_gnn_set_params(tf_gnn* model, const float* host_params, int count); /* resets the optimizer */
int tf_gnn_get_params(tf_gnn* model, float* host_params, int count);
/* weights[nnz]: the network's weight of every stored entry, 0 on the diagonal. */
int tf_gnn_infer(tf_gnn* model, int n, int64_t nnz, const int* rp, const int* ci, const double* v, double* weights);
/* loss = mean over off-diagonal entries of (w - target)^2 (target[nnz] per stored
 * entry, device); host_grad (param_count floats, may be NULL) its gradient. */
int tf_gnn_loss_grad(tf_gnn* model, int n, int64_t nnz, const int* rp, const int* ci, const double* v,
                     const float* target, double* loss, float* host_grad);
/* one Adam step on that loss; loss receives the value before the step. */
int tf_gnn_train_step(tf_gnn* model, int n, int64_t nnz, const int* rp, const int* ci, const double* v,
                      const float* target, float learning_rate, double* loss);
/* Training targets: `sweeps` weighted-Jacobi sweeps on A X = 0 from the n x k
 * row-major block X0, then out[p] = (x_i.x_j)^2 / ((x_i.x_i)(x_j.x_j)) for
 * every stored entry p = (i, j) (0 on the diagonal). */
int tf_affinities(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v, int k, int sweeps,
                  double omega, const double* X0, float* out);
/* The network an AMG handle uses when options.learned != 0 (not owned; NULL: none). */
int tf_amg_set_affinity_model(tf_amg* amg, tf_gnn* model);
/* After setup: whether the learned hierarchy is in use, and the convergence
 * estimates (tf_amg_convergence_estimate) of both (-1 when not measured). */
int tf_amg_learned_info(tf_amg* amg, int* used, double* rho_learned, double* rho_plain);
/* Convergence of the current cycle as a CG preconditioner: the mean residual
 * reduction per iteration (|r_10|/|r_0|)^(1/10) of 10 PCG iterations on
 * A x = b, x0 = 0, b fixed pseudo-random (smaller is better). */
int tf_amg_convergence_estimate(tf_amg* amg, double* rho);

int tf_solver_apply_segmented(tf_solver* solver, int segments, const int* host_offsets, const double* b, double* x,
                              tf_solve_result* results);
/* The same segmented CG in one persistent launch: the initial residuals,
 * every iteration with its whole AMG cycle, each segment's stopping test and
 * the final residuals run as phases of one cooperative grid separated by
 * grid-wide barriers, with no host round trip until the solve ends
 * (Cg.cuh:123-189 per segment).  Results equal tf_solver_apply_segmented's.
 * TF_E_INVALID when the device cannot launch cooperatively or the factored
 * coarsest level has more than 1024 rows. */
int tf_solver_apply_segmented_persistent(tf_solver* solver, int segments, const int* host_offsets, const double* b,
                                         double* x, tf_solve_result* results);
/* Seconds of the last setup / apply measured with HIP events on the solver stream. */
int tf_solver_timings(tf_solver* solver, double* setup_seconds, double* apply_seconds);

/* ---- sliced-ELL row products: IterationContext::multiply / residual and the
 * Jacobi update (IterationContext.cu:26-32,75,155) over a storage made for
 * 64-lane waves: slices of 64 rows, entries slice-major, 16-bit column
 * offsets when every |column - row| <= 32767, one-byte row lengths.  Results
 * equal the CSR entry points' bit for bit.  The handle keeps FP64 and FP32
 * values and belongs to the context it was created on (which must outlive it). */
typedef struct tf_sell tf_sell;
/* TF_E_INVALID (no handle) if a row has more than 255 entries or the padding
 * would exceed max_padding * nnz. */
int tf_sell_create(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v, double max_padding,
                   tf_sell** sell);
int tf_sell_destroy(tf_sell* sell);
/* index_bits 16 or 32; stored_entries includes padding; reordered: rows sorted by length in 256-row windows */
int tf_sell_info(const tf_sell* sell, int* index_bits, int64_t* stored_entries, int* reordered);
int tf_sell_spmv_f64(tf_ctx* ctx, const tf_sell* sell, const double* x, double* y);
int tf_sell_spmv_f32(tf_ctx* ctx, const tf_sell* sell, const float* x, float* y);
int tf_sell_residual_f64(tf_ctx* ctx, const tf_sell* sell, const double* b, const double* x, double* r);
int tf_sell_residual_f32(tf_ctx* ctx, const tf_sell* sell, const float* b, const float* x, float* r);
int tf_sell_jacobi_f64(tf_ctx* ctx, const tf_sell* sell, const double* inv, double w, const double* b,
                       const double* x_in, double* x_out);
int tf_sell_jacobi_f32(tf_ctx* ctx, const tf_sell* sell, const float* inv, float w, const float* b,
                       const float* x_in, float* x_out);

/* ---- structured-stencil row products: the same three operations over a
 * storage without column indices, for operators whose rows all take their
 * columns from one list of at most 16 offsets from the row (structured-grid
 * and finite-difference matrices): a 16-bit term mask per row and one dense
 * value array per offset; for a bit-for-bit symmetric operator only the
 * offsets >= 0 are stored and a(i, i-d) is read as a(i-d, i).  No reference
 * counterpart: the products replace IterationContext::multiply / residual
 * (IterationContext.cu:26-32,75,155) for such operators, with results equal
 * to the CSR entry points' bit for bit.  The CSR must be canonical (sorted,
 * unique columns per row).  The handle keeps FP64 and FP32 values and
 * belongs to the context it was created on. */
typedef struct tf_stencil tf_stencil;
/* TF_E_INVALID (no handle) if the matrix is not canonical or its rows use
 * more than 16 distinct offsets. */
int tf_stencil_create(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v,
                      tf_stencil** stencil);
int tf_stencil_destroy(tf_stencil* stencil);
/* terms: distinct offsets; arrays: value arrays stored; mirrored: 1 when the
 * negative offsets read the positive ones' arrays */
int tf_stencil_info(const tf_stencil* stencil, int* terms, int* arrays, int* mirrored);
int tf_stencil_spmv_f64(tf_ctx* ctx, const tf_stencil* stencil, const double* x, double* y);
int tf_stencil_spmv_f32(tf_ctx* ctx, const tf_stencil* stencil, const float* x, float* y);
int tf_stencil_residual_f64(tf_ctx* ctx, const tf_stencil* stencil, const double* b, const double* x, double* r);
int tf_stencil_residual_f32(tf_ctx* ctx, const tf_stencil* stencil, const float* b, const float* x, float* r);
int tf_stencil_jacobi_f64(tf_ctx* ctx, const tf_stencil* stencil, const double* inv, double w, const double* b,
                          const double* x_in, double* x_out);
int tf_stencil_jacobi_f32(tf_ctx* ctx, const tf_stencil* stencil, const float* inv, float w, const float* b,
                          const float* x_in, float* x_out);

/* ---- IO: io::readMatrixMarket (include/IO/MatrixMarket.cuh, src/IO/MatrixMarket.cu:35-122) ----
 * Host CSR of a Matrix Market file (rows in file order, duplicates kept,
 * symmetric storage mirrored).  Call with NULL arrays to get rows, cols and
 * nnz, then again with host arrays of rows+1, nnz and nnz elements.
 * Malformed input returns TF_E_INVALID with the reader's message. */
int tf_read_matrix_market(const char* path, int* rows, int* cols, int64_t* nnz, int* host_rp, int* host_ci,
                          double* host_v);

#ifdef __cplusplus
}
#endif
#endif
