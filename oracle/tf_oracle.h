/* CPU oracle for the aggregation-AMG Krylov path: C interface.
 *
 * TEST INFRASTRUCTURE ONLY. This library is a host FP64 re-execution of the
 * reference method. Only the test suite, the smoke check and the CPU-baseline
 * leg of the benchmark load it, and only as the checker. The product library
 * (thermofem_amd/libthermofem_hip.so) never links, loads or calls it.
 *
 * Conventions: CSR arrays are int32 row offsets (n+1), int32 columns and
 * double values.  Functions that produce CSR write into caller buffers sized by
 * the stated upper bound and return the produced nnz (>= 0) or a negative code.
 * tfo_last_error() returns the message of the last failure on this thread.
 */
#ifndef TF_ORACLE_H
#define TF_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

const char* tfo_last_error(void);

/* Stage functions (bit-exact restatements). */
int64_t tfo_canonical(int n, const int* rp, const int* ci, const double* v, int* out_rp, int* out_ci, double* out_v);
int tfo_validate_rowmax(int n, const int* rp, const int* ci, const double* v, double* rowmax);
int64_t tfo_strength(int n, const int* rp, const int* ci, const double* v, double theta, int* g_rp, int* g_ci,
                     double* g_w); /* bound: 2*nnz */
int tfo_match(int n, const int* g_rp, const int* g_ci, const double* g_w, int rounds, int* map);
int64_t tfo_contract(int n, const int* g_rp, const int* g_ci, const double* g_w, const int* map, int* out_rp,
                     int* out_ci, double* out_w); /* out_rp sized count+1, bound nnz */
int tfo_aggregate(int n, const int* g_rp, const int* g_ci, const double* g_w, int rounds, int absorb, int* map,
                  int* offsets, int* members, int64_t* histogram); /* returns coarse rows */
int64_t tfo_galerkin(int n, const int* rp, const int* ci, const double* v, const int* map, int* out_rp, int* out_ci,
                     double* out_v); /* bound nnz */
int tfo_lu_factor(int n, const int* rp, const int* ci, const double* v, double* lu, int* pivots); /* 1 ok, 0 singular */
void tfo_lu_solve(int n, const double* lu, const int* pivots, const double* b, double* x);
void tfo_spmv(int n, const int* rp, const int* ci, const double* v, const double* x, double* y);
void tfo_spmv_f32(int n, const int* rp, const int* ci, const float* v, const float* x, float* y);

/* Hierarchy, cycles and solvers. level_precision: 64 or 32 (cycle arithmetic). */
typedef struct tfo_amg_options {
    int max_levels, coarse_rows, matching_rounds, pre_sweeps, post_sweeps, coarse_sweeps;
    double jacobi_weight, strength_threshold;
    int absorb_singletons, cycle; /* cycle: 0 V, 1 W, 2 F */
    int level_precision;
} tfo_amg_options;

void* tfo_amg_create(const tfo_amg_options* options);
void tfo_amg_destroy(void* amg);
int tfo_amg_setup(void* amg, int n, const int* rp, const int* ci, const double* v);
int tfo_amg_levels(void* amg);
int tfo_amg_level_info(void* amg, int level, int64_t* rows, int64_t* nnz, int64_t* histogram);
int tfo_amg_level_matrix(void* amg, int level, int* rp, int* ci, double* v);
int tfo_amg_level_transfer(void* amg, int level, int* map, int* offsets, int* members);
const char* tfo_amg_stop_reason(void* amg);
const char* tfo_amg_coarse_solver(void* amg);
int tfo_amg_apply(void* amg, const double* rhs, double* x);
int tfo_amg_visits(void* amg, int64_t* visits);

typedef struct tfo_solve_result {
    int status; /* 0 Converged, 1 MaxIterations, 2 Breakdown, 3 NonFinite */
    int iterations;
    double residual_norm, relative_residual;
} tfo_solve_result;

/* kind: 0 CG, 1 BiCGStab. amg may be NULL (identity preconditioner). */
int tfo_krylov_solve(int kind, void* amg, int n, const int* rp, const int* ci, const double* v, const double* b,
                     double* x, int max_iterations, double rtol, double atol, int refresh, tfo_solve_result* result);
int tfo_jacobi(int n, const int* rp, const int* ci, const double* v, int sweeps, double weight, const double* b,
               double* x);
int tfo_threads(void);

#ifdef __cplusplus
}
#endif
#endif
