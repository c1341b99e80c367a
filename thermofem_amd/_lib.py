"""ctypes binding of the C ABI in include/thermofem_hip.h.

The library is the in-tree libthermofem_hip.so built by thermofem_amd/build.py.
There is no fallback: a missing or unloadable library raises ImportError.
"""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libthermofem_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "thermofem_hip.h")

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
f64 = C.c_double
f32 = C.c_float
sz = C.c_size_t
I32P = C.POINTER(C.c_int)
I64P = C.POINTER(C.c_int64)
F64P = C.POINTER(C.c_double)


class AmgOptions(C.Structure):
    _fields_ = [("max_levels", i32), ("coarse_rows", i32), ("matching_rounds", i32), ("pre_sweeps", i32),
                ("post_sweeps", i32), ("coarse_sweeps", i32), ("jacobi_weight", f64),
                ("strength_threshold", f64), ("absorb_singletons", i32), ("cycle", i32),
                ("level_precision", i32), ("coarse_tail", i32), ("learned", i32)]


class KrylovOptions(C.Structure):
    _fields_ = [("max_iterations", i32), ("relative_tolerance", f64), ("absolute_tolerance", f64),
                ("residual_refresh_interval", i32), ("use_graph", i32)]


class SolveResultC(C.Structure):
    _fields_ = [("status", i32), ("iterations", i32), ("residual_norm", f64), ("relative_residual", f64)]


_SIGS = {
    "tf_last_error": (C.c_char_p, []),
    "tf_device_count": (i32, [I32P]),
    "tf_device_info": (i32, [i32, C.c_char_p, sz, I32P]),
    "tf_sell_create": (i32, [vp, i32, i64, vp, vp, vp, f64, C.POINTER(vp)]),
    "tf_sell_destroy": (i32, [vp]),
    "tf_sell_info": (i32, [vp, I32P, I64P, I32P]),
    "tf_sell_spmv_f64": (i32, [vp, vp, vp, vp]),
    "tf_sell_spmv_f32": (i32, [vp, vp, vp, vp]),
    "tf_sell_residual_f64": (i32, [vp, vp, vp, vp, vp]),
    "tf_sell_residual_f32": (i32, [vp, vp, vp, vp, vp]),
    "tf_sell_jacobi_f64": (i32, [vp, vp, vp, f64, vp, vp, vp]),
    "tf_sell_jacobi_f32": (i32, [vp, vp, vp, f32, vp, vp, vp]),
    "tf_stencil_create": (i32, [vp, i32, i64, vp, vp, vp, C.POINTER(vp)]),
    "tf_stencil_destroy": (i32, [vp]),
    "tf_stencil_info": (i32, [vp, I32P, I32P, I32P]),
    "tf_stencil_spmv_f64": (i32, [vp, vp, vp, vp]),
    "tf_stencil_spmv_f32": (i32, [vp, vp, vp, vp]),
    "tf_stencil_residual_f64": (i32, [vp, vp, vp, vp, vp]),
    "tf_stencil_residual_f32": (i32, [vp, vp, vp, vp, vp]),
    "tf_stencil_jacobi_f64": (i32, [vp, vp, vp, f64, vp, vp, vp]),
    "tf_stencil_jacobi_f32": (i32, [vp, vp, vp, f32, vp, vp, vp]),
    "tf_read_matrix_market": (i32, [C.c_char_p, I32P, I32P, I64P, I32P, I32P, F64P]),
    "tf_get_device": (i32, [I32P]),
    "tf_set_device": (i32, [i32]),
    "tf_ctx_create": (i32, [i32, C.POINTER(vp)]),
    "tf_ctx_destroy": (i32, [vp]),
    "tf_ctx_synchronize": (i32, [vp]),
    "tf_ctx_stream": (vp, [vp]),
    "tf_ctx_device": (i32, [vp]),
    "tf_event_create": (i32, [C.POINTER(vp)]),
    "tf_event_destroy": (i32, [vp]),
    "tf_event_record": (i32, [vp, vp]),
    "tf_event_elapsed_ms": (i32, [vp, vp, C.POINTER(f32)]),
    "tf_alloc": (i32, [i32, i32, sz, C.POINTER(vp)]),
    "tf_free": (i32, [i32, vp]),
    "tf_device_view": (i32, [i32, vp, C.POINTER(vp)]),
    "tf_memcpy": (i32, [vp, vp, sz]),
    "tf_memcpy_async": (i32, [vp, vp, vp, sz]),
    "tf_fill_zero": (i32, [vp, vp, sz]),
    "tf_spmv_csr_f64": (i32, [vp, i32, i64, vp, vp, vp, vp, vp]),
    "tf_spmv_csr_f32": (i32, [vp, i32, i64, vp, vp, vp, vp, vp]),
    "tf_residual_csr_f64": (i32, [vp, i32, i64, vp, vp, vp, vp, vp, vp]),
    "tf_residual_csr_f32": (i32, [vp, i32, i64, vp, vp, vp, vp, vp, vp]),
    "tf_spmm_csr_f64": (i32, [vp, i32, i64, vp, vp, vp, i32, vp, vp]),
    "tf_spmm_csr_f32": (i32, [vp, i32, i64, vp, vp, vp, i32, vp, vp]),
    "tf_asum_f64": (i32, [vp, i32, vp, F64P]),
    "tf_asum_f32": (i32, [vp, i32, vp, C.POINTER(f32)]),
    "tf_dense_transpose_f64": (i32, [vp, i32, i32, vp, vp]),
    "tf_dense_transpose_f32": (i32, [vp, i32, i32, vp, vp]),
    "tf_gemv_f64": (i32, [vp, i32, i32, vp, vp, vp]),
    "tf_gemv_f32": (i32, [vp, i32, i32, vp, vp, vp]),
    "tf_gemm_f64": (i32, [vp, i32, i32, i32, vp, vp, vp]),
    "tf_gemm_f32": (i32, [vp, i32, i32, i32, vp, vp, vp]),
    "tf_csr_add_f64": (i32, [vp, i32, i32, i64, vp, vp, vp, i64, vp, vp, vp, f64, vp, C.POINTER(vp), C.POINTER(vp),
                             I64P]),
    "tf_csr_transpose_f64": (i32, [vp, i32, i32, i64, vp, vp, vp, vp, C.POINTER(vp), C.POINTER(vp)]),
    "tf_csr_multiply_f64": (i32, [vp, i32, i32, i32, i64, vp, vp, vp, i64, vp, vp, vp, vp, C.POINTER(vp),
                                  C.POINTER(vp), I64P]),
    "tf_jacobi_update_f64": (i32, [vp, i32, f64, vp, vp, vp, vp]),
    "tf_jacobi_update_f32": (i32, [vp, i32, f32, vp, vp, vp, vp]),
    "tf_jacobi_fused_f64": (i32, [vp, i32, i64, vp, vp, vp, vp, f64, vp, vp, vp]),
    "tf_jacobi_fused_f32": (i32, [vp, i32, i64, vp, vp, vp, vp, f32, vp, vp, vp]),
    "tf_inverse_diagonal_f64": (i32, [vp, i32, vp, vp, vp, vp, I32P]),
    "tf_inverse_diagonal_f32": (i32, [vp, i32, vp, vp, vp, vp, I32P]),
    "tf_axpy_f64": (i32, [vp, i32, f64, vp, vp]),
    "tf_axpy_f32": (i32, [vp, i32, f32, vp, vp]),
    "tf_scal_f64": (i32, [vp, i32, f64, vp]),
    "tf_scal_f32": (i32, [vp, i32, f32, vp]),
    "tf_dot_f64": (i32, [vp, i32, vp, vp, F64P]),
    "tf_dot_f32": (i32, [vp, i32, vp, vp, C.POINTER(f32)]),
    "tf_nrm2_f64": (i32, [vp, i32, vp, F64P]),
    "tf_nrm2_f32": (i32, [vp, i32, vp, C.POINTER(f32)]),
    "tf_convert_f64_f32": (i32, [vp, sz, vp, vp]),
    "tf_convert_f32_f64": (i32, [vp, sz, vp, vp]),
    "tf_convert_i64_i32": (i32, [vp, sz, vp, vp]),
    "tf_convert_i32_i64": (i32, [vp, sz, vp, vp]),
    "tf_copy_f64": (i32, [vp, sz, vp, vp]),
    "tf_scan_excl_i32": (i32, [vp, i32, vp, vp]),
    "tf_scan_excl_i32_noalloc": (i32, [vp, i32, vp, vp]),
    "tf_csr_check": (i32, [vp, i32, i64, vp, vp, vp, I32P]),
    "tf_csr_validate_structure": (i32, [vp, i32, i32, i64, vp, vp, I32P]),
    "tf_csr_canonicalize": (i32, [vp, i32, i64, vp, vp, vp, vp, C.POINTER(vp), C.POINTER(vp), I64P]),
    "tf_agg_validate_rowmax": (i32, [vp, i32, vp, vp, vp, vp, I32P]),
    "tf_agg_strength": (i32, [vp, i32, i64, vp, vp, vp, f64, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), I64P]),
    "tf_agg_match_round": (i32, [vp, i32, vp, vp, vp, vp, vp]),
    "tf_agg_pair": (i32, [vp, i32, vp, vp, vp, i32, vp, I32P]),
    "tf_agg_contract": (i32, [vp, i32, vp, vp, vp, vp, i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), I64P]),
    "tf_agg_aggregate": (i32, [vp, i32, vp, vp, vp, i32, i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), I32P,
                               I64P]),
    "tf_galerkin": (i32, [vp, i32, vp, vp, vp, vp, i32, vp, vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), I64P]),
    "tf_restrict_f64": (i32, [vp, i32, vp, vp, vp, vp]),
    "tf_restrict_f32": (i32, [vp, i32, vp, vp, vp, vp]),
    "tf_prolong_f64": (i32, [vp, i32, vp, vp, vp]),
    "tf_prolong_f32": (i32, [vp, i32, vp, vp, vp]),
    "tf_coarse_lu_factor_f64": (i32, [vp, i32, vp, vp, vp, vp, vp, vp, I32P]),
    "tf_coarse_lu_solve_f64": (i32, [vp, i32, vp, vp, vp, vp]),
    "tf_coarse_lu_solve_f32": (i32, [vp, i32, vp, vp, vp, vp]),
    "tf_coarse_inv_apply_f64": (i32, [vp, i32, vp, vp, vp]),
    "tf_coarse_inv_apply_f32": (i32, [vp, i32, vp, vp, vp]),
    "tf_amg_create": (i32, [i32, C.POINTER(AmgOptions), C.POINTER(vp)]),
    "tf_amg_destroy": (i32, [vp]),
    "tf_amg_setup": (i32, [vp, i32, i64, vp, vp, vp, i32]),
    "tf_amg_apply": (i32, [vp, vp, vp, i32]),
    "tf_amg_apply_block": (i32, [vp, i32, vp, vp]),
    "tf_amg_block_visits": (i32, [vp, C.POINTER(C.c_int64)]),
    "tf_amg_level_count": (i32, [vp]),
    "tf_amg_level_info": (i32, [vp, i32, I64P, I64P, I64P]),
    "tf_amg_level_matrix": (i32, [vp, i32, vp, vp, vp]),
    "tf_amg_level_transfer": (i32, [vp, i32, vp, vp, vp]),
    "tf_amg_stop_reason": (C.c_char_p, [vp]),
    "tf_amg_coarse_solver": (C.c_char_p, [vp]),
    "tf_amg_visits": (i32, [vp, I64P]),
    "tf_solver_create": (i32, [i32, i32, C.POINTER(KrylovOptions), C.POINTER(AmgOptions), C.POINTER(vp)]),
    "tf_solver_destroy": (i32, [vp]),
    "tf_solver_setup": (i32, [vp, i32, i64, vp, vp, vp, i32]),
    "tf_solver_apply": (i32, [vp, vp, vp, i32, C.POINTER(SolveResultC)]),
    "tf_solver_apply_block": (i32, [vp, i32, vp, vp, C.POINTER(SolveResultC)]),
    "tf_solver_apply_segmented": (i32, [vp, i32, vp, vp, vp, C.POINTER(SolveResultC)]),
    "tf_solver_apply_segmented_persistent": (i32, [vp, i32, vp, vp, vp, C.POINTER(SolveResultC)]),
    "tf_gnn_param_count": (i32, []),
    "tf_gnn_create": (i32, [i32, C.POINTER(vp)]),
    "tf_gnn_destroy": (i32, [vp]),
    "tf_gnn_set_params": (i32, [vp, vp, i32]),
    "tf_gnn_get_params": (i32, [vp, vp, i32]),
    "tf_gnn_infer": (i32, [vp, i32, i64, vp, vp, vp, vp]),
    "tf_gnn_loss_grad": (i32, [vp, i32, i64, vp, vp, vp, vp, C.POINTER(f64), vp]),
    "tf_gnn_train_step": (i32, [vp, i32, i64, vp, vp, vp, vp, C.c_float, C.POINTER(f64)]),
    "tf_affinities": (i32, [vp, i32, i64, vp, vp, vp, i32, i32, f64, vp, vp]),
    "tf_amg_set_affinity_model": (i32, [vp, vp]),
    "tf_amg_learned_info": (i32, [vp, C.POINTER(i32), C.POINTER(f64), C.POINTER(f64)]),
    "tf_amg_convergence_estimate": (i32, [vp, C.POINTER(f64)]),
    "tf_solver_create_with_amg": (i32, [i32, C.POINTER(KrylovOptions), vp, C.POINTER(vp)]),
    "tf_solver_amg": (vp, [vp]),
    "tf_solver_timings": (i32, [vp, F64P, F64P]),
}

_lib = None


class TfError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def header_functions():
    """Names of every function the C ABI header declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tf_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Load the HIP library (raises ImportError when it is absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run python -c 'import __graft_entry__; __graft_entry__.build()'")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().tf_last_error()
        raise TfError(rc, msg.decode() if msg else f"error {rc}")
    return rc
