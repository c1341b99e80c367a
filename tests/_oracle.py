"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and to the
reference-built checker (oracle/_ref/libref_oracle.so, present only where the
reference tree was available at build time).  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_oracle.so")

I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
I64P = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class AmgOptions(C.Structure):
    _fields_ = [("max_levels", C.c_int), ("coarse_rows", C.c_int), ("matching_rounds", C.c_int),
                ("pre_sweeps", C.c_int), ("post_sweeps", C.c_int), ("coarse_sweeps", C.c_int),
                ("jacobi_weight", C.c_double), ("strength_threshold", C.c_double),
                ("absorb_singletons", C.c_int), ("cycle", C.c_int), ("level_precision", C.c_int)]


class SolveResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("residual_norm", C.c_double),
                ("relative_residual", C.c_double)]


def amg_options(max_levels=20, coarse_rows=64, matching_rounds=8, pre_sweeps=1, post_sweeps=1, coarse_sweeps=8,
                jacobi_weight=0.75, strength_threshold=0.25, absorb_singletons=True, cycle=0, level_precision=64):
    return AmgOptions(max_levels, coarse_rows, matching_rounds, pre_sweeps, post_sweeps, coarse_sweeps,
                      jacobi_weight, strength_threshold, int(absorb_singletons), cycle, level_precision)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])
        L = C.CDLL(ORACLE_SO)
        L.tfo_last_error.restype = C.c_char_p
        L.tfo_canonical.restype = C.c_int64
        L.tfo_canonical.argtypes = [C.c_int, I32P, I32P, F64P, I32P, I32P, F64P]
        L.tfo_validate_rowmax.argtypes = [C.c_int, I32P, I32P, F64P, F64P]
        L.tfo_strength.restype = C.c_int64
        L.tfo_strength.argtypes = [C.c_int, I32P, I32P, F64P, C.c_double, I32P, I32P, F64P]
        L.tfo_match.argtypes = [C.c_int, I32P, I32P, F64P, C.c_int, I32P]
        L.tfo_contract.restype = C.c_int64
        L.tfo_contract.argtypes = [C.c_int, I32P, I32P, F64P, I32P, I32P, I32P, F64P]
        L.tfo_aggregate.argtypes = [C.c_int, I32P, I32P, F64P, C.c_int, C.c_int, I32P, I32P, I32P, I64P]
        L.tfo_galerkin.restype = C.c_int64
        L.tfo_galerkin.argtypes = [C.c_int, I32P, I32P, F64P, I32P, I32P, I32P, F64P]
        L.tfo_lu_factor.argtypes = [C.c_int, I32P, I32P, F64P, F64P, I32P]
        L.tfo_lu_solve.argtypes = [C.c_int, F64P, I32P, F64P, F64P]
        L.tfo_spmv.argtypes = [C.c_int, I32P, I32P, F64P, F64P, F64P]
        L.tfo_spmv_f32.argtypes = [C.c_int, I32P, I32P, F32P, F32P, F32P]
        L.tfo_amg_create.restype = C.c_void_p
        L.tfo_amg_create.argtypes = [C.POINTER(AmgOptions)]
        L.tfo_amg_destroy.argtypes = [C.c_void_p]
        L.tfo_amg_setup.argtypes = [C.c_void_p, C.c_int, I32P, I32P, F64P]
        L.tfo_amg_levels.argtypes = [C.c_void_p]
        L.tfo_amg_level_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64), I64P]
        L.tfo_amg_level_matrix.argtypes = [C.c_void_p, C.c_int, I32P, I32P, F64P]
        L.tfo_amg_level_transfer.argtypes = [C.c_void_p, C.c_int, I32P, I32P, I32P]
        L.tfo_amg_stop_reason.restype = C.c_char_p
        L.tfo_amg_stop_reason.argtypes = [C.c_void_p]
        L.tfo_amg_coarse_solver.restype = C.c_char_p
        L.tfo_amg_coarse_solver.argtypes = [C.c_void_p]
        L.tfo_amg_apply.argtypes = [C.c_void_p, F64P, F64P]
        L.tfo_amg_visits.argtypes = [C.c_void_p, I64P]
        L.tfo_krylov_solve.argtypes = [C.c_int, C.c_void_p, C.c_int, I32P, I32P, F64P, F64P, F64P, C.c_int,
                                       C.c_double, C.c_double, C.c_int, C.POINTER(SolveResult)]
        L.tfo_jacobi.argtypes = [C.c_int, I32P, I32P, F64P, C.c_int, C.c_double, F64P, F64P]
        _lib = L
    return _lib


def _err(L):
    return RuntimeError(L.tfo_last_error().decode())


def canonical(rp, ci, v):
    L = lib(); n = len(rp) - 1; m = len(v)
    orp = np.zeros(n + 1, np.int32); oci = np.zeros(max(m, 1), np.int32); ov = np.zeros(max(m, 1))
    k = L.tfo_canonical(n, rp, ci, v, orp, oci, ov)
    if k < 0:
        raise _err(L)
    return orp, oci[:k].copy(), ov[:k].copy()


def rowmax(rp, ci, v):
    L = lib(); n = len(rp) - 1
    out = np.zeros(n)
    if L.tfo_validate_rowmax(n, rp, ci, v, out) < 0:
        raise _err(L)
    return out


def strength(rp, ci, v, theta):
    L = lib(); n = len(rp) - 1; m = 2 * len(v) + 1
    orp = np.zeros(n + 1, np.int32); oci = np.zeros(m, np.int32); ov = np.zeros(m)
    k = L.tfo_strength(n, rp, ci, v, theta, orp, oci, ov)
    if k < 0:
        raise _err(L)
    return orp, oci[:k].copy(), ov[:k].copy()


def match(g, rounds):
    L = lib(); rp, ci, w = g; n = len(rp) - 1
    m = np.zeros(n, np.int32)
    c = L.tfo_match(n, rp, ci, w, rounds, m)
    if c < 0:
        raise _err(L)
    return m, c


def contract(g, mp):
    L = lib(); rp, ci, w = g; n = len(rp) - 1
    count = int(mp.max()) + 1
    orp = np.zeros(count + 1, np.int32); oci = np.zeros(len(w) + 1, np.int32); ow = np.zeros(len(w) + 1)
    k = L.tfo_contract(n, rp, ci, w, np.ascontiguousarray(mp, np.int32), orp, oci, ow)
    if k < 0:
        raise _err(L)
    return orp, oci[:k].copy(), ow[:k].copy()


def aggregate(g, rounds, absorb):
    L = lib(); rp, ci, w = g; n = len(rp) - 1
    mp = np.zeros(n, np.int32); off = np.zeros(n + 1, np.int32); mem = np.zeros(n, np.int32)
    hist = np.zeros(4, np.int64)
    nc = L.tfo_aggregate(n, rp, ci, w, rounds, int(absorb), mp, off, mem, hist)
    if nc < 0:
        raise _err(L)
    return mp, off[:nc + 1].copy(), mem, hist


def galerkin(a, mp):
    L = lib(); rp, ci, v = a; n = len(rp) - 1
    nc = int(mp.max()) + 1
    orp = np.zeros(nc + 1, np.int32); oci = np.zeros(len(v) + 1, np.int32); ov = np.zeros(len(v) + 1)
    k = L.tfo_galerkin(n, rp, ci, v, np.ascontiguousarray(mp, np.int32), orp, oci, ov)
    if k < 0:
        raise _err(L)
    return orp, oci[:k].copy(), ov[:k].copy()


def lu_factor(a):
    L = lib(); rp, ci, v = a; n = len(rp) - 1
    lu = np.zeros(n * n); piv = np.zeros(n, np.int32)
    ok = L.tfo_lu_factor(n, rp, ci, v, lu, piv)
    if ok < 0:
        raise _err(L)
    return bool(ok), lu, piv


def lu_solve(lu, piv, b):
    L = lib(); n = len(b); x = np.zeros(n)
    L.tfo_lu_solve(n, lu, piv, np.ascontiguousarray(b, np.float64), x)
    return x


def spmv(a, x):
    L = lib(); rp, ci, v = a; n = len(rp) - 1
    y = np.zeros(n)
    L.tfo_spmv(n, np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32),
               np.ascontiguousarray(v, np.float64), np.ascontiguousarray(x, np.float64), y)
    return y


def spmv_f32(a, x):
    L = lib(); rp, ci, v = a; n = len(rp) - 1
    y = np.zeros(n, np.float32)
    L.tfo_spmv_f32(n, rp, ci, np.ascontiguousarray(v, np.float32), np.ascontiguousarray(x, np.float32), y)
    return y


class Amg:
    """Host AMG hierarchy + cycle (Amg.cuh:81-161)."""

    def __init__(self, **opts):
        self.L = lib()
        self.opts = amg_options(**opts)
        self.h = self.L.tfo_amg_create(C.byref(self.opts))
        if not self.h:
            raise _err(self.L)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.tfo_amg_destroy(self.h)
            self.h = None

    def setup(self, a):
        rp, ci, v = a
        if self.L.tfo_amg_setup(self.h, len(rp) - 1, rp, ci, v) < 0:
            raise _err(self.L)
        return self

    def levels(self):
        return self.L.tfo_amg_levels(self.h)

    def level_info(self, d):
        r = C.c_int64(); z = C.c_int64(); h = np.zeros(4, np.int64)
        self.L.tfo_amg_level_info(self.h, d, C.byref(r), C.byref(z), h)
        return r.value, z.value, h

    def level_matrix(self, d):
        r, z, _ = self.level_info(d)
        rp = np.zeros(r + 1, np.int32); ci = np.zeros(max(z, 1), np.int32); v = np.zeros(max(z, 1))
        self.L.tfo_amg_level_matrix(self.h, d, rp, ci, v)
        return rp, ci[:z], v[:z]

    def level_transfer(self, d):
        r, _, _ = self.level_info(d)
        mp = np.zeros(r, np.int32); off = np.zeros(r + 1, np.int32); mem = np.zeros(r, np.int32)
        nc = self.L.tfo_amg_level_transfer(self.h, d, mp, off, mem)
        return mp, off[:nc + 1], mem

    def stop_reason(self):
        return self.L.tfo_amg_stop_reason(self.h).decode()

    def coarse_solver(self):
        return self.L.tfo_amg_coarse_solver(self.h).decode()

    def apply(self, b):
        x = np.zeros(len(b))
        if self.L.tfo_amg_apply(self.h, np.ascontiguousarray(b, np.float64), x) < 0:
            raise _err(self.L)
        return x

    def visits(self):
        v = np.zeros(self.levels(), np.int64)
        self.L.tfo_amg_visits(self.h, v)
        return v


def krylov(kind, amg, a, b, x0=None, maxit=1000, rtol=1e-8, atol=0.0, refresh=50):
    """kind 'cg' or 'bicgstab'; returns (x, status, iterations, residual, relative)."""
    L = lib(); rp, ci, v = a; n = len(rp) - 1
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    res = SolveResult()
    rc = L.tfo_krylov_solve(0 if kind == "cg" else 1, amg.h if amg is not None else None, n, rp, ci, v,
                            np.ascontiguousarray(b, np.float64), x, maxit, rtol, atol, refresh, C.byref(res))
    if rc < 0:
        raise _err(L)
    return x, res.status, res.iterations, res.residual_norm, res.relative_residual


def jacobi(a, sweeps, weight, b, x):
    L = lib(); rp, ci, v = a; n = len(rp) - 1
    x = np.array(x, np.float64)
    if L.tfo_jacobi(n, rp, ci, v, sweeps, weight, np.ascontiguousarray(b, np.float64), x) < 0:
        raise _err(L)
    return x


# ---- reference-built checker (present only where the reference tree was) ----

def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_SO)
        R.tfr_last_error.restype = C.c_char_p
        for name in ("tfr_canonical", "tfr_galerkin", "tfr_contract", "tfr_strength", "tfr_match",
                     "tfr_aggregate", "tfr_read_matrix_market"):
            getattr(R, name).restype = C.c_int64
        R.tfr_canonical.argtypes = [C.c_int, I32P, I32P, F64P, I32P, I32P, F64P]
        R.tfr_strength.argtypes = [C.c_int, I32P, I32P, F64P, C.c_double, I32P, I32P, F64P]
        R.tfr_match.argtypes = [C.c_int, I32P, I32P, F64P, C.c_int, I32P]
        R.tfr_contract.argtypes = [C.c_int, I32P, I32P, F64P, I32P, I32P, I32P, F64P]
        R.tfr_aggregate.argtypes = [C.c_int, I32P, I32P, F64P, C.c_int, C.c_int, I32P, I32P, I32P, I64P]
        R.tfr_galerkin.argtypes = [C.c_int, I32P, I32P, F64P, C.c_int, I32P, I32P, I32P, I32P, I32P, F64P]
        _ref = R
    return _ref


def _rerr(R):
    return RuntimeError(R.tfr_last_error().decode())


def ref_canonical(rp, ci, v):
    R = ref(); n = len(rp) - 1; m = len(v)
    orp = np.zeros(n + 1, np.int32); oci = np.zeros(max(m, 1), np.int32); ov = np.zeros(max(m, 1))
    k = R.tfr_canonical(n, rp, ci, v, orp, oci, ov)
    if k < 0:
        raise _rerr(R)
    return orp, oci[:k].copy(), ov[:k].copy()


def ref_strength(rp, ci, v, theta):
    R = ref(); n = len(rp) - 1; m = 2 * len(v) + 1
    orp = np.zeros(n + 1, np.int32); oci = np.zeros(m, np.int32); ov = np.zeros(m)
    k = R.tfr_strength(n, rp, ci, v, theta, orp, oci, ov)
    if k < 0:
        raise _rerr(R)
    return orp, oci[:k].copy(), ov[:k].copy()


def ref_aggregate(g, rounds, absorb):
    R = ref(); rp, ci, w = g; n = len(rp) - 1
    mp = np.zeros(n, np.int32); off = np.zeros(n + 1, np.int32); mem = np.zeros(n, np.int32)
    hist = np.zeros(4, np.int64)
    nc = R.tfr_aggregate(n, rp, ci, w, rounds, int(absorb), mp, off, mem, hist)
    if nc < 0:
        raise _rerr(R)
    return mp, off[:nc + 1].copy(), mem, hist


def ref_galerkin(a, t):
    R = ref(); rp, ci, v = a; n = len(rp) - 1
    mp, off, mem = t[0], t[1], t[2]
    nc = len(off) - 1
    orp = np.zeros(nc + 1, np.int32); oci = np.zeros(len(v) + 1, np.int32); ov = np.zeros(len(v) + 1)
    k = R.tfr_galerkin(n, rp, ci, v, nc, mp, off, mem, orp, oci, ov)
    if k < 0:
        raise _rerr(R)
    return orp, oci[:k].copy(), ov[:k].copy()
