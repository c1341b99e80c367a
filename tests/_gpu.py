"""GPU-side helpers for the parity tests: thin wrappers that run one C-ABI
entry point on device copies of host arrays and bring the result back."""
import ctypes as C

import numpy as np

from thermofem_amd._lib import check, lib
from thermofem_amd.device import Context, DeviceArray

_ctx = None


def ctx():
    global _ctx
    if _ctx is None:
        _ctx = Context(0)
    return _ctx


def up(a, dtype=None):
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype))
    return DeviceArray.from_host(a)


def _adopt_csr(rp_ptr, ci_ptr, v_ptr, n, nnz):
    rp = DeviceArray.adopt(rp_ptr, n + 1, np.int32)
    ci = DeviceArray.adopt(ci_ptr, nnz, np.int32)
    v = DeviceArray.adopt(v_ptr, nnz, np.float64)
    return rp.get(), ci.get(), v.get()


def canonical(a):
    rp, ci, v = (up(a[0], np.int32), up(a[1], np.int32), up(a[2], np.float64))
    n = len(a[0]) - 1
    orp = DeviceArray(n + 1, np.int32)
    oci, ov, nnz = C.c_void_p(), C.c_void_p(), C.c_int64()
    check(lib().tf_csr_canonicalize(ctx().h, n, len(a[1]), rp.ptr, ci.ptr, v.ptr, orp.ptr, C.byref(oci), C.byref(ov),
                                    C.byref(nnz)))
    return orp.get(), DeviceArray.adopt(oci, nnz.value, np.int32).get(), DeviceArray.adopt(ov, nnz.value,
                                                                                          np.float64).get()


def strength(a, theta):
    rp, ci, v = (up(a[0], np.int32), up(a[1], np.int32), up(a[2], np.float64))
    n = len(a[0]) - 1
    g = [C.c_void_p(), C.c_void_p(), C.c_void_p()]
    nnz = C.c_int64()
    check(lib().tf_agg_strength(ctx().h, n, len(a[1]), rp.ptr, ci.ptr, v.ptr, theta, C.byref(g[0]), C.byref(g[1]),
                                C.byref(g[2]), C.byref(nnz)))
    return _adopt_csr(g[0], g[1], g[2], n, nnz.value)


def aggregate(g, rounds, absorb):
    rp, ci, w = (up(g[0], np.int32), up(g[1], np.int32), up(g[2], np.float64))
    n = len(g[0]) - 1
    mp, off, mem = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nc = C.c_int()
    hist = (C.c_int64 * 4)()
    check(lib().tf_agg_aggregate(ctx().h, n, rp.ptr, ci.ptr, w.ptr, rounds, int(absorb), C.byref(mp), C.byref(off),
                                 C.byref(mem), C.byref(nc), hist))
    return (DeviceArray.adopt(mp, n, np.int32).get(), DeviceArray.adopt(off, nc.value + 1, np.int32).get(),
            DeviceArray.adopt(mem, n, np.int32).get(), np.array(list(hist), np.int64))


def galerkin(a, t):
    mp, off, mem = t[0], t[1], t[2]
    n = len(a[0]) - 1
    nc = len(off) - 1
    rp, ci, v = (up(a[0], np.int32), up(a[1], np.int32), up(a[2], np.float64))
    dm, doff, dmem = up(mp, np.int32), up(off, np.int32), up(mem, np.int32)
    o = [C.c_void_p(), C.c_void_p(), C.c_void_p()]
    nnz = C.c_int64()
    check(lib().tf_galerkin(ctx().h, n, rp.ptr, ci.ptr, v.ptr, dm.ptr, nc, doff.ptr, dmem.ptr, C.byref(o[0]),
                            C.byref(o[1]), C.byref(o[2]), C.byref(nnz)))
    return _adopt_csr(o[0], o[1], o[2], nc, nnz.value)


def restrict(t, fine, dtype=np.float64):
    off, mem = t[1], t[2]
    nc = len(off) - 1
    f = up(fine, dtype)
    c = DeviceArray(nc, dtype)
    doff, dmem = up(off, np.int32), up(mem, np.int32)  # keep alive across the launch
    fn = lib().tf_restrict_f64 if dtype == np.float64 else lib().tf_restrict_f32
    check(fn(ctx().h, nc, doff.ptr, dmem.ptr, f.ptr, c.ptr))
    ctx().sync()
    return c.get()


def prolong(t, coarse, fine, dtype=np.float64):
    mp = t[0]
    f = up(fine, dtype)
    c = up(coarse, dtype)
    dmp = up(mp, np.int32)  # keep alive across the launch
    fn = lib().tf_prolong_f64 if dtype == np.float64 else lib().tf_prolong_f32
    check(fn(ctx().h, len(mp), dmp.ptr, c.ptr, f.ptr))
    ctx().sync()
    return f.get()


def spmv(a, x, dtype=np.float64):
    n = len(a[0]) - 1
    rp, ci, v = up(a[0], np.int32), up(a[1], np.int32), up(a[2], dtype)
    dx = up(x, dtype)
    y = DeviceArray(n, dtype)
    fn = lib().tf_spmv_csr_f64 if dtype == np.float64 else lib().tf_spmv_csr_f32
    check(fn(ctx().h, n, len(a[1]), rp.ptr, ci.ptr, v.ptr, dx.ptr, y.ptr))
    ctx().sync()
    return y.get()


def spmm(a, X, dtype=np.float64):
    """Y = A X for a row-major n x k block."""
    n = len(a[0]) - 1
    X = np.ascontiguousarray(X, dtype)
    k = X.shape[1]
    rp, ci, v = up(a[0], np.int32), up(a[1], np.int32), up(a[2], dtype)
    dx = up(X.ravel(), dtype)
    y = DeviceArray(n * k, dtype)
    fn = lib().tf_spmm_csr_f64 if dtype == np.float64 else lib().tf_spmm_csr_f32
    check(fn(ctx().h, n, len(a[1]), rp.ptr, ci.ptr, v.ptr, k, dx.ptr, y.ptr))
    ctx().sync()
    return y.get().reshape(n, k)


def lu_factor(a):
    n = len(a[0]) - 1
    rp, ci, v = up(a[0], np.int32), up(a[1], np.int32), up(a[2], np.float64)
    lu = DeviceArray(n * n, np.float64)
    piv = DeviceArray(n, np.int32)
    inv = DeviceArray(n * n, np.float64)
    sing = C.c_int()
    check(lib().tf_coarse_lu_factor_f64(ctx().h, n, rp.ptr, ci.ptr, v.ptr, lu.ptr, piv.ptr, inv.ptr, C.byref(sing)))
    return sing.value == 0, lu, piv, inv


class Sell:
    """A sliced-ELL handle (tf_sell_*) built from a host CSR matrix."""

    def __init__(self, a, max_padding=1.0):
        self.n = len(a[0]) - 1
        self.rp, self.ci, self.v = up(a[0], np.int32), up(a[1], np.int32), up(a[2], np.float64)
        self.h = C.c_void_p()
        check(lib().tf_sell_create(ctx().h, self.n, len(a[1]), self.rp.ptr, self.ci.ptr, self.v.ptr, max_padding,
                                   C.byref(self.h)))

    def info(self):
        bits, stored, reordered = C.c_int(), C.c_int64(), C.c_int()
        check(lib().tf_sell_info(self.h, C.byref(bits), C.byref(stored), C.byref(reordered)))
        return bits.value, stored.value, bool(reordered.value)

    def spmv(self, x, dtype=np.float64):
        dx, y = up(x, dtype), DeviceArray(self.n, dtype)
        fn = lib().tf_sell_spmv_f64 if dtype == np.float64 else lib().tf_sell_spmv_f32
        check(fn(ctx().h, self.h, dx.ptr, y.ptr))
        ctx().sync()
        return y.get()

    def residual(self, b, x, dtype=np.float64):
        db, dx, r = up(b, dtype), up(x, dtype), DeviceArray(self.n, dtype)
        fn = lib().tf_sell_residual_f64 if dtype == np.float64 else lib().tf_sell_residual_f32
        check(fn(ctx().h, self.h, db.ptr, dx.ptr, r.ptr))
        ctx().sync()
        return r.get()

    def jacobi(self, inv, w, b, x, dtype=np.float64):
        di, db, dx, out = up(inv, dtype), up(b, dtype), up(x, dtype), DeviceArray(self.n, dtype)
        fn = lib().tf_sell_jacobi_f64 if dtype == np.float64 else lib().tf_sell_jacobi_f32
        check(fn(ctx().h, self.h, di.ptr, w, db.ptr, dx.ptr, out.ptr))
        ctx().sync()
        return out.get()

    def close(self):
        if self.h:
            check(lib().tf_sell_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stencil:
    """A structured-stencil handle (tf_stencil_*) built from a host CSR matrix."""

    def __init__(self, a):
        self.n = len(a[0]) - 1
        self.rp, self.ci, self.v = up(a[0], np.int32), up(a[1], np.int32), up(a[2], np.float64)
        self.h = C.c_void_p()
        check(lib().tf_stencil_create(ctx().h, self.n, len(a[1]), self.rp.ptr, self.ci.ptr, self.v.ptr,
                                      C.byref(self.h)))

    def info(self):
        terms, arrays, mirrored = C.c_int(), C.c_int(), C.c_int()
        check(lib().tf_stencil_info(self.h, C.byref(terms), C.byref(arrays), C.byref(mirrored)))
        return terms.value, arrays.value, bool(mirrored.value)

    def spmv(self, x, dtype=np.float64):
        dx, y = up(x, dtype), DeviceArray(self.n, dtype)
        fn = lib().tf_stencil_spmv_f64 if dtype == np.float64 else lib().tf_stencil_spmv_f32
        check(fn(ctx().h, self.h, dx.ptr, y.ptr))
        ctx().sync()
        return y.get()

    def residual(self, b, x, dtype=np.float64):
        db, dx, r = up(b, dtype), up(x, dtype), DeviceArray(self.n, dtype)
        fn = lib().tf_stencil_residual_f64 if dtype == np.float64 else lib().tf_stencil_residual_f32
        check(fn(ctx().h, self.h, db.ptr, dx.ptr, r.ptr))
        ctx().sync()
        return r.get()

    def jacobi(self, inv, w, b, x, dtype=np.float64):
        di, db, dx, out = up(inv, dtype), up(b, dtype), up(x, dtype), DeviceArray(self.n, dtype)
        fn = lib().tf_stencil_jacobi_f64 if dtype == np.float64 else lib().tf_stencil_jacobi_f32
        check(fn(ctx().h, self.h, di.ptr, w, db.ptr, dx.ptr, out.ptr))
        ctx().sync()
        return out.get()

    def close(self):
        if self.h:
            check(lib().tf_stencil_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
