"""Python mirror of the reference's solver interface on top of the C ABI.

Names and meanings follow include/Solver/Interfaces.cuh (SolveStatus,
SolveResult), include/Preconditioner/Amg.cuh (AmgOptions, Amg diagnostics),
include/Solver/Cg.cuh and include/Solver/BiCgStab.cuh (options, solve).
Contract violations raise ValueError (std::invalid_argument in the
reference); numerical failure is reported through SolveStatus.
"""
import ctypes as C
import enum
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import TfError, check, lib
from .device import DeviceArray


class Cycle(enum.IntEnum):
    V = 0
    W = 1
    F = 2


class SolveStatus(enum.IntEnum):
    Converged = 0
    MaxIterations = 1
    Breakdown = 2
    NonFinite = 3


@dataclass
class SolveResult:
    status: SolveStatus
    iterations: int
    residual_norm: float
    relative_residual: float


@dataclass
class AmgOptions:
    max_levels: int = 20
    coarse_rows: int = 64
    matching_rounds: int = 8
    pre_sweeps: int = 1
    post_sweeps: int = 1
    coarse_sweeps: int = 8
    jacobi_weight: float = 0.75
    strength_threshold: float = 0.25
    absorb_singletons: bool = True
    cycle: Cycle = Cycle.V
    level_precision: int = 64  # 32: FP32 cycle under the FP64 caller
    coarse_tail: int = 0  # coarse levels as one one-workgroup program: 0 timed at setup, 1 always, 2 never
    learned: int = 0  # matching weights: 0 strength, 1 learned affinities, 2 learned where PCG converges faster with it

    def c(self):
        return _lib.AmgOptions(self.max_levels, self.coarse_rows, self.matching_rounds, self.pre_sweeps,
                               self.post_sweeps, self.coarse_sweeps, self.jacobi_weight, self.strength_threshold,
                               int(self.absorb_singletons), int(self.cycle), self.level_precision,
                               self.coarse_tail, self.learned)


@dataclass
class KrylovOptions:
    max_iterations: int = 1000
    relative_tolerance: float = 1e-8
    absolute_tolerance: float = 0.0
    residual_refresh_interval: int = 50
    use_graph: bool = True

    def c(self):
        return _lib.KrylovOptions(self.max_iterations, self.relative_tolerance, self.absolute_tolerance,
                                  self.residual_refresh_interval, int(self.use_graph))


CgOptions = KrylovOptions
BiCgStabOptions = KrylovOptions


def _raise(e):
    if e.code in (-1, -4):
        raise ValueError(str(e)) from None
    raise e


def _csr_args(a):
    rp, ci, v = (np.ascontiguousarray(a[0], np.int32), np.ascontiguousarray(a[1], np.int32),
                 np.ascontiguousarray(a[2], np.float64))
    return rp, ci, v


class _AmgView:
    """Diagnostics of an AMG hierarchy (Amg.cuh:163-172)."""

    def __init__(self, handle):
        self._h = handle

    def level_count(self):
        return lib().tf_amg_level_count(self._h)

    def level_statistics(self):
        out = []
        for d in range(self.level_count()):
            r, z, h = C.c_int64(), C.c_int64(), (C.c_int64 * 4)()
            check(lib().tf_amg_level_info(self._h, d, C.byref(r), C.byref(z), h))
            out.append((r.value, z.value, list(h)))
        return out

    def stop_reason(self):
        return lib().tf_amg_stop_reason(self._h).decode()

    def coarse_solver(self):
        return lib().tf_amg_coarse_solver(self._h).decode()

    def last_cycle_visits(self):
        v = (C.c_int64 * self.level_count())()
        check(lib().tf_amg_visits(self._h, v))
        return list(v)

    def set_affinity_model(self, model):
        """Network (gnn.AffinityModel, or None) whose weights the matching uses
        when the options ask for learned affinities; keep it alive meanwhile."""
        self._model = model
        check(lib().tf_amg_set_affinity_model(self._h, model.handle if model is not None else None))

    def learned_info(self):
        """(learned hierarchy in use, its contraction factor, the plain one's); -1: not measured."""
        u, a, b = C.c_int(), C.c_double(), C.c_double()
        check(lib().tf_amg_learned_info(self._h, C.byref(u), C.byref(a), C.byref(b)))
        return bool(u.value), a.value, b.value

    def convergence_estimate(self):
        r = C.c_double()
        check(lib().tf_amg_convergence_estimate(self._h, C.byref(r)))
        return r.value

    def level_matrix(self, d):
        r, z, _ = self.level_statistics()[d]
        rp = np.zeros(r + 1, np.int32); ci = np.zeros(max(z, 1), np.int32); v = np.zeros(max(z, 1))
        check(lib().tf_amg_level_matrix(self._h, d, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                                        v.ctypes.data_as(C.c_void_p)))
        return rp, ci[:z], v[:z]

    def level_transfer(self, d):
        stats = self.level_statistics()
        r = stats[d][0]; nc = stats[d + 1][0]
        mp = np.zeros(r, np.int32); off = np.zeros(nc + 1, np.int32); mem = np.zeros(r, np.int32)
        check(lib().tf_amg_level_transfer(self._h, d, mp.ctypes.data_as(C.c_void_p),
                                          off.ctypes.data_as(C.c_void_p), mem.ctypes.data_as(C.c_void_p)))
        return mp, off, mem


class Amg(_AmgView):
    """Aggregation AMG preconditioner (Amg<double,int>, or the FP32-cycle adapter)."""

    def __init__(self, options=None, device=0):
        self.options = options or AmgOptions()
        h = C.c_void_p()
        o = self.options.c()
        try:
            check(lib().tf_amg_create(device, C.byref(o), C.byref(h)))
        except TfError as e:
            _raise(e)
        super().__init__(h)
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None):
            lib().tf_amg_destroy(self.handle)
            self.handle = None

    def setup(self, a):
        rp, ci, v = _csr_args(a)
        try:
            check(lib().tf_amg_setup(self.handle, len(rp) - 1, len(ci), rp.ctypes.data_as(C.c_void_p),
                                     ci.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p), 1))
        except TfError as e:
            _raise(e)
        self.n = len(rp) - 1
        return self

    def apply(self, rhs):
        rhs = np.ascontiguousarray(rhs, np.float64)
        x = np.zeros_like(rhs)
        try:
            check(lib().tf_amg_apply(self.handle, rhs.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p), 1))
        except TfError as e:
            _raise(e)
        return x

    def apply_block(self, rhs):
        """One cycle on each column of a row-major n x k block (host arrays)."""
        rhs = np.ascontiguousarray(rhs, np.float64)
        n, k = rhs.shape
        dr = DeviceArray.from_host(rhs.ravel())
        dz = DeviceArray(n * k, np.float64)
        try:
            check(lib().tf_amg_apply_block(self.handle, k, dr.ptr, dz.ptr))
        except TfError as e:
            _raise(e)
        return dz.get().reshape(n, k)

    def last_block_visits(self):
        v = (C.c_int64 * self.level_count())()
        check(lib().tf_amg_block_visits(self.handle, v))
        return list(v)


class _Krylov:
    kind = 0

    def __init__(self, options=None, preconditioner="default", device=0):
        """preconditioner: an AmgOptions, None (Identity) or 'default'."""
        self.options = options or KrylovOptions()
        if isinstance(preconditioner, str):
            preconditioner = AmgOptions() if self.kind == 1 else None
        self.amg_options = preconditioner
        h = C.c_void_p()
        o = self.options.c()
        ao = preconditioner.c() if preconditioner is not None else None
        try:
            check(lib().tf_solver_create(device, self.kind, C.byref(o), C.byref(ao) if ao is not None else None,
                                         C.byref(h)))
        except TfError as e:
            _raise(e)
        self.handle = h
        self.device = device

    @classmethod
    def with_amg(cls, amg, options=None):
        """A solver preconditioned by an existing Amg (tf_solver_create_with_amg):
        the hierarchy is shared, so setting up either re-builds it for both."""
        self = cls.__new__(cls)
        self.options = options or KrylovOptions()
        self.amg_options = amg.options
        self._amg = amg  # the solver borrows the hierarchy: keep it alive
        h = C.c_void_p()
        o = self.options.c()
        try:
            check(lib().tf_solver_create_with_amg(self.kind, C.byref(o), amg.handle, C.byref(h)))
        except TfError as e:
            _raise(e)
        self.handle = h
        self.device = 0
        return self

    def __del__(self):
        if getattr(self, "handle", None):
            lib().tf_solver_destroy(self.handle)
            self.handle = None

    def setup(self, a):
        rp, ci, v = _csr_args(a)
        self.n = len(rp) - 1
        try:
            check(lib().tf_solver_setup(self.handle, self.n, len(ci), rp.ctypes.data_as(C.c_void_p),
                                        ci.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p), 1))
        except TfError as e:
            _raise(e)
        return self

    def setup_device(self, n, nnz, rp, ci, v):
        """Setup from device arrays (DeviceArray pointers)."""
        self.n = n
        check(lib().tf_solver_setup(self.handle, n, nnz, rp, ci, v, 0))
        return self

    def apply(self, b, x0=None):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64)
        r = _lib.SolveResultC()
        try:
            check(lib().tf_solver_apply(self.handle, b.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p), 1,
                                        C.byref(r)))
        except TfError as e:
            _raise(e)
        return x, SolveResult(SolveStatus(r.status), r.iterations, r.residual_norm, r.relative_residual)

    def apply_device(self, b, x):
        """Solve with device vectors (DeviceArray); x holds the guess and the result."""
        r = _lib.SolveResultC()
        check(lib().tf_solver_apply(self.handle, b.ptr, x.ptr, 0, C.byref(r)))
        return SolveResult(SolveStatus(r.status), r.iterations, r.residual_norm, r.relative_residual)

    def apply_block(self, b, x0=None):
        """CG on the k columns of a row-major n x k block at once (host arrays):
        returns (x, [SolveResult per column])."""
        b = np.ascontiguousarray(b, np.float64)
        n, k = b.shape
        x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64)
        db = DeviceArray.from_host(b.ravel())
        dx = DeviceArray.from_host(x.ravel())
        res = self.apply_block_device(k, db, dx)
        return dx.get().reshape(n, k), res

    def apply_block_device(self, k, b, x):
        """Block solve with device blocks (DeviceArray, row-major n x k)."""
        r = (_lib.SolveResultC * k)()
        try:
            check(lib().tf_solver_apply_block(self.handle, k, b.ptr, x.ptr, r))
        except TfError as e:
            _raise(e)
        return [SolveResult(SolveStatus(q.status), q.iterations, q.residual_norm, q.relative_residual) for q in r]

    def apply_segmented(self, offsets, b, x0=None, persistent=False):
        """CG on the independent diagonal blocks [offsets[s], offsets[s+1]) of
        the set-up matrix at once (host arrays): (x, [SolveResult per segment]).
        persistent: the whole solve in one cooperative launch."""
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64)
        db, dx = DeviceArray.from_host(b), DeviceArray.from_host(x)
        res = self.apply_segmented_device(offsets, db, dx, persistent)
        return dx.get(), res

    def apply_segmented_device(self, offsets, b, x, persistent=False):
        off = np.ascontiguousarray(offsets, np.int32)
        S = len(off) - 1
        r = (_lib.SolveResultC * max(S, 1))()
        fn = lib().tf_solver_apply_segmented_persistent if persistent else lib().tf_solver_apply_segmented
        try:
            check(fn(self.handle, S, off.ctypes.data_as(C.c_void_p), b.ptr, x.ptr, r))
        except TfError as e:
            _raise(e)
        return [SolveResult(SolveStatus(q.status), q.iterations, q.residual_norm, q.relative_residual)
                for q in r[:S]]

    def amg(self):
        h = lib().tf_solver_amg(self.handle)
        return _AmgView(C.c_void_p(h)) if h else None

    def timings(self):
        s, a = C.c_double(), C.c_double()
        check(lib().tf_solver_timings(self.handle, C.byref(s), C.byref(a)))
        return s.value, a.value


class Cg(_Krylov):
    """Preconditioned CG (Cg.cuh); default preconditioner Identity."""
    kind = 0


class BiCgStab(_Krylov):
    """Right-preconditioned BiCGStab (BiCgStab.cuh); default preconditioner Amg."""
    kind = 1
