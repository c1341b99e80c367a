"""Segmented batch solve: independent systems stored as one block-diagonal
matrix, CG on every segment at once with its own scalars and stopping test.

CPU: the batch generator (cube edges ~ U{16..96} from seed 3) and the
block-diagonal assembly.

GPU: with the identity preconditioner each segment's CG is the textbook CG
of that system alone, so it is compared with the oracle's CG on the system
by itself (status, iterations within 1, solution within 1e-9); with AMG
every segment reaches 1e-8 in the FP64 residual recomputed on the host and
its solution agrees with a single-system GPU solve within 1e-6; edge cases: an empty segment, a
one-row segment, a zero right-hand side, an iteration limit, a matrix that
couples two segments (rejected) and BiCGStab (rejected).  The persistent
path (the whole solve in one cooperative launch) is held to the per-kernel
path: the same statuses and iteration counts and the same iterates, for
every case above.
"""
import numpy as np
import pytest

import _oracle as O
from thermofem_amd import problems as P


def test_batch_edges_and_assembly():
    edges = (16 + np.floor(81 * P.uniform(3, 32))).astype(int)
    assert edges.min() >= 16 and edges.max() <= 96 and len(set(edges.tolist())) > 10
    systems = [P.poisson1d(3), P.poisson2d(2), P.poisson1d(1)]
    (rp, ci, v), off = P.block_diagonal(systems)
    assert off.tolist() == [0, 3, 7, 8]
    assert len(ci) == sum(len(s[1]) for s in systems)
    rows = np.repeat(np.arange(8), np.diff(rp))
    seg = np.searchsorted(off, rows, side="right") - 1
    assert np.all((ci >= off[seg]) & (ci < off[seg + 1]))
    a, b, off2, e = P.heat3d_batch(count=3, lo=3, hi=5)
    assert off2[-1] == len(b) == len(a[0]) - 1 == int((e ** 3).sum())


def _solver():
    from thermofem_amd.solver import AmgOptions, BiCgStab, Cg, KrylovOptions, SolveStatus
    return AmgOptions, BiCgStab, Cg, KrylovOptions, SolveStatus


def _segment(a, off, s):
    rp, ci, v = a
    lo, hi = off[s], off[s + 1]
    srp = (rp[lo:hi + 1] - rp[lo]).astype(np.int32)
    return srp, (ci[rp[lo]:rp[hi]] - lo).astype(np.int32), v[rp[lo]:rp[hi]]


@pytest.mark.gpu
def test_identity_segments_match_oracle_cg():
    AmgOptions, BiCgStab, Cg, KrylovOptions, SolveStatus = _solver()
    # well-conditioned systems (sigma 0.25), so that unpreconditioned CG
    # iteration counts are not at the mercy of rounding
    systems = ([P.heat3d_lognormal(e, seed=20 + e, sigma=0.25)[0] for e in (3, 7, 5, 9)] +
               [P.poisson1d(1), P.poisson2d(6)])
    a, off = P.block_diagonal(systems)
    b = P.uniform(8, off[-1]) - 0.25
    s = Cg(KrylovOptions(relative_tolerance=1e-10), None).setup(a)
    x, res = s.apply_segmented(off, b)
    for k in range(len(systems)):
        sa = _segment(a, off, k)
        bs = b[off[k]:off[k + 1]]
        xh, st, it, _, _ = O.krylov("cg", None, sa, bs, rtol=1e-10)
        print(f"segment {k}: it {res[k].iterations} oracle {it}", flush=True)
        assert res[k].status == SolveStatus.Converged and st == 0
        assert abs(res[k].iterations - it) <= 1
        assert np.linalg.norm(x[off[k]:off[k + 1]] - xh) <= 1e-9 * np.linalg.norm(xh)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
def test_amg_segments_converge(precision):
    AmgOptions, BiCgStab, Cg, KrylovOptions, SolveStatus = _solver()
    a, b, off, edges = P.heat3d_batch(count=6, lo=4, hi=14)
    b[off[2]:off[3]] = 0.0  # a zero right-hand side converges at iteration 0
    amg = AmgOptions(coarse_rows=256, pre_sweeps=2, post_sweeps=2, jacobi_weight=2 / 3, level_precision=precision)
    s = Cg(KrylovOptions(relative_tolerance=1e-8), amg).setup(a)
    x, res = s.apply_segmented(off, b)
    for k in range(len(edges)):
        sa = _segment(a, off, k)
        bs, xs = b[off[k]:off[k + 1]], x[off[k]:off[k + 1]]
        print(f"segment {k} edge {edges[k]}: {res[k].status.name} it {res[k].iterations}", flush=True)
        assert res[k].status == SolveStatus.Converged
        if k == 2:
            assert res[k].iterations == 0 and not xs.any()
            continue
        assert np.linalg.norm(bs - O.spmv(sa, xs)) <= 1.0001e-8 * np.linalg.norm(bs)
        one = Cg(KrylovOptions(relative_tolerance=1e-8), amg).setup(sa)
        x1, r1 = one.apply(bs)
        # the batch hierarchy coarsens every segment in step (a small system
        # alone may be solved by the coarse LU directly), so iteration counts
        # differ; the solutions agree to the tolerance
        print(f"  single-system solve: it {r1.iterations}", flush=True)
        assert np.linalg.norm(xs - x1) <= 1e-6 * np.linalg.norm(x1)


@pytest.mark.gpu
def test_segments_edge_cases():
    AmgOptions, BiCgStab, Cg, KrylovOptions, SolveStatus = _solver()
    systems = [P.poisson1d(40), P.poisson1d(1), P.poisson2d(12)]
    a, off = P.block_diagonal(systems)
    off = np.array([0, 40, 40, 41, 185], np.int32)  # an empty segment between the first two systems
    b = P.uniform(4, off[-1]) + 0.1
    s = Cg(KrylovOptions(relative_tolerance=1e-8, max_iterations=3), None).setup(a)
    x, res = s.apply_segmented(off, b)
    assert res[1].status == SolveStatus.Converged and res[1].iterations == 0
    assert res[2].status == SolveStatus.Converged and res[2].iterations == 1
    assert np.isclose(x[40], b[40] / 2.0, rtol=1e-15)
    assert res[0].status == SolveStatus.MaxIterations and res[0].iterations == 3
    xh, st, it, rn, _ = O.krylov("cg", None, _segment(a, np.array([0, 40, 41, 185]), 0), b[:40], maxit=3)
    assert st == 1 and np.isclose(res[0].residual_norm, rn, rtol=1e-9)
    # a coupled matrix is rejected
    coupled = (a[0], a[1].copy(), a[2])
    coupled[1][0] = 45
    with pytest.raises(ValueError):
        Cg(KrylovOptions(), None).setup(coupled).apply_segmented(off, b)
    with pytest.raises(ValueError):
        BiCgStab(KrylovOptions(), None).setup(a).apply_segmented(off, b)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
def test_non_finite_segment_does_not_spread(precision):
    """One segment's right-hand side holds a NaN (and another's an Inf): those
    segments stop NonFinite at once, every other segment converges as if
    alone (the shared cycle only ever sees zeros in the stopped segments,
    and the coarse inverse skips its zero couplings)."""
    AmgOptions, BiCgStab, Cg, KrylovOptions, SolveStatus = _solver()
    a, b, off, edges = P.heat3d_batch(count=5, lo=6, hi=12)
    b_bad = b.copy()
    b_bad[off[1] + 3] = np.nan
    b_bad[off[3]] = np.inf
    amg = AmgOptions(coarse_rows=256, pre_sweeps=2, post_sweeps=2, jacobi_weight=2 / 3, level_precision=precision)
    s = Cg(KrylovOptions(relative_tolerance=1e-8), amg).setup(a)
    x_ok, res_ok = s.apply_segmented(off, b)
    x, res = s.apply_segmented(off, b_bad)
    for k in range(len(edges)):
        seg = slice(off[k], off[k + 1])
        print(f"segment {k}: {res[k].status.name} it {res[k].iterations}", flush=True)
        if k in (1, 3):
            assert res[k].status == SolveStatus.NonFinite
            continue
        assert res[k].status == SolveStatus.Converged
        assert res[k].iterations == res_ok[k].iterations
        assert np.array_equal(x[seg], x_ok[seg])


def _same_results(r1, r2):
    return all(a.status == b.status and a.iterations == b.iterations for a, b in zip(r1, r2))


def _close(x1, x2, tol=1e-12):
    return np.linalg.norm(x1 - x2) <= tol * max(np.linalg.norm(x2), 1e-300)


@pytest.mark.gpu
def test_persistent_identity_matches_oracle_cg():
    AmgOptions, BiCgStab, Cg, KrylovOptions, SolveStatus = _solver()
    systems = ([P.heat3d_lognormal(e, seed=20 + e, sigma=0.25)[0] for e in (3, 7, 5, 9)] +
               [P.poisson1d(1), P.poisson2d(6)])
    a, off = P.block_diagonal(systems)
    b = P.uniform(8, off[-1]) - 0.25
    s = Cg(KrylovOptions(relative_tolerance=1e-10), None).setup(a)
    x0, r0 = s.apply_segmented(off, b)
    x, res = s.apply_segmented(off, b, persistent=True)
    for k in range(len(systems)):
        sa = _segment(a, off, k)
        bs = b[off[k]:off[k + 1]]
        xh, st, it, _, _ = O.krylov("cg", None, sa, bs, rtol=1e-10)
        print(f"segment {k}: persistent it {res[k].iterations} launched {r0[k].iterations} oracle {it}", flush=True)
        assert res[k].status == SolveStatus.Converged and st == 0
        assert abs(res[k].iterations - it) <= 1
        assert np.linalg.norm(x[off[k]:off[k + 1]] - xh) <= 1e-9 * np.linalg.norm(xh)
    assert _same_results(res, r0) and _close(x, x0)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("cycle", [0, 1, 2])
def test_persistent_amg_matches_launched(precision, cycle):
    AmgOptions, BiCgStab, Cg, KrylovOptions, SolveStatus = _solver()
    from thermofem_amd.solver import Cycle
    a, b, off, edges = P.heat3d_batch(count=6, lo=4, hi=14)
    b[off[2]:off[3]] = 0.0
    amg = AmgOptions(coarse_rows=256, pre_sweeps=2, post_sweeps=2, jacobi_weight=2 / 3, level_precision=precision,
                     cycle=Cycle(cycle))
    s = Cg(KrylovOptions(relative_tolerance=1e-8), amg).setup(a)
    x0, r0 = s.apply_segmented(off, b)
    x, res = s.apply_segmented(off, b, persistent=True)
    print(f"persistent {[q.iterations for q in res]} launched {[q.iterations for q in r0]}, "
          f"diff {np.abs(x - x0).max():.3e}", flush=True)
    assert _same_results(res, r0) and _close(x, x0)
    for k in range(len(edges)):
        assert res[k].status == SolveStatus.Converged
        sa = _segment(a, off, k)
        bs, xs = b[off[k]:off[k + 1]], x[off[k]:off[k + 1]]
        assert np.linalg.norm(bs - O.spmv(sa, xs)) <= 1.0001e-8 * max(np.linalg.norm(bs), 1e-300)
    # a second call on the same layout reuses the plan and gives the same result
    x2, res2 = s.apply_segmented(off, b, persistent=True)
    assert _same_results(res2, res) and np.array_equal(x2, x)


@pytest.mark.gpu
def test_persistent_edge_cases():
    AmgOptions, BiCgStab, Cg, KrylovOptions, SolveStatus = _solver()
    systems = [P.poisson1d(40), P.poisson1d(1), P.poisson2d(12)]
    a, _ = P.block_diagonal(systems)
    off = np.array([0, 40, 40, 41, 185], np.int32)
    b = P.uniform(4, off[-1]) + 0.1
    s = Cg(KrylovOptions(relative_tolerance=1e-8, max_iterations=3), None).setup(a)
    x0, r0 = s.apply_segmented(off, b)
    x, res = s.apply_segmented(off, b, persistent=True)
    assert res[1].status == SolveStatus.Converged and res[1].iterations == 0
    assert res[2].status == SolveStatus.Converged and res[2].iterations == 1
    assert res[0].status == SolveStatus.MaxIterations and res[0].iterations == 3
    assert _same_results(res, r0) and _close(x, x0)
    # a zero iteration limit: only the initial and final residuals
    s0 = Cg(KrylovOptions(relative_tolerance=1e-8, max_iterations=0), None).setup(a)
    xz, rz = s0.apply_segmented(off, b, persistent=True)
    assert np.array_equal(xz, np.zeros_like(b)) and rz[0].status == SolveStatus.MaxIterations
    with pytest.raises(ValueError):
        BiCgStab(KrylovOptions(), None).setup(a).apply_segmented(off, b, persistent=True)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
def test_persistent_non_finite_segment_does_not_spread(precision):
    AmgOptions, BiCgStab, Cg, KrylovOptions, SolveStatus = _solver()
    a, b, off, edges = P.heat3d_batch(count=5, lo=6, hi=12)
    b_bad = b.copy()
    b_bad[off[1] + 3] = np.nan
    b_bad[off[3]] = np.inf
    amg = AmgOptions(coarse_rows=256, pre_sweeps=2, post_sweeps=2, jacobi_weight=2 / 3, level_precision=precision)
    s = Cg(KrylovOptions(relative_tolerance=1e-8), amg).setup(a)
    x_ok, res_ok = s.apply_segmented(off, b, persistent=True)
    x, res = s.apply_segmented(off, b_bad, persistent=True)
    for k in range(len(edges)):
        seg = slice(off[k], off[k + 1])
        print(f"segment {k}: {res[k].status.name} it {res[k].iterations}", flush=True)
        if k in (1, 3):
            assert res[k].status == SolveStatus.NonFinite
            continue
        assert res[k].status == SolveStatus.Converged
        assert res[k].iterations == res_ok[k].iterations
        assert np.array_equal(x[seg], x_ok[seg])
