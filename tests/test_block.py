"""Multi-right-hand-side path: the P1 tetrahedral heat operator (M/dt + K on
Kuhn tets), SpMM, the block AMG cycle and block CG.

CPU: the FEM generator against what the discretisation fixes (nonzero count
of the Freudenthal mesh, symmetry, positive definiteness, the mass matrix
integrating to the cube's volume, the stiffness matrix annihilating
constants and, with constant conductivity, linear functions at interior
nodes).

GPU: every SpMM column equals the oracle's sequential SpMV of that column
bit for bit; every column of a block cycle equals the single-vector cycle
of that column bit for bit (V/W/F, FP64 and FP32 levels, with and without
the one-workgroup coarse tail on the single-vector side, LU and Jacobi
coarsest levels); block CG columns match single-vector CG solves (status,
iterations within 1, solutions within 1e-6) and the oracle's FP64 CG, with
a zero column and an iteration limit as edge cases.
"""
import numpy as np
import pytest

import _oracle as O
from thermofem_amd import problems as P


def _freudenthal_nnz(m):
    # each node couples to itself and along +-{e1, e2, e3, e1+e2, e1+e3, e2+e3, e1+e2+e3}
    edges = 3 * (m - 1) * m * m + 3 * (m - 1) ** 2 * m + (m - 1) ** 3
    return m ** 3 + 2 * edges


def _dense(a):
    rp, ci, v = a
    n = len(rp) - 1
    d = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(rp))
    np.add.at(d, (rows, ci), v)
    return d


@pytest.mark.parametrize("nc", [1, 2, 5])
def test_fem_structure(nc):
    rp, ci, v = P.fem_p1_heat(nc)
    m = nc + 1
    assert len(rp) == m ** 3 + 1 and len(ci) == _freudenthal_nnz(m)
    rows = np.repeat(np.arange(m ** 3), np.diff(rp))
    assert np.all(np.diff(ci)[np.diff(rows) == 0] > 0)
    d = _dense((rp, ci, v))
    assert np.array_equal(d, d.T)
    assert np.linalg.eigvalsh(d).min() > 0


def test_fem_full_size_count():
    """64^3 cells: 274,625 nodes and 4,018,753 nonzeros."""
    assert _freudenthal_nnz(65) == 4018753


def test_fem_mass_and_stiffness():
    """M = A(dt=1) - A(dt=inf) integrates to the volume; K = A(dt=inf)
    annihilates constants, and with constant k linear functions at interior
    nodes (the P1 patch test)."""
    nc = 4
    m = nc + 1
    K = _dense(P.fem_p1_heat(nc, dt=np.inf, sigma=0.0))
    M = _dense(P.fem_p1_heat(nc, dt=1.0, sigma=0.0)) - K
    one = np.ones(m ** 3)
    assert abs(one @ M @ one - 1.0) < 1e-14
    assert np.abs(K @ one).max() < 1e-14
    idx = np.arange(m ** 3)
    i, j, k = idx % m, (idx // m) % m, idx // (m * m)
    interior = (i > 0) & (i < nc) & (j > 0) & (j < nc) & (k > 0) & (k < nc)
    for f in (i / nc, j / nc, k / nc, (i + 2 * j - 3 * k) / nc):
        assert np.abs((K @ f)[interior]).max() < 1e-14
    # a lumped row of M: 1/(20) vol * (sum over tets at the node of 5) -> interior nodes carry h^3
    assert np.allclose((M @ one)[interior], (1.0 / nc) ** 3, rtol=1e-12)
    # variable conductivity still annihilates constants
    Kv = _dense(P.fem_p1_heat(nc, dt=np.inf))
    assert np.abs(Kv @ one).max() < 1e-14


def test_block_rhs_layout():
    b = P.block_rhs(5, 3, seed=6)
    assert b.shape == (5, 3)
    assert np.array_equal(b.ravel(), P.uniform(6, 15))


# ---------------------------------------------------------------- GPU


def _solver():
    from thermofem_amd.solver import Amg, AmgOptions, BiCgStab, Cg, Cycle, KrylovOptions, SolveStatus
    return Amg, AmgOptions, BiCgStab, Cg, Cycle, KrylovOptions, SolveStatus


@pytest.mark.gpu
@pytest.mark.parametrize("k", [1, 3, 8, 64, 100])
def test_spmm_columns_bit_exact(k):
    import _gpu as G
    a = P.fem_p1_heat(8)
    n = len(a[0]) - 1
    X = P.block_rhs(n, k, seed=7) - 0.5
    Y = G.spmm(a, X)
    for j in range(k):
        assert np.array_equal(Y[:, j], O.spmv(a, X[:, j])), j
    Y32 = G.spmm(a, X.astype(np.float32), np.float32)
    a32 = (a[0], a[1], a[2].astype(np.float32))
    for j in range(0, k, max(1, k // 7)):
        assert np.array_equal(Y32[:, j], O.spmv_f32(a32, X[:, j].astype(np.float32))), j


@pytest.mark.gpu
@pytest.mark.parametrize("k", [33, 64])
def test_spmm_empty_and_uneven_rows(k):
    """Rows of different lengths (longer than a wave's 64-entry staging
    window too) and empty rows, also as the first and the last row."""
    import _gpu as G
    rng = np.random.default_rng(5)
    n = 203
    lens = rng.integers(0, 40, n)
    lens[[0, 1, 2, 57, 58, 131, n - 1]] = 0
    lens[[10, 11, 12, 200]] = [150, 64, 65, 129]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(n, int(m), replace=False)) for m in lens]).astype(np.int32)
    v = rng.standard_normal(int(rp[-1]))
    a = (rp, ci, v)
    X = rng.standard_normal((n, k))
    Y = G.spmm(a, X)
    for j in range(k):
        assert np.array_equal(Y[:, j], O.spmv(a, X[:, j])), j
    Y32 = G.spmm(a, X.astype(np.float32), np.float32)
    a32 = (rp, ci, v.astype(np.float32))
    for j in range(k):
        assert np.array_equal(Y32[:, j], O.spmv_f32(a32, X[:, j].astype(np.float32))), j


@pytest.mark.gpu
@pytest.mark.parametrize("tail", [1, 2])
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("cycle", [0, 1, 2])
def test_block_cycle_equals_single_cycles(cycle, precision, tail):
    """Column j of the block cycle == Amg.apply(column j) bit for bit."""
    Amg, AmgOptions, _, _, Cycle, _, _ = _solver()
    a, b = P.poisson2d_config(96, seed=0)
    n = len(a[0]) - 1
    k = 6
    R = np.stack([b * (j + 1) - 0.3 * j for j in range(k)], axis=1)
    if precision == 32:
        R = R.astype(np.float32).astype(np.float64)
    amg = Amg(AmgOptions(coarse_rows=64, cycle=Cycle(cycle), level_precision=precision, pre_sweeps=2,
                         post_sweeps=1, jacobi_weight=2 / 3, coarse_tail=tail)).setup(a)
    Z = amg.apply_block(R)
    visits = amg.last_block_visits()
    print(f"block cycle {cycle} fp{precision} tail {tail} levels {amg.level_count()}", flush=True)
    for j in range(k):
        z = amg.apply(R[:, j])
        assert np.array_equal(Z[:, j], z), j
    assert visits == amg.last_cycle_visits()
    assert Z.shape == (n, k)


@pytest.mark.gpu
def test_block_cycle_jacobi_coarsest():
    """Coarsening stopped above the LU bound: the coarsest level runs Jacobi sweeps."""
    Amg, AmgOptions, _, _, Cycle, _, _ = _solver()
    a = P.poisson1d(300)
    R = np.stack([np.sin(0.01 * (j + 1) * np.arange(300)) for j in range(3)], axis=1)
    amg = Amg(AmgOptions(max_levels=2, coarse_rows=4)).setup(a)
    assert amg.coarse_solver() != "pivoted GPU LU"
    Z = amg.apply_block(R)
    for j in range(3):
        assert np.array_equal(Z[:, j], amg.apply(R[:, j]))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
def test_block_cg_matches_single_solves(precision):
    Amg, AmgOptions, BiCgStab, Cg, Cycle, KrylovOptions, SolveStatus = _solver()
    a = P.fem_p1_heat(12)
    n = len(a[0]) - 1
    k = 8
    B = P.block_rhs(n, k)
    B[:, 5] = 0.0  # zero right-hand side: converged at iteration 0 with x = 0
    amg = AmgOptions(coarse_rows=256, pre_sweeps=2, post_sweeps=2, jacobi_weight=2 / 3, level_precision=precision)
    s = Cg(KrylovOptions(relative_tolerance=1e-8), amg).setup(a)
    X, res = s.apply_block(B)
    host = O.Amg(coarse_rows=256, pre_sweeps=2, post_sweeps=2, jacobi_weight=2 / 3).setup(a)
    for j in range(k):
        x1, r1 = s.apply(B[:, j])
        print(f"block cg fp{precision} col {j}: {res[j].status.name} it={res[j].iterations} "
              f"single it={r1.iterations}", flush=True)
        assert res[j].status == SolveStatus.Converged
        if j == 5:
            assert res[j].iterations == 0 and not X[:, j].any()
            continue
        assert abs(res[j].iterations - r1.iterations) <= 1
        assert np.linalg.norm(B[:, j] - O.spmv(a, X[:, j])) <= 1.0001e-8 * np.linalg.norm(B[:, j])
        assert np.linalg.norm(X[:, j] - x1) <= 1e-6 * np.linalg.norm(x1)
        if j < 2:
            xh, st, it, _, _ = O.krylov("cg", host, a, B[:, j])
            assert st == 0 and np.linalg.norm(X[:, j] - xh) <= 1e-6 * np.linalg.norm(xh)


@pytest.mark.gpu
def test_block_cg_iteration_limit_and_guess():
    Amg, AmgOptions, BiCgStab, Cg, Cycle, KrylovOptions, SolveStatus = _solver()
    a = P.fem_p1_heat(6)
    n = len(a[0]) - 1
    B = P.block_rhs(n, 3, seed=9)
    s = Cg(KrylovOptions(relative_tolerance=1e-12, max_iterations=2), AmgOptions(coarse_rows=64)).setup(a)
    X, res = s.apply_block(B)
    for j in range(3):
        x1, r1 = s.apply(B[:, j])
        assert res[j].status == r1.status == SolveStatus.MaxIterations
        assert res[j].iterations == 2
        assert np.isclose(res[j].relative_residual, r1.relative_residual, rtol=1e-9)
    # an exact initial guess: converged at iteration 0
    s2 = Cg(KrylovOptions(relative_tolerance=1e-8), AmgOptions(coarse_rows=64)).setup(a)
    Xs, _ = s2.apply_block(B)
    X2, res2 = s2.apply_block(B, x0=Xs)
    assert all(r.iterations == 0 and r.status == SolveStatus.Converged for r in res2)
    assert np.array_equal(X2, Xs)


@pytest.mark.gpu
def test_block_solve_requires_cg():
    Amg, AmgOptions, BiCgStab, Cg, Cycle, KrylovOptions, SolveStatus = _solver()
    a = P.poisson1d(16)
    s = BiCgStab(KrylovOptions(), AmgOptions(coarse_rows=4)).setup(a)
    with pytest.raises(ValueError):
        s.apply_block(np.ones((16, 2)))
