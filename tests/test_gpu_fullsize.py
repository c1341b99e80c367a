"""GPU parity at the configurations' stated sizes.

- 128^3 log-normal heat (n = 2,097,152): the whole hierarchy (every level's
  matrix, transfer and aggregate histogram) bit-exact against the oracle;
  the FP32 cycle bit-exact against the oracle's FP32 cycle; chosen (FP32
  cycle) and plain (FP64) AMG-PCG reach 1e-8 in the FP64 residual recomputed
  on the host, with the solution error against the oracle's FP64 PCG solve
  printed and bounded.
- 96^3 upwind convection-diffusion (n = 884,736): hierarchy bit-exact;
  BiCGStab with the FP32 V-cycle against the oracle's FP64 solve.
- P1 tets on 64^3 cells (n = 274,625), 64 right-hand sides: every SpMM
  column bit-exact against the oracle's sequential SpMV (FP64 and FP32); the
  block CG solve reaches 1e-8 in every column, two columns against the
  oracle's CG.
- 32 heat systems with edges U{16..96} (9,195,298 rows): the segmented solve
  reaches 1e-8 in every segment's host-recomputed FP64 residual, launched
  per kernel and as one persistent launch (same iterations and iterates).
"""
import numpy as np
import pytest

import _gpu as G
import _oracle as O
from thermofem_amd import problems as P
from thermofem_amd.solver import Amg, AmgOptions, BiCgStab, Cg, KrylovOptions, SolveStatus

pytestmark = pytest.mark.gpu

C2 = dict(coarse_rows=256, pre_sweeps=2, post_sweeps=2, jacobi_weight=2 / 3)


def log(msg):
    print(msg, flush=True)


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


_c2 = {}


def c2():
    if not _c2:
        a, b = P.heat3d_lognormal(128, seed=1)
        _c2["a"], _c2["b"] = a, b
        _c2["host"] = O.Amg(**C2).setup(a)
    return _c2["a"], _c2["b"], _c2["host"]


def test_c2_hierarchy_bit_exact_128():
    a, _, host = c2()
    log(f"c2 128^3: oracle hierarchy {host.levels()} levels")
    dev = Amg(AmgOptions(**C2)).setup(a)
    assert dev.level_count() == host.levels() >= 8
    stats = dev.level_statistics()
    for d in range(host.levels()):
        log(f"c2 level {d}: rows {stats[d][0]} nnz {stats[d][1]}")
        assert same(dev.level_matrix(d), host.level_matrix(d)), d
        if d + 1 < host.levels():
            assert same(dev.level_transfer(d), host.level_transfer(d)), d
        hr, hz, hh = host.level_info(d)
        assert (stats[d][0], stats[d][1], list(stats[d][2])) == (hr, hz, list(hh)), d
    assert dev.stop_reason() == host.stop_reason() == "coarse size"
    assert dev.coarse_solver() == host.coarse_solver() == "pivoted GPU LU"


def test_c2_fp32_cycle_bit_exact_128():
    """One V(2,2) cycle with FP32 levels on the 128^3 right-hand side: the
    same roundings in the same order as the oracle's FP32 cycle (pinned by an
    independent float32 restatement), at every one of the 2,097,152 rows."""
    a, b, _ = c2()
    host = O.Amg(level_precision=32, **C2).setup(a)
    dev = Amg(AmgOptions(level_precision=32, **C2)).setup(a)
    x = dev.apply(b)
    xh = host.apply(b)
    log(f"c2 fp32 cycle: max |diff| {np.max(np.abs(x - xh)):.3e}")
    assert np.array_equal(x, xh)
    assert list(dev.last_cycle_visits()) == [1] * dev.level_count()


@pytest.mark.parametrize("precision", [32, 64])
def test_c2_pcg_full_size(precision):
    a, b, host = c2()
    if "xh" not in _c2:
        xh, st, it, _, rel = O.krylov("cg", host, a, b)
        assert st == 0
        _c2["xh"], _c2["it"] = xh, it
        log(f"c2 oracle FP64 PCG: {it} iterations, rel {rel:.3e}")
    xh, it = _c2["xh"], _c2["it"]
    s = Cg(KrylovOptions(relative_tolerance=1e-8), AmgOptions(level_precision=precision, **C2)).setup(a)
    x, res = s.apply(b)
    rel = np.linalg.norm(b - O.spmv(a, x)) / np.linalg.norm(b)
    err = np.linalg.norm(x - xh) / np.linalg.norm(xh)
    log(f"c2 PCG level precision {precision}: {res.iterations} iterations, host-recomputed rel {rel:.3e}, "
        f"error vs oracle FP64 solution {err:.3e}")
    assert res.status == SolveStatus.Converged
    assert rel <= 1.0001e-8
    assert abs(res.iterations - it) <= 2
    assert err <= 1e-6


def test_c3_hierarchy_and_fp32_bicgstab_96():
    a, b = P.convection_diffusion3d(96, 30.0)
    opts = dict(coarse_rows=256, jacobi_weight=2 / 3)
    host = O.Amg(**opts).setup(a)
    dev = Amg(AmgOptions(**opts)).setup(a)
    log(f"c3 96^3: {host.levels()} levels")
    assert dev.level_count() == host.levels()
    for d in range(host.levels()):
        assert same(dev.level_matrix(d), host.level_matrix(d)), d
        if d + 1 < host.levels():
            assert same(dev.level_transfer(d), host.level_transfer(d)), d
    xh, st, it, _, _ = O.krylov("bicgstab", host, a, b)
    assert st == 0
    s = BiCgStab(KrylovOptions(relative_tolerance=1e-8), AmgOptions(level_precision=32, **opts)).setup(a)
    x, res = s.apply(b)
    rel = np.linalg.norm(b - O.spmv(a, x)) / np.linalg.norm(b)
    err = np.linalg.norm(x - xh) / np.linalg.norm(xh)
    log(f"c3 BiCGStab FP32 V: {res.iterations} iterations (oracle FP64 {it}), rel {rel:.3e}, error {err:.3e}")
    assert res.status == SolveStatus.Converged and rel <= 1.0001e-8
    assert err <= 1e-6


_c4 = {}


def c4():
    if not _c4:
        _c4["a"] = P.fem_p1_heat(64)
        n = len(_c4["a"][0]) - 1
        _c4["B"] = P.block_rhs(n, 64)
    return _c4["a"], _c4["B"]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_c4_spmm_columns_bit_exact_full_size(dtype):
    a, B = c4()
    n = len(a[0]) - 1
    assert (n, len(a[1]), B.shape) == (274625, 4018753, (n, 64))
    X = B.astype(dtype)
    Y = G.spmm(a, X, dtype)
    for j in range(64):
        ref = O.spmv(a, X[:, j]) if dtype == np.float64 else O.spmv_f32(a, X[:, j])
        assert np.array_equal(Y[:, j], ref), j
    log(f"c4 spmm {np.dtype(dtype).name}: 64 columns bit-exact")


def test_c4_block_cg_full_size():
    a, B = c4()
    s = Cg(KrylovOptions(relative_tolerance=1e-8), AmgOptions(level_precision=32, **C2)).setup(a)
    X, res = s.apply_block(B)
    normb = np.linalg.norm(B, axis=0)
    worst = 0.0
    for j in range(64):
        assert res[j].status == SolveStatus.Converged, j
        worst = max(worst, np.linalg.norm(B[:, j] - O.spmv(a, X[:, j])) / normb[j])
    log(f"c4 block CG: iterations {sorted(set(r.iterations for r in res))}, worst host rel {worst:.3e}")
    assert worst <= 1.0001e-8
    host = O.Amg(**C2).setup(a)
    for j in (0, 63):
        xh, st, it, _, _ = O.krylov("cg", host, a, B[:, j])
        assert st == 0 and abs(it - res[j].iterations) <= 2
        assert np.linalg.norm(X[:, j] - xh) <= 1e-6 * np.linalg.norm(xh)


def test_c4_batch_32_systems_full_size():
    a, b, off, edges = P.heat3d_batch()
    assert len(edges) == 32 and min(edges) >= 16 and max(edges) <= 96
    s = Cg(KrylovOptions(relative_tolerance=1e-8), AmgOptions(level_precision=32, **C2)).setup(a)
    x, res = s.apply_segmented(off, b)
    r = b - O.spmv(a, x)
    worst = 0.0
    for k in range(32):
        seg = slice(off[k], off[k + 1])
        assert res[k].status == SolveStatus.Converged, k
        worst = max(worst, np.linalg.norm(r[seg]) / np.linalg.norm(b[seg]))
    log(f"batch of 32 (rows {off[-1]}): iterations {min(q.iterations for q in res)}.."
        f"{max(q.iterations for q in res)}, worst host rel {worst:.3e}")
    assert worst <= 1.0001e-8
    # the whole solve as one persistent launch: the same iterations and
    # iterates as the per-kernel path
    xp, rp = s.apply_segmented(off, b, persistent=True)
    log(f"persistent launch: iterations {[q.iterations for q in rp]}")
    for k in range(32):
        assert rp[k].status == SolveStatus.Converged and rp[k].iterations == res[k].iterations, k
    assert np.linalg.norm(xp - x) <= 1e-12 * np.linalg.norm(x)
