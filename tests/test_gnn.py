"""Learned edge affinities: the numpy restatement (oracle/gnn_oracle.py)
pinned by central finite differences and closed forms on the CPU; the HIP
forward, backward, Adam and affinity kernels against it on the GPU; and the
learned hierarchy inside AMG (fallback forced by uninformative weights,
gain of the trained weights over the strength weights).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import gnn_oracle as GO  # noqa: E402

from thermofem_amd import problems as P  # noqa: E402


def small_problem():
    return P.q1_diffusion(4, P.anisotropic_tensors(1, 5), 16)


def small_graph():
    a = small_problem()
    return a, GO.Graph(*a)


def test_bf16_rounding_kats():
    x = np.array([1.0, 1 + 2.0 ** -7, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -2.5, 3.0e38, 1 + 2.0 ** -8 + 2.0 ** -16],
                 np.float32)
    want = [1.0, 1 + 2.0 ** -7, 1.0, 1 + 2.0 ** -6, -2.5, None, 1 + 2.0 ** -7]
    got = GO.bf16(x)
    for g, w in zip(got, want):
        if w is not None:
            assert g == w
    # 3e38 rounds to the nearest bfloat16 (0x7F62 -> 3.0026e38), still finite
    assert np.isfinite(got[5]) and abs(got[5] - 3.0e38) / 3.0e38 < 2 ** -8


def test_oracle_gradient_matches_finite_differences():
    """Central differences of the unquantized loss on 80 parameters."""
    from thermofem_amd.gnn import init_params
    a, G = small_graph()
    theta = init_params(3).astype(np.float64)
    # biases nonzero so that every block of the network is exercised
    theta += 0.05 * (2 * P.uniform(77, len(theta)) - 1)
    target = P.uniform(78, len(G.i))
    _, g = GO.loss_grad(theta, G, target, quantize=False)
    idx = np.unique((P.uniform(79, 80) * len(theta)).astype(int))
    h = 1e-6
    for q in idx:
        e = np.zeros_like(theta)
        e[q] = h
        fd = (GO.loss(theta + e, G, target, False) - GO.loss(theta - e, G, target, False)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1e-3, abs(g[q])) + 1e-10, q


def test_oracle_network_is_symmetric_and_features_closed_form():
    from thermofem_amd.gnn import init_params
    a, G = small_graph()
    w = GO.forward(init_params(1).astype(np.float64), G)
    assert np.allclose(w, w[G.twin], rtol=0, atol=0)
    assert np.all((w > 0) & (w < 1))
    # features of the isotropic 27-point Q1 operator: interior rows have 26
    # neighbours and off-diagonal sum 8/3 - 0 = |d| (zero row sum)
    b = P.q1_diffusion(5, np.eye(3)[None, None, None].repeat(1, 0), 16)
    Gb = GO.Graph(*b)
    interior = Gb.deg == 26
    assert interior.any()
    assert np.allclose(Gb.f[interior, 0], 1.0, rtol=1e-12)
    assert np.allclose(Gb.f[interior, 2], 1.0)


def test_oracle_affinity_closed_forms():
    a = small_problem()
    n = len(a[0]) - 1
    rows = np.repeat(np.arange(n), np.diff(a[0]))
    # one vector, no sweeps: every affinity is 1
    c1 = GO.affinity(*a, np.ones((n, 1)) + P.uniform(3, n)[:, None], sweeps=0)
    assert np.allclose(c1[rows != a[1]], 1.0, rtol=1e-14) and np.all(c1[rows == a[1]] == 0)
    X0 = 2 * P.uniform(4, n * 8).reshape(n, 8) - 1
    c = GO.affinity(*a, X0, sweeps=10)
    G = GO.Graph(*a)
    cg = c[G.pos]
    assert np.all((cg >= 0) & (cg <= 1 + 1e-12))
    assert np.allclose(cg, cg[G.twin], rtol=1e-12)  # symmetric in i, j


def test_training_problems_have_symmetric_patterns():
    from thermofem_amd.gnn import training_problem
    for q in (0, 1):
        a = training_problem(q)
        G = GO.Graph(*a)
        assert len(G.i) > 0 and np.array_equal(G.i[G.twin], G.j)


# ---------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_affinities_match_oracle():
    from thermofem_amd.gnn import affinity_targets, test_vectors
    a = P.q1_diffusion(10, P.anisotropic_tensors(1, 8), 16)
    n = len(a[0]) - 1
    got = affinity_targets(a)
    want = GO.affinity(*a, test_vectors(n))
    assert np.allclose(got, want.astype(np.float32), rtol=2e-6, atol=1e-7)


@pytest.mark.gpu
def test_forward_and_gradient_match_oracle():
    from thermofem_amd.gnn import AffinityModel, init_params
    a = P.jump_diffusion(9, seed=3, block=4)[0]
    G = GO.Graph(*a)
    theta = init_params(5)
    theta = (theta + 0.05 * (2 * P.uniform(9, len(theta)) - 1)).astype(np.float32)
    m = AffinityModel(theta)
    assert np.array_equal(m.params(), theta)
    w = m.infer(a)
    wo = GO.forward(theta.astype(np.float64), G)
    assert np.all(w[np.repeat(np.arange(G.n), np.diff(a[0])) == a[1]] == 0)
    assert np.allclose(w[G.pos], wo, rtol=0, atol=5e-6)
    target = np.zeros(len(a[1]), np.float32)
    target[G.pos] = P.uniform(10, len(G.i))
    loss, g = m.loss_grad(a, target)
    lo, go = GO.loss_grad(theta.astype(np.float64), G, target[G.pos].astype(np.float64))
    assert abs(loss - lo) <= 1e-5 * lo
    assert np.linalg.norm(g - go) <= 2e-4 * np.linalg.norm(go)


@pytest.mark.gpu
def test_training_lowers_the_loss():
    from thermofem_amd.gnn import AffinityModel, affinity_targets, init_params
    a = P.q1_diffusion(12, P.anisotropic_tensors(1, 21), 16)
    t = affinity_targets(a)
    m = AffinityModel(init_params(0))
    first = m.train_step(a, t, 5e-3)
    for _ in range(40):
        last = m.train_step(a, t, 5e-3)
    print(f"loss {first:.5f} -> {last:.5f}", flush=True)
    assert last < 0.7 * first


def adversarial_params():
    """A network that scores an edge high when its coupling is weak:
    z_0 = relu(g_1) (the smaller relative coupling), w = sigmoid(5 - 10 z_0);
    every other weight zero."""
    from thermofem_amd.gnn import SHAPES, param_count
    theta = np.zeros(param_count(), np.float32)
    at = {}
    o = 0
    for name, shape in SHAPES:
        at[name] = o
        o += int(np.prod(shape))
    H = 32
    theta[at["We"] + 0 * (2 * H + 3) + 2 * H + 1] = 1.0
    theta[at["wo"] + 0] = -10.0
    theta[at["bo"]] = 5.0
    return theta


@pytest.mark.gpu
def test_adversarial_affinities_fall_back_to_plain():
    """With weights that prefer the weakest couplings the matching pairs
    across the material jumps and along zero entries; with learned = 2 PCG
    converges faster on the plain hierarchy, which is kept, level for level."""
    from thermofem_amd.gnn import AffinityModel
    from thermofem_amd.solver import Amg, AmgOptions
    a = P.jump_diffusion(24, seed=4, block=6)[0]
    model = AffinityModel(adversarial_params())
    w = model.infer(a)
    assert w.max() > 0.99  # zero couplings score ~ sigmoid(5)
    amg = Amg(AmgOptions(coarse_rows=64, learned=2))
    amg.set_affinity_model(model)
    amg.setup(a)
    used, rl, rp = amg.learned_info()
    print(f"adversarial: used {used} estimate learned {rl:.4f} plain {rp:.4f}", flush=True)
    assert not used and rp < rl
    plain = Amg(AmgOptions(coarse_rows=64)).setup(a)
    assert amg.level_count() == plain.level_count()
    for d in range(plain.level_count()):
        assert all(np.array_equal(x, y) for x, y in zip(amg.level_matrix(d), plain.level_matrix(d)))
    forced = Amg(AmgOptions(coarse_rows=64, learned=1))
    forced.set_affinity_model(model)
    forced.setup(a)
    assert forced.learned_info()[0]
    assert forced.convergence_estimate() == pytest.approx(rl, rel=1e-9)


@pytest.mark.gpu
def test_trained_network_gains_on_rotated_anisotropy():
    """The committed weights (trained on the GPU by these kernels) against the
    strength weights on a held-out rotated anisotropic problem (48^3 cells,
    seed 7, FP32 V(2,2) cycle, omega 0.45): fewer PCG iterations to 1e-8; and
    on a four-material jump cube, where the learned hierarchy does worse,
    learned = 2 keeps the plain one."""
    from thermofem_amd.gnn import AffinityModel
    from thermofem_amd.solver import AmgOptions, Cg, KrylovOptions, SolveStatus
    model = AffinityModel.trained()

    def solve(a, b, mode, omega):
        s = Cg(KrylovOptions(relative_tolerance=1e-8, max_iterations=2000),
               AmgOptions(coarse_rows=256, pre_sweeps=2, post_sweeps=2, jacobi_weight=omega, level_precision=32,
                          learned=mode))
        view = s.amg()
        if mode:
            view.set_affinity_model(model)
        s.setup(a)
        x, r = s.apply(b)
        assert r.status == SolveStatus.Converged
        assert np.linalg.norm(b - csr_mv(a, x)) <= 1.0001e-8 * np.linalg.norm(b)
        return r.iterations, view.learned_info()

    a, b = P.anisotropic_diffusion(48, seed=7)
    plain, _ = solve(a, b, 0, 0.45)
    learned, _ = solve(a, b, 1, 0.45)
    print(f"anisotropic 48^3: plain {plain} learned {learned} iterations", flush=True)
    assert learned < plain
    a, b = P.jump_diffusion(48, seed=9)
    plain, _ = solve(a, b, 0, 2 / 3)
    auto, (used, rl, rp) = solve(a, b, 2, 2 / 3)
    print(f"jump 48^3: plain {plain} auto {auto} used {used} ({rl:.4f} vs {rp:.4f})", flush=True)
    assert not used and auto == plain


def csr_mv(a, x):
    rp, ci, v = a
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return np.bincount(rows, weights=v * x[ci], minlength=len(rp) - 1)
