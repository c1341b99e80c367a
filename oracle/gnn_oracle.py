"""CPU oracle of the learned edge-affinity network and of the affinity targets.

TEST INFRASTRUCTURE ONLY: the parity checker for the HIP forward/backward
kernels (tests/, smoke(), the bench's CPU leg).  The product library never
imports it.  There is no reference counterpart: the reference aggregates on
the strength measure alone (src/Multigrid/GpuAggregation.cu:339-364 builds
the weights that the matching at :402-431 consumes); the network replaces
those weights.

Everything is float64 numpy.  The network (H = 32 hidden, FN = 4 node and
FE = 3 edge features) on a matrix A with a symmetric pattern, edges p = (i, j)
for every stored off-diagonal entry:

  node features  f_i = [sum_j |a_ij| / |d_i|, max_j |a_ij| / |d_i|, deg_i / 26, 1]
  edge features  s_p = (|a_ij| + |a_ji|) / 2, m_i = max_j |a_ij|,
                 g_p = [s_p / sqrt(|d_i d_j|), min(s_p/m_i, s_p/m_j), max(s_p/m_i, s_p/m_j)]
  h0_i  = relu(W0 f_i + b0)
  layer l = 1, 2:
     m_p  = relu(Wm [h_j ; g_p] + bm)
     M_i  = mean over the edges of row i of m_p
     h_i  = relu(Wu [h_i ; M_i] + bu)
  z_p = relu(We [h_i + h_j ; |h_i - h_j| ; g_p] + be)
  w_p = sigmoid(wo . z_p + bo)                      (symmetric in i, j)

d_i = sum of the stored diagonal entries of row i.  Weights are used rounded
to bfloat16 (round to nearest even) when `quantize` is set, as the GPU does;
loss = mean over edges of (w_p - t_p)^2.  Gradients are taken with respect
to the weights the forward pass used (the master weights receive them).
"""
import numpy as np

H, FN, FE = 32, 4, 3
LAYERS = 2

# parameter blocks in storage order: name, shape
SHAPES = [("W0", (H, FN)), ("b0", (H,))]
for _l in range(LAYERS):
    SHAPES += [(f"Wm{_l}", (H, H + FE)), (f"bm{_l}", (H,)), (f"Wu{_l}", (H, 2 * H)), (f"bu{_l}", (H,))]
SHAPES += [("We", (H, 2 * H + FE)), ("be", (H,)), ("wo", (H,)), ("bo", (1,))]
COUNT = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, at = {}, 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        out[name] = np.asarray(theta[at:at + size], np.float64).reshape(shape)
        at += size
    return out


def pack(params):
    return np.concatenate([np.asarray(params[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (rounded & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(f), np.float64(np.nan), out)


class Graph:
    """Edges of a CSR matrix with a symmetric pattern (off-diagonal entries)."""

    def __init__(self, rp, ci, v):
        rp = np.asarray(rp, np.int64)
        ci = np.asarray(ci, np.int64)
        v = np.asarray(v, np.float64)
        n = len(rp) - 1
        rows = np.repeat(np.arange(n), np.diff(rp))
        diag = np.bincount(rows[rows == ci], weights=v[rows == ci], minlength=n)
        off = rows != ci
        self.n = n
        self.pos = np.nonzero(off)[0]          # matrix positions of the edges
        self.i = rows[off]
        self.j = ci[off]
        a = np.abs(v[off])
        # twin edge (j, i) of every edge
        key = self.i * n + self.j
        order = np.argsort(key)
        tkey = self.j * n + self.i
        loc = np.searchsorted(key[order], tkey)
        if np.any(loc >= len(key)) or np.any(key[order][np.minimum(loc, len(key) - 1)] != tkey):
            raise ValueError("pattern is not symmetric")
        self.twin = order[loc]
        s = 0.5 * (a + a[self.twin])
        absd = np.abs(diag)
        m = np.zeros(n)
        np.maximum.at(m, self.i, a)
        deg = np.bincount(self.i, minlength=n).astype(np.float64)
        self.deg = deg
        rowsum = np.bincount(self.i, weights=a, minlength=n)
        self.f = np.stack([rowsum / absd, m / absd, deg / 26.0, np.ones(n)], axis=1)
        ri = s / m[self.i]
        rj = s / m[self.j]
        self.g = np.stack([s / np.sqrt(absd[self.i] * absd[self.j]), np.minimum(ri, rj), np.maximum(ri, rj)], axis=1)


def _relu(x):
    return np.maximum(x, 0.0)


def forward(theta, G, quantize=True, keep=False):
    P = unpack(theta)
    if quantize:
        P = {k: bf16(v) for k, v in P.items()}
    c = {}
    pre0 = G.f @ P["W0"].T + P["b0"]
    h = _relu(pre0)
    c["pre0"] = pre0
    c["h"] = [h]
    for l in range(LAYERS):
        xm = np.concatenate([h[G.j], G.g], axis=1)
        mpre = xm @ P[f"Wm{l}"].T + P[f"bm{l}"]
        m = _relu(mpre)
        M = np.zeros((G.n, H))
        np.add.at(M, G.i, m)
        M /= np.maximum(G.deg, 1.0)[:, None]
        xu = np.concatenate([h, M], axis=1)
        upre = xu @ P[f"Wu{l}"].T + P[f"bu{l}"]
        h = _relu(upre)
        c.setdefault("xm", []).append(xm)
        c.setdefault("mpre", []).append(mpre)
        c.setdefault("xu", []).append(xu)
        c.setdefault("upre", []).append(upre)
        c["h"].append(h)
    hi, hj = h[G.i], h[G.j]
    xe = np.concatenate([hi + hj, np.abs(hi - hj), G.g], axis=1)
    zpre = xe @ P["We"].T + P["be"]
    z = _relu(zpre)
    s = z @ P["wo"] + P["bo"][0]
    w = 1.0 / (1.0 + np.exp(-s))
    c.update(xe=xe, zpre=zpre, z=z, w=w, P=P, hi=hi, hj=hj)
    return (w, c) if keep else w


def loss(theta, G, target, quantize=True):
    w = forward(theta, G, quantize)
    return float(np.mean((w - target) ** 2))


def loss_grad(theta, G, target, quantize=True):
    """(loss, gradient) by reverse-mode differentiation of forward()."""
    w, c = forward(theta, G, quantize, keep=True)
    P = c["P"]
    E = len(w)
    g = {k: np.zeros_like(v) for k, v in P.items()}
    dw = 2.0 * (w - target) / E
    ds = dw * w * (1.0 - w)
    g["wo"] = c["z"].T @ ds
    g["bo"] = np.array([ds.sum()])
    dz = np.outer(ds, P["wo"]) * (c["zpre"] > 0)
    g["We"] = dz.T @ c["xe"]
    g["be"] = dz.sum(axis=0)
    dxe = dz @ P["We"]
    sgn = np.sign(c["hi"] - c["hj"])
    dhi = dxe[:, :H] + dxe[:, H:2 * H] * sgn
    dhj = dxe[:, :H] - dxe[:, H:2 * H] * sgn
    dh = np.zeros((G.n, H))
    np.add.at(dh, G.i, dhi)
    np.add.at(dh, G.j, dhj)
    for l in reversed(range(LAYERS)):
        dupre = dh * (c["upre"][l] > 0)
        g[f"Wu{l}"] = dupre.T @ c["xu"][l]
        g[f"bu{l}"] = dupre.sum(axis=0)
        dxu = dupre @ P[f"Wu{l}"]
        dh = dxu[:, :H].copy()
        dM = dxu[:, H:] / np.maximum(G.deg, 1.0)[:, None]
        dm = dM[G.i] * (c["mpre"][l] > 0)
        g[f"Wm{l}"] = dm.T @ c["xm"][l]
        g[f"bm{l}"] = dm.sum(axis=0)
        dxm = dm @ P[f"Wm{l}"]
        np.add.at(dh, G.j, dxm[:, :H])
    dpre0 = dh * (c["pre0"] > 0)
    g["W0"] = dpre0.T @ G.f
    g["b0"] = dpre0.sum(axis=0)
    return float(np.mean((w - target) ** 2)), pack(g)


def affinity(rp, ci, v, X0, sweeps=10, omega=2.0 / 3.0):
    """Affinities of the stored entries from `sweeps` weighted-Jacobi sweeps on
    A x = 0 applied to the columns of X0 (n x k):
    c_ij = (x_i . x_j)^2 / ((x_i . x_i)(x_j . x_j)), 0 on diagonal positions.
    Each sweep is x <- x + (omega * (1/d)) (0 - A x), row sums of rounded
    products in stored order (the roundings of the GPU's block Jacobi sweep)."""
    rp = np.asarray(rp, np.int64)
    ci = np.asarray(ci, np.int64)
    v = np.asarray(v, np.float64)
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    diag = np.bincount(rows[rows == ci], weights=v[rows == ci], minlength=n)
    X = np.array(X0, np.float64)
    wd = omega * (1.0 / diag)
    for _ in range(sweeps):
        AX = np.zeros_like(X)
        np.add.at(AX, rows, v[:, None] * X[ci])  # sequential in stored order
        X = X + wd[:, None] * (0.0 - AX)
    nrm = np.einsum("ij,ij->i", X, X)
    dots = np.einsum("pj,pj->p", X[rows], X[ci])
    out = dots * dots / (nrm[rows] * nrm[ci])
    out[rows == ci] = 0.0
    return out
