"""Learned edge affinities: the network of tf_gnn.hip behind a small Python
interface, its training loop and the trained weights.

The network scores every off-diagonal entry of a matrix with a symmetric
pattern; AMG setup with AmgOptions(learned=1 or 2) and a model attached
(Amg.set_affinity_model) hands those scores to the pairwise matching in
place of the strength weights (src/Multigrid/GpuAggregation.cu:339-364 makes
the weights that :402-431 matches on; the reference has no learned part).
Training targets are the affinities of smoothed test vectors
(affinity_targets).  Everything here runs through the C ABI; there is no CPU
path.
"""
import ctypes as C
import os

import numpy as np

from . import problems as P
from ._lib import TfError, check, lib
from .device import Context, DeviceArray, DeviceCsr

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "gnn_affinity_weights.npy")

# parameter blocks in storage order (name, shape); H = 32 hidden, 4 node and
# 3 edge features, two message-passing layers
H, FN, FE = 32, 4, 3
SHAPES = [("W0", (H, FN)), ("b0", (H,))]
for _l in range(2):
    SHAPES += [(f"Wm{_l}", (H, H + FE)), (f"bm{_l}", (H,)), (f"Wu{_l}", (H, 2 * H)), (f"bu{_l}", (H,))]
SHAPES += [("We", (H, 2 * H + FE)), ("be", (H,)), ("wo", (H,)), ("bo", (1,))]


def param_count():
    return int(lib().tf_gnn_param_count())


def init_params(seed=0):
    """Weights U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from the SplitMix64 stream of
    `seed`, biases zero."""
    u = P.uniform(seed, sum(int(np.prod(s)) for _, s in SHAPES))
    out, at = [], 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        if name.startswith("W") or name == "wo":
            out.append((2 * u[at:at + size] - 1) / np.sqrt(shape[-1]))
        else:
            out.append(np.zeros(size))
        at += size
    return np.concatenate(out).astype(np.float32)


def _csr_on_device(a):
    if isinstance(a, DeviceCsr):
        return a, a.n, a.nnz
    d = DeviceCsr(*a)
    return d, d.n, d.nnz


class AffinityModel:
    """The network on one GPU (tf_gnn)."""

    def __init__(self, params=None, device=0):
        h = C.c_void_p()
        check(lib().tf_gnn_create(device, C.byref(h)))
        self.handle = h
        self.device = device
        if params is not None:
            self.set_params(params)

    def __del__(self):
        if getattr(self, "handle", None):
            lib().tf_gnn_destroy(self.handle)
            self.handle = None

    @classmethod
    def trained(cls, device=0):
        """The committed trained weights (loaded without pickle)."""
        return cls(np.load(WEIGHTS, allow_pickle=False), device)

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        check(lib().tf_gnn_set_params(self.handle, p.ctypes.data_as(C.c_void_p), len(p)))

    def params(self):
        p = np.zeros(param_count(), np.float32)
        check(lib().tf_gnn_get_params(self.handle, p.ctypes.data_as(C.c_void_p), len(p)))
        return p

    def infer(self, a):
        """Weight of every stored entry (host float64 array of nnz)."""
        d, n, nnz = _csr_on_device(a)
        w = DeviceArray(max(nnz, 1), np.float64)
        check(lib().tf_gnn_infer(self.handle, n, nnz, *d.args(), w.ptr))
        return w.get()[:nnz]

    def loss_grad(self, a, target):
        """(loss, gradient) against per-entry targets (device array or host)."""
        d, n, nnz = _csr_on_device(a)
        t = target if isinstance(target, DeviceArray) else DeviceArray.from_host(np.asarray(target, np.float32))
        loss = C.c_double()
        g = np.zeros(param_count(), np.float32)
        check(lib().tf_gnn_loss_grad(self.handle, n, nnz, *d.args(), t.ptr, C.byref(loss),
                                     g.ctypes.data_as(C.c_void_p)))
        return loss.value, g

    def train_step(self, a, target, lr=1e-3):
        d, n, nnz = _csr_on_device(a)
        t = target if isinstance(target, DeviceArray) else DeviceArray.from_host(np.asarray(target, np.float32))
        loss = C.c_double()
        check(lib().tf_gnn_train_step(self.handle, n, nnz, *d.args(), t.ptr, C.c_float(lr), C.byref(loss)))
        return loss.value


def test_vectors(n, k=8, seed=11):
    """Start vectors of the affinity targets: U[-1, 1), row-major n x k."""
    return 2.0 * P.uniform(seed, n * k).reshape(n, k) - 1.0


def affinity_targets(a, k=8, sweeps=10, omega=2.0 / 3.0, seed=11, ctx=None, device_out=False, sharpen=1):
    """Affinity of every stored entry after `sweeps` weighted-Jacobi sweeps on
    A x = 0 from k seeded random vectors (float32 per entry, 0 on the
    diagonal), raised to the power `sharpen` (spreads the strongly coupled
    end of [0, 1], where the values the matching must rank lie close)."""
    d, n, nnz = _csr_on_device(a)
    ctx = ctx or Context(0)
    X0 = DeviceArray.from_host(test_vectors(n, k, seed).ravel())
    out = DeviceArray(max(nnz, 1), np.float32)
    check(lib().tf_affinities(ctx.h, n, nnz, *d.args(), k, sweeps, omega, X0.ptr, out.ptr))
    if sharpen != 1:
        out = DeviceArray.from_host(out.get() ** np.float32(sharpen))
    return out if device_out else out.get()[:nnz]


def training_problem(index, seed=1000):
    """Problem `index` of the training set: Q1 diffusion on nc^3 cells with nc
    ~ U{32..64}; even indices rotated anisotropic tensors, odd indices
    four-material jump coefficients (blocks of 8^3 cells)."""
    s = seed + 10 * index
    nc = 32 + int(np.floor(33 * P.uniform(s, 1)[0]))
    if index % 2 == 0:
        nb = (nc + 15) // 16
        return P.q1_diffusion(nc, P.anisotropic_tensors(nb, s + 1), 16)
    return P.jump_diffusion(nc, seed=s + 1, block=8)[0]


def train(count=200, epochs=4, lr=2e-3, seed=0, sharpen=1, log=print):
    """Train from init_params(seed) on `count` training problems (targets
    affinity_targets(..., sharpen)); returns the weights and the mean loss
    per epoch."""
    model = AffinityModel(init_params(seed))
    ctx = Context(0)
    data = []
    for q in range(count):
        a = training_problem(q)
        d = DeviceCsr(*a)
        t = affinity_targets(d, ctx=ctx, device_out=True, sharpen=sharpen)
        data.append((d, t))
        if q % 20 == 0:
            log(f"training set: problem {q} n {d.n} nnz {d.nnz}")
    history = []
    for e in range(epochs):
        order = np.argsort(P.uniform(seed + 100 + e, count))
        total = 0.0
        for q in order:
            d, t = data[q]
            total += model.train_step(d, t, lr)
        history.append(total / count)
        log(f"epoch {e}: mean loss {history[-1]:.6f}")
    return model.params(), history


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="train the affinity network on the GPU")
    # the committed weights: 60 problems, 4 epochs (longer training lowers the
    # regression loss but not the PCG iteration counts of the hierarchies)
    ap.add_argument("--count", type=int, default=60)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--sharpen", type=float, default=1.0)
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args()
    w, hist = train(args.count, args.epochs, args.lr, sharpen=args.sharpen, log=lambda m: print(m, flush=True))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, w, allow_pickle=False)
    print(f"saved {args.out}: {len(w)} weights, losses {hist}", flush=True)
