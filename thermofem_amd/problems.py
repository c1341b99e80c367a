"""Synthetic linear systems for the benchmark configurations.

Every generator builds canonical CSR (int32 offsets/columns, float64 values,
columns ascending) on the host with numpy.  Random fields come from a
counter-based SplitMix64 stream (x_k = seed + (k+1)*0x9E3779B97F4A7C15
through the SplitMix64 finalizer, u = (z >> 11) * 2^-53), with Box-Muller
normals, so every platform draws the same bits.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, count, offset=0):
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, count, offset=0):
    """U[0,1) doubles."""
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, count):
    """N(0,1) by Box-Muller from two consecutive uniform streams."""
    u1 = uniform(seed, count, 0)
    u2 = uniform(seed, count, count)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def _compress(cols, vals, mask):
    """(n, k) candidate columns (ascending per row) -> CSR."""
    counts = mask.sum(axis=1)
    rp = np.zeros(cols.shape[0] + 1, np.int64)
    np.cumsum(counts, out=rp[1:])
    return rp.astype(np.int32), cols[mask].astype(np.int32), vals[mask].astype(np.float64)


def poisson1d(n):
    i = np.arange(n)
    cols = np.stack([i - 1, i, i + 1], axis=1)
    vals = np.tile(np.array([-1.0, 2.0, -1.0]), (n, 1))
    mask = (cols >= 0) & (cols < n)
    return _compress(cols, vals, mask)


def poisson2d(nx, ny=None):
    """5-point [4; -1 x 4] on an nx x ny interior grid, homogeneous Dirichlet."""
    ny = nx if ny is None else ny
    n = nx * ny
    idx = np.arange(n)
    ix, iy = idx % nx, idx // nx
    cols = np.stack([idx - nx, idx - 1, idx, idx + 1, idx + nx], axis=1)
    vals = np.tile(np.array([-1.0, -1.0, 4.0, -1.0, -1.0]), (n, 1))
    mask = np.stack([iy > 0, ix > 0, np.ones(n, bool), ix < nx - 1, iy < ny - 1], axis=1)
    return _compress(cols, vals, mask)


def poisson2d_config(nx=256, seed=0):
    """2D 5-point Laplacian, 256^2 interior grid, RHS U[0,1) from seed 0."""
    a = poisson2d(nx)
    return a, uniform(seed, nx * nx)


def heat3d_lognormal(nx=128, seed=1, sigma=1.0):
    """Cell-centred FV steady heat conduction on the unit cube, nx^3 cells.

    k_cell = exp(sigma * N(0,1)) i.i.d.; interior faces use the harmonic mean
    2 kL kR / (kL + kR); faces x=0 (T=0) and x=1 (T=1) are Dirichlet with
    coefficient 2 k_cell (half-cell distance), the T=1 part moving to b; all
    other faces are insulated.  Row index i + nx*(j + nx*k).  Coefficients
    are scaled by 1/h (uniform scaling of the system).  Returns ((rp, ci, v), b).
    """
    n = nx ** 3
    k = np.exp(sigma * normal(seed, n))
    idx = np.arange(n)
    ix = idx % nx
    iy = (idx // nx) % nx
    iz = idx // (nx * nx)
    s = [nx * nx, nx, 1]
    coords = [iz, iy, ix]
    cols = np.empty((n, 7), np.int64)
    vals = np.zeros((n, 7))
    mask = np.zeros((n, 7), bool)
    diag = np.zeros(n)
    # columns ascending: -z, -y, -x, diag, +x, +y, +z
    order = [(0, -1), (1, -1), (2, -1), None, (2, +1), (1, +1), (0, +1)]
    for slot, o in enumerate(order):
        if o is None:
            cols[:, slot] = idx
            mask[:, slot] = True
            continue
        axis, sign = o
        nb = idx + sign * s[axis]
        valid = (coords[axis] > 0) if sign < 0 else (coords[axis] < nx - 1)
        nbc = np.where(valid, nb, idx)
        kf = 2.0 * k * k[nbc] / (k + k[nbc])
        cols[:, slot] = nbc
        vals[:, slot] = np.where(valid, -kf, 0.0)
        mask[:, slot] = valid
        diag += np.where(valid, kf, 0.0)
    b = np.zeros(n)
    left = ix == 0
    right = ix == nx - 1
    diag += np.where(left, 2.0 * k, 0.0) + np.where(right, 2.0 * k, 0.0)
    b += np.where(right, 2.0 * k * 1.0, 0.0)
    vals[:, 3] = diag
    return _compress(cols, vals, mask), b


def spmv_bytes(n, nnz, value_size):
    """Algorithmic bytes of one CSR SpMV: values+columns, offsets, x once, y."""
    return nnz * (value_size + 4) + 4 * (n + 1) + 2 * n * value_size
