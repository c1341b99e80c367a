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


def fourier_velocity(nx, seed, modes=8, kmax=4):
    """Velocity at the nodes ((i+1)h, (j+1)h, (k+1)h), h = 1/(nx+1), of an nx^3
    interior grid: u(x) = sum_m a_m sin(2 pi (k_m . x) + phi_m) over `modes`
    modes, integer wavenumbers k_m in {1..kmax}^3, amplitudes a_m ~ N(0,1)^3,
    phases ~ U[0, 2 pi).  Draw order: the 3*modes wavenumbers, then the
    3*modes amplitudes (one normal stream), then the modes phases.
    Returns (n, 3), node index i + nx*(j + nx*k).
    """
    kv = (np.floor(kmax * uniform(seed, 3 * modes)).astype(np.int64) + 1).reshape(modes, 3)
    amp = normal(seed + 1, 3 * modes).reshape(modes, 3)
    phase = 2.0 * np.pi * uniform(seed + 2, modes)
    h = 1.0 / (nx + 1)
    t = (np.arange(nx) + 1) * h
    z, y, x = np.meshgrid(t, t, t, indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    u = np.zeros((nx ** 3, 3))
    for m in range(modes):
        s = np.sin(2.0 * np.pi * (pts @ kv[m].astype(np.float64)) + phase[m])
        u += s[:, None] * amp[m][None, :]
    return u


def convection_diffusion3d(nx=96, peclet=10.0, seed=4, rhs_seed=5):
    """Steady convection-diffusion -lap(c) + u.grad(c) = f on the unit cube,
    7-point finite differences on an nx^3 interior node grid (kappa = 1,
    homogeneous Dirichlet: neighbours outside the grid are dropped), first-
    order upwind advection.  The velocity is fourier_velocity(nx, seed)
    scaled so that the largest cell Peclet number max_d |u_d| h / kappa
    equals `peclet`.  Rows are scaled by h^2: with q_d = h u_d the row holds
    6 + sum_d |q_d| on the diagonal, -1 - |q_d| at the upwind neighbour of
    axis d and -1 at the downwind one.  Non-symmetric M-matrix with a
    symmetric pattern; b ~ U[0,1) from rhs_seed.  Returns ((rp, ci, v), b).
    """
    n = nx ** 3
    u = fourier_velocity(nx, seed)
    q = u * (peclet / np.max(np.abs(u)))
    idx = np.arange(n)
    ix = idx % nx
    iy = (idx // nx) % nx
    iz = idx // (nx * nx)
    s = [nx * nx, nx, 1]
    coords = [iz, iy, ix]
    comp = [2, 1, 0]  # velocity component along each stride axis (z, y, x)
    cols = np.empty((n, 7), np.int64)
    vals = np.zeros((n, 7))
    mask = np.zeros((n, 7), bool)
    diag = np.full(n, 6.0)
    order = [(0, -1), (1, -1), (2, -1), None, (2, +1), (1, +1), (0, +1)]
    for slot, o in enumerate(order):
        if o is None:
            cols[:, slot] = idx
            mask[:, slot] = True
            continue
        axis, sign = o
        qd = q[:, comp[axis]]
        valid = (coords[axis] > 0) if sign < 0 else (coords[axis] < nx - 1)
        # upwind side: -axis neighbour when q_d > 0, +axis neighbour when q_d < 0
        upwind = (qd > 0) if sign < 0 else (qd < 0)
        cols[:, slot] = np.where(valid, idx + sign * s[axis], idx)
        vals[:, slot] = np.where(upwind, -1.0 - np.abs(qd), -1.0)
        mask[:, slot] = valid
    for d in range(3):
        diag += np.abs(q[:, d])
    vals[:, 3] = diag
    return _compress(cols, vals, mask), uniform(rhs_seed, n)


# Kuhn (Freudenthal) subdivision of the unit cube into 6 tetrahedra: tet p
# walks from corner (0,0,0) to (1,1,1) along the axes in the order of the
# permutation p; every tet has volume 1/6.
_KUHN = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def fem_p1_heat(nc=64, dt=1e-3, sigma=0.5, seed=2):
    """P1 finite elements for one backward-Euler step of the heat equation,
    A = M / dt + K, on the unit cube split into nc^3 cells x 6 Kuhn tets
    ((nc+1)^3 nodes, natural boundary conditions).  k_tet = exp(sigma N(0,1))
    per tet (tet index 6 * cell + p, cell = i + nc*(j + nc*k)).  Element
    matrices: K_e = k_tet vol grad(l_a).grad(l_b), M_e = vol/20 (1 + delta_ab),
    vol = h^3/6.  Entries are summed in the order cell, tet, local row, local
    column.  SPD.  Node index i + (nc+1)*(j + (nc+1)*k).  Returns (rp, ci, v).
    """
    h = 1.0 / nc
    m = nc + 1
    n = m ** 3
    kt = np.exp(sigma * normal(seed, 6 * nc ** 3))
    cell = np.arange(nc ** 3)
    ci_, cj, ck = cell % nc, (cell // nc) % nc, cell // (nc * nc)
    base = ci_ + m * (cj + m * ck)
    stride = np.array([1, m, m * m])
    # barycentric gradients (times h) of the path tet in its own axis frame:
    # l0 = 1 - x_a, l1 = x_a - x_b, l2 = x_b - x_c, l3 = x_c
    gram = np.array([[1, -1, 0, 0], [-1, 2, -1, 0], [0, -1, 2, -1], [0, 0, -1, 1]], np.float64)
    vol = h ** 3 / 6.0
    mass = vol / 20.0 * (np.ones((4, 4)) + np.eye(4))
    rows, cols, vals = [], [], []
    nodes = np.empty((nc ** 3, 6, 4), np.int64)
    for p, perm in enumerate(_KUHN):
        off = np.zeros(4, np.int64)
        for q in range(3):
            off[q + 1] = off[q] + stride[perm[q]]
        nodes[:, p, :] = base[:, None] + off[None, :]
    stiff = kt.reshape(nc ** 3, 6)[:, :, None, None] * (vol / h ** 2) * gram[None, None]
    elem = stiff + (mass / dt)[None, None]
    rows = np.repeat(nodes[:, :, :, None], 4, axis=3).ravel()
    cols = np.repeat(nodes[:, :, None, :], 4, axis=2).ravel()
    key = rows * n + cols
    uniq, inverse = np.unique(key, return_inverse=True)
    v = np.bincount(inverse.ravel(), weights=elem.ravel(), minlength=len(uniq))
    r = (uniq // n).astype(np.int64)
    c = (uniq % n).astype(np.int32)
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(r, minlength=n), out=rp[1:])
    return rp.astype(np.int32), c, v


def block_rhs(n, k, seed=6):
    """k right-hand sides U[0,1), row-major n x k (entry (i, j) at i*k + j)."""
    return uniform(seed, n * k).reshape(n, k)


def block_diagonal(systems):
    """Concatenate CSR matrices [(rp, ci, v), ...] into one block-diagonal CSR
    and the row offsets of its segments."""
    rps, cis, vs, off = [np.zeros(1, np.int64)], [], [], [0]
    nnz = 0
    for rp, ci, v in systems:
        n = len(rp) - 1
        rps.append(np.asarray(rp[1:], np.int64) + nnz)
        cis.append(np.asarray(ci, np.int64) + off[-1])
        vs.append(np.asarray(v, np.float64))
        nnz += len(ci)
        off.append(off[-1] + n)
    return ((np.concatenate(rps).astype(np.int32), np.concatenate(cis).astype(np.int32), np.concatenate(vs)),
            np.array(off, np.int32))


def heat3d_batch(count=32, lo=16, hi=96, seed=3):
    """`count` independent log-normal heat problems (heat3d_lognormal) with
    cube edges ~ U{lo..hi} drawn from `seed`; system s uses field seed
    100 + s.  Returns ((rp, ci, v) block diagonal, b, offsets, edges)."""
    edges = (lo + np.floor((hi - lo + 1) * uniform(seed, count))).astype(np.int64)
    systems, rhs = [], []
    for s, e in enumerate(edges):
        a, b = heat3d_lognormal(int(e), seed=100 + s)
        systems.append(a)
        rhs.append(b)
    a, off = block_diagonal(systems)
    return a, np.concatenate(rhs), off, edges


def _q1_reference_stiffness():
    """T[d, e, a, b] = integral over the unit cube of dN_a/dx_d * dN_b/dx_e for
    the 8 trilinear shape functions (local node a = ax + 2 ay + 4 az); the
    element matrix of a constant tensor K is sum_de K[d, e] T[d, e]."""
    bits = np.array([[(a >> q) & 1 for q in range(3)] for a in range(8)])
    sgn = np.where(bits == 1, 1.0, -1.0)
    T = np.zeros((3, 3, 8, 8))
    for d in range(3):
        for e in range(3):
            for a in range(8):
                for b in range(8):
                    prod = 1.0
                    for q in range(3):
                        same = bits[a, q] == bits[b, q]
                        if q == d and q == e:
                            f = 1.0 if same else -1.0          # int phi_a' phi_b'
                        elif q == d:
                            f = sgn[a, q] / 2.0                # int phi_a' phi_b
                        elif q == e:
                            f = sgn[b, q] / 2.0                # int phi_a phi_b'
                        else:
                            f = 1.0 / 3.0 if same else 1.0 / 6.0
                        prod *= f
                    T[d, e, a, b] = prod
    return T


def euler_rotation(alpha, beta, gamma):
    """R = Rz(alpha) Ry(beta) Rz(gamma)."""
    def rz(t):
        c, s = np.cos(t), np.sin(t)
        return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])

    def ry(t):
        c, s = np.cos(t), np.sin(t)
        return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])
    return rz(alpha) @ ry(beta) @ rz(gamma)


def q1_diffusion(nc, tensors, block=16):
    """Trilinear (Q1) finite elements for -div(K grad u) = f on the unit cube
    with nc^3 cells, homogeneous Dirichlet on the boundary: unknowns are the
    (nc-1)^3 interior nodes, 27-point rows with the cross terms of a full
    tensor.  tensors[bz, by, bx] is the 3x3 SPD tensor of the block of
    `block`^3 cells containing the cell.  Element matrices are the exact
    integrals (no quadrature), scaled by 1/h (a uniform factor).  Node index
    i + m (j + m k) with m = nc - 1 over the interior nodes, columns ascending.
    Returns (rp, ci, v)."""
    T = _q1_reference_stiffness()
    nb = tensors.shape[0]
    Ke = np.einsum("zyxde,deab->zyxab", tensors, T).reshape(-1, 8, 8)
    c = np.arange(nc)
    bz, by, bx = np.meshgrid(c // block, c // block, c // block, indexing="ij")
    cell_block = (bz * nb + by) * nb + bx  # (nc, nc, nc)
    m1 = nc + 1
    S = np.zeros((m1, m1, m1, 27))
    for a in range(8):
        az, ay, ax = (a >> 2) & 1, (a >> 1) & 1, a & 1
        for b in range(8):
            dz, dy, dx = ((b >> 2) & 1) - az, ((b >> 1) & 1) - ay, (b & 1) - ax
            slot = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
            S[az:az + nc, ay:ay + nc, ax:ax + nc, slot] += Ke[cell_block, a, b]
    m = nc - 1
    n = m ** 3
    idx = np.arange(n)
    ix, iy, iz = idx % m, (idx // m) % m, idx // (m * m)
    cols = np.empty((n, 27), np.int64)
    vals = np.empty((n, 27))
    mask = np.empty((n, 27), bool)
    for slot in range(27):
        dz, dy, dx = slot // 9 - 1, (slot // 3) % 3 - 1, slot % 3 - 1
        jx, jy, jz = ix + dx, iy + dy, iz + dz
        ok = (jx >= 0) & (jx < m) & (jy >= 0) & (jy < m) & (jz >= 0) & (jz < m)
        cols[:, slot] = np.where(ok, jx + m * (jy + m * jz), 0)
        vals[:, slot] = S[iz + 1, iy + 1, ix + 1, slot]
        mask[:, slot] = ok
    return _compress(cols, vals, mask)


def anisotropic_tensors(nb, seed, eps_lo=1e-4, eps_hi=1.0):
    """nb^3 tensors R diag(1, eps, eps) R^T: Euler angles ~ U[0, 2 pi) from
    seed (alpha, beta, gamma per block), eps log-uniform in [eps_lo, eps_hi]
    from seed + 1."""
    count = nb ** 3
    ang = 2.0 * np.pi * uniform(seed, 3 * count).reshape(count, 3)
    le = np.log(eps_lo) + (np.log(eps_hi) - np.log(eps_lo)) * uniform(seed + 1, count)
    eps = np.exp(le)
    out = np.empty((count, 3, 3))
    for q in range(count):
        R = euler_rotation(*ang[q])
        out[q] = R @ np.diag([1.0, eps[q], eps[q]]) @ R.T
    return out.reshape(nb, nb, nb, 3, 3)


def anisotropic_diffusion(nc=128, seed=7, block=16):
    """Rotated anisotropic diffusion: K = R diag(1, eps, eps) R^T per block of
    16^3 cells (anisotropic_tensors), Q1 elements, (nc-1)^3 unknowns.
    Returns ((rp, ci, v), b) with b ~ U[0,1) from seed + 2."""
    nb = (nc + block - 1) // block
    a = q1_diffusion(nc, anisotropic_tensors(nb, seed), block)
    return a, uniform(seed + 2, (nc - 1) ** 3)


def jump_diffusion(nc=128, seed=9, block=16):
    """Four isotropic materials k in {1e-3, 1e-1, 1e1, 1e3} assigned to the
    blocks of 16^3 cells from seed, Q1 elements.  Returns ((rp, ci, v), b)."""
    nb = (nc + block - 1) // block
    mat = np.floor(4 * uniform(seed, nb ** 3)).astype(np.int64)
    k = 10.0 ** (-3.0 + 2.0 * mat)
    tensors = (k[:, None, None] * np.eye(3)[None]).reshape(nb, nb, nb, 3, 3)
    a = q1_diffusion(nc, tensors, block)
    return a, uniform(seed + 2, (nc - 1) ** 3)


def spmv_bytes(n, nnz, value_size):
    """Algorithmic bytes of one CSR SpMV: values+columns, offsets, x once, y."""
    return nnz * (value_size + 4) + 4 * (n + 1) + 2 * n * value_size
