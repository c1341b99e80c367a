"""Sliced-ELL row products (tf_sell_*): bit-identical to the oracle's row
sums (sequential sums of rounded products in stored order), for SpMV,
residual and the fused Jacobi sweep, in FP64 and FP32, through both column
encodings (16-bit offsets and 32-bit columns), with empty rows, row counts
that are not a multiple of the 64-row slice, and the representability
limits (rows over 255 entries, padding bound)."""
import numpy as np
import pytest

import _gpu as G
import _oracle as O
from thermofem_amd import problems as P
from thermofem_amd._lib import TfError

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def _f32_residual(a, b, x):
    y = O.spmv_f32(a, x)
    return np.float32(b.astype(np.float32)) - y


def _inv_diag(a):
    rp, ci, v = a
    n = len(rp) - 1
    d = np.zeros(n)
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            if ci[k] == i:
                d[i] += v[k]
    return 1.0 / d


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_sell_spmv_bit_exact(which):
    a = P.poisson2d(256) if which == "c1" else P.heat3d_lognormal(128, seed=1)[0]
    n = len(a[0]) - 1
    x = P.uniform(11, n) - 0.5
    log(f"sell spmv {which} n={n}")
    s = G.Sell(a, max_padding=0.25)
    bits, stored, reordered = s.info()
    assert bits == 16 and len(a[1]) <= stored <= 1.25 * len(a[1]) and not reordered
    assert np.array_equal(s.spmv(x), O.spmv(a, x))
    assert np.array_equal(s.spmv(x, np.float32), O.spmv_f32(a, x))
    s.close()


def test_sell_residual_and_jacobi_bit_exact():
    a, b = P.heat3d_lognormal(32, seed=2)
    n = len(a[0]) - 1
    x0 = P.uniform(4, n)
    inv = _inv_diag(a)
    s = G.Sell(a)
    log(f"sell residual/jacobi n={n}")
    assert np.array_equal(s.residual(b, x0), b - O.spmv(a, x0))
    assert np.array_equal(s.jacobi(inv, 2.0 / 3.0, b, x0), O.jacobi(a, 1, 2.0 / 3.0, b, x0))
    # FP32: the same operations in float32, with the oracle's FP32 row sums
    b32, x32, inv32 = b.astype(np.float32), x0.astype(np.float32), inv.astype(np.float32)
    ax = O.spmv_f32(a, x0)
    assert np.array_equal(s.residual(b, x0, np.float32), b32 - ax)
    w = np.float32(2.0 / 3.0)
    expected = x32 + (w * inv32) * (b32 - ax)
    assert np.array_equal(s.jacobi(inv, w, b, x0, np.float32), expected)


def test_sell_wide_columns_ragged_rows_and_empty_rows():
    """Columns far from the row force 32-bit columns; rows of 0..40 entries,
    n not a multiple of 64, unsorted duplicates kept in stored order."""
    rng = np.random.default_rng(7)
    n = 70001
    lens = rng.integers(0, 41, n)
    lens[rng.integers(0, n, 500)] = 0
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    ci = rng.integers(0, n, rp[-1]).astype(np.int32)
    v = rng.standard_normal(rp[-1])
    a = (rp, ci, v)
    x = rng.standard_normal(n)
    s = G.Sell(a, max_padding=4.0)
    bits, stored, reordered = s.info()
    log(f"sell wide n={n} bits={bits} stored={stored} nnz={rp[-1]} reordered={reordered}")
    assert bits == 32 and not reordered
    assert np.array_equal(s.spmv(x), O.spmv(a, x))
    assert np.array_equal(s.spmv(x, np.float32), O.spmv_f32(a, x))
    # the same rows ordered by length inside 256-row windows (padding bound 0.3)
    t = G.Sell(a, max_padding=0.3)
    bits, stored, reordered = t.info()
    log(f"sell wide sorted stored={stored} reordered={reordered}")
    assert bits == 32 and reordered and stored <= 1.3 * rp[-1]
    assert np.array_equal(t.spmv(x), O.spmv(a, x))
    assert np.array_equal(t.spmv(x, np.float32), O.spmv_f32(a, x))
    b = rng.standard_normal(n)
    inv = rng.uniform(0.5, 2.0, n)
    assert np.array_equal(t.residual(b, x), b - O.spmv(a, x))
    assert np.array_equal(t.jacobi(inv, 0.75, b, x), x + (0.75 * inv) * (b - O.spmv(a, x)))


def test_sell_hierarchy_level_bit_exact():
    """A Galerkin coarse matrix (aggregation numbering) of the 2D Poisson problem."""
    a = P.poisson2d(128)
    host = O.Amg(coarse_rows=64).setup(a)
    ac = host.level_matrix(1)
    n = len(ac[0]) - 1
    x = P.uniform(3, n)
    s = G.Sell(ac, max_padding=0.3)
    bits, stored, reordered = s.info()
    log(f"sell level-1 n={n} bits={bits} stored={stored} nnz={len(ac[1])} reordered={reordered}")
    assert bits == 16 and stored <= 1.3 * len(ac[1])
    assert np.array_equal(s.spmv(x), O.spmv(ac, x))
    assert np.array_equal(s.spmv(x, np.float32), O.spmv_f32(ac, x))


def test_sell_rejects_unrepresentable():
    n = 300
    lens = np.full(n, 3)
    lens[5] = 256
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    ci = (np.arange(rp[-1]) % n).astype(np.int32)
    a = (rp, ci, np.ones(rp[-1]))
    with pytest.raises(TfError, match="255"):
        G.Sell(a)
    lens[5] = 200  # representable, but the slice of row 5 pads far beyond 10 %
    rp[1:] = np.cumsum(lens)
    a = (rp, (np.arange(rp[-1]) % n).astype(np.int32), np.ones(rp[-1]))
    with pytest.raises(TfError):
        G.Sell(a, max_padding=0.1)
    s = G.Sell(a, max_padding=100.0)
    x = np.linspace(-1, 1, n)
    assert np.array_equal(s.spmv(x), O.spmv(a, x))
