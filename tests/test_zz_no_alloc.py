"""The allocation-free exclusive scan (tf_scan_excl_i32_noalloc) against the
allocating one (tf_scan_excl_i32) through the C ABI: two calls in a row, the
second smaller, both equal to the allocating scan and to numpy's cumulative
sum; a length past the fixed scratch's limit is rejected before any launch."""
import numpy as np
import pytest

import _gpu as G
from thermofem_amd._lib import TfError, check, lib
from thermofem_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


def _scan(fn, x):
    d = G.up(x, np.int32)
    out = DeviceArray(len(x), np.int32)
    check(fn(G.ctx().h, len(x), d.ptr, out.ptr))
    G.ctx().sync()
    return out.get()


def test_noalloc_scan_matches_allocating_scan():
    rng = np.random.default_rng(21)
    for n in (5_000_017, 3_001):  # second call smaller: the fixed scratch keeps stale totals
        x = rng.integers(0, 7, n).astype(np.int32)
        print(f"noalloc scan n={n}", flush=True)
        got = _scan(lib().tf_scan_excl_i32_noalloc, x)
        ref = _scan(lib().tf_scan_excl_i32, x)
        assert np.array_equal(got, ref)
        assert np.array_equal(got, np.concatenate([[0], np.cumsum(x)[:-1]]).astype(np.int32))


def test_noalloc_scan_rejects_lengths_past_its_scratch():
    d = G.up(np.zeros(4, np.int32))
    with pytest.raises(TfError):
        check(lib().tf_scan_excl_i32_noalloc(G.ctx().h, (1 << 28) + 1, d.ptr, d.ptr))
