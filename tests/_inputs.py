"""Deterministic test inputs shared by the CPU and GPU suites.

``sample(n)`` restates the seeded generator of test/amg_native_test.cpp:28-52
(std::mt19937 seeded with 173+n, equal-weight edges, asymmetric edges, an
isolated row every 13th).  ``MT19937`` is the standard 32-bit Mersenne Twister
(the algorithm std::mt19937 is required to implement); its 10000th output
from the default seed is pinned in the tests against the value the C++
standard mandates (4123659995).
"""
import numpy as np


class MT19937:
    def __init__(self, seed=5489):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.index = 624

    def _twist(self):
        mt = self.mt
        for i in range(624):
            y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7FFFFFFF)
            v = mt[(i + 397) % 624] ^ (y >> 1)
            if y & 1:
                v ^= 0x9908B0DF
            mt[i] = v
        self.index = 0

    def __call__(self):
        if self.index >= 624:
            self._twist()
        y = self.mt[self.index]
        self.index += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def sample(n):
    """CSR of amg_native_test.cpp:28-52 as (rp, ci, v) numpy arrays."""
    rng = MT19937(173 + n)
    rp, ci, v = [0], [], []
    for i in range(n):
        row = []
        diag = 2.0
        for j in range(n):
            # C++ evaluates the && chain left to right and short-circuits, so
            # random() is drawn only when the first three terms hold.
            if i != j and i % 13 != 0 and j % 13 != 0 and (rng() % 11 == 0 or abs(i - j) == 1):
                value = -float(1 + rng() % 4)
                row.append((j, value))
                diag -= value
        row.append((i, diag))
        row.sort()
        for j, value in row:
            ci.append(j)
            v.append(value)
        rp.append(len(v))
    return (np.array(rp, dtype=np.int32), np.array(ci, dtype=np.int32), np.array(v, dtype=np.float64))


def poisson1d(n):
    """Tridiagonal [-1 2 -1] (amg_native_test.cpp:303-323)."""
    rp, ci, v = [0], [], []
    for i in range(n):
        if i:
            ci.append(i - 1); v.append(-1.0)
        ci.append(i); v.append(2.0)
        if i + 1 < n:
            ci.append(i + 1); v.append(-1.0)
        rp.append(len(v))
    return (np.array(rp, dtype=np.int32), np.array(ci, dtype=np.int32), np.array(v, dtype=np.float64))


def shuffled_with_duplicates(rp, ci, v, seed=7):
    """Same matrix with rows permuted internally and some entries split into
    two duplicates (canonicalisation must sum them back)."""
    rng = np.random.default_rng(seed)
    n = len(rp) - 1
    nrp, nci, nv = [0], [], []
    for i in range(n):
        cols = list(ci[rp[i]:rp[i + 1]]); vals = list(v[rp[i]:rp[i + 1]])
        out = []
        for c, x in zip(cols, vals):
            if rng.random() < 0.3:
                out.append((c, x * 0.25)); out.append((c, x * 0.75))
            else:
                out.append((c, x))
        order = rng.permutation(len(out))
        for k in order:
            nci.append(out[k][0]); nv.append(out[k][1])
        nrp.append(len(nv))
    return (np.array(nrp, dtype=np.int32), np.array(nci, dtype=np.int32), np.array(nv, dtype=np.float64))
