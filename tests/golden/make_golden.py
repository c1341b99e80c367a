"""Writes tests/golden/aggregation_golden.json from the reference's own host
aggregation oracle (test/support/NativeAggregationOracle.cu), compiled from
the reference sources by `make -C oracle ref` into oracle/_ref.

Inputs are the reference test generator sample(n) (amg_native_test.cpp:28-52)
and the 2D 5-point Laplacian; each case stores the strength graph, the
aggregation (map, member CSR, size histogram) and the Galerkin coarse CSR the
reference produces.  Run in a container that has the reference tree:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import _inputs as I  # noqa: E402
import _oracle as O  # noqa: E402
from thermofem_amd import problems as P  # noqa: E402


def main():
    assert O.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    cases = []
    grid = [("sample", n, th) for n in (1, 2, 4, 17, 61, 128) for th in (0.25,)]
    grid += [("sample", n, th) for n in (17, 128) for th in (0.0, 1.0)]
    grid += [("poisson2d", 16, 0.25)]
    for kind, n, theta in grid:
        raw = I.sample(n) if kind == "sample" else P.poisson2d(n)
        a = O.ref_canonical(*raw)
        g = O.ref_strength(*a, theta)
        for rounds in (1, 8):
            for absorb in (False, True):
                t = O.ref_aggregate(g, rounds, absorb)
                c = O.ref_galerkin(a, t)
                cases.append({
                    "input": kind, "n": n, "theta": theta, "rounds": rounds, "absorb": absorb,
                    "g_rp": g[0].tolist(), "g_ci": g[1].tolist(), "g_w": g[2].tolist(),
                    "map": t[0].tolist(), "offsets": t[1].tolist(), "members": t[2].tolist(),
                    "histogram": t[3].tolist(),
                    "c_rp": c[0].tolist(), "c_ci": c[1].tolist(), "c_v": c[2].tolist(),
                })
    out = os.path.join(HERE, "aggregation_golden.json")
    with open(out, "w") as f:
        json.dump({"source": "reference host oracle test/support/NativeAggregationOracle.cu (compiled)",
                   "cases": cases}, f, separators=(",", ":"))
    print(out, len(cases), "cases", os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
