"""Matrix Market reader (io::readMatrixMarket, src/IO/MatrixMarket.cu:35-122)
and the solve-matrix CLI (example/solve-matrix.cu, docs/matrix-market.md).

CPU: the library's tf_read_matrix_market against the reference's own KATs
(test/matrix_market_test.cpp:8-44) and, where oracle/_ref was built from the
reference sources, against the reference reader on the same files (valid and
malformed, random general and symmetric inputs with duplicates, comments and
blank lines).  GPU: the CLI on the reference's poisson.mtx fixture and on a
larger generated matrix, for every preconditioner / cycle choice."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
from thermofem_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "thermofem_amd", "bin", "solve-matrix")
FIXTURE = os.path.join(ROOT, "tests", "golden", "poisson.mtx")

MALFORMED = [
    "",
    "%%MatrixMarket matrix array real general\n2 2\n",
    "%%MatrixMarket matrix coordinate complex general\n2 2 0\n",
    "%%MatrixMarket matrix coordinate real symmetric\n2 3 0\n",
    "%%MatrixMarket matrix coordinate real general\n-1 2 0\n",
    "%%MatrixMarket matrix coordinate real general\n2147483648 2 0\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 2\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 3 2\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 nan\n",
    "%%MatrixMarket matrix coordinate integer general\n2 2 1\n1 1 1.5\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 2 extra\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 0\n1 1 2\n",
    "%%MatrixMarket matrix coordinate real symmetric\n2 2 2\n1 2 1\n2 1 1\n",
    # further malformed forms (extra header word, dimension fields, overflow)
    "%%MatrixMarket matrix coordinate real general extra\n1 1 0\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 0 0\n",
    "%%MatrixMarket matrix coordinate real general\n2 2\n",
    "%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1e999\n",
    "%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 inf\n",
    "%%MatrixMarket matrix coordinate real general\n1 1 1\n1.0 1 2\n",
    "%%matrixmarket matrix coordinate real general\n1 1 0\n",
]


def read_tf(path):
    L = _lib.lib()
    rows, cols, nnz = C.c_int(), C.c_int(), C.c_int64()
    rc = L.tf_read_matrix_market(path.encode(), C.byref(rows), C.byref(cols), C.byref(nnz), None, None, None)
    if rc != 0:
        raise RuntimeError(L.tf_last_error().decode())
    rp = np.zeros(rows.value + 1, np.int32)
    ci = np.zeros(max(nnz.value, 1), np.int32)
    v = np.zeros(max(nnz.value, 1))
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    rc = L.tf_read_matrix_market(path.encode(), C.byref(rows), C.byref(cols), C.byref(nnz), P(rp, C.c_int),
                                 P(ci, C.c_int), P(v, C.c_double))
    assert rc == 0
    return rows.value, cols.value, rp, ci[:nnz.value], v[:nnz.value]


def read_ref(path):
    R = O.ref()
    R.tfr_read_matrix_market.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p,
                                         C.c_void_p, C.c_void_p]
    rows, cols = C.c_int(), C.c_int()
    nnz = R.tfr_read_matrix_market(path.encode(), C.byref(rows), C.byref(cols), None, None, None)
    if nnz < 0:
        raise RuntimeError(R.tfr_last_error().decode())
    rp = np.zeros(rows.value + 1, np.int32)
    ci = np.zeros(max(nnz, 1), np.int32)
    v = np.zeros(max(nnz, 1))
    assert R.tfr_read_matrix_market(path.encode(), C.byref(rows), C.byref(cols), rp.ctypes.data, ci.ctypes.data,
                                    v.ctypes.data) == nnz
    return rows.value, cols.value, rp, ci[:nnz], v[:nnz]


def write(tmp_path, name, text):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w") as f:
        f.write(text)
    return p


def random_file(rng, symmetric, integer):
    n = int(rng.integers(1, 12))
    m = int(rng.integers(1, 4 * n))
    lines = ["%%MatrixMarket matrix coordinate {} {}".format("integer" if integer else "real",
                                                             "symmetric" if symmetric else "general"),
             "% generated", ""]
    entries = []
    for _ in range(m):
        i, j = int(rng.integers(1, n + 1)), int(rng.integers(1, n + 1))
        if symmetric and i < j:
            i, j = j, i
        val = int(rng.integers(-9, 10)) if integer else float(rng.normal())  # plain Python scalars
        entries.append(f"  {i}\t{j} {val!r}" if not integer else f"{i} {j} {val}")
        if rng.random() < 0.2:
            entries.append("% interleaved comment")
        if rng.random() < 0.1:
            entries.append("   ")
    lines.append(f"{n} {n} {m}")
    lines += entries
    return "\n".join(lines) + "\n"


def test_reader_kats():
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        p = write(d, "g.mtx", "%%MatrixMarket matrix coordinate real general\n% comment\n3 3 4\n3 1 2\n1 2 4\n"
                              "1 2 -1\n3 3 5\n")
        rows, cols, rp, ci, v = read_tf(p)
        assert (rows, cols) == (3, 3)
        assert rp.tolist() == [0, 2, 2, 4] and ci.tolist() == [1, 1, 0, 2] and v.tolist() == [4, -1, 2, 5]
        p = write(d, "s.mtx", "%%MatrixMarket matrix coordinate integer symmetric\n2 2 3\n1 1 2\n2 1 -1\n2 2 3\n")
        _, _, rp, ci, v = read_tf(p)
        assert rp.tolist() == [0, 2, 4] and ci.tolist() == [0, 1, 0, 1] and v.tolist() == [2, -1, -1, 3]
        p = write(d, "e.mtx", "%%MatrixMarket matrix coordinate real general\n0 0 0\n")
        assert read_tf(p)[2].tolist() == [0]
        rows, cols, rp, ci, v = read_tf(FIXTURE)
        assert (rows, cols, len(v)) == (4, 4, 10)
        dense = np.zeros((4, 4))
        for i in range(4):
            for k in range(rp[i], rp[i + 1]):
                dense[i, ci[k]] += v[k]
        assert np.array_equal(dense, 2 * np.eye(4) - np.eye(4, k=1) - np.eye(4, k=-1))


@pytest.mark.parametrize("k", range(len(MALFORMED)))
def test_reader_rejects_malformed(tmp_path, k):
    p = write(tmp_path, "bad.mtx", MALFORMED[k])
    with pytest.raises(RuntimeError):
        read_tf(p)
    if O.ref_available():
        with pytest.raises(RuntimeError):
            read_ref(p)


def test_reader_missing_file_is_an_error(tmp_path):
    with pytest.raises(RuntimeError, match="Cannot open"):
        read_tf(os.path.join(str(tmp_path), "absent.mtx"))


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")
def test_reader_matches_reference_reader(tmp_path):
    rng = np.random.default_rng(11)
    files = [FIXTURE]
    for t in range(60):
        files.append(write(tmp_path, f"r{t}.mtx", random_file(rng, symmetric=bool(t % 2), integer=bool(t % 3 == 0))))
    for p in files:
        a, b = read_tf(p), read_ref(p)
        assert a[:2] == b[:2], p
        for x, y in zip(a[2:], b[2:]):
            assert np.array_equal(x, y), p


def test_cli_rejects_bad_arguments_before_using_the_device(tmp_path):
    for args in ([], [FIXTURE, "amg", "x"], [FIXTURE, "ilu"], [os.path.join(str(tmp_path), "absent.mtx")],
                 [FIXTURE, "amg", "v", "extra"]):
        r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1, (args, r.stdout, r.stderr)
        assert r.stderr.startswith("error="), r.stderr
    p = write(tmp_path, "rect.mtx", "%%MatrixMarket matrix coordinate real general\n2 3 1\n1 1 1\n")
    r = subprocess.run([CLI, p], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "square" in r.stderr


def parse_kv(out):
    kv = {}
    for line in out.splitlines():
        if "=" in line and not line.startswith("level="):
            k, _, val = line.partition("=")
            kv[k] = val
    return kv


@pytest.mark.gpu
@pytest.mark.parametrize("args", [["amg", "v"], ["amg", "w"], ["amg", "f"], ["identity"]])
def test_cli_solves_reference_fixture(args):
    print("solve-matrix", FIXTURE, *args, flush=True)
    r = subprocess.run([CLI, FIXTURE, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout, r.stderr)
    kv = parse_kv(r.stdout)
    assert kv["validation"] == "PASS" and kv["status"] == "Converged"
    assert kv["rows"] == "4" and kv["nnz"] == "10"
    assert float(kv["host_relative_residual"]) <= 1.01e-8
    if args[0] == "amg":
        assert kv["levels"] == "1" and kv["coarse_solver"] == "pivoted GPU LU"
        assert kv["stop_reason"] == "coarse size" and kv["cycle"] == args[1]


@pytest.mark.gpu
def test_cli_solves_generated_poisson(tmp_path):
    from thermofem_amd import problems as P
    rp, ci, v = P.poisson2d(48)
    n = len(rp) - 1
    lines = ["%%MatrixMarket matrix coordinate real symmetric", f"{n} {n} {(len(v) + n) // 2}"]
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            if ci[k] <= i:
                lines.append(f"{i + 1} {int(ci[k]) + 1} {float(v[k])!r}")
    p = write(tmp_path, "p2d.mtx", "\n".join(lines) + "\n")
    for args in (["amg", "v"], ["amg", "w"], ["amg", "f"]):
        print("solve-matrix p2d", *args, flush=True)
        r = subprocess.run([CLI, p, *args], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (r.stdout, r.stderr)
        kv = parse_kv(r.stdout)
        assert kv["validation"] == "PASS" and int(kv["levels"]) >= 3
        assert float(kv["operator_complexity"]) >= 1.0
