"""The C++20 drop-in API (thermofem_amd/include/thermofem/*.hpp) exercised by
tests/cpp/api_test.cpp: suites of checks whose inputs and expected values are
those of the reference's tests (solver_test.cpp, cg_test.cpp,
multigrid_operations_test.cpp, amg_native_test.cpp, data_structure_test.cpp,
matrix_market_test.cpp, linear_algebra_test.cpp), run over the HIP library.
The option, CSR-validation and Matrix Market suites need no device and run on
the CPU; the full program runs on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "build", "api_test")


def _binary():
    if not os.path.exists(BIN):
        from thermofem_amd.build import build_programs
        build_programs()
    return BIN


def test_cpp_api_host_cases():
    r = subprocess.run([_binary(), "--cpu-only"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-4000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "0 failed checks" in r.stdout


@pytest.mark.gpu
def test_cpp_api_all_cases():
    print("running", BIN, flush=True)
    r = subprocess.run([_binary()], capture_output=True, text=True, timeout=600)
    print(r.stdout[-6000:], r.stderr[-6000:])
    assert r.returncode == 0
    assert "0 failed checks" in r.stdout
