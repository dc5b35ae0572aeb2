"""The drop-in C++ API (Polynomial / Ciphertext / Yashe over the C ABI)
running the reference's own test programs: src/test_mul.cu and the
PolySuite / YasheSuite cases of src/test.cu (tests/cpp/test_api.cpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "build", "test_api")


def test_reference_test_programs_through_cpp_api():
    assert os.path.exists(BIN), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=600)
    print(p.stdout)
    lines = [l for l in p.stdout.splitlines() if l.startswith(("PASS", "FAIL", "ERROR"))]
    assert len([l for l in lines if l.startswith("PASS")]) >= 12, p.stdout + p.stderr
    assert p.returncode == 0, p.stdout + p.stderr


SHIM = os.path.join(ROOT, "tests", "cpp", "build", "test_reference_shim")


def test_reference_device_layer_through_shim():
    # the reference's CUDAFunctions / callCRT / callICRT signatures, bound to
    # the library by cuyashe_amd/host/reference_shim.cpp, run src/test_mul.cu's
    # two known answers through callCRT -> callPolynomialMul -> callICRT
    assert os.path.exists(SHIM), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    p = subprocess.run([SHIM], capture_output=True, text=True, timeout=300)
    print(p.stdout)
    lines = [l for l in p.stdout.splitlines() if l.startswith(("PASS", "FAIL", "ERROR"))]
    assert len([l for l in lines if l.startswith("PASS")]) == 4, p.stdout + p.stderr
    assert p.returncode == 0, p.stdout + p.stderr
