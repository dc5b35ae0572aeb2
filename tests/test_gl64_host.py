"""Host build of the Goldilocks helpers the kernels use
(cuyashe_amd/csrc/gl64.hpp), checked against 128-bit arithmetic: every
twiddle exponent of the shift-only DFTs (mul_pow2, e in [-192, 400)) on the
field's edge values and random elements, and mul/add/sub/reduce128."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_gl64_helpers_match_128bit_arithmetic(tmp_path):
    exe = tmp_path / "gl64_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "cuyashe_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "gl64_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")
