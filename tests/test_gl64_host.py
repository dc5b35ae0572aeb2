"""Host build of the Goldilocks helpers the kernels use
and small-prime helpers (cuyashe_amd/csrc/gl64.hpp, modp.hpp), checked
against 128-bit arithmetic: every twiddle exponent of the shift-only DFTs
(mul_pow2, e in [-192, 400)) on the field's edge values and random elements,
mul/add/sub/sub_full/reduce128/reduce_wide, the Montgomery product and
reduction, the small-prime finish and the small-prime reduction of CRT sums
(mod_small)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.parametrize("cxx", ["g++", CLANG])
def test_gl64_helpers_match_128bit_arithmetic(tmp_path, cxx):
    # g++ takes the plain-C carry helpers, clang++ the __builtin_addc/subc
    # path the device build uses
    if not os.path.exists(cxx) and cxx != "g++":
        pytest.skip(f"{cxx} not present")
    exe = tmp_path / "gl64_check"
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "cuyashe_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "gl64_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")
