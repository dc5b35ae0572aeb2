"""The library's hipcc flags keep packed FP32 instructions out of device code.

`-Xclang -target-feature -Xclang -packed-fp32-ops` is passed on every hipcc
line (cuyashe_amd/build.py). hipcc's host pass warns that it does not know
the feature; the device pass honours it. A float2 multiply-add, which gfx950
compiles to v_pk_fma_f32 by default, must compile to unpacked instructions
with the library's flags. CPU only (hipcc cross-compiles).
"""
import os
import shutil
import subprocess

import pytest

from cuyashe_amd import build

HIPCC = build.HIPCC

SRC = r"""
#include <hip/hip_runtime.h>
__global__ void k(float2* a, const float2* b, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    float2 x = a[i], y = b[i];
    a[i] = make_float2(x.x * y.x + x.x, x.y * y.y + x.y);
  }
}
"""


def _device_asm(tmp_path, flags):
    src = tmp_path / "k.hip"
    src.write_text(SRC)
    out = tmp_path / "k.s"
    subprocess.run([HIPCC, *flags, "--cuda-device-only", "-S", "-o", str(out), str(src)], check=True,
                   capture_output=True, timeout=300)
    return out.read_text()


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not installed")
def test_packed_fp32_kept_out_of_device_code(tmp_path):
    flags = [f for f in build.HIP_FLAGS if f not in ("-fPIC",)]
    with_flag = _device_asm(tmp_path, flags)
    assert "v_fma" in with_flag or "v_fmac" in with_flag
    assert "v_pk_" not in with_flag
    # the same kernel without the feature flag does use packed FP32, so the
    # check above is not vacuous
    plain = [f for f in flags if f not in ("-Xclang", "-target-feature", "-packed-fp32-ops")]
    assert "v_pk_fma_f32" in _device_asm(tmp_path, plain)
