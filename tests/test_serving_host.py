"""Host-side pieces of the serving loop (no GPU): the plaintext checker and
the request generator."""
import numpy as np

from cuyashe_amd.scheme import negacyclic_mul_mod
from cuyashe_amd.serving import KINDS, c5_rings, negacyclic_mul_mod_t, poisson_requests


def test_fft_negacyclic_matches_direct():
    rng = np.random.default_rng(0)
    for n in (2, 8, 64, 2048, 16384):
        a = rng.integers(0, 1024, size=n)
        b = rng.integers(0, 1024, size=n)
        if n <= 2048:
            assert np.array_equal(negacyclic_mul_mod_t(a, b, 1024), negacyclic_mul_mod(a, b, 1024))
        # x^(n-1) * x = -1: the wrap carries the sign
        e = np.zeros(n, dtype=np.int64)
        e[n - 1] = 1
        x = np.zeros(n, dtype=np.int64)
        x[1 % n] = 1
        want = np.zeros(n, dtype=np.int64)
        want[0] = 1023 if n > 1 else 1
        if n > 1:
            assert np.array_equal(negacyclic_mul_mod_t(e, x, 1024), want)


def test_poisson_requests_shape():
    rings = c5_rings()
    assert len(rings) == 12 and {n for n, _ in rings} == {2048, 4096, 8192, 16384}
    reqs = poisson_requests(20000, 5000.0, rings, 1024, 32, seed=3)
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) >= 0) and arr[0] == 0.0
    # mean gap 1/qps
    assert abs(np.mean(np.diff(arr)) * 5000.0 - 1.0) < 0.05
    frac = {k: sum(r.kind == k for r in reqs) / len(reqs) for k in KINDS}
    assert abs(frac["add"] - 0.6) < 0.02 and abs(frac["mulks"] - 0.3) < 0.02 and abs(frac["encdec"] - 0.1) < 0.02
    for r in reqs[:200]:
        if r.kind == "encdec":
            assert r.msg.shape == (r.ring[0],) and r.msg.max() < 1024
        else:
            assert 0 <= r.a < 32 and 0 <= r.b < 32


def test_fft_negacyclic_batched_rows():
    # the checker runs one FFT over a stack of requests: rows are independent
    rng = np.random.default_rng(1)
    a = rng.integers(0, 1024, size=(5, 256))
    b = rng.integers(0, 1024, size=(5, 256))
    got = negacyclic_mul_mod_t(a, b, 1024)
    for k in range(5):
        assert np.array_equal(got[k], negacyclic_mul_mod(a[k], b[k], 1024))


def test_request_record_matches_c_layout(tmp_path):
    # the numpy record handed to yh_server_run has yh_request's layout
    import shutil
    import subprocess
    from pathlib import Path
    import pytest
    from cuyashe_amd._lib import REQUEST_DTYPE
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    inc = Path(__file__).resolve().parents[1] / "include"
    src = tmp_path / "layout.cpp"
    fields = REQUEST_DTYPE.names
    src.write_text('#include <cstdio>\n#include <cstddef>\n#include "yashe_hip.h"\nint main() {\n'
                   '  std::printf("%zu", sizeof(yh_request));\n' +
                   "".join(f'  std::printf(" %zu", offsetof(yh_request, {f}));\n' for f in fields) +
                   '  return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run([gxx, "-I", str(inc), str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert out[0] == REQUEST_DTYPE.itemsize
    assert out[1:] == [REQUEST_DTYPE.fields[f][1] for f in fields]
