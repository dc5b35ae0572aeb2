"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU):
the wire decoder, bignum, Z_q and ring planner of the library
(tests/cpp/sanitize_host.cpp), the Goldilocks helpers (tests/cpp/gl64_check.cpp)
and the oracle's C restatement (oracle/cpuref.c) at the optimisation level it
ships with. Every report aborts the program (-fno-sanitize-recover=all)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")


def _runtime(name):
    return subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()


def test_host_code_under_sanitizers(tmp_path):
    exe = tmp_path / "sanitize_host"
    subprocess.run(["g++", "-O1", "-std=c++17", *SAN, os.path.join(ROOT, "tests", "cpp", "sanitize_host.cpp"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=ENV)
    assert out.returncode == 0 and out.stdout.strip().startswith("OK"), out.stdout + out.stderr


def test_gl64_helpers_under_sanitizers(tmp_path):
    exe = tmp_path / "gl64_check"
    subprocess.run(["g++", "-O1", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "cuyashe_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "gl64_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=ENV)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr


ORACLE_SCRIPT = r'''
import random
import numpy as np
from oracle import cpuref
rng = random.Random(3)
for n, q in [(256, 655615111), (16, (1 << 127) - 1), (64, (1 << 255) - 19), (32, (1 << 511) - 187),
             (8, 77287149995192912462927307869)]:
    a = [rng.randrange(q) for _ in range(n)]
    b = [rng.randrange(q) for _ in range(n)]
    a[0] = 0
    cpuref.ring_mul(a, b, n, q)
    cpuref.poly_mul_plain(a[: n // 2 + 1], b[:3])
    QL = (q.bit_length() + 63) // 64
    g = cpuref.yashe_mul_limbs(cpuref.to_limbs(a, QL), cpuref.to_limbs([q - 1] * n, QL), n, q, 1024)
    lw = cpuref.num_digits(q, 32)
    G = np.stack([cpuref.to_limbs([rng.randrange(q) for _ in range(n)], QL) for _ in range(lw)])
    cpuref.keyswitch_limbs(g, G, n, q, 32)
    primes = cpuref.gen_primes(n, q.bit_length())
    cpuref.icrt(cpuref.crt(a, primes), primes)
cpuref.poly_inv([rng.randrange((1 << 127) - 1) for _ in range(16)], 16, (1 << 127) - 1)
cpuref.ntt(list(range(64)))
print("OK")
'''


def test_oracle_under_sanitizers(tmp_path):
    so = tmp_path / "libcpuref_san.so"
    subprocess.run(["gcc", "-O3", "-march=x86-64-v3", "-std=gnu11", "-fPIC", "-shared", *SAN,
                    os.path.join(ROOT, "oracle", "cpuref.c"), "-o", str(so)], check=True)
    env = dict(ENV, CUYASHE_ORACLE_LIB=str(so), LD_PRELOAD=f"{_runtime('libasan.so')}:{_runtime('libubsan.so')}",
               PYTHONPATH=ROOT)
    out = subprocess.run([sys.executable, "-c", ORACLE_SCRIPT], capture_output=True, text=True, timeout=300, env=env,
                         cwd=ROOT)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout[-2000:] + out.stderr[-4000:]
