"""Build recipe for the gfx950 shared library (and the host C++ API).

Runs hipcc directly (no CMake, no JIT cache) and writes the libraries into
``cuyashe_amd/lib`` so they travel with the source tree.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

ARCH = "gfx950"
# Packed FP32 forms are kept out of every HIP object (the kernels are
# integer-only, the flag guards any float code the compiler might vectorise).
HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

LIB_HIP = os.path.join(LIBDIR, "libyashe_hip.so")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


# Translation units of the library: the C ABI's host logic and one per kernel
# family (compiled in parallel, then linked; each object carries its own
# gfx950 code object).
HIP_SRCS = ["k_ks_fused_lq2.hip", "k_ks_fused_lq1.hip", "k_prod.hip", "k_batch.hip", "k_keyswitch.hip", "k_recon.hip",
            "k_misc.hip", "yashe_hip.hip", "serve.hip"]
OBJDIR = os.path.join(ROOT, "_build", "hip")


def build_hip(verbose: bool = True, jobs: int = None) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = jobs or max(1, min(len(HIP_SRCS), os.cpu_count() or 1))
    objs, procs, pending = [], [], list(HIP_SRCS)
    while pending or procs:
        while pending and len(procs) < jobs:
            src = pending.pop(0)
            obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
            objs.append(obj)
            cmd = [HIPCC, *HIP_FLAGS, "-c", "-I", os.path.join(ROOT, "include"), "-o", obj, os.path.join(CSRC, src)]
            print("+", " ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd)))
        src, p = procs.pop(0)
        if p.wait() != 0:
            for _, q in procs:
                q.wait()
            raise subprocess.CalledProcessError(p.returncode, src)
    tmp = LIB_HIP + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB_HIP)
    return LIB_HIP


LIB_UBENCH = os.path.join(LIBDIR, "libyashe_ubench.so")


def build_ubench() -> str:
    """The transform-arithmetic microbenchmark (a library of its own, read
    by bench.py; no product kernel depends on it)."""
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB_UBENCH + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "ubench.hip")])
    os.replace(tmp, LIB_UBENCH)
    return LIB_UBENCH


HOST_SRCS = ["polynomial.cpp", "yashe.cpp"]
LIB_HOST = os.path.join(LIBDIR, "libcuyashe_host.so")
TEST_API = os.path.join(ROOT, "tests", "cpp", "build", "test_api")
TEST_SHIM = os.path.join(ROOT, "tests", "cpp", "build", "test_reference_shim")


def build_host() -> str:
    """The drop-in C++ API (Polynomial / Ciphertext / Yashe) over the C ABI:
    plain C++ (g++), linked against libyashe_hip.so."""
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(HERE, "host", s) for s in HOST_SRCS]
    tmp = LIB_HOST + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", os.path.join(ROOT, "include"), *srcs,
          "-L", LIBDIR, "-lyashe_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, LIB_HOST)
    os.makedirs(os.path.dirname(TEST_API), exist_ok=True)
    # the reference's device-layer interface over the ABI, and its test program
    _run(["g++", "-O2", "-std=c++17", "-Wall", os.path.join(ROOT, "tests", "cpp", "test_reference_shim.cpp"),
          os.path.join(HERE, "host", "reference_shim.cpp"), "-L", LIBDIR, "-lyashe_hip",
          "-Wl,-rpath," + LIBDIR, "-Wl,-rpath,$ORIGIN/../../../cuyashe_amd/lib", "-o", TEST_SHIM])
    _run(["g++", "-O2", "-std=c++17", "-Wall", os.path.join(ROOT, "tests", "cpp", "test_api.cpp"),
          "-L", LIBDIR, "-lcuyashe_host", "-lyashe_hip", "-Wl,-rpath," + LIBDIR, "-Wl,-rpath,$ORIGIN/../../../cuyashe_amd/lib",
          "-o", TEST_API])
    return LIB_HOST


def build_all():
    return [build_hip(), build_host(), build_ubench()]


if __name__ == "__main__":
    build_all()
