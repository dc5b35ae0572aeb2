"""Oracle for the YASHE polynomial-ring hot path (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of bench.py
may import this module; the product never does. The heavy arithmetic (CRT,
Stockham NTT over P = 2^64 - 2^32 + 1, ICRT) is the C restatement in
``oracle/cpuref.c``; this module composes it on Python integers the way the
reference composes its operators:

* ``poly_mul_plain``  Polynomial::operator* (src/polynomial.cpp:52-60) ->
  common_multiplication (src/common.h:136-226) -> callPolynomialMul
  (src/cuda_functions.cu:617-738) -> icrt.
* ``ring_reduce``     Polynomial::reduce host path (src/cuda_functions.cu:1061-1078)
  followed by ``%= q`` (src/polynomial.h:593-610, non-negative result).
* ``yashe_mul``       Ciphertext::operator* (src/ciphertext.cpp:38-73): the
  rounding of :55-63 applied to t * g, g = c1*c2 mod x^n+1 over Z (the
  reference drops t and reduces mod q first; see the deviation list).
* ``worddecomp`` / ``keyswitch``  src/ciphertext.cpp:81-130 with w = 2^w_bits
  as a base, mask w - 1, shift w_bits, l_w = ceil(bits(q) / w_bits).
* ``decrypt``         Yashe::decrypt (src/yashe.cpp:143-201), every coefficient.
* ``poly_inv``        Polynomial::InvMod (src/polynomial.h:1338-1367).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import List, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CUYASHE_ORACLE_LIB points at another build of the same C file (the
# sanitizer test loads an -fsanitize=address,undefined build).
LIB_PATH = os.environ.get("CUYASHE_ORACLE_LIB") or os.path.join(HERE, "build", "libcpuref.so")
P = (1 << 64) - (1 << 32) + 1

_lib = None
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.or_gmul.restype = ctypes.c_uint64
        L.or_gmul.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.or_ntt.restype = ctypes.c_int
        L.or_ntt.argtypes = [_u64p, ctypes.c_int, ctypes.c_int]
        L.or_crt.restype = None
        L.or_crt.argtypes = [_u64p, ctypes.c_int, ctypes.c_int, _u64p, ctypes.c_int, _u64p]
        L.or_icrt.restype = ctypes.c_int
        L.or_icrt.argtypes = [_u64p, ctypes.c_int, _u64p, ctypes.c_int, _u64p, ctypes.c_int]
        L.or_gen_primes.restype = ctypes.c_int
        L.or_gen_primes.argtypes = [ctypes.c_int, ctypes.c_int, _u64p, ctypes.c_int]
        L.or_poly_mul_plain.restype = ctypes.c_int
        L.or_poly_mul_plain.argtypes = [_u64p, ctypes.c_int, _u64p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        _u64p, ctypes.c_int, _u64p, ctypes.c_int]
        L.or_poly_inv.restype = ctypes.c_int
        L.or_poly_inv.argtypes = [_u64p, ctypes.c_int, _u64p, ctypes.c_int, _u64p]
        L.or_yashe_mul.restype = ctypes.c_int
        L.or_yashe_mul.argtypes = [_u64p, _u64p, ctypes.c_int, _u64p, ctypes.c_int, ctypes.c_uint64, _u64p]
        L.or_keyswitch.restype = ctypes.c_int
        L.or_keyswitch.argtypes = [_u64p, _u64p, ctypes.c_int, _u64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u64p]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


# ------------------------------------------------------------------ conversions

def limbs_needed(values: Sequence[int]) -> int:
    m = max((abs(int(v)) for v in values), default=0)
    return max(1, (m.bit_length() + 63) // 64)


def to_limbs(values: Sequence[int], L: int) -> np.ndarray:
    """Limb-major [L][len(values)] uint64 array of non-negative integers."""
    out = np.zeros((L, len(values)), dtype=np.uint64)
    mask = (1 << 64) - 1
    for j, v in enumerate(values):
        v = int(v)
        if v < 0 or v >> (64 * L):
            raise ValueError("value does not fit")
        for l in range(L):
            out[l, j] = (v >> (64 * l)) & mask
    return out


def from_limbs(arr: np.ndarray) -> List[int]:
    arr = np.asarray(arr, dtype=np.uint64)
    L, n = arr.shape
    out = []
    for j in range(n):
        v = 0
        for l in range(L - 1, -1, -1):
            v = (v << 64) | int(arr[l, j])
        out.append(v)
    return out


# ------------------------------------------------------------------ primitives

def gmul(a: int, b: int) -> int:
    return int(lib().or_gmul(a, b))


def ntt(vec: Sequence[int], inverse: bool = False) -> List[int]:
    """Natural-order cyclic transform over P, omega = 7^((P-1)/N)."""
    a = np.ascontiguousarray(np.array([int(v) % P for v in vec], dtype=np.uint64))
    rc = lib().or_ntt(_ptr(a), len(a), 1 if inverse else 0)
    if rc:
        raise ValueError("bad NTT length")
    return [int(v) for v in a]


def crt(coefs: Sequence[int], primes: Sequence[int]) -> List[List[int]]:
    L = limbs_needed(coefs)
    A = np.ascontiguousarray(to_limbs(coefs, L))
    pr = np.ascontiguousarray(np.array(primes, dtype=np.uint64))
    out = np.zeros((len(primes), len(coefs)), dtype=np.uint64)
    lib().or_crt(_ptr(A), L, len(coefs), _ptr(pr), len(primes), _ptr(out))
    return [[int(v) for v in row] for row in out]


def icrt(res: Sequence[Sequence[int]], primes: Sequence[int], out_limbs: int = None) -> List[int]:
    R = len(primes)
    n = len(res[0])
    M = 1
    for p in primes:
        M *= int(p)
    LO = out_limbs or max(1, (M.bit_length() + 63) // 64)
    r = np.ascontiguousarray(np.array(res, dtype=np.uint64).reshape(R, n))
    pr = np.ascontiguousarray(np.array(primes, dtype=np.uint64))
    out = np.zeros((LO, n), dtype=np.uint64)
    rc = lib().or_icrt(_ptr(r), n, _ptr(pr), R, _ptr(out), LO)
    if rc:
        raise ValueError("ICRT output does not fit")
    return from_limbs(out)


def gen_primes(terms: int, qbits: int) -> List[int]:
    buf = np.zeros(256, dtype=np.uint64)
    R = lib().or_gen_primes(int(terms), int(qbits), _ptr(buf), 256)
    if R < 0:
        raise ValueError("too many primes")
    return [int(v) for v in buf[:R]]


def ntt_length(deg_a: int, deg_b: int, min_N: int = 1) -> int:
    """NTT length of common_multiplication (src/common.h:139-145): the power
    of two >= deg a + deg b, at least the current CUDAFunctions::N; every
    operand coefficient must have a slot."""
    N = 1
    while N < deg_a + deg_b:
        N <<= 1
    N = max(N, min_N, 2)  # at least 2 (a constant times a constant)
    while N < deg_a + 1 or N < deg_b + 1:
        N <<= 1
    return N


def poly_mul_plain(a: Sequence[int], b: Sequence[int], N: int = None, qbits: int = None) -> List[int]:
    """Exact product of non-negative polynomials mod x^N - 1 (cyclic wrap when
    deg a + deg b >= N), as the reference's operator* computes it."""
    a = [int(v) for v in a]
    b = [int(v) for v in b]
    if N is None:
        N = ntt_length(len(a) - 1, len(b) - 1)
    if qbits is None:
        qbits = max(max(a).bit_length(), max(b).bit_length(), 1)
    terms = max(1, min(len(a), len(b)))
    primes = gen_primes(terms, qbits)
    L = max(limbs_needed(a), limbs_needed(b))
    A = np.ascontiguousarray(to_limbs(a, L))
    B = np.ascontiguousarray(to_limbs(b, L))
    M = 1
    for p in primes:
        M *= p
    LO = max(1, (M.bit_length() + 63) // 64)
    pr = np.ascontiguousarray(np.array(primes, dtype=np.uint64))
    out = np.zeros((LO, N), dtype=np.uint64)
    rc = lib().or_poly_mul_plain(_ptr(A), len(a), _ptr(B), len(b), L, N, _ptr(pr), len(primes), _ptr(out), LO)
    if rc:
        raise ValueError(f"oracle product failed ({rc})")
    return from_limbs(out)


# ------------------------------------------------------------------ scalar operations

ADD, SUB, MUL = 0, 1, 2  # add_mode_t (src/settings.h:36)


def residue_scalar(res: Sequence[Sequence[int]], primes: Sequence[int], op: int, s) -> List[List[int]]:
    """polynomialOPInteger (src/cuda_functions.cu:477-548) on residues
    [R][spacing]: ADD / SUB the scalar's residue on coefficient 0 of each
    row, MUL every residue, all mod p_i. s is the scalar (an int, reduced
    mod each p_i as Integer::crt does, src/integer.cpp:62-82) or the list of
    its residues, one per prime (the reference's integer_array)."""
    out = []
    for i, (row, p) in enumerate(zip(res, primes)):
        sp = (int(s[i]) if isinstance(s, (list, tuple)) else int(s)) % p
        r = [int(v) for v in row]
        if op == MUL:
            r = [v * sp % p for v in r]
        elif op == ADD:
            r[0] = (r[0] + sp) % p
        else:
            r[0] = (r[0] - sp) % p
        out.append(r)
    return out


def ring_scalar(a: Sequence[int], op: int, s: int, q: int) -> List[int]:
    """The scalar operators on a ring element (src/polynomial.h:459-573,
    Integer::operator* src/integer.cpp:167-222) followed by % q, the
    arithmetic src/test.cu:630-790 checks against NTL's ZZ_p."""
    r = [int(v) for v in a]
    if op == MUL:
        return [v * int(s) % q for v in r]
    r[0] = (r[0] + int(s)) % q if op == ADD else (r[0] - int(s)) % q
    return r


# ------------------------------------------------------------------ ring operations

def fold(c: Sequence[int], n: int) -> List[int]:
    """c_i - c_{i+n} for i < n over Z (reduce(), src/cuda_functions.cu:1066-1072)."""
    c = list(c) + [0] * max(0, 2 * n - len(c))
    return [int(c[i]) - int(c[i + n]) for i in range(n)]


def ring_reduce(c: Sequence[int], n: int, q: int) -> List[int]:
    return [v % q for v in fold(c, n)]


def ring_mul(a: Sequence[int], b: Sequence[int], n: int, q: int) -> List[int]:
    """a*b; reduce(); %= q  -- the sequence of multiplyByPolynomial
    (src/test.cu:472-524)."""
    prod = poly_mul_plain(a, b, N=2 * n, qbits=q.bit_length())
    return ring_reduce(prod, n, q)


def round_tq(x: int, t: int, q: int) -> int:
    """Rounding of src/ciphertext.cpp:55-63 on t*x: NTL floor division, non-
    negative remainder, round up when 2*rem > q, result mod q."""
    quot, rem = divmod(t * x, q)  # Python divmod == NTL / and % for q > 0
    return (quot + 1) % q if 2 * rem > q else quot % q


def yashe_mul(c1: Sequence[int], c2: Sequence[int], n: int, q: int, t: int) -> List[int]:
    prod = poly_mul_plain(c1, c2, N=2 * n, qbits=q.bit_length())
    g = fold(prod, n)
    return [round_tq(v, t, q) for v in g]


def num_digits(q: int, w_bits: int) -> int:
    return (q.bit_length() + w_bits - 1) // w_bits


def worddecomp(c: Sequence[int], w_bits: int, lw: int) -> List[List[int]]:
    """D_k(c) = floor(c / w^k) mod w for k < l_w (src/ciphertext.cpp:109-130
    with the mask w-1 and shift w_bits)."""
    mask = (1 << w_bits) - 1
    return [[(int(v) >> (w_bits * k)) & mask for v in c] for k in range(lw)]


def keyswitch(c: Sequence[int], gammas: Sequence[Sequence[int]], n: int, q: int, w_bits: int) -> List[int]:
    """sum_k D_k(c) * gamma_k, then reduce() (src/ciphertext.cpp:81-107)."""
    lw = len(gammas)
    digits = worddecomp(c, w_bits, lw)
    acc = [0] * (2 * n)
    for k in range(lw):
        if not any(digits[k]):
            continue
        prod = poly_mul_plain(digits[k], gammas[k], N=2 * n, qbits=max(q.bit_length(), w_bits))
        for j in range(2 * n):
            acc[j] += prod[j]
    return ring_reduce(acc, n, q)


def decrypt(f: Sequence[int], c: Sequence[int], n: int, q: int, t: int) -> List[int]:
    """m = [round(t/q * [f*c]_q)]_t per coefficient (src/yashe.cpp:167-197)."""
    g = ring_mul(f, c, n, q)
    out = []
    for v in g:
        quot, rem = divmod(v * t, q)
        out.append((quot + 1) % t if 2 * rem > q else quot % t)
    return out


def poly_inv(f: Sequence[int], n: int, q: int) -> List[int]:
    """f^-1 in Z_q[x]/(x^n + 1), q prime (Polynomial::InvMod)."""
    QL = max(1, (q.bit_length() + 63) // 64)
    F = np.zeros((n, QL), dtype=np.uint64)
    mask = (1 << 64) - 1
    for j in range(n):
        v = int(f[j]) % q if j < len(f) else 0
        for l in range(QL):
            F[j, l] = (v >> (64 * l)) & mask
    F = np.ascontiguousarray(F)
    Q = np.ascontiguousarray(np.array([(q >> (64 * l)) & mask for l in range(QL)], dtype=np.uint64))
    out = np.zeros((n, QL), dtype=np.uint64)
    rc = lib().or_poly_inv(_ptr(F), n, _ptr(Q), QL, _ptr(out))
    if rc:
        raise ValueError("not invertible")
    return [sum(int(out[j, l]) << (64 * l) for l in range(QL)) for j in range(n)]


# ------------------------------------------------------------------ composite operations in C
# The same compositions as yashe_mul / keyswitch above, run entirely in the C
# restatement (or_yashe_mul / or_keyswitch) on limb-major arrays; the ctypes
# calls release the interpreter lock, so threads run them in parallel.

def _q_array(q: int):
    QL = max(1, (q.bit_length() + 63) // 64)
    return np.ascontiguousarray(np.array([(q >> (64 * l)) & ((1 << 64) - 1) for l in range(QL)], dtype=np.uint64)), QL


def yashe_mul_limbs(c1: np.ndarray, c2: np.ndarray, n: int, q: int, t: int) -> np.ndarray:
    """[QL][n] limbs of [round(t c1 c2 / q)]_q (c1, c2: [QL][n] limbs)."""
    Q, QL = _q_array(q)
    a = np.ascontiguousarray(c1, dtype=np.uint64)
    b = np.ascontiguousarray(c2, dtype=np.uint64)
    out = np.zeros((QL, n), dtype=np.uint64)
    rc = lib().or_yashe_mul(_ptr(a), _ptr(b), n, _ptr(Q), QL, int(t), _ptr(out))
    if rc:
        raise ValueError(f"or_yashe_mul failed ({rc})")
    return out


def keyswitch_limbs(c: np.ndarray, gammas: np.ndarray, n: int, q: int, w_bits: int) -> np.ndarray:
    """[QL][n] limbs of [sum_k D_k(c) gamma_k]_(x^n+1, q) (gammas: [lw][QL][n])."""
    Q, QL = _q_array(q)
    a = np.ascontiguousarray(c, dtype=np.uint64)
    g = np.ascontiguousarray(gammas, dtype=np.uint64)
    out = np.zeros((QL, n), dtype=np.uint64)
    rc = lib().or_keyswitch(_ptr(a), _ptr(g), n, _ptr(Q), QL, int(w_bits), int(g.shape[0]), _ptr(out))
    if rc:
        raise ValueError(f"or_keyswitch failed ({rc})")
    return out
