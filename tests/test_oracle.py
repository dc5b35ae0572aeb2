"""Pins the oracle (CPU) before it is trusted as the parity reference.

Each oracle function is checked against data the reference's own tests hold
(the test_mul.cu KATs, test_bnt.cu, test_reduce.cu), against closed forms
(brute-force DFT, schoolbook products, explicit folds, exact rational
rounding) and with planted mistakes that one of those checks must catch.
"""
import random

import numpy as np

import pytest

from oracle import cpuref

P = cpuref.P


def schoolbook(a, b):
    out = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


def negacyclic(a, b, n):
    out = [0] * n
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            k = i + j
            out[k % n] += -x * y if k >= n else x * y
    return out


# ------------------------------------------------------------------ Goldilocks / NTT

def test_gmul_matches_uint128_mod_p(kats):
    for x, y, z in kats["s_mul"]["cases"]:
        assert cpuref.gmul(int(x), int(y)) == int(z)
    rng = random.Random(1)
    for _ in range(200):
        x, y = rng.getrandbits(64), rng.getrandbits(64)
        assert cpuref.gmul(x, y) == x * y % P


def test_ntt_is_the_dft(kats):
    for case in kats["ntt_dft"]["cases"]:
        x = [int(v) for v in case["x"]]
        assert cpuref.ntt(x) == [int(v) for v in case["X"]]
        assert pow(7, (P - 1) // case["N"], P) == int(case["wN"])


def test_ntt_brute_force_and_roundtrip():
    rng = random.Random(2)
    for N in (2, 4, 8, 32, 256):
        x = [rng.randrange(P) for _ in range(N)]
        w = pow(7, (P - 1) // N, P)
        want = [sum(x[j] * pow(w, j * k, P) for j in range(N)) % P for k in range(N)]
        assert cpuref.ntt(x) == want
        assert cpuref.ntt(cpuref.ntt(x), inverse=True) == x


def test_ntt_convolution_theorem():
    rng = random.Random(3)
    N = 64
    a = [rng.randrange(1 << 20) for _ in range(N)]
    b = [rng.randrange(1 << 20) for _ in range(N)]
    A, B = cpuref.ntt(a), cpuref.ntt(b)
    c = cpuref.ntt([x * y % P for x, y in zip(A, B)], inverse=True)
    want = [0] * N
    for i in range(N):
        for j in range(N):
            want[(i + j) % N] += a[i] * b[j]
    assert c == [v % P for v in want]


# ------------------------------------------------------------------ CRT / ICRT

def test_crt_closed_form_and_roundtrip():
    rng = random.Random(4)
    primes = cpuref.gen_primes(64, 127)
    M = 1
    for p in primes:
        M *= p
    xs = [rng.randrange(M) for _ in range(50)] + [0, 1, M - 1]
    res = cpuref.crt(xs, primes)
    for i, p in enumerate(primes):
        assert res[i] == [x % p for x in xs]
    assert cpuref.icrt(res, primes) == xs


def test_test_bnt_crt_icrt(kats):
    k = kats["test_bnt_crt"]
    X, q = int(k["X"]), int(k["q"])
    primes = cpuref.gen_primes(1, q.bit_length())
    M = 1
    for p in primes:
        M *= p
    assert M > 2 * q * q
    assert cpuref.icrt(cpuref.crt([X], primes), primes) == [X]


def test_gen_primes_rule():
    for terms, qbits in ((8, 96), (4096, 127), (8192, 255)):
        primes = cpuref.gen_primes(terms, qbits)
        assert primes == sorted(primes, reverse=True) and len(set(primes)) == len(primes)
        M = 1
        for p in primes:
            M *= p
            assert terms * (p - 1) ** 2 < P
            assert all(p % d for d in range(2, int(p ** 0.5) + 1))
        assert M > 2 * terms * (1 << (2 * qbits))


# ------------------------------------------------------------------ products

def test_kat_test_mul_1_cyclic_wrap(kats):
    k = kats["test_mul_1"]
    a = [int(v) for v in k["a"]]
    b = [int(v) for v in k["b"]]
    N = cpuref.ntt_length(len(a) - 1, len(b) - 1)
    assert N == 16
    got = cpuref.poly_mul_plain(a, b, N=N, qbits=int(k["q"]).bit_length())
    assert got == [int(v) for v in k["expected"]]


def test_kat_test_mul_2_exact(kats):
    k = kats["test_mul_2"]
    a = [int(v) for v in k["a"]]
    b = [int(v) for v in k["b"]]
    N = cpuref.ntt_length(len(a) - 1, len(b) - 1)
    got = cpuref.poly_mul_plain(a, b, N=N, qbits=int(k["q"]).bit_length())
    assert got[:15] == [int(v) for v in k["expected"]] and got[15] == 0


def test_plain_product_schoolbook():
    rng = random.Random(5)
    for n, qb in ((8, 64), (64, 127), (256, 255)):
        a = [rng.getrandbits(qb) for _ in range(n)]
        b = [rng.getrandbits(qb) for _ in range(n)]
        got = cpuref.poly_mul_plain(a, b, N=2 * n, qbits=qb)
        assert got[:2 * n - 1] == schoolbook(a, b)


def test_ring_ops_closed_forms(kats):
    k = kats["test_multiplyByPolynomial"]
    q, n = int(k["q"]), k["n"]
    for case in k["cases"]:
        a = [int(v) for v in case["a"]]
        b = [int(v) for v in case["b"]]
        assert cpuref.ring_mul(a, b, n, q) == [int(v) for v in case["expected"]]
    k = kats["test_phiReduce"]
    for case in k["cases"]:
        assert cpuref.ring_reduce([int(v) for v in case["c"]], k["n"], int(k["q"])) == \
            [int(v) for v in case["expected"]]
    k = kats["test_reduce_fold"]
    for inp, exp in zip(k["inputs"], k["expected"]):
        assert cpuref.ring_reduce(inp, k["n"], int(k["q"])) == exp


def test_several_multiplications_chain(kats):
    k = kats["test_severalMultiplications"]
    q, n = int(k["q"]), k["n"]
    a = [int(v) for v in k["a"]]
    c = cpuref.ring_mul(a, [int(v) for v in k["b"]], n, q)
    assert c == [int(v) for v in k["chain"][0]]
    for want in k["chain"][1:]:
        c = cpuref.ring_mul(c, a, n, q)
        assert c == [int(v) for v in want]


def test_yashe_mul_rounding_closed_form():
    rng = random.Random(6)
    n, q, t = 16, (1 << 127) - 1, 1024
    c1 = [rng.randrange(q) for _ in range(n)]
    c2 = [rng.randrange(q) for _ in range(n)]
    g = negacyclic(c1, c2, n)
    from fractions import Fraction
    want = []
    for v in g:
        x = Fraction(t * v, q)
        fl = x.numerator // x.denominator
        # round half down (2*rem > q rounds up; q odd so no ties)
        r = fl + 1 if x - fl > Fraction(1, 2) else fl
        want.append(r % q)
    assert cpuref.yashe_mul(c1, c2, n, q, t) == want


def test_keyswitch_identity():
    # With gamma_k = f * w^k (no noise) the keyswitch returns c*f mod (x^n+1, q):
    # sum_k D_k(c) w^k = c.
    rng = random.Random(7)
    n, q, wb = 16, (1 << 127) - 1, 32
    lw = cpuref.num_digits(q, wb)
    f = [rng.randrange(q) for _ in range(n)]
    gammas = []
    for k in range(lw):
        gammas.append([v * (1 << (wb * k)) % q for v in f])
    c = [rng.randrange(q) for _ in range(n)]
    assert cpuref.keyswitch(c, gammas, n, q, wb) == [v % q for v in negacyclic(c, f, n)]


def test_worddecomp_recomposes():
    rng = random.Random(8)
    q = (1 << 255) - 19
    for wb in (8, 16, 32):
        lw = cpuref.num_digits(q, wb)
        c = [rng.randrange(q) for _ in range(20)]
        d = cpuref.worddecomp(c, wb, lw)
        assert [sum(d[k][j] << (wb * k) for k in range(lw)) for j in range(20)] == c
        assert all(0 <= x < (1 << wb) for row in d for x in row)


def test_poly_inverse():
    rng = random.Random(9)
    n, q = 32, (1 << 127) - 1
    f = [rng.randrange(q) for _ in range(n)]
    fi = cpuref.poly_inv(f, n, q)
    one = [v % q for v in negacyclic(f, fi, n)]
    assert one == [1] + [0] * (n - 1)


def test_decrypt_closed_form():
    # c = delta*m + (small noise) with f = 1 decrypts to m.
    rng = random.Random(10)
    n, q, t = 16, (1 << 127) - 1, 1024
    delta = q // t
    m = [rng.randrange(t) for _ in range(n)]
    c = [(delta * v + rng.randrange(-50, 50)) % q for v in m]
    f = [1] + [0] * (n - 1)
    assert cpuref.decrypt(f, c, n, q, t) == m


# ------------------------------------------------------------------ planted mistakes
# Each check above must fail on a plausible oracle error.

def test_planted_mistakes_are_caught(kats, monkeypatch):
    k = kats["test_mul_1"]
    a = [int(v) for v in k["a"]]
    b = [int(v) for v in k["b"]]
    want = [int(v) for v in k["expected"]]
    # (1) dropping the cyclic wrap (linear product truncated to N)
    lin = schoolbook(a, b)[:16]
    assert lin != want
    # (2) a wrong sign in the fold
    c = [3, 5, 7, 11]
    good = cpuref.ring_reduce(c, 2, 101)
    bad = [(c[i] + c[i + 2]) % 101 for i in range(2)]
    assert good != bad
    # (3) a transposed operand in the digit decomposition
    d = cpuref.worddecomp([(1 << 40) + 5], 32, 2)
    assert d == [[5], [256]] and d[::-1] != d
    # (4) floor only, without rounding up when 2 * rem > q
    q, t = (1 << 61) - 1, 1024
    x = next(v for v in range(q // 2, q // 2 + 5000) if 2 * (t * v % q) > q)
    assert cpuref.round_tq(x, t, q) == (t * x // q + 1) % q
    assert cpuref.round_tq(x, t, q) != (t * x // q) % q
    # (5) a scalar add applied to every coefficient instead of coefficient 0
    primes = cpuref.gen_primes(4, 64)
    res = cpuref.crt([5, 6, 7, 8], primes)
    wrong = [[(v + 3) % p for v in row] for row, p in zip(res, primes)]
    assert cpuref.residue_scalar(res, primes, cpuref.ADD, 3) != wrong


# ------------------------------------------------------------------ scalar operations

def test_residue_scalar_through_crt_icrt():
    # polynomialOPInteger on residues is the integer operation mod M: the
    # (pinned) ICRT of the modified residues gives a_0 +- s, a_j s mod M
    rng = random.Random(11)
    q = (1 << 127) - 1
    primes = cpuref.gen_primes(16, q.bit_length())
    M = 1
    for p in primes:
        M *= p
    a = [rng.randrange(q) for _ in range(16)]
    res = cpuref.crt(a, primes)
    for s in (0, 1, rng.randrange(q), rng.randrange(1 << 400)):
        back = cpuref.icrt(cpuref.residue_scalar(res, primes, cpuref.MUL, s), primes)
        assert back == [v * s % M for v in a]
        back = cpuref.icrt(cpuref.residue_scalar(res, primes, cpuref.ADD, s), primes)
        assert back == [(a[0] + s) % M] + a[1:]
        back = cpuref.icrt(cpuref.residue_scalar(res, primes, cpuref.SUB, s), primes)
        assert back == [(a[0] - s) % M] + a[1:]


def test_ring_scalar_through_the_product():
    # a scalar product is the ring product with the constant polynomial s
    # (the pinned plain product, fold and mod q); adding s to coefficient 0
    # distributes over a ring product: (a + s) b = a b + s b
    rng = random.Random(12)
    n, q = 16, (1 << 127) - 1
    a = [rng.randrange(q) for _ in range(n)]
    b = [rng.randrange(q) for _ in range(n)]
    for s in (0, 1, q - 1, rng.randrange(q)):
        const = [s % q] + [0] * (n - 1)
        assert cpuref.ring_scalar(a, cpuref.MUL, s, q) == cpuref.ring_mul(a, const, n, q)
        lhs = cpuref.ring_mul(cpuref.ring_scalar(a, cpuref.ADD, s, q), b, n, q)
        ab = cpuref.ring_mul(a, b, n, q)
        assert lhs == [(u + s * v) % q for u, v in zip(ab, b)]
        lhs = cpuref.ring_mul(cpuref.ring_scalar(a, cpuref.SUB, s, q), b, n, q)
        assert lhs == [(u - s * v) % q for u, v in zip(ab, b)]


# ------------------------------------------------------------------ compositions in C

@pytest.mark.parametrize("n,q,t,wb", [(16, (1 << 127) - 1, 1024, 32), (64, (1 << 255) - 19, 1 << 16, 32),
                                      (32, (1 << 511) - 187, 1024, 16), (8, 77287149995192912462927307869, 35951, 32)])
def test_c_compositions_equal_python(n, q, t, wb):
    # or_yashe_mul / or_keyswitch (used by the CPU baseline) against the
    # Python compositions pinned above, including all-(q-1) coefficients
    rng = random.Random(n + wb)
    QL = (q.bit_length() + 63) // 64
    c1 = [rng.randrange(q) for _ in range(n)]
    c2 = [q - 1] * n
    lw = cpuref.num_digits(q, wb)
    gam = [[rng.randrange(q) for _ in range(n)] for _ in range(lw)]
    g = cpuref.yashe_mul(c1, c2, n, q, t)
    gc = cpuref.yashe_mul_limbs(cpuref.to_limbs(c1, QL), cpuref.to_limbs(c2, QL), n, q, t)
    assert cpuref.from_limbs(gc) == g
    G = np.stack([cpuref.to_limbs(x, QL) for x in gam])
    assert cpuref.from_limbs(cpuref.keyswitch_limbs(gc, G, n, q, wb)) == cpuref.keyswitch(g, gam, n, q, wb)


def test_c_compositions_thread_safe():
    # the transform tables are per thread: concurrent products of different
    # lengths give the single-threaded results
    import concurrent.futures
    q = (1 << 127) - 1
    rng = random.Random(5)
    cases = []
    for n in (16, 32, 64, 128, 256, 512) * 2:
        c1 = cpuref.to_limbs([rng.randrange(q) for _ in range(n)], 2)
        c2 = cpuref.to_limbs([rng.randrange(q) for _ in range(n)], 2)
        cases.append((c1, c2, n))
    want = [cpuref.yashe_mul_limbs(a, b, n, q, 1024) for a, b, n in cases]
    with concurrent.futures.ThreadPoolExecutor(max_workers=6) as ex:
        got = list(ex.map(lambda c: cpuref.yashe_mul_limbs(c[0], c[1], c[2], q, 1024), cases))
    assert all(np.array_equal(g, w) for g, w in zip(got, want))
