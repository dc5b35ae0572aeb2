"""CPU checks of the algebra behind the reconstruction epilogues.

The device never forms the reconstructed integer c = sum_i y_i M_i - k M for
the mod-q, scale-and-round and decrypt epilogues; it folds the sum through
per-modulus constants instead:
  c mod q           = (sum y_i (M_i mod q) + k (q - M mod q)) mod q
  round(t c / q)    = sum y_i alpha_i - k alpha_M + floor(S / q) + [2 (S mod q) > q]
with t M_i = q alpha_i + beta_i, t M = q alpha_M + beta_M,
S = sum y_i beta_i - k beta_M. These tests restate the folds with Python
integers on the library's own modulus sets and compare them with the
oracle's definitions (c mod q, and the rounding of src/ciphertext.cpp:55-63
as restated by the oracle). The GPU parity tests check the device code.
"""
import random

import pytest

import cuyashe_amd as ya
from oracle import cpuref

P = cpuref.P


def reconstruction(primes, c):
    """y_i = (c mod m_i) (M/m_i)^-1 mod m_i and k with c = sum y_i M_i - k M."""
    M = 1
    for m in primes:
        M *= m
    ys = []
    for m in primes:
        Mi = M // m
        ys.append((c % m) * pow(Mi, -1, m) % m)
    total = sum(y * (M // m) for y, m in zip(ys, primes))
    assert (total - c) % M == 0
    k = (total - c) // M
    return M, ys, k


def fold_mod_q(primes, ys, k, M, q):
    acc = sum(y * ((M // m) % q) for y, m in zip(ys, primes)) + k * (q - M % q)
    return acc % q


def fold_mod_q_split(primes, ys, k, M, q):
    """As fold_mod_q with the y of modulus 0 (P) split into 32-bit halves:
    y_0 = y_0l + 2^32 y_0h, the high half's constant 2^32 (M_0 mod q) mod q."""
    X = [(M // m) % q for m in primes]
    y0l, y0h = ys[0] & 0xFFFFFFFF, ys[0] >> 32
    XR = (X[0] << 32) % q
    acc = y0l * X[0] + y0h * XR + sum(y * x for y, x in zip(ys[1:], X[1:])) + k * (q - M % q)
    assert acc < (1 << (64 * ((q.bit_length() + 63) // 64) + 34))
    return acc % q


def fold_scale_split(primes, ys, k, M, q, t):
    """The scale-and-round fold with y_0 split: the high half uses
    beta_0' = 2^32 beta_0 mod q and alpha_0' = 2^32 alpha_0 + floor(2^32 beta_0 / q)."""
    alphas = [(t * (M // m)) // q for m in primes]
    betas = [(t * (M // m)) % q for m in primes]
    aM, bM = (t * M) // q, (t * M) % q
    y0l, y0h = ys[0] & 0xFFFFFFFF, ys[0] >> 32
    b0p = (betas[0] << 32) % q
    a0p = ((alphas[0] << 32) + ((betas[0] << 32) // q)) % q
    S1 = y0l * betas[0] + y0h * b0p + sum(y * b for y, b in zip(ys[1:], betas[1:])) + k * (q - bM)
    Q1, r = divmod(S1, q)
    up = 1 if 2 * r > q else 0
    acc = y0l * (alphas[0] % q) + y0h * a0p + sum(y * (a % q) for y, a in zip(ys[1:], alphas[1:])) \
        + k * (q - ((aM + 1) % q)) + Q1 + up
    return acc % q


def fold_scale(primes, ys, k, M, q, t, off_by_one=False):
    alphas = [(t * (M // m)) // q for m in primes]
    betas = [(t * (M // m)) % q for m in primes]
    aM, bM = (t * M) // q, (t * M) % q
    S1 = sum(y * b for y, b in zip(ys, betas)) + k * (q - bM)  # S + k q >= 0
    Q1, r = divmod(S1, q)
    up = 1 if 2 * r > q else 0
    neg_a1 = q - ((aM + (0 if off_by_one else 1)) % q)
    acc = sum(y * (a % q) for y, a in zip(ys, alphas)) + k * neg_a1 + Q1 + up
    return acc % q


SHAPES = [(4096, (1 << 127) - 1, 1024), (8192, (1 << 255) - 19, 1 << 16), (64, 77287149995192912462927307869, 35951),
          (16, (1 << 511) - 187, 1024)]


@pytest.mark.parametrize("n,q,t", SHAPES)
def test_folded_epilogues_match_definitions(n, q, t):
    info, primes, ks_primes = ya.ring_plan(n, q, t, 32, keyswitch=True)
    rng = random.Random(n ^ q.bit_length())
    for mods in (primes, ks_primes):
        M = 1
        for m in mods:
            M *= m
        B = M // 4
        cases = [0, 1, -1, B, -B, q, -q, q // 2, -(q // 2), (q + 1) // 2]
        cases += [rng.randrange(-B, B + 1) for _ in range(200)]
        for c in cases:
            M_, ys, k = reconstruction(mods, c)
            assert M_ == M
            assert fold_mod_q(mods, ys, k, M, q) == c % q
            assert fold_mod_q_split(mods, ys, k, M, q) == c % q
            if mods is primes:
                want = cpuref.round_tq(c, t, q)
                assert fold_scale(mods, ys, k, M, q, t) == want, c
                assert fold_scale_split(mods, ys, k, M, q, t) == want, c


def test_folded_scale_planted_mistake_is_caught():
    n, q, t = 4096, (1 << 127) - 1, 1024
    _, primes, _ = ya.ring_plan(n, q, t, 32, keyswitch=True)
    M = 1
    for m in primes:
        M *= m
    rng = random.Random(5)
    wrong = 0
    for _ in range(50):
        c = rng.randrange(-(M // 4), M // 4)
        _, ys, k = reconstruction(primes, c)
        if k and fold_scale(primes, ys, k, M, q, t, off_by_one=True) != cpuref.round_tq(c, t, q):
            wrong += 1
    assert wrong > 0
