"""GPU parity of the YASHE multiply, the keyswitch and the fused
mul+keyswitch at the largest rings (src/ciphertext.cpp:38-130):
n = 8192 with q = 2^255 - 19 and n = 16384 with q = 2^127 - 1 and
q = 2^511 - 187, two ciphertext pairs each, bit-exact against the oracle.

Per ring the oracle computes each pair's scaled product g once and its
keyswitch once: the GPU multiply is checked against g, the GPU keyswitch of
g against the oracle keyswitch, and the fused GPU operation against the same
keyswitch (so the three operations share the oracle's work).
"""
import concurrent.futures
import random

import pytest

import cuyashe_amd as ya
from oracle import cpuref

pytestmark = pytest.mark.gpu

Q127 = (1 << 127) - 1
Q255 = (1 << 255) - 19
Q511 = (1 << 511) - 187

_cache = {}


def rand_poly(rng, n, q):
    return [rng.randrange(q) for _ in range(n)]


def _oracle(n, q, t, wb, C1, C2, gammas):
    """(scaled products, their keyswitches) for every pair, one thread per
    pair (the oracle's C code runs without the interpreter lock)."""
    def one(i):
        g = cpuref.yashe_mul(C1[i], C2[i], n, q, t)
        return g, cpuref.keyswitch(g, gammas, n, q, wb)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(C1)) as ex:
        res = list(ex.map(one, range(len(C1))))
    return [r[0] for r in res], [r[1] for r in res]


def _case(n, q):
    key = (n, q)
    if key not in _cache:
        t, wb, batch = 1024, 32, 2
        rng = random.Random(n ^ (q & 0xFFFF))
        lw = cpuref.num_digits(q, wb)
        gammas = [rand_poly(rng, n, q) for _ in range(lw)]
        C1 = [rand_poly(rng, n, q) for _ in range(batch)]
        C2 = [rand_poly(rng, n, q) for _ in range(batch)]
        C2[1] = [q - 1] * n  # the largest operand values
        g, ks = _oracle(n, q, t, wb, C1, C2, gammas)
        _cache[key] = (t, wb, gammas, C1, C2, g, ks)
    return _cache[key]


SHAPES = [(8192, Q255), (16384, Q127), (16384, Q511)]


@pytest.mark.timeout(900)
@pytest.mark.parametrize("n,q", SHAPES)
def test_large_ring_mul_keyswitch(gpu_ctx, n, q):
    t, wb, gammas, C1, C2, want_mul, want_ks = _case(n, q)
    batch = len(C1)
    ring = ya.Ring(gpu_ctx, n, q, t=t, w_bits=wb)
    assert ring.lw == len(gammas)
    evk = ring.evk(gpu_ctx.to_device(ring.pack(gammas)))
    d1, d2 = gpu_ctx.to_device(ring.pack(C1)), gpu_ctx.to_device(ring.pack(C2))
    out = gpu_ctx.alloc(batch * ring.L * n * 8)
    # YASHE multiply
    ring.yashe_mul(d1, d2, out, batch)
    assert ring.unpack(out.download((batch, ring.L, n))) == want_mul
    # keyswitch of the oracle's products
    dg = gpu_ctx.to_device(ring.pack(want_mul))
    ring.keyswitch(evk, dg, out, batch)
    assert ring.unpack(out.download((batch, ring.L, n))) == want_ks
    # fused mul + keyswitch, at the default batch threshold and with the
    # scale-and-round inside the digit transforms from every batch size
    ring.mul_keyswitch(evk, d1, d2, out, batch)
    assert ring.unpack(out.download((batch, ring.L, n))) == want_ks
    try:
        gpu_ctx.set_fuse_min_batch(1)
        ring.mul_keyswitch(evk, d1, d2, out, batch)
        assert ring.unpack(out.download((batch, ring.L, n))) == want_ks
    finally:
        gpu_ctx.set_fuse_min_batch(0)
    evk.close()
    ring.close()


# ------------------------------------------------------------------ plain products above the longest transform

@pytest.mark.parametrize("q,da,db,min_N", [
    (Q127, 16383, 16383, 1),        # N = 32768: both halves, the top of the linear product
    (Q511, 16383, 9000, 1),         # N = 32768, uneven degrees, 511-bit q
    (Q127, 16383, 16383, 65536),    # N = 65536: zero coefficients above 32766
    (Q127, 5000, 6000, 32768),      # degree sum below 16384: the cyclic half alone
    (Q127, 16383, 0, 32768),        # times a constant
])
def test_plain_longer_than_transform(gpu_ctx, q, da, db, min_N):
    # Polynomial::operator* of two degree-16383 operands
    # (src/cuda_functions.cu:617-738 at CUDAFunctions::N = 32768)
    rng = random.Random(da * 7 + db + min_N)
    ring = ya.Ring(gpu_ctx, 16384, q)
    a, b = rand_poly(rng, da + 1, q), rand_poly(rng, db + 1, q)
    a[-1], b[-1] = q - 1, q - 1  # full degree, largest coefficients
    N = cpuref.ntt_length(da, db, min_N)
    assert N > 16384
    A = gpu_ctx.to_device(ya.ints_to_array(a, ring.L))
    B = gpu_ctx.to_device(ya.ints_to_array(b, ring.L))
    out = gpu_ctx.alloc(ring.LW * N * 8)
    assert ring.poly_mul_plain(A, da, B, db, out, ring.LW, min_N) == N
    got = ya.array_to_ints(out.download((ring.LW, N)))
    assert got == cpuref.poly_mul_plain(a, b, N=N, qbits=q.bit_length())
    assert not any(got[da + db + 1:])
    ring.close()


def test_plain_longer_than_transform_small_ring(gpu_ctx):
    # a ring below the longest transform has no negacyclic half of that
    # length: its operands (degree < n) never reach x^16384
    rng = random.Random(5)
    q = Q127
    ring = ya.Ring(gpu_ctx, 4096, q)
    a, b = rand_poly(rng, 4096, q), rand_poly(rng, 4096, q)
    A = gpu_ctx.to_device(ya.ints_to_array(a, ring.L))
    B = gpu_ctx.to_device(ya.ints_to_array(b, ring.L))
    out = gpu_ctx.alloc(ring.LW * 32768 * 8)
    assert ring.poly_mul_plain(A, 4095, B, 4095, out, ring.LW, 32768) == 32768
    got = ya.array_to_ints(out.download((ring.LW, 32768)))
    assert got == cpuref.poly_mul_plain(a, b, N=32768, qbits=q.bit_length())
    with pytest.raises(ya.YashError, match="2\\^20"):
        ring.poly_mul_plain(A, 4095, B, 4095, out, ring.LW, 1 << 21)
    ring.close()
