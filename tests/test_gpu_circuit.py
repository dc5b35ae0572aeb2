"""Key generation on the device, the scheme layer, and a homomorphic circuit
captured into a HIP graph and replayed.

Bit-exact against the oracle where it is affordable; everything else through
decryption (Dec(Enc(m1) Enc(m2)) = m1 m2 mod t) and replay == eager.
"""
import random

import numpy as np
import pytest

import cuyashe_amd as ya
from cuyashe_amd import workloads
from cuyashe_amd.scheme import Yashe, negacyclic_mul_mod
from oracle import cpuref

pytestmark = pytest.mark.gpu

Q127 = (1 << 127) - 1
Q255 = (1 << 255) - 19
Q_POLY = 77287149995192912462927307869  # src/test.cu:45


def rand_poly(rng, n, q):
    return [rng.randrange(q) for _ in range(n)]


def put(ctx, ring, polys):
    return ctx.to_device(ring.pack(polys))


def get(buf, ring, batch):
    return ring.unpack(buf.download((batch, ring.L, ring.n)))


@pytest.mark.parametrize("n,q", [(2, Q_POLY), (8, Q_POLY), (64, Q127), (256, Q255)])
def test_ring_inverse_matches_oracle(gpu_ctx, n, q):
    rng = random.Random(n)
    ring = ya.Ring(gpu_ctx, n, q)
    f = rand_poly(rng, n, q)
    out = gpu_ctx.alloc(ring.L * n * 8)
    ring.inverse(put(gpu_ctx, ring, [f]), out)
    assert get(out, ring, 1)[0] == cpuref.poly_inv(f, n, q)
    ring.close()


def test_ring_inverse_large_is_an_inverse(gpu_ctx):
    # f = t f' + 1 as in key generation, at the largest ring of the circuits
    rng = np.random.default_rng(3)
    n, q, t = 16384, (1 << 511) - 187, 1 << 16
    ring = ya.Ring(gpu_ctx, n, q)
    fp = rng.integers(-1, 2, size=n)
    f = [(t * int(x) + (j == 0)) % q for j, x in enumerate(fp)]
    df = put(gpu_ctx, ring, [f])
    inv = gpu_ctx.alloc(ring.L * n * 8)
    ring.inverse(df, inv)
    one = gpu_ctx.alloc(ring.L * n * 8)
    ring.ring_mul(df, inv, one, 1)
    assert get(one, ring, 1)[0] == [1] + [0] * (n - 1)
    ring.close()


def test_ring_inverse_rejects_zero(gpu_ctx):
    ring = ya.Ring(gpu_ctx, 16, Q127)
    out = gpu_ctx.alloc(ring.L * 16 * 8)
    with pytest.raises(ya.YashError):
        ring.inverse(put(gpu_ctx, ring, [[0] * 16]), out)
    ring.close()


def test_ring_mul_bcast_matches_oracle(gpu_ctx):
    rng = random.Random(5)
    n, q, batch = 256, Q255, 3
    ring = ya.Ring(gpu_ctx, n, q)
    a = rand_poly(rng, n, q)
    B = [rand_poly(rng, n, q) for _ in range(batch)]
    out = gpu_ctx.alloc(batch * ring.L * n * 8)
    ring.ring_mul_bcast(put(gpu_ctx, ring, [a]), put(gpu_ctx, ring, B), out, batch)
    got = get(out, ring, batch)
    for i in range(batch):
        assert got[i] == cpuref.ring_mul(a, B[i], n, q)
    ring.close()


@pytest.mark.parametrize("n,q,t", [(1024, Q127, 1024), (4096, Q127, 1024), (2048, Q255, 1 << 16)])
def test_scheme_roundtrip_add_mul(gpu_ctx, n, q, t):
    Y = Yashe(gpu_ctx, n, q, t, 32, seed=n).generate_keys()
    rng = np.random.default_rng(n + 1)
    m1 = rng.integers(0, t, size=(4, n))
    m2 = rng.integers(0, t, size=(4, n))
    c1, c2 = Y.encrypt(m1), Y.encrypt(m2)
    assert np.array_equal(Y.decrypt(c1, 4), m1)
    s = gpu_ctx.alloc(4 * Y.elt_bytes)
    Y.add(c1, c2, s, 4)
    assert np.array_equal(Y.decrypt(s, 4), (m1 + m2) % t)
    # mul without keyswitch decrypts under f^2 (src/yashe.cpp:167-197)
    p = gpu_ctx.alloc(4 * Y.elt_bytes)
    Y.ring.yashe_mul(c1, c2, p, 4)
    k = gpu_ctx.alloc(4 * Y.elt_bytes)
    Y.mul_keyswitch(c1, c2, k, 4)
    want = np.stack([negacyclic_mul_mod(m1[i], m2[i], t) for i in range(4)])
    assert np.array_equal(Y.decrypt(p, 4, aftermul=True), want)
    assert np.array_equal(Y.decrypt(k, 4), want)
    Y.ring.close()


_circuit = workloads.depth3_circuit
_plain_circuit = workloads.depth3_plain


def test_graph_circuit_depth3(gpu_ctx):
    """128 ciphertexts, n = 8192, q = 2^255 - 19, t = 2^16, w = 2^32: the
    depth-3 multiplication tree with 256 additions, captured once and
    replayed; every replay equals the eager run and decrypts to the plain
    circuit over Z_t[x]/(x^n + 1)."""
    n, q, t = 8192, Q255, 1 << 16
    Y = Yashe(gpu_ctx, n, q, t, 32, seed=41).generate_keys()
    rng = np.random.default_rng(42)
    x = rng.integers(0, t, size=(128, n))
    X = Y.encrypt(x)
    e = Y.elt_bytes
    B = {k: gpu_ctx.alloc(v * e) for k, v in workloads.CIRCUIT_BUFFERS.items()}
    _circuit(Y, X, **B)
    eager = B["Z"].download((16, Y.L, n)).copy()
    p0 = get(B["Pm"], Y.ring, 1)[0]
    assert np.array_equal(Y.decrypt(B["Z"], 16), _plain_circuit(x, t))
    # one node of every level bit-exact against the oracle (mul + keyswitch
    # of the level's GPU inputs), and the additions feeding them
    xs = get(X, Y.ring, 2)
    a = get(B["A"], Y.ring, 128)
    assert a[0] == [(u + v) % q for u, v in zip(xs[0], xs[1])]
    gam = get(Y.gamma, Y.ring, Y.lw)
    assert p0 == cpuref.keyswitch(cpuref.yashe_mul(a[0], a[64], n, q, t), gam, n, q, 32)
    b = get(B["B"], Y.ring, 64)
    assert b[0] == [(u + v) % q for u, v in zip(p0, xs[0])]
    r0 = get(B["Rm"], Y.ring, 1)[0]
    assert r0 == cpuref.keyswitch(cpuref.yashe_mul(b[0], b[32], n, q, t), gam, n, q, 32)
    s_ = get(B["S"], Y.ring, 32)
    u0 = get(B["U"], Y.ring, 1)[0]
    assert u0 == cpuref.keyswitch(cpuref.yashe_mul(s_[0], s_[16], n, q, t), gam, n, q, 32)

    gpu_ctx.graph_begin()
    _circuit(Y, X, **B)
    g = gpu_ctx.graph_end()
    other = gpu_ctx.alloc(e)
    for rep in range(3):
        for buf in B.values():
            buf_bytes = buf.nbytes
            ya.lib().yh_memset(gpu_ctx.h, buf.ptr, 0xA5, buf_bytes)
        gpu_ctx.sync()
        g.launch()
        gpu_ctx.sync()
        assert np.array_equal(B["Z"].download((16, Y.L, n)), eager), rep
        # eager work between replays
        Y.ring.ring_mul(X.view(rep * e, e), X.view((rep + 1) * e, e), other, 1)
        gpu_ctx.sync()
    g.close()
    Y.ring.close()


def test_graph_capture_refuses_allocation(gpu_ctx):
    # a capture that would have to grow scratch fails instead of allocating
    ring = ya.Ring(gpu_ctx, 64, Q127)
    a = put(gpu_ctx, ring, [[1] * 64] * 4)
    out = gpu_ctx.alloc(4 * ring.L * 64 * 8)
    gpu_ctx.graph_begin()
    with pytest.raises(ya.YashError, match="capture"):
        ring.ring_mul(a, a, out, 4)
    with pytest.raises(ya.YashError, match="capture"):
        gpu_ctx.sync()
    gpu_ctx.graph_end().close()  # nothing was recorded
    # the context works eagerly afterwards
    ring.ring_mul(a, a, out, 4)
    want = cpuref.ring_mul([1] * 64, [1] * 64, 64, Q127)
    assert get(out, ring, 4) == [want] * 4
    ring.close()
