"""GPU parity of the scalar operations (yh_ring_scalar, yh_residue_scalar)
against the oracle, and the lifetime rules of captured graphs and batch
plans (a ring, key or plan that a graph still references can be neither
reallocated nor destroyed)."""
import random

import numpy as np
import pytest

import cuyashe_amd as ya
from oracle import cpuref

pytestmark = pytest.mark.gpu

Q61 = (1 << 61) - 1
Q127 = (1 << 127) - 1
Q255 = (1 << 255) - 19
Q511 = (1 << 511) - 187
Q_POLY = 77287149995192912462927307869  # src/test.cu:45


def rand_poly(rng, n, q):
    return [rng.randrange(q) for _ in range(n)]


def put(ctx, ring, polys):
    return ctx.to_device(ring.pack(polys))


def get(buf, ring, batch):
    return ring.unpack(buf.download((batch, ring.L, ring.n)))


# ------------------------------------------------------------------ scalar operations

@pytest.mark.parametrize("n,q", [(1024, Q127), (4096, Q127), (8, Q_POLY), (256, Q61), (512, Q255), (64, Q511)])
def test_ring_scalar_matches_oracle(gpu_ctx, n, q):
    rng = random.Random(n * 7 + q % 1000)
    ring = ya.Ring(gpu_ctx, n, q)
    batch = 3
    A = [rand_poly(rng, n, q) for _ in range(batch)]
    A[1] = [q - 1] * n  # every coefficient at the top of the range
    da = put(gpu_ctx, ring, A)
    out = gpu_ctx.alloc(batch * ring.L * n * 8)
    scalars = [0, 1, 2, q - 1, q, q + 5, rng.randrange(q), rng.randrange(1 << 700)]  # wider than q too
    for op in (ya.YH_ADD, ya.YH_SUB, ya.YH_MUL):
        for s in scalars:
            ring.scalar(da, op, s, out, batch)
            got = get(out, ring, batch)
            for i in range(batch):
                assert got[i] == cpuref.ring_scalar(A[i], op, s, q), (op, s, i)
    # in place (out aliases the input)
    s = rng.randrange(q)
    ring.scalar(da, ya.YH_MUL, s, da, batch)
    assert get(da, ring, batch) == [cpuref.ring_scalar(a, cpuref.MUL, s, q) for a in A]
    ring.close()


@pytest.mark.parametrize("n,q,spacing", [(1024, Q127, 2048), (8, Q_POLY, 16), (4096, Q255, 4096)])
def test_residue_scalar_matches_oracle(gpu_ctx, n, q, spacing):
    rng = random.Random(spacing + n)
    ring = ya.Ring(gpu_ctx, n, q)
    primes = ring.primes()
    coefs = rand_poly(rng, n, q)
    res = cpuref.crt(coefs, primes)
    res = [row + [0] * (spacing - n) for row in res]
    d_res = gpu_ctx.to_device(np.array(res, dtype=np.uint64))
    out = gpu_ctx.alloc(len(primes) * spacing * 8)
    mask = (1 << 64) - 1
    for op in (ya.YH_ADD, ya.YH_SUB, ya.YH_MUL):
        for s in (0, 3, rng.randrange(q), rng.randrange(1 << 300)):
            # the scalar as Integer's residues, and as unreduced words (the
            # kernel takes each word mod its prime, as the reference's % does)
            for words in ([s % p for p in primes], [(s % p + p * rng.randrange(1 << 30)) & mask if i else s & mask
                                                   for i, p in enumerate(primes)]):
                d_s = gpu_ctx.to_device(np.array(words, dtype=np.uint64))
                ring.residue_scalar(out, d_res, spacing, op, d_s)
                got = out.download((len(primes), spacing))
                want = cpuref.residue_scalar(res, primes, op, [w % p for w, p in zip(words, primes)])
                assert [[int(v) for v in row] for row in got] == want, (op, s)
    ring.close()


BN_T = np.dtype([("alloc", np.int32), ("used", np.int32), ("sign", np.int32), ("pad", np.int32),
                 ("dp", np.uint64)])


def _bn_array(ctx, values, alloc):
    """Device coefficients in the reference's bn_t layout: (headers, words)."""
    n = len(values)
    words = np.zeros((n, alloc), dtype=np.uint64)
    for j, v in enumerate(values):
        for l in range(alloc):
            words[j, l] = (v >> (64 * l)) & ((1 << 64) - 1)
    d_words = ctx.to_device(words)
    hdr = np.zeros(n, dtype=BN_T)
    hdr["alloc"] = alloc
    hdr["used"] = [max(1, (int(v).bit_length() + 63) // 64) if v else 0 for v in values]
    hdr["dp"] = [d_words.ptr + 8 * alloc * j for j in range(n)]
    return ctx.to_device(hdr), d_words


@pytest.mark.parametrize("n,q,ncoef", [(16, Q_POLY, 16), (1024, Q127, 1000), (256, Q511, 256)])
def test_crt_icrt_bn_layout(gpu_ctx, n, q, ncoef):
    # callCRT / callICRT over the reference's bn_t coefficients
    # (src/cuda_bn.cu:970-1045): residues equal the oracle's CRT, and the
    # ICRT writes the coefficients back with their used word counts
    rng = random.Random(n + ncoef)
    ring = ya.Ring(gpu_ctx, n, q)
    primes = ring.primes()
    M = ring.modulus()
    alloc = 32  # STD_BNT_WORDS_ALLOC
    vals = [rng.randrange(M) for _ in range(ncoef)]
    vals[0], vals[1] = 0, M - 1
    hdr, words = _bn_array(gpu_ctx, vals, alloc)
    res = gpu_ctx.alloc(len(primes) * n * 8)
    ring.crt_bn(hdr, ncoef, n, res)
    got = res.download((len(primes), n))
    want = cpuref.crt(vals, primes)
    for i in range(len(primes)):
        assert [int(v) for v in got[i][:ncoef]] == want[i] and not got[i][ncoef:].any(), i
    out_hdr, out_words = _bn_array(gpu_ctx, [rng.randrange(1 << 64) for _ in range(ncoef)], alloc)
    ring.icrt_bn(res, n, ncoef, out_hdr)
    h = out_hdr.download((ncoef,), dtype=BN_T)
    w = out_words.download((ncoef, alloc))
    back = [sum(int(w[j, l]) << (64 * l) for l in range(int(h["used"][j]))) for j in range(ncoef)]
    assert back == vals
    assert all(int(h["used"][j]) == (vals[j].bit_length() + 63) // 64 for j in range(ncoef))
    assert not h["sign"].any()
    ring.close()


def test_scalar_in_graph_replays(gpu_ctx):
    # a scalar operation captured into a graph keeps its scalar (passed by
    # value) and replays to the eager result
    rng = random.Random(3)
    n, q = 4096, Q127
    ring = ya.Ring(gpu_ctx, n, q)
    A = [rand_poly(rng, n, q) for _ in range(2)]
    da = put(gpu_ctx, ring, A)
    out = gpu_ctx.alloc(2 * ring.L * n * 8)
    s = rng.randrange(q)
    gpu_ctx.graph_begin()
    ring.scalar(da, ya.YH_MUL, s, out, 2)
    g = gpu_ctx.graph_end()
    want = [cpuref.ring_scalar(a, cpuref.MUL, s, q) for a in A]
    for rep in range(3):
        ring.scalar(da, ya.YH_ADD, rep + 1, out, 2)  # eager work in between overwrites out
        gpu_ctx.sync()
        g.launch()
        gpu_ctx.sync()
        assert get(out, ring, 2) == want, rep
    g.close()
    ring.close()


def test_graph_replays_fused_mul_keyswitch(gpu_ctx):
    # the scale-and-round inside the digit transforms (k_ks_scale_digits),
    # captured and replayed: every replay equals the eager run and the oracle
    rng = random.Random(22)
    n, q, t = 1024, Q127, 1024
    ring = ya.Ring(gpu_ctx, n, q, t=t, w_bits=32)
    gam = [rand_poly(rng, n, q) for _ in range(ring.lw)]
    evk = ring.evk(put(gpu_ctx, ring, gam))
    C1 = [rand_poly(rng, n, q) for _ in range(3)]
    C2 = [rand_poly(rng, n, q) for _ in range(3)]
    d1, d2 = put(gpu_ctx, ring, C1), put(gpu_ctx, ring, C2)
    out = gpu_ctx.alloc(3 * ring.L * n * 8)
    try:
        gpu_ctx.set_fuse_min_batch(1)
        ring.mul_keyswitch(evk, d1, d2, out, 3)
        eager = get(out, ring, 3)
        assert eager[1] == cpuref.keyswitch(cpuref.yashe_mul(C1[1], C2[1], n, q, t), gam, n, q, 32)
        gpu_ctx.graph_begin()
        ring.mul_keyswitch(evk, d1, d2, out, 3)
        g = gpu_ctx.graph_end()
        for rep in range(3):
            ya.lib().yh_memset(gpu_ctx.h, out.ptr, 0, out.nbytes)
            gpu_ctx.sync()
            g.launch()
            gpu_ctx.sync()
            assert get(out, ring, 3) == eager, rep
        g.close()
    finally:
        gpu_ctx.set_fuse_min_batch(0)
    evk.close()
    ring.close()


# ------------------------------------------------------------------ lifetimes

def test_graph_holds_ring_scratch_and_key(gpu_ctx):
    rng = random.Random(21)
    n, q, t = 1024, Q127, 1024
    ring = ya.Ring(gpu_ctx, n, q, t=t, w_bits=32)
    gam = [rand_poly(rng, n, q) for _ in range(ring.lw)]
    evk = ring.evk(put(gpu_ctx, ring, gam))
    C1 = [rand_poly(rng, n, q) for _ in range(4)]
    C2 = [rand_poly(rng, n, q) for _ in range(4)]
    d1, d2 = put(gpu_ctx, ring, C1), put(gpu_ctx, ring, C2)
    out = gpu_ctx.alloc(4 * ring.L * n * 8)
    ring.mul_keyswitch(evk, d1, d2, out, 1)  # sizes scratch for one pair
    eager = get(out, ring, 1)
    gpu_ctx.graph_begin()
    ring.mul_keyswitch(evk, d1, d2, out, 1)
    g = gpu_ctx.graph_end()
    # a larger eager batch would reallocate the scratch the graph uses
    with pytest.raises(ya.YashError, match="captured graph"):
        ring.mul_keyswitch(evk, d1, d2, out, 4)
    with pytest.raises(ya.YashError, match="referenced"):
        ring.close()
    with pytest.raises(ya.YashError, match="referenced"):
        evk.close()
    for rep in range(2):
        ya.lib().yh_memset(gpu_ctx.h, out.ptr, 0, out.nbytes)
        gpu_ctx.sync()
        g.launch()
        gpu_ctx.sync()
        assert get(out, ring, 1) == eager, rep
    g.close()
    # released: the ring grows its scratch again and the results are right
    ring.mul_keyswitch(evk, d1, d2, out, 4)
    got = get(out, ring, 4)
    assert got[0] == eager[0]
    assert got[3] == cpuref.keyswitch(cpuref.yashe_mul(C1[3], C2[3], n, q, t), gam, n, q, 32)
    evk.close()
    ring.close()


def test_batch_plan_holds_rings(gpu_ctx):
    rng = np.random.default_rng(5)
    ring = ya.Ring(gpu_ctx, 1024, Q127)
    ha = np.stack([ring.pack([[int(v) for v in rng.integers(0, 1 << 62, size=1024)]])[0] for _ in range(2)])
    A = gpu_ctx.to_device(ha)
    OUT = gpu_ctx.alloc(ha.nbytes)
    e = ring.L * 1024 * 8
    items = [(ring, A.view(k * e, e), A.view(k * e, e), OUT.view(k * e, e)) for k in range(2)]
    b = ya.Batch(gpu_ctx, items)
    with pytest.raises(ya.YashError, match="referenced"):
        ring.close()
    gpu_ctx.graph_begin()
    b.run()
    g = gpu_ctx.graph_end()
    with pytest.raises(ya.YashError, match="captured graph"):
        b.close()
    g.launch()
    gpu_ctx.sync()
    want = ring.unpack(ha)
    got = ring.unpack(OUT.download(ha.shape))
    for k in range(2):
        assert got[k] == cpuref.ring_mul(want[k], want[k], 1024, Q127)
    g.close()
    b.close()
    ring.close()
