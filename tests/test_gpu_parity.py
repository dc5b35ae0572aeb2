"""GPU parity: the HIP library (through its C ABI) against the oracle.

Integer work, so every comparison is bit-exact.
"""
import random

import numpy as np
import pytest

import cuyashe_amd as ya
from oracle import cpuref

pytestmark = pytest.mark.gpu

P = cpuref.P
Q61 = (1 << 61) - 1  # one 64-bit limb
Q127 = (1 << 127) - 1
Q255 = (1 << 255) - 19
Q_POLY = 77287149995192912462927307869  # src/test.cu:45


def rand_poly(rng, n, q):
    return [rng.randrange(q) for _ in range(n)]


def upload_polys(ctx, ring, polys):
    return ctx.to_device(ring.pack(polys))


def download_polys(buf, ring, batch):
    return ring.unpack(buf.download((batch, ring.L, ring.n)))


# ------------------------------------------------------------------ NTT primitive

@pytest.mark.parametrize("N", [16, 32, 64, 256, 1024, 4096, 8192, 16384])
def test_ntt_matches_oracle(gpu_ctx, N):
    rng = random.Random(N)
    count = 3
    xs = [[rng.randrange(P) for _ in range(N)] for _ in range(count)]
    arr = np.array(xs, dtype=np.uint64)
    buf = gpu_ctx.to_device(arr)
    gpu_ctx.ntt(buf, N, count)
    got = buf.download((count, N))
    for c in range(count):
        assert [int(v) for v in got[c]] == cpuref.ntt(xs[c])
    gpu_ctx.ntt(buf, N, count, inverse=True)
    back = buf.download((count, N))
    assert np.array_equal(back, arr)


def test_ntt_golden_dft(gpu_ctx, kats):
    for case in kats["ntt_dft"]["cases"]:
        x = np.array([int(v) for v in case["x"]], dtype=np.uint64)
        buf = gpu_ctx.to_device(x)
        gpu_ctx.ntt(buf, case["N"], 1)
        assert [int(v) for v in buf.download(case["N"])] == [int(v) for v in case["X"]]


def test_pointwise_mulmod(gpu_ctx, kats):
    cases = kats["s_mul"]["cases"]
    a = np.array([int(x) % P for x, _, _ in cases], dtype=np.uint64)
    b = np.array([int(y) % P for _, y, _ in cases], dtype=np.uint64)
    da, db = gpu_ctx.to_device(a), gpu_ctx.to_device(b)
    gpu_ctx.pointwise(da, db, len(cases))
    got = da.download(len(cases))
    assert [int(v) for v in got] == [int(x) % P * (int(y) % P) % P for x, y, _ in cases]


# ------------------------------------------------------------------ CRT / ICRT

@pytest.mark.parametrize("q", [Q127, Q255])
def test_crt_icrt(gpu_ctx, q):
    rng = random.Random(q & 0xFFFF)
    ring = ya.Ring(gpu_ctx, 64, q, t=1024, w_bits=32)
    primes = ring.primes()
    M = ring.modulus()
    n = 300
    xs = [rng.randrange(q) for _ in range(n - 3)] + [0, 1, q - 1]
    arr = ya.ints_to_array(xs, ring.L)
    d = gpu_ctx.to_device(arr)
    spacing = 512
    res = gpu_ctx.alloc(ring.R * spacing * 8)
    ring.crt(d, ring.L, n, spacing, res)
    got = res.download((ring.R, spacing))
    want = cpuref.crt(xs, primes)
    for i in range(ring.R):
        assert [int(v) for v in got[i, :n]] == want[i]
        assert not got[i, n:].any()
    out = gpu_ctx.alloc(ring.LW * n * 8)
    ring.icrt(res, spacing, n, out, ring.LW)
    assert ya.array_to_ints(out.download((ring.LW, n))) == xs
    # arbitrary residues: the ICRT closed form in [0, M)
    rr = [[rng.randrange(p) for _ in range(spacing)] for p in primes]
    res.upload(np.array(rr, dtype=np.uint64))
    ring.icrt(res, spacing, spacing, out if spacing <= n else gpu_ctx.alloc(ring.LW * spacing * 8), ring.LW)
    out2 = gpu_ctx.alloc(ring.LW * spacing * 8)
    ring.icrt(res, spacing, spacing, out2, ring.LW)
    got2 = ya.array_to_ints(out2.download((ring.LW, spacing)))
    assert got2 == cpuref.icrt(rr, primes, ring.LW)
    assert all(0 <= v < M for v in got2)
    ring.close()


def test_bnt_crt_icrt_golden(gpu_ctx, kats):
    k = kats["test_bnt_crt"]
    X, q = int(k["X"]), int(k["q"])
    ring = ya.Ring(gpu_ctx, 2, q)
    d = gpu_ctx.to_device(ya.ints_to_array([X], ring.L))
    res = gpu_ctx.alloc(ring.R * 8)
    ring.crt(d, ring.L, 1, 1, res)
    out = gpu_ctx.alloc(ring.LW * 8)
    ring.icrt(res, 1, 1, out, ring.LW)
    assert ya.array_to_ints(out.download((ring.LW, 1))) == [X]
    ring.close()


def test_residue_addsub(gpu_ctx):
    rng = random.Random(11)
    ring = ya.Ring(gpu_ctx, 64, Q127)
    primes = ring.primes()
    S = 64
    a = [[rng.randrange(p) for _ in range(S)] for p in primes]
    b = [[rng.randrange(p) for _ in range(S)] for p in primes]
    da = gpu_ctx.to_device(np.array(a, dtype=np.uint64))
    db = gpu_ctx.to_device(np.array(b, dtype=np.uint64))
    dc = gpu_ctx.alloc(ring.R * S * 8)
    for sub in (False, True):
        ring.residue_addsub(dc, da, db, S, sub)
        got = dc.download((ring.R, S))
        for i, p in enumerate(primes):
            want = [((x - y) if sub else (x + y)) % p for x, y in zip(a[i], b[i])]
            assert [int(v) for v in got[i]] == want
    ring.close()


# ------------------------------------------------------------------ plain product

def _plain(gpu_ctx, ring, a, b, min_N=1):
    da = gpu_ctx.to_device(ya.ints_to_array(a, ring.L))
    db = gpu_ctx.to_device(ya.ints_to_array(b, ring.L))
    N = cpuref.ntt_length(len(a) - 1, len(b) - 1, min_N)
    out = gpu_ctx.alloc(ring.LW * N * 8)
    got_N = ring.poly_mul_plain(da, len(a) - 1, db, len(b) - 1, out, ring.LW, min_N)
    assert got_N == N
    return ya.array_to_ints(out.download((ring.LW, N)))


def test_plain_kat_test_mul(gpu_ctx, kats):
    # src/test_mul.cu: the ring is n = 8 over q (Polynomial::BuildNthCyclotomic(&phi, 8));
    # the first operands have 9 coefficients, so use a ring wide enough to hold them.
    k1, k2 = kats["test_mul_1"], kats["test_mul_2"]
    q = int(k1["q"])
    ring = ya.Ring(gpu_ctx, 16, q)
    got = _plain(gpu_ctx, ring, [int(v) for v in k1["a"]], [int(v) for v in k1["b"]])
    assert got == [int(v) for v in k1["expected"]]
    got = _plain(gpu_ctx, ring, [int(v) for v in k2["a"]], [int(v) for v in k2["b"]])
    assert got[:15] == [int(v) for v in k2["expected"]] and got[15] == 0
    ring.close()


def test_plain_simple_multiply(gpu_ctx, kats):
    k = kats["test_simpleMultiplyByPolynomial"]
    ring = ya.Ring(gpu_ctx, 8, int(k["q"]))
    got = _plain(gpu_ctx, ring, k["a"], k["b"], min_N=16)  # CUDAFunctions::init(2*degree)
    assert got[:15] == k["expected"]
    ring.close()


@pytest.mark.parametrize("n,q", [(8, Q_POLY), (64, Q127), (1024, Q127), (4096, Q127), (2048, Q255), (1024, Q61)])
def test_plain_random(gpu_ctx, n, q):
    rng = random.Random(n)
    ring = ya.Ring(gpu_ctx, n, q)
    a, b = rand_poly(rng, n, q), rand_poly(rng, n, q)
    got = _plain(gpu_ctx, ring, a, b)
    assert got == cpuref.poly_mul_plain(a, b, N=2 * n if n > 1 else 2, qbits=q.bit_length())
    ring.close()


# ------------------------------------------------------------------ ring product / reduce

def test_ring_mul_golden(gpu_ctx, kats):
    k = kats["test_multiplyByPolynomial"]
    q, n = int(k["q"]), k["n"]
    ring = ya.Ring(gpu_ctx, n, q)
    A = [[int(v) for v in c["a"]] for c in k["cases"]]
    B = [[int(v) for v in c["b"]] for c in k["cases"]]
    da, db = upload_polys(gpu_ctx, ring, A), upload_polys(gpu_ctx, ring, B)
    out = gpu_ctx.alloc(len(A) * ring.L * n * 8)
    ring.ring_mul(da, db, out, len(A))
    assert download_polys(out, ring, len(A)) == [[int(v) for v in c["expected"]] for c in k["cases"]]
    # severalMultiplications chain (src/test.cu:823-900)
    k = kats["test_severalMultiplications"]
    a = upload_polys(gpu_ctx, ring, [[int(v) for v in k["a"]]])
    c = upload_polys(gpu_ctx, ring, [[int(v) for v in k["b"]]])
    tmp = gpu_ctx.alloc(ring.L * n * 8)
    ring.ring_mul(c, a, tmp)  # a*b == b*a
    assert download_polys(tmp, ring, 1)[0] == [int(v) for v in k["chain"][0]]
    for want in k["chain"][1:]:
        ring.ring_mul(tmp, a, c)
        tmp, c = c, tmp
        assert download_polys(tmp, ring, 1)[0] == [int(v) for v in want]
    ring.close()


@pytest.mark.parametrize("n,q", [(2, Q_POLY), (4, Q_POLY), (8, Q127), (16, Q127), (32, Q127), (64, Q127),
                                 (1024, Q127), (4096, Q127), (8192, Q255), (4096, (1 << 511) - 187),
                                 (16384, Q127), (256, 655615111), (4096, Q61)])
def test_ring_mul_random(gpu_ctx, n, q):
    rng = random.Random(n * 7 + q % 1000)
    ring = ya.Ring(gpu_ctx, n, q)
    batch = 2
    A = [rand_poly(rng, n, q) for _ in range(batch)]
    B = [rand_poly(rng, n, q) for _ in range(batch)]
    da, db = upload_polys(gpu_ctx, ring, A), upload_polys(gpu_ctx, ring, B)
    out = gpu_ctx.alloc(batch * ring.L * n * 8)
    ring.ring_mul(da, db, out, batch)
    got = download_polys(out, ring, batch)
    for i in range(batch):
        assert got[i] == cpuref.ring_mul(A[i], B[i], n, q), f"pair {i}"
    ring.close()


@pytest.mark.parametrize("n", [8, 4096])
def test_ring_mul_extremes(gpu_ctx, n):
    # all-(q-1) operands reach the largest |coefficient| n (q-1)^2; zeros and ones
    q = Q127
    ring = ya.Ring(gpu_ctx, n, q)
    cases = [([q - 1] * n, [q - 1] * n), ([0] * n, [q - 1] * n), ([1] + [0] * (n - 1), [q - 1] * n),
             ([q - 1] * n, [1] * n)]
    A = [a for a, _ in cases]
    B = [b for _, b in cases]
    da, db = upload_polys(gpu_ctx, ring, A), upload_polys(gpu_ctx, ring, B)
    out = gpu_ctx.alloc(len(cases) * ring.L * n * 8)
    ring.ring_mul(da, db, out, len(cases))
    got = download_polys(out, ring, len(cases))
    for i, (a, b) in enumerate(cases):
        assert got[i] == cpuref.ring_mul(a, b, n, q), f"case {i}"
    ring.close()


def test_ring_reduce_golden(gpu_ctx, kats):
    k = kats["test_phiReduce"]
    q, n = int(k["q"]), k["n"]
    ring = ya.Ring(gpu_ctx, n, q)
    C = [[int(v) for v in c["c"]] + [0] for c in k["cases"]]  # 2n coefficients
    arr = np.stack([ya.ints_to_array(c, ring.L) for c in C])
    d = gpu_ctx.to_device(arr)
    out = gpu_ctx.alloc(len(C) * ring.L * n * 8)
    ring.ring_reduce(d, ring.L, out, len(C))
    assert download_polys(out, ring, len(C)) == [[int(v) for v in c["expected"]] for c in k["cases"]]
    ring.close()
    k = kats["test_reduce_fold"]
    ring = ya.Ring(gpu_ctx, k["n"], int(k["q"]))
    for inp, exp in zip(k["inputs"], k["expected"]):
        c = inp + [0] * (2 * k["n"] - len(inp))
        d = gpu_ctx.to_device(ya.ints_to_array(c, ring.L)[None])
        out = gpu_ctx.alloc(ring.L * k["n"] * 8)
        ring.ring_reduce(d, ring.L, out, 1)
        assert download_polys(out, ring, 1)[0] == exp
    ring.close()
    k = kats["test_reduce_modn"]
    ring = ya.Ring(gpu_ctx, 2, int(k["q"]))
    c = [int(k["value"]), 0, 0, 0]
    L_in = 2
    d = gpu_ctx.to_device(ya.ints_to_array(c, L_in)[None])
    out = gpu_ctx.alloc(ring.L * 2 * 8)
    ring.ring_reduce(d, L_in, out, 1)
    assert download_polys(out, ring, 1)[0][0] == int(k["expected"])
    ring.close()


def test_ring_addsub(gpu_ctx):
    rng = random.Random(12)
    for q in (Q127, Q255):
        n = 256
        ring = ya.Ring(gpu_ctx, n, q)
        A = [rand_poly(rng, n, q) for _ in range(3)] + [[q - 1] * n]
        B = [rand_poly(rng, n, q) for _ in range(3)] + [[q - 1] * n]
        da, db = upload_polys(gpu_ctx, ring, A), upload_polys(gpu_ctx, ring, B)
        out = gpu_ctx.alloc(4 * ring.L * n * 8)
        for sub in (False, True):
            ring.ring_addsub(da, db, out, 4, sub)
            got = download_polys(out, ring, 4)
            for i in range(4):
                assert got[i] == [((x - y) if sub else (x + y)) % q for x, y in zip(A[i], B[i])]
        ring.close()


# ------------------------------------------------------------------ YASHE operations

@pytest.mark.parametrize("n,q,t", [(4, Q_POLY, 35951), (64, Q127, 1024), (4096, Q127, 1024), (1024, Q255, 1 << 16),
                                   (64, Q61, 17), (4096, Q61, 1024)])
def test_yashe_mul(gpu_ctx, n, q, t):
    rng = random.Random(n + t)
    ring = ya.Ring(gpu_ctx, n, q, t=t, w_bits=32)
    batch = 2
    C1 = [rand_poly(rng, n, q) for _ in range(batch)]
    C2 = [rand_poly(rng, n, q) for _ in range(batch)]
    d1, d2 = upload_polys(gpu_ctx, ring, C1), upload_polys(gpu_ctx, ring, C2)
    out = gpu_ctx.alloc(batch * ring.L * n * 8)
    ring.yashe_mul(d1, d2, out, batch)
    got = download_polys(out, ring, batch)
    for i in range(batch):
        assert got[i] == cpuref.yashe_mul(C1[i], C2[i], n, q, t)
    ring.close()


@pytest.mark.parametrize("n,q,wb", [(4, Q_POLY, 32), (64, Q127, 32), (4096, Q127, 32), (256, Q255, 16),
                                    (1024, Q127, 24), (4096, Q61, 32), (256, Q61, 16)])
def test_keyswitch(gpu_ctx, n, q, wb):
    rng = random.Random(n * 3 + wb)
    ring = ya.Ring(gpu_ctx, n, q, t=1024, w_bits=wb)
    lw = ring.lw
    assert lw == cpuref.num_digits(q, wb)
    gammas = [rand_poly(rng, n, q) for _ in range(lw)]
    dg = gpu_ctx.to_device(ring.pack(gammas))
    evk = ring.evk(dg)
    batch = 2
    C = [rand_poly(rng, n, q) for _ in range(batch)]
    dc = upload_polys(gpu_ctx, ring, C)
    out = gpu_ctx.alloc(batch * ring.L * n * 8)
    ring.keyswitch(evk, dc, out, batch)
    got = download_polys(out, ring, batch)
    for i in range(batch):
        assert got[i] == cpuref.keyswitch(C[i], gammas, n, q, wb)
    evk.close()
    ring.close()


def _crt_max_residue(primes, q):
    """An integer below q whose residue mod each of the leading primes is
    (p-1)/2, the largest centred residue (as many primes as fit below q)."""
    x, m = 0, 1
    for p in primes:
        if m * p >= q:
            break
        # x' = x (mod m), x' = (p-1)/2 (mod p)
        t = ((p - 1) // 2 - x) * pow(m, -1, p) % p
        x, m = x + m * t, m * p
    return x


@pytest.mark.parametrize("n,wb", [(16, 32), (4096, 32), (1024, 16)])
def test_keyswitch_extremes(gpu_ctx, n, wb):
    # The digit transforms are shared by every keyswitch modulus, so the
    # multiply-accumulate of centred digits (d - 2^(wb-1)) must stay inside
    # the Goldilocks bound: all-(q-1) ciphertexts give digits near 2^wb - 1
    # (centred: near +2^(wb-1)), all-zero ones digits 0 (centred: -2^(wb-1)),
    # and keys whose residues are the largest centred value make every term of
    # the last coefficient's negacyclic sum the same sign.
    q = Q127
    rng = random.Random(n + wb)
    ring = ya.Ring(gpu_ctx, n, q, t=1024, w_bits=wb)
    lw = ring.lw
    G = _crt_max_residue(ring.ks_primes()[1:], q)
    cases = [([q - 1] * n, [[G] * n] * lw),
             ([0] * n, [[G] * n] * lw),
             ([q - 1] * n, [[q - 1] * n] * lw),
             ([q - 1] * n, [rand_poly(rng, n, q) for _ in range(lw)]),
             ([0] * (n - 1) + [q - 1], [[G] * n] * lw)]
    for ci, (c, gammas) in enumerate(cases):
        evk = ring.evk(gpu_ctx.to_device(ring.pack(gammas)))
        dc = upload_polys(gpu_ctx, ring, [c])
        out = gpu_ctx.alloc(ring.L * n * 8)
        ring.keyswitch(evk, dc, out, 1)
        assert download_polys(out, ring, 1)[0] == cpuref.keyswitch(c, gammas, n, q, wb), f"case {ci}"
        evk.close()
    ring.close()


def test_mul_keyswitch_fused_equals_steps(gpu_ctx):
    rng = random.Random(13)
    n, q, t = 4096, Q127, 1024
    ring = ya.Ring(gpu_ctx, n, q, t=t, w_bits=32)
    gammas = [rand_poly(rng, n, q) for _ in range(ring.lw)]
    evk = ring.evk(gpu_ctx.to_device(ring.pack(gammas)))
    batch = 3
    C1 = [rand_poly(rng, n, q) for _ in range(batch)]
    C2 = [rand_poly(rng, n, q) for _ in range(batch)]
    d1, d2 = upload_polys(gpu_ctx, ring, C1), upload_polys(gpu_ctx, ring, C2)
    fused = gpu_ctx.alloc(batch * ring.L * n * 8)
    ring.mul_keyswitch(evk, d1, d2, fused, batch)
    mid = gpu_ctx.alloc(batch * ring.L * n * 8)
    steps = gpu_ctx.alloc(batch * ring.L * n * 8)
    ring.yashe_mul(d1, d2, mid, batch)
    ring.keyswitch(evk, mid, steps, batch)
    got = download_polys(fused, ring, batch)
    assert got == download_polys(steps, ring, batch)
    # one op against the oracle end to end; alone == in the batch
    want0 = cpuref.keyswitch(cpuref.yashe_mul(C1[0], C2[0], n, q, t), gammas, n, q, 32)
    assert got[0] == want0
    single = gpu_ctx.alloc(ring.L * n * 8)
    ring.mul_keyswitch(evk, d1, d2, single, 1)
    assert download_polys(single, ring, 1)[0] == want0
    evk.close()
    ring.close()


# Rings whose scale-and-round can run inside the digit transforms
# (k_ks_scale_digits: n >= 1024, q of one or two limbs, w = 2^32; one limb
# of 33 and 61 bits, two limbs of 65 bits (three digits, the third of one
# bit), 90, 97 and 100 bits, and full 128 bits, and 127 bits at n = 2048 and
# 4096) and rings that always take the separate reconstruction and
# digit kernels (n < 1024, w = 2^16, a four-limb q): mul+keyswitch equals the
# multiply followed by the keyswitch, and one operation equals the oracle, on
# random and all-(q - 1) operands, fused from every batch size and never.
@pytest.mark.parametrize("n,q,wb", [
    (16, Q127, 32), (256, (1 << 100) - 15, 32), (1024, Q61, 32), (1024, (1 << 100) - 15, 32),
    (1024, (1 << 90) - 33, 32), (1024, Q_POLY, 32), (2048, Q127, 32), (4096, Q127, 32),
    (1024, (1 << 33) - 9, 32), (1024, (1 << 65) - 49, 32), (1024, (1 << 128) - 159, 32),
    (1024, Q127, 16), (512, Q255, 32)])
def test_mul_keyswitch_fused_shapes(gpu_ctx, n, q, wb):
    rng = random.Random(n + wb)
    t = 1024
    ring = ya.Ring(gpu_ctx, n, q, t=t, w_bits=wb)
    gammas = [rand_poly(rng, n, q) for _ in range(ring.lw)]
    evk = ring.evk(gpu_ctx.to_device(ring.pack(gammas)))
    C1 = [rand_poly(rng, n, q), [q - 1] * n, rand_poly(rng, n, q)]
    C2 = [rand_poly(rng, n, q), [q - 1] * n, [0] * n]
    batch = len(C1)
    d1, d2 = upload_polys(gpu_ctx, ring, C1), upload_polys(gpu_ctx, ring, C2)
    mid = gpu_ctx.alloc(batch * ring.L * n * 8)
    steps = gpu_ctx.alloc(batch * ring.L * n * 8)
    ring.yashe_mul(d1, d2, mid, batch)
    ring.keyswitch(evk, mid, steps, batch)
    want_steps = download_polys(steps, ring, batch)
    wants = [cpuref.keyswitch(cpuref.yashe_mul(C1[i], C2[i], n, q, t), gammas, n, q, wb) for i in (0, 1)]
    out = gpu_ctx.alloc(batch * ring.L * n * 8)
    try:
        # fused from every batch size, then never fused
        for min_batch in (1, 1 << 30):
            gpu_ctx.set_fuse_min_batch(min_batch)
            ring.mul_keyswitch(evk, d1, d2, out, batch)
            got = download_polys(out, ring, batch)
            assert got == want_steps, f"min_batch {min_batch}"
            for i in (0, 1):
                assert got[i] == wants[i], f"operation {i}, min_batch {min_batch}"
            assert got[2] == [0] * n
    finally:
        gpu_ctx.set_fuse_min_batch(0)
    evk.close()
    ring.close()


def test_mul_keyswitch_list_equals_arrays(gpu_ctx):
    # yh_yashe_mul_keyswitch_list (the serving loop's entry: per-element
    # addresses, out of order) equals the array form, with the scale-and-round
    # fused into the digit transforms and with separate kernels
    rng = random.Random(31)
    n, q, t, batch = 1024, Q127, 1024, 5
    ring = ya.Ring(gpu_ctx, n, q, t=t, w_bits=32)
    gammas = [rand_poly(rng, n, q) for _ in range(ring.lw)]
    evk = ring.evk(gpu_ctx.to_device(ring.pack(gammas)))
    C1 = [rand_poly(rng, n, q) for _ in range(batch)]
    C2 = [rand_poly(rng, n, q) for _ in range(batch)]
    d1, d2 = upload_polys(gpu_ctx, ring, C1), upload_polys(gpu_ctx, ring, C2)
    ref = gpu_ctx.alloc(batch * ring.L * n * 8)
    ring.mul_keyswitch(evk, d1, d2, ref, batch)
    want = download_polys(ref, ring, batch)
    elt = ring.L * n * 8
    perm = [3, 0, 4, 1, 2]  # element k of the lists is pair perm[k], written to slot k
    out = gpu_ctx.alloc(batch * elt)
    l1 = gpu_ctx.to_device(np.array([d1.ptr + p * elt for p in perm], dtype=np.uint64))
    l2 = gpu_ctx.to_device(np.array([d2.ptr + p * elt for p in perm], dtype=np.uint64))
    lo = gpu_ctx.to_device(np.array([out.ptr + k * elt for k in range(batch)], dtype=np.uint64))
    try:
        for min_batch in (1, 1 << 30):
            gpu_ctx.set_fuse_min_batch(min_batch)
            ya.lib().yh_memset(gpu_ctx.h, out.ptr, 0, out.nbytes)
            ring.mul_keyswitch_list(evk, l1, l2, lo, batch)
            got = download_polys(out, ring, batch)
            assert got == [want[p] for p in perm], min_batch
    finally:
        gpu_ctx.set_fuse_min_batch(0)
    evk.close()
    ring.close()


def test_fuse_min_batch_rejects_negative(gpu_ctx):
    with pytest.raises(ya.YashError, match="bad argument"):
        gpu_ctx.set_fuse_min_batch(-1)
    gpu_ctx.set_fuse_min_batch(0)


def test_decrypt_matches_oracle(gpu_ctx):
    rng = random.Random(14)
    n, q, t = 1024, Q127, 1024
    ring = ya.Ring(gpu_ctx, n, q, t=t, w_bits=32)
    f = rand_poly(rng, n, q)
    C = [rand_poly(rng, n, q) for _ in range(2)]
    df = upload_polys(gpu_ctx, ring, [f])
    dc = upload_polys(gpu_ctx, ring, C)
    dm = gpu_ctx.alloc(2 * n * 8)
    ring.decrypt(df, dc, dm, 2)
    got = dm.download((2, n))
    for i in range(2):
        assert [int(v) for v in got[i]] == cpuref.decrypt(f, C[i], n, q, t)
    ring.close()


@pytest.mark.parametrize("da,db", [(0, 0), (8, 0), (0, 7), (16, 0), (7, 1), (31, 1), (1, 1)])
def test_plain_edge_degrees(gpu_ctx, da, db):
    # constants and power-of-two degrees: the output length rule of the library
    # (yh_poly_mul_plain_length) sizes every buffer
    rng = random.Random(da * 100 + db)
    q = Q127
    ring = ya.Ring(gpu_ctx, 64, q)
    a = rand_poly(rng, da + 1, q)
    b = rand_poly(rng, db + 1, q)
    got = _plain(gpu_ctx, ring, a, b)
    N = cpuref.ntt_length(da, db)
    assert got == cpuref.poly_mul_plain(a, b, N=N, qbits=q.bit_length())
    ring.close()
