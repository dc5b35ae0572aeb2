"""Synthetic workloads of the benchmark configurations (seeded, host-side).

* uneven products: n from {1024, 2048, 4096, 8192} with weights
  0.4/0.3/0.2/0.1, q = 2^127 - 1 or 2^255 - 19 (50/50), operands uniform
  mod q;
* operands are generated as limb arrays ([count][L][n] uint64), the layout
  of every ring element on the device.
"""
from __future__ import annotations

import random
from typing import Dict, List, Sequence, Tuple

import numpy as np

from ._lib import Context, Ring

Q127 = (1 << 127) - 1
Q255 = (1 << 255) - 19
Q511 = (1 << 511) - 187  # prime


def uniform_mod_q(rng: np.random.Generator, count: int, n: int, q: int) -> np.ndarray:
    """[count][L][n] limbs uniform in [0, q): uniform bits below 2^bits(q),
    values >= q (probability < 2^-62 for these moduli) replaced by 0."""
    L = (q.bit_length() + 63) // 64
    top_bits = q.bit_length() - 64 * (L - 1)
    a = rng.integers(0, 1 << 63, size=(count, L, n), dtype=np.uint64) * np.uint64(2) + \
        rng.integers(0, 2, size=(count, L, n), dtype=np.uint64)
    if top_bits < 64:
        a[:, L - 1, :] &= np.uint64((1 << top_bits) - 1)
    qtop = np.uint64(q >> (64 * (L - 1)))
    for c, j in zip(*np.nonzero(a[:, L - 1, :] >= qtop)):
        v = sum(int(a[c, l, j]) << (64 * l) for l in range(L))
        if v >= q:
            a[c, :, j] = 0
    return a


def uneven_spec(count: int, seed: int) -> List[Tuple[int, int]]:
    """(n, q) of `count` products: n weighted .4/.3/.2/.1 over 1024..8192,
    q 127- or 255-bit."""
    r = random.Random(seed)
    ns = r.choices([1024, 2048, 4096, 8192], weights=[0.4, 0.3, 0.2, 0.1], k=count)
    return [(n, Q127 if r.random() < 0.5 else Q255) for n in ns]


def make_batch(ctx: Context, spec: Sequence[Tuple[int, int]], seed: int):
    """Device operands for the products of `spec`, grouped by ring (one
    buffer per ring). Returns (rings, groups, items):
      rings[(n, q)]  -> Ring
      groups[(n, q)] -> (A, B, OUT, host A, host B, product indices)
      items          -> (ring, a, b, out) per product, in spec order."""
    rng = np.random.default_rng(seed)
    rings: Dict[Tuple[int, int], Ring] = {}
    order: Dict[Tuple[int, int], List[int]] = {}
    for k, key in enumerate(spec):
        order.setdefault(key, []).append(k)
    items = [None] * len(spec)
    groups = {}
    for key, idx in order.items():
        n, q = key
        ring = Ring(ctx, n, q)
        rings[key] = ring
        ha = uniform_mod_q(rng, len(idx), n, q)
        hb = uniform_mod_q(rng, len(idx), n, q)
        A, B = ctx.to_device(ha), ctx.to_device(hb)
        OUT = ctx.alloc(ha.nbytes)
        e = ring.L * n * 8
        for m, k in enumerate(idx):
            items[k] = (ring, A.view(m * e, e), B.view(m * e, e), OUT.view(m * e, e))
        groups[key] = (A, B, OUT, ha, hb, idx)
    return rings, groups, items


# ---------------------------------------------------------------- circuit

CIRCUIT_BUFFERS = dict(A=128, Pm=64, B=64, Rm=32, S=32, U=16, V=16, Z=16)


def depth3_circuit(Y, X, A, Pm, B, Rm, S, U, V, Z):
    """128 ciphertexts -> depth-3 tree of 64 + 32 + 16 mul+keyswitch nodes and
    256 additions (Y: a scheme.Yashe with keys; X: [128][L][n] ciphertexts;
    the other buffers hold CIRCUIT_BUFFERS[name] ciphertexts each):

        a_i = x_i + x_{i+1 mod 128}        128 adds
        p_i = a_i a_{i+64}, i < 64          64 mul+ks
        b_i = p_i + x_i                     64 adds
        r_i = b_i b_{i+32}, i < 32          32 mul+ks
        s_i = r_i + r_{i+1 mod 32}          32 adds
        u_i = s_i s_{i+16}, i < 16          16 mul+ks
        v_i = u_i + x_i                     16 adds
        z_i = v_i + v_{i+1 mod 16}          16 adds
    """
    e = Y.elt_bytes
    ring = Y.ring
    ring.ring_addsub(X.view(0, 127 * e), X.view(e, 127 * e), A.view(0, 127 * e), 127)
    ring.ring_addsub(X.view(127 * e, e), X.view(0, e), A.view(127 * e, e), 1)
    ring.mul_keyswitch(Y.evk, A.view(0, 64 * e), A.view(64 * e, 64 * e), Pm, 64)
    ring.ring_addsub(Pm, X.view(0, 64 * e), B, 64)
    ring.mul_keyswitch(Y.evk, B.view(0, 32 * e), B.view(32 * e, 32 * e), Rm, 32)
    ring.ring_addsub(Rm.view(0, 31 * e), Rm.view(e, 31 * e), S.view(0, 31 * e), 31)
    ring.ring_addsub(Rm.view(31 * e, e), Rm.view(0, e), S.view(31 * e, e), 1)
    ring.mul_keyswitch(Y.evk, S.view(0, 16 * e), S.view(16 * e, 16 * e), U, 16)
    ring.ring_addsub(U, X.view(0, 16 * e), V, 16)
    ring.ring_addsub(V.view(0, 15 * e), V.view(e, 15 * e), Z.view(0, 15 * e), 15)
    ring.ring_addsub(V.view(15 * e, e), V.view(0, e), Z.view(15 * e, e), 1)


def depth3_plain(x: np.ndarray, t: int) -> np.ndarray:
    """The same circuit on the messages, over Z_t[x]/(x^n + 1)."""
    from .scheme import negacyclic_mul_mod

    def mul(a, b):
        return negacyclic_mul_mod(a, b, t)

    a = (x + np.roll(x, -1, axis=0)) % t
    p = np.stack([mul(a[i], a[i + 64]) for i in range(64)])
    b = (p + x[:64]) % t
    r = np.stack([mul(b[i], b[i + 32]) for i in range(32)])
    s = (r + np.roll(r, -1, axis=0)) % t
    u = np.stack([mul(s[i], s[i + 16]) for i in range(16)])
    v = (u + x[:16]) % t
    return (v + np.roll(v, -1, axis=0)) % t
