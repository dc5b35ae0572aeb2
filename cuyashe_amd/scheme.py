"""YASHE key generation, encryption and decryption on device-resident data.

Mirrors the reference's Yashe class (src/yashe.cpp, src/yashe.h) for Python
callers such as the serving loop: every ring operation runs through the C ABI
of libyashe_hip.so, and only the small random samples (ternary and discrete
Gaussian polynomials, messages) are drawn on the host and uploaded once.

    keys:    f = t f' + 1 (f' ternary, f invertible mod q),
             h = t g f^-1,                          src/yashe.cpp:37-69
             gamma_k = f w^k + e_k + h s_k, k < lw   src/yashe.cpp:70-93
    encrypt: c = h s + e + floor(q/t) m             src/yashe.cpp:104-142
             (floor(q/t) m by the device scalar product, yh_ring_scalar)
    decrypt: m = [round(t/q [f c]_q)]_t (f f after a mul without
             keyswitch)                             src/yashe.cpp:167-197

The samplers are not bit-compatible with the reference's (it draws from
rand()/NTL's RNG); what matters for parity is the arithmetic on the samples.
Errors follow the discrete Gaussian D_{Z, sigma} with sigma = 8 cut at
+-48 (6 sigma): P(x) proportional to exp(-x^2 / (2 sigma^2)) on the integers
of [-48, 48], drawn by inverse transform of that table (the reference's
settings, src/yashe.h:29-33; its own sampler draws uniform values mod 7,
src/distribution.cpp:4-63).
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from ._lib import (Context, DeviceBuffer, Ring, YashError, ints_to_array, WIRE_CIPHERTEXT, WIRE_KEYSWITCH_KEY,
                   WIRE_PUBLIC_KEY, WIRE_SECRET_KEY, YH_MUL, wire_blob_size, wire_peek)

GAUSS_SIGMA = 8.0   # src/yashe.h:29-33: discrete Gaussian with sigma 8
GAUSS_BOUND = 48    # truncated at 6 sigma
KEYGEN_ATTEMPTS = 17  # draws of f before key generation gives up (as the C++ Yashe)


def discrete_gaussian_cdf(sigma: float = GAUSS_SIGMA, bound: int = GAUSS_BOUND) -> np.ndarray:
    """Cumulative distribution of D_{Z, sigma} restricted to [-bound, bound]
    (entry i for the value i - bound); the last entry is 1."""
    x = np.arange(-bound, bound + 1, dtype=np.float64)
    w = np.exp(-(x * x) / (2.0 * sigma * sigma))
    c = np.cumsum(w)
    return c / c[-1]


_GAUSS_CDF = discrete_gaussian_cdf()


def sample_discrete_gaussian(rng: np.random.Generator, shape) -> np.ndarray:
    """Integers distributed as D_{Z, 8} on [-48, 48] (inverse transform)."""
    u = rng.random(size=shape)
    idx = np.searchsorted(_GAUSS_CDF, u, side="right")
    return np.minimum(idx, 2 * GAUSS_BOUND).astype(np.int64) - GAUSS_BOUND


class Yashe:
    def __init__(self, ctx: Context, n: int, q: int, t: int, w_bits: int = 32, seed: int = 0,
                 ring: Optional[Ring] = None):
        self.ctx = ctx
        self.ring = ring if ring is not None else Ring(ctx, n, q, t=t, w_bits=w_bits)
        self.n, self.q, self.t = int(n), int(q), int(t)
        self.L = self.ring.L
        self.lw = self.ring.lw
        self.delta = self.q // self.t
        self.rng = np.random.default_rng(seed)
        self._ternary = ints_to_array([self.q - 1, 0, 1], self.L)            # index v + 1
        self._gauss = ints_to_array([v % self.q for v in range(-GAUSS_BOUND, GAUSS_BOUND + 1)], self.L)
        self.elt_bytes = self.L * self.n * 8
        self.f = self.ff = self.h = self.evk = None

    # ------------------------------------------------------------ sampling
    def ternary(self, batch: int) -> np.ndarray:
        return self.rng.integers(-1, 2, size=(batch, self.n), dtype=np.int64)

    def gaussian(self, batch: int) -> np.ndarray:
        return sample_discrete_gaussian(self.rng, (batch, self.n))

    def _lookup(self, table: np.ndarray, idx: np.ndarray) -> np.ndarray:
        """[batch][L][n] limbs of table[idx]."""
        return np.ascontiguousarray(np.transpose(table[:, idx], (1, 0, 2)))

    def pack_small(self, v: np.ndarray, kind: str) -> np.ndarray:
        if kind == "ternary":
            return self._lookup(self._ternary, v + 1)
        if kind == "gauss":
            return self._lookup(self._gauss, v + GAUSS_BOUND)
        raise ValueError(kind)

    def pack_message(self, m: np.ndarray) -> np.ndarray:
        """Messages m in [0, t) as ring elements [batch][L][n] (t < 2^64)."""
        m = np.atleast_2d(np.asarray(m, dtype=np.int64))
        if m.min(initial=0) < 0 or m.max(initial=0) >= self.t:
            raise YashError("messages must lie in [0, t)")
        out = np.zeros((m.shape[0], self.L, self.n), dtype=np.uint64)
        out[:, 0, :] = m.astype(np.uint64)
        return out

    # ------------------------------------------------------------ keys
    def generate_keys(self):
        n, q, t = self.n, self.q, self.t
        ring = self.ring
        for attempt in range(KEYGEN_ATTEMPTS):
            fp = self.ternary(1)[0]
            f_vals = [(t * int(x) + (1 if j == 0 else 0)) % q for j, x in enumerate(fp)]
            d_f = self.ctx.to_device(ints_to_array(f_vals, self.L)[None])
            finv = self.ctx.alloc(self.elt_bytes)
            try:
                ring.inverse(d_f, finv)
                break
            except YashError as e:
                # f not invertible: draw again (src/yashe.cpp:45-55); any other
                # failure (allocation, capture) is not the sample's fault
                if "not invertible" not in str(e) or attempt == KEYGEN_ATTEMPTS - 1:
                    raise
        g = self.ternary(1)[0]
        d_tg = self.ctx.to_device(ints_to_array([(t * int(x)) % q for x in g], self.L)[None])
        self.h = self.ctx.alloc(self.elt_bytes)
        ring.ring_mul(d_tg, finv, self.h, 1)
        self.f = d_f
        self.ff = self.ctx.alloc(self.elt_bytes)
        ring.ring_mul(d_f, d_f, self.ff, 1)
        if self.lw:
            lw = self.lw
            S = self.ctx.to_device(self.pack_small(self.ternary(lw), "ternary"))
            gam = self.ctx.alloc(lw * self.elt_bytes)
            ring.ring_mul_bcast(self.h, S, gam, lw)
            E = self.ctx.to_device(self.pack_small(self.gaussian(lw), "gauss"))
            ring.ring_addsub(gam, E, gam, lw)
            fw = np.stack([ints_to_array([(v << (ring.w_bits * k)) % q for v in f_vals], self.L) for k in range(lw)])
            FW = self.ctx.to_device(fw)
            ring.ring_addsub(gam, FW, gam, lw)
            self.evk = ring.evk(gam)
            self.gamma = gam
        return self

    # ------------------------------------------------------------ key files
    def save_keys(self, path: str, secret: bool = True):
        """Writes h, the keyswitch key gamma_0..gamma_{lw-1} and (secret=True)
        f as concatenated wire-format blobs of this ring."""
        ring = self.ring
        blobs = [ring.export(WIRE_PUBLIC_KEY, self.h, 1)]
        if self.lw:
            blobs.append(ring.export(WIRE_KEYSWITCH_KEY, self.gamma, self.lw))
        if secret:
            blobs.append(ring.export(WIRE_SECRET_KEY, self.f, 1))
        with open(path, "wb") as fh:
            for b in blobs:
                fh.write(b)

    def load_keys(self, path: str):
        """Reads a key file written by save_keys for the same ring: h, the
        resident keyswitch key (rebuilt from gamma_k) and f if present."""
        with open(path, "rb") as fh:
            data = fh.read()
        off = 0
        ring = self.ring
        while off < len(data):
            size = wire_blob_size(data, off)
            blob = data[off:off + size]
            kind = wire_peek(blob[:96]).kind
            if kind == WIRE_PUBLIC_KEY:
                self.h = ring.import_(WIRE_PUBLIC_KEY, blob)
            elif kind == WIRE_KEYSWITCH_KEY:
                self.gamma = ring.import_(WIRE_KEYSWITCH_KEY, blob)
                self.evk = ring.evk(self.gamma)
            elif kind == WIRE_SECRET_KEY:
                self.f = ring.import_(WIRE_SECRET_KEY, blob)
                self.ff = self.ctx.alloc(self.elt_bytes)
                ring.ring_mul(self.f, self.f, self.ff, 1)
            else:
                raise YashError(f"unexpected element kind {kind} in key file")
            off += size
        return self

    def export_ciphertexts(self, c: DeviceBuffer, batch: int) -> bytes:
        return self.ring.export(WIRE_CIPHERTEXT, c, batch)

    def import_ciphertexts(self, data: bytes) -> DeviceBuffer:
        return self.ring.import_(WIRE_CIPHERTEXT, data)

    # ------------------------------------------------------------ ciphertexts
    def encrypt(self, m: np.ndarray, out: Optional[DeviceBuffer] = None) -> DeviceBuffer:
        """Encrypt a [batch][n] array of messages in [0, t) into device
        ciphertexts [batch][L][n]."""
        m = np.atleast_2d(np.asarray(m, dtype=np.int64))
        batch = m.shape[0]
        ring = self.ring
        out = out if out is not None else self.ctx.alloc(batch * self.elt_bytes)
        S = self.ctx.to_device(self.pack_small(self.ternary(batch), "ternary"))
        ring.ring_mul_bcast(self.h, S, out, batch)
        E = self.ctx.to_device(self.pack_small(self.gaussian(batch), "gauss"))
        ring.ring_addsub(out, E, out, batch)
        DM = self.ctx.to_device(self.pack_message(m))
        ring.scalar(DM, YH_MUL, self.delta, DM, batch)
        ring.ring_addsub(out, DM, out, batch)
        return out

    def decrypt(self, c: DeviceBuffer, batch: int = 1, aftermul: bool = False) -> np.ndarray:
        dm = self.ctx.alloc(batch * self.n * 8)
        self.ring.decrypt(self.ff if aftermul else self.f, c, dm, batch)
        return dm.download((batch, self.n))

    def add(self, a: DeviceBuffer, b: DeviceBuffer, out: DeviceBuffer, batch: int = 1):
        self.ring.ring_addsub(a, b, out, batch)

    def mul_keyswitch(self, a: DeviceBuffer, b: DeviceBuffer, out: DeviceBuffer, batch: int = 1):
        self.ring.mul_keyswitch(self.evk, a, b, out, batch)


def negacyclic_mul_mod(a: np.ndarray, b: np.ndarray, t: int) -> np.ndarray:
    """a b mod (x^n + 1, t) for message polynomials (exact in int64 while
    n (t - 1)^2 < 2^63)."""
    n = a.shape[-1]
    assert n * (t - 1) ** 2 < (1 << 63)
    full = np.convolve(np.mod(a, t).astype(np.int64), np.mod(b, t).astype(np.int64))
    out = full[:n].copy()
    out[: n - 1] -= full[n:]
    return np.mod(out, t)
