"""Continuous-batching request loop for homomorphic operations.

Requests (homomorphic add, multiply+keyswitch, or encrypt-then-decrypt) on
rings of several sizes arrive over time. Keys and a pool of ciphertexts stay
resident in HBM per ring. The loop itself is native (yh_server_run in the
library): one host thread

  * queues arrived requests per (ring, operation); a queue with fewer than
    `depth` groups in flight forms its next group from its oldest requests
    (up to max_batch) while the HBM reserved for work in flight (per
    request: the library's footprint of its operation, plus its output slot
    and its decryption) stays within a budget and the ring has a free slot,
    so groups grow with the load;
  * launches each group as ONE batched operation over pointer lists (no
    gather copies) on that ring's own stream, so groups of different rings
    overlap on the device, followed by a batched decryption of the results
    and their copy to the host;
  * retires groups whose completion event has fired, stamping each request
    with the device time its result was complete;
  * keeps admitting new arrivals while earlier groups run.

This module sets up the keys and pools, hands requests to that loop, and
checks every decryption against the plaintext computation afterwards.
Results do not depend on the mix: a request's ciphertext is computed by the
same per-element arithmetic alone or in any batch, and encryption noise is
drawn on the device from a generator keyed by the request id.

The reference has no request loop; its operations are the Ciphertext
operators (src/ciphertext.cpp:38-107) and Yashe::encrypt/decrypt
(src/yashe.cpp:104-197), one call and one cudaDeviceSynchronize at a time.
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import REQUEST_DTYPE, Context, ServeRing, ServeStats, _vp, check, lib
from .scheme import Yashe
from .workloads import Q127, Q255, Q511

KINDS = ("add", "mulks", "encdec")


@dataclasses.dataclass
class Request:
    rid: int
    kind: str                       # "add" | "mulks" | "encdec"
    ring: Tuple[int, int]           # (n, q)
    a: int = -1                     # resident ciphertext indices (add, mulks)
    b: int = -1
    msg: Optional[np.ndarray] = None  # encdec message, [n] in [0, t)
    arrival: float = 0.0            # seconds after the start of the run
    done: Optional[float] = None    # result ready (device clock, same origin)
    ok: Optional[bool] = None       # decryption matched the plaintext result
    output: Optional[np.ndarray] = None  # result ciphertext limbs (keep_outputs)

    @property
    def latency(self) -> Optional[float]:
        return None if self.done is None else self.done - self.arrival


def negacyclic_mul_mod_t(a: np.ndarray, b: np.ndarray, t: int) -> np.ndarray:
    """a b mod (x^n + 1, t) along the last axis by a length-2n real FFT,
    exact while the convolution values stay far below 2^52 (n (t-1)^2 <=
    2^40 here); the rounding margin is checked."""
    n = a.shape[-1]
    assert n * (t - 1) ** 2 < (1 << 40)
    fa = np.fft.rfft(np.mod(a, t).astype(np.float64), 2 * n)
    fb = np.fft.rfft(np.mod(b, t).astype(np.float64), 2 * n)
    c = np.fft.irfft(fa * fb, 2 * n)
    ci = np.rint(c)
    assert np.max(np.abs(c - ci)) < 0.25
    ci = ci.astype(np.int64)
    return np.mod(ci[..., :n] - ci[..., n:], t)


class _RingState:
    def __init__(self, ctx: Context, key, t: int, w_bits: int, pool: int, slots: int, seed: int):
        n, q = key
        self.key = key
        self.Y = Yashe(ctx, n, q, t, w_bits, seed=seed).generate_keys()
        self.ring = self.Y.ring
        self.n = n
        self.L = self.Y.L
        self.slots = slots
        rng = np.random.default_rng(seed + 1)
        self.pool_m = rng.integers(0, t, size=(pool, n))
        self.pool = self.Y.encrypt(self.pool_m)
        self.ring.sync()


class Server:
    """Continuous-batching loop over resident keys and ciphertexts of several
    rings (see the module docstring)."""

    def __init__(self, ctx: Context, rings: Sequence[Tuple[int, int]], t: int = 1024, w_bits: int = 32,
                 pool: int = 32, slots: int = 64, hbm_budget: int = 4 << 30, max_batch: int = 64,
                 seed: int = 0, keep_outputs: bool = False, depth: int = 2):
        self.ctx = ctx
        self.t = t
        self.max_batch = max_batch
        self.keep_outputs = keep_outputs
        self.seed = seed
        self.pool_size = pool
        self.rings: Dict[Tuple[int, int], _RingState] = {}
        self._index: Dict[Tuple[int, int], int] = {}
        for k, key in enumerate(rings):
            self.rings[key] = _RingState(ctx, key, t, w_bits, pool, slots, seed + 1000 * k)
            self._index[key] = k
        desc = (ServeRing * len(self.rings))()
        for k, rs in enumerate(self.rings.values()):
            desc[k] = ServeRing(rs.ring.h, rs.Y.evk.h if rs.Y.evk is not None else None, rs.Y.f.ptr, rs.Y.h.ptr,
                                rs.pool.ptr, pool, slots)
        h = _vp()
        check(lib().yh_server_create(ctx.h, desc, len(self.rings), hbm_budget, max_batch, depth, seed,
                                     ctypes.byref(h)), "yh_server_create")
        self.h = h.value
        self._budget = hbm_budget

    # ---------------------------------------------------------------- resources
    @property
    def budget(self) -> int:
        return self._budget

    @budget.setter
    def budget(self, v: int):
        check(lib().yh_server_set_budget(self.h, int(v)), "yh_server_set_budget")
        self._budget = int(v)

    def footprint(self, r: Request) -> int:
        return int(lib().yh_server_footprint(self.h, self._index[r.ring], KINDS.index(r.kind)))

    def close(self):
        """Releases the loop, then the rings' keys and rings."""
        if self.h:
            check(lib().yh_server_destroy(self.h), "yh_server_destroy")
            self.h = None
        for rs in self.rings.values():
            rs.ring.sync()
            if rs.Y.evk is not None:
                rs.Y.evk.close()
            rs.ring.close()
        self.rings = {}

    # ---------------------------------------------------------------- checks
    def _check(self, reqs: List[Request], results: List[np.ndarray]):
        """Compare each decryption with the plaintext computation (mul+keyswitch
        results in batches of one FFT per ring)."""
        t = self.t
        mul: Dict[Tuple[int, int], List[int]] = {}
        for m, r in enumerate(reqs):
            rs = self.rings[r.ring]
            if r.kind == "encdec":
                r.ok = bool(np.array_equal(results[m].astype(np.int64), np.asarray(r.msg, dtype=np.int64)))
            elif r.kind == "add":
                want = np.mod(rs.pool_m[r.a] + rs.pool_m[r.b], t)
                r.ok = bool(np.array_equal(results[m].astype(np.int64), want))
            else:
                mul.setdefault(r.ring, []).append(m)
        for key, idx in mul.items():
            rs = self.rings[key]
            for c0 in range(0, len(idx), 128):
                chunk = idx[c0:c0 + 128]
                a = rs.pool_m[[reqs[m].a for m in chunk]]
                b = rs.pool_m[[reqs[m].b for m in chunk]]
                want = negacyclic_mul_mod_t(a, b, t)
                for row, m in enumerate(chunk):
                    reqs[m].ok = bool(np.array_equal(results[m].astype(np.int64), want[row]))

    # ---------------------------------------------------------------- loop
    def run(self, requests: Sequence[Request], idle_sleep: Optional[float] = None) -> dict:
        """Serves the requests (arrival times relative to this call) in the
        native loop, then checks every decryption. idle_sleep is unused (the
        loop polls its completion events)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            if r.kind not in KINDS or r.ring not in self.rings:
                raise ValueError(f"request {r.rid}: unknown kind or ring")
            if self.footprint(r) > self.budget:
                raise ValueError(f"request {r.rid} needs more HBM than the whole budget")
        count = len(reqs)
        arr = np.zeros(count, dtype=REQUEST_DTYPE)
        sizes = [self.rings[r.ring].n for r in reqs]
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        results = np.zeros(int(offs[-1]), dtype=np.uint64)
        msgs = np.zeros(int(offs[-1]), dtype=np.uint64)
        outputs = None
        if self.keep_outputs:
            osz = [self.rings[r.ring].L * self.rings[r.ring].n for r in reqs]
            ooffs = np.concatenate([[0], np.cumsum(osz)]).astype(np.int64)
            outputs = np.zeros(int(ooffs[-1]), dtype=np.uint64)
        for m, r in enumerate(reqs):
            arr[m]["arrival"] = r.arrival
            arr[m]["rid"] = r.rid
            arr[m]["kind"] = KINDS.index(r.kind)
            arr[m]["ring"] = self._index[r.ring]
            arr[m]["a"], arr[m]["b"] = r.a, r.b
            if r.kind == "encdec":
                msgs[offs[m]:offs[m + 1]] = np.asarray(r.msg, dtype=np.int64)
                arr[m]["msg"] = msgs.ctypes.data + 8 * int(offs[m])
            arr[m]["result"] = results.ctypes.data + 8 * int(offs[m])
            if outputs is not None:
                arr[m]["output"] = outputs.ctypes.data + 8 * int(ooffs[m])
        st = ServeStats()
        check(lib().yh_server_run(self.h, arr.ctypes.data, count, ctypes.byref(st)), "yh_server_run")
        res = [results[offs[m]:offs[m + 1]] for m in range(count)]
        for m, r in enumerate(reqs):
            r.done = float(arr[m]["done"])
            if outputs is not None:
                rs = self.rings[r.ring]
                r.output = outputs[ooffs[m]:ooffs[m + 1]].reshape(rs.L, rs.n).copy()
        self._check(reqs, res)
        lat = np.array([r.latency for r in reqs])
        by_kind = {}
        for kind in KINDS:
            lk = np.array([r.latency for r in reqs if r.kind == kind])
            if len(lk):
                by_kind[kind] = {"count": int(len(lk)), "p50_ms": round(float(np.percentile(lk, 50)) * 1e3, 4),
                                 "p99_ms": round(float(np.percentile(lk, 99)) * 1e3, 4)}
        return {
            "requests": count,
            "ops_per_s": round(count / st.wall_s, 1) if st.wall_s > 0 else 0.0,
            "wall_s": round(st.wall_s, 4),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 4),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 4),
            "by_kind": by_kind,
            "all_decrypts_ok": all(r.ok for r in reqs),
            "groups": int(st.groups),
            "mean_batch": round(float(st.mean_batch), 2),
            "max_inflight_groups": int(st.max_inflight),
            "peak_reserved_bytes": int(st.peak_reserved),
            "budget_bytes": int(self.budget),
        }


def c5_rings() -> List[Tuple[int, int]]:
    return [(n, q) for n in (2048, 4096, 8192, 16384) for q in (Q127, Q255, Q511)]


def poisson_requests(count: int, qps: float, rings: Sequence[Tuple[int, int]], t: int, pool: int,
                     seed: int = 0, mix=(0.6, 0.3, 0.1)) -> List[Request]:
    """Poisson arrivals at `qps`; ring uniform over `rings`; kinds by `mix`
    (add, mul+keyswitch, encrypt/decrypt); operands uniform over the pool."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, size=count)
    arrivals = np.cumsum(gaps) - gaps[0]
    kinds = rng.choice(len(KINDS), size=count, p=list(mix))
    ring_idx = rng.integers(0, len(rings), size=count)
    out = []
    for k in range(count):
        ring = rings[ring_idx[k]]
        kind = KINDS[kinds[k]]
        r = Request(rid=k, kind=kind, ring=ring, arrival=float(arrivals[k]))
        if kind == "encdec":
            r.msg = rng.integers(0, t, size=ring[0])
        else:
            r.a, r.b = (int(v) for v in rng.integers(0, pool, size=2))
        out.append(r)
    return out
