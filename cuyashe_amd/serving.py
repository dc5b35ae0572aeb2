"""Continuous-batching request loop for homomorphic operations.

Requests (homomorphic add, multiply+keyswitch, or encrypt-then-decrypt) on
rings of several sizes arrive over time. Keys and a pool of ciphertexts stay
resident in HBM per ring. The loop

  * admits arrived requests in arrival order while the HBM reserved for work
    in flight (per request: the library's footprint of its operation, plus
    its output slot) stays within a budget and the ring has a free slot;
  * groups admitted requests by (ring, operation) and launches each group as
    ONE batched operation over pointer lists (no gather copies) on that
    ring's own stream, so groups of different rings overlap on the device,
    followed by a batched decryption of the results;
  * retires groups whose completion event has fired: each result is
    decrypted and checked against the plaintext computation, and the
    request's latency is its arrival-to-result time on the device clock;
  * keeps admitting new arrivals while earlier groups run.

Results do not depend on the mix: a request's ciphertext is computed by the
same per-element arithmetic alone or in any batch, and encryption noise is
drawn from a generator seeded by the request id.

The reference has no request loop; its operations are the Ciphertext
operators (src/ciphertext.cpp:38-107) and Yashe::encrypt/decrypt
(src/yashe.cpp:104-197), one call and one cudaDeviceSynchronize at a time.
"""
from __future__ import annotations

import collections
import concurrent.futures
import dataclasses
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import OP_ADD, OP_MUL_KS, YH_MUL, Context, DeviceBuffer, Event, check, lib
from .scheme import Yashe, sample_discrete_gaussian
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
    """a b mod (x^n + 1, t) by a length-2n real FFT, exact while the
    convolution values stay far below 2^52 (n (t-1)^2 <= 2^40 here); the
    rounding margin is checked."""
    n = a.shape[-1]
    assert n * (t - 1) ** 2 < (1 << 40)
    fa = np.fft.rfft(np.mod(a, t).astype(np.float64), 2 * n)
    fb = np.fft.rfft(np.mod(b, t).astype(np.float64), 2 * n)
    c = np.fft.irfft(fa * fb, 2 * n)
    ci = np.rint(c)
    assert np.max(np.abs(c - ci)) < 0.25
    ci = ci.astype(np.int64)
    return np.mod(ci[:n] - ci[n:], t)


class _RingState:
    def __init__(self, ctx: Context, key, t: int, w_bits: int, pool: int, slots: int, seed: int):
        n, q = key
        self.key = key
        self.Y = Yashe(ctx, n, q, t, w_bits, seed=seed).generate_keys()
        self.ring = self.Y.ring
        self.ring.use_own_stream()
        self.n = n
        self.elt = self.Y.elt_bytes
        rng = np.random.default_rng(seed + 1)
        self.pool_m = rng.integers(0, t, size=(pool, n))
        self.pool = self.Y.encrypt(self.pool_m)
        self.ring.sync()
        self.slots = slots
        self.out = ctx.alloc(slots * self.elt)          # result ciphertexts
        self.free = collections.deque(range(slots))
        self.fp = {"add": self.ring.footprint(OP_ADD), "mulks": self.ring.footprint(OP_MUL_KS),
                   "encdec": 3 * self.elt}

    def out_ptr(self, s):
        return self.out.ptr + s * self.elt

    def pool_ptr(self, i):
        return self.pool.ptr + i * self.elt


class _Group:
    def __init__(self, rs, kind, reqs, slots, tables, ev, reserved):
        self.rs, self.kind, self.reqs, self.slots = rs, kind, reqs, slots
        self.tables, self.ev, self.reserved = tables, ev, reserved


class Server:
    """Continuous-batching loop over resident keys and ciphertexts of several
    rings (see the module docstring)."""

    def __init__(self, ctx: Context, rings: Sequence[Tuple[int, int]], t: int = 1024, w_bits: int = 32,
                 pool: int = 32, slots: int = 64, hbm_budget: int = 4 << 30, max_batch: int = 64,
                 seed: int = 0, keep_outputs: bool = False):
        self.ctx = ctx
        self.t = t
        self.max_batch = max_batch
        self.budget = hbm_budget
        self.keep_outputs = keep_outputs
        self.seed = seed
        self.rings: Dict[Tuple[int, int], _RingState] = {}
        for k, key in enumerate(rings):
            self.rings[key] = _RingState(ctx, key, t, w_bits, pool, slots, seed + 1000 * k)
        self.pool_size = pool
        self._tables: List[DeviceBuffer] = []  # free pointer tables (4 x max_batch pointers)
        self._events: List[Event] = []
        self._stage_free: Dict[tuple, List[DeviceBuffer]] = {}
        self.start_ev = Event(ctx)
        self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=2)
        self._checks: List[concurrent.futures.Future] = []

    # ---------------------------------------------------------------- resources
    def _table(self) -> DeviceBuffer:
        return self._tables.pop() if self._tables else self.ctx.alloc(4 * self.max_batch * 8)

    def close(self):
        """Releases the rings' keys and rings (after the loop has drained)."""
        self._pool.shutdown(wait=True)
        for rs in self.rings.values():
            rs.ring.sync()
            if rs.Y.evk is not None:
                rs.Y.evk.close()
            rs.ring.close()
        self.rings = {}

    def _event(self) -> Event:
        return self._events.pop() if self._events else Event(self.ctx)

    def footprint(self, r: Request) -> int:
        rs = self.rings[r.ring]
        return rs.fp[r.kind] + rs.elt + rs.n * 8

    # ---------------------------------------------------------------- launch / retire
    def _staging(self, rs: _RingState, kind: str) -> DeviceBuffer:
        """A per-ring staging buffer of max_batch elements, reused only after
        the group holding it has retired."""
        free = self._stage_free.setdefault((rs.key, kind), [])
        if free:
            return free.pop()
        return self.ctx.alloc(self.max_batch * rs.elt)

    def _encdec_inputs(self, rs: _RingState, reqs: List[Request]):
        """Host samples of the encryptions (seeded by request id)."""
        Y = rs.Y
        S, E = [], []
        for r in reqs:
            g = np.random.default_rng((self.seed, r.rid))
            S.append(g.integers(-1, 2, size=rs.n))
            E.append(sample_discrete_gaussian(g, rs.n))
        return (Y.pack_small(np.stack(S), "ternary"), Y.pack_small(np.stack(E), "gauss"),
                Y.pack_message(np.stack([r.msg for r in reqs])))

    def _launch(self, rs: _RingState, kind: str, reqs: List[Request], reserved: int) -> _Group:
        k = len(reqs)
        mb = self.max_batch
        slots = [rs.free.popleft() for _ in range(k)]
        ring = rs.ring
        tab = self._table()
        outs = [rs.out_ptr(s) for s in slots]
        mstage = self._staging(rs, "M")  # decrypted results of the group, contiguous
        msgs = [mstage.ptr + m * rs.n * 8 for m in range(k)]
        host = np.zeros(4 * mb, dtype=np.uint64)
        stage = [("M", mstage)]
        if kind == "encdec":
            # c = h s + e + floor(q/t) m, written straight into the output slots
            S, E, DM = self._encdec_inputs(rs, reqs)
            bufs = {name: self._staging(rs, name) for name in ("S", "E", "DM", "HS")}
            bufs["S"].upload(S)
            bufs["E"].upload(E)
            bufs["DM"].upload(DM)
            stage += list(bufs.items())
            ring.scalar(bufs["DM"], YH_MUL, rs.Y.delta, bufs["DM"], k)  # floor(q/t) m mod q
            ring.ring_mul_bcast(rs.Y.h, bufs["S"], bufs["HS"], k)
            host[:k] = [bufs["HS"].ptr + m * rs.elt for m in range(k)]
            host[mb:mb + k] = [bufs["E"].ptr + m * rs.elt for m in range(k)]
            host[2 * mb:2 * mb + k] = outs
            host[3 * mb:3 * mb + k] = [bufs["DM"].ptr + m * rs.elt for m in range(k)]
            tab.upload(host)
            l0, l1, l2, l3 = (tab.view(j * mb * 8, mb * 8) for j in range(4))
            ring.addsub_list(l0, l1, l2, k)   # h s + e
            ring.addsub_list(l2, l3, l2, k)   # + floor(q/t) m
            ev_op = self._event()
            ev_op.record(ring)
            host[3 * mb:3 * mb + k] = msgs
            tab2 = self._table()
            tab2.upload(host)
            ring.decrypt_list(rs.Y.f, tab2.view(2 * mb * 8, mb * 8), tab2.view(3 * mb * 8, mb * 8), k)
            tables = [tab, tab2]
        else:
            host[:k] = [rs.pool_ptr(r.a) for r in reqs]
            host[mb:mb + k] = [rs.pool_ptr(r.b) for r in reqs]
            host[2 * mb:2 * mb + k] = outs
            host[3 * mb:3 * mb + k] = msgs
            tab.upload(host)
            la, lb, lo, lm = (tab.view(j * mb * 8, mb * 8) for j in range(4))
            if kind == "add":
                ring.addsub_list(la, lb, lo, k)
            else:
                ring.mul_keyswitch_list(rs.Y.evk, la, lb, lo, k)
            ev_op = self._event()
            ev_op.record(ring)
            ring.decrypt_list(rs.Y.f, lo, lm, k)  # the client's check, batched
            tables = [tab]
        ev = self._event()
        ev.record(ring)
        g = _Group(rs, kind, reqs, slots, tables, ev, reserved)
        g.ev_op = ev_op
        g.stage = stage
        g.outs = outs
        g.mstage = mstage
        return g

    def _dtoh(self, ptr: int, shape, dtype=np.uint64) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        check(lib().yh_memcpy_dtoh(self.ctx.h, out.ctypes.data, ptr, out.nbytes), "yh_memcpy_dtoh")
        return out

    def _check(self, rs: _RingState, reqs: List[Request], got: np.ndarray):
        """Compare each decryption with the plaintext computation."""
        t = self.t
        for m, r in enumerate(reqs):
            if r.kind == "encdec":
                want = r.msg
            elif r.kind == "add":
                want = np.mod(rs.pool_m[r.a] + rs.pool_m[r.b], t)
            else:
                want = negacyclic_mul_mod_t(rs.pool_m[r.a], rs.pool_m[r.b], t)
            r.ok = bool(np.array_equal(got[m].astype(np.int64), want))

    def _retire(self, g: _Group):
        rs = g.rs
        done = self.start_ev.elapsed_ms(g.ev_op) * 1e-3
        got = self._dtoh(g.mstage.ptr, (len(g.reqs), rs.n))  # one copy per group
        for m, r in enumerate(g.reqs):
            r.done = done
            if self.keep_outputs:
                r.output = self._dtoh(g.outs[m], (rs.Y.L, rs.n))
        # the plaintext checks run beside the loop
        self._checks.append(self._pool.submit(self._check, rs, g.reqs, got))
        for s in g.slots:
            rs.free.append(s)
        self._tables.extend(g.tables)
        self._events.extend([g.ev, g.ev_op])
        for name, buf in g.stage:
            self._stage_free.setdefault((rs.key, name), []).append(buf)

    # ---------------------------------------------------------------- loop
    def run(self, requests: Sequence[Request], idle_sleep: float = 20e-6) -> dict:
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            if r.kind not in KINDS or r.ring not in self.rings:
                raise ValueError(f"request {r.rid}: unknown kind or ring")
            if self.footprint(r) > self.budget:
                raise ValueError(f"request {r.rid} needs more HBM than the whole budget")
        pending: collections.deque = collections.deque()
        inflight: List[_Group] = []
        used = 0
        peak_used = 0
        i = 0
        groups = 0
        batch_sizes = []
        self.ctx.sync()
        for rs in self.rings.values():
            rs.ring.sync()
        self.start_ev.record()
        t0 = time.perf_counter()
        while i < len(reqs) or pending or inflight:
            now = time.perf_counter() - t0
            while i < len(reqs) and reqs[i].arrival <= now:
                pending.append(reqs[i])
                i += 1
            # retire finished groups
            still = []
            for g in inflight:
                if g.ev.done():
                    self._retire(g)
                    used -= g.reserved
                else:
                    still.append(g)
            inflight = still
            # admit in arrival order while memory and slots allow
            batch: Dict[Tuple[Tuple[int, int], str], List[Request]] = collections.OrderedDict()
            reserved: Dict[Tuple[Tuple[int, int], str], int] = {}
            free_slots = {key: len(rs.free) for key, rs in self.rings.items()}
            while pending:
                r = pending[0]
                fp = self.footprint(r)
                key = (r.ring, r.kind)
                if used + fp > self.budget:
                    break  # strict arrival order: wait for memory
                if free_slots[r.ring] == 0 or len(batch.get(key, ())) >= self.max_batch:
                    break
                pending.popleft()
                used += fp
                free_slots[r.ring] -= 1
                batch.setdefault(key, []).append(r)
                reserved[key] = reserved.get(key, 0) + fp
            peak_used = max(peak_used, used)
            for (ring_key, kind), rl in batch.items():
                inflight.append(self._launch(self.rings[ring_key], kind, rl, reserved[(ring_key, kind)]))
                groups += 1
                batch_sizes.append(len(rl))
            if not batch and (not inflight or i < len(reqs)):
                # nothing to launch: wait for the next arrival or completion
                time.sleep(idle_sleep)
        wall = time.perf_counter() - t0
        for f in self._checks:
            f.result()
        self._checks = []
        lat = np.array([r.latency for r in reqs])
        by_kind = {}
        for kind in KINDS:
            lk = np.array([r.latency for r in reqs if r.kind == kind])
            if len(lk):
                by_kind[kind] = {"count": int(len(lk)), "p50_ms": round(float(np.percentile(lk, 50)) * 1e3, 4),
                                 "p99_ms": round(float(np.percentile(lk, 99)) * 1e3, 4)}
        return {
            "requests": len(reqs),
            "ops_per_s": round(len(reqs) / wall, 1),
            "wall_s": round(wall, 4),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 4),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 4),
            "by_kind": by_kind,
            "all_decrypts_ok": all(r.ok for r in reqs),
            "groups": groups,
            "mean_batch": round(float(np.mean(batch_sizes)), 2) if batch_sizes else 0.0,
            "peak_reserved_bytes": int(peak_used),
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
