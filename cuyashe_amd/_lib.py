"""ctypes binding of ``libyashe_hip.so`` (the C ABI in include/yashe_hip.h).

There is no fallback: if the gfx950 library is not built, or no GPU is
visible, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CUYASHE_AMD_LIB points at another build of the same library (e.g. one
# compiled with different tuning macros).
LIB_PATH = os.environ.get("CUYASHE_AMD_LIB") or os.path.join(HERE, "lib", "libyashe_hip.so")

YH_OK = 0
YH_FORWARD = 1
YH_INVERSE = 0
YH_ADD = 0
YH_SUB = 1
YH_MUL = 2

OP_RING_MUL, OP_YASHE_MUL, OP_KEYSWITCH, OP_MUL_KS, OP_ADD = range(5)
KERNEL_PROD, KERNEL_KEYSWITCH, KERNEL_ICRT, KERNEL_BATCH, KERNEL_KS_DIGITS = range(5)
WIRE_CIPHERTEXT, WIRE_KEYSWITCH_KEY, WIRE_PUBLIC_KEY, WIRE_SECRET_KEY, WIRE_PLAINTEXT = range(1, 6)

_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p


class YashError(RuntimeError):
    pass


class RingInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("q_limbs", ctypes.c_int), ("q_bits", ctypes.c_int), ("R", ctypes.c_int),
        ("R_ks", ctypes.c_int), ("w_bits", ctypes.c_int), ("lw", ctypes.c_int), ("work_limbs", ctypes.c_int),
        ("m_bits", ctypes.c_int), ("m_ks_bits", ctypes.c_int), ("t", ctypes.c_uint64), ("prime_max", ctypes.c_uint64),
    ]


class WireHeader(ctypes.Structure):
    _fields_ = [("version", ctypes.c_int), ("kind", ctypes.c_int), ("n", ctypes.c_int), ("q_limbs", ctypes.c_int),
                ("t", ctypes.c_uint64), ("w_bits", ctypes.c_int), ("count", ctypes.c_int),
                ("q", ctypes.c_uint64 * 8)]

    def modulus(self) -> int:
        return sum(int(self.q[l]) << (64 * l) for l in range(self.q_limbs))


class BatchItem(ctypes.Structure):
    _fields_ = [("ring", ctypes.c_void_p), ("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("out", ctypes.c_void_p)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("wg_threads", ctypes.c_int), ("wgs_per_cu", ctypes.c_int), ("queues", ctypes.c_int),
                ("waves", ctypes.c_int)]

    def as_dict(self):
        return {"wg_threads": self.wg_threads, "wgs_per_cu": self.wgs_per_cu, "queues": self.queues,
                "waves": self.waves}


class ServeRing(ctypes.Structure):
    _fields_ = [("ring", ctypes.c_void_p), ("evk", ctypes.c_void_p), ("d_f", ctypes.c_void_p),
                ("d_h", ctypes.c_void_p), ("d_pool", ctypes.c_void_p), ("pool", ctypes.c_int), ("slots", ctypes.c_int)]


class ServeStats(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("groups", ctypes.c_long), ("mean_batch", ctypes.c_double),
                ("peak_reserved", ctypes.c_size_t), ("max_inflight", ctypes.c_int)]


class ServeRequest(ctypes.Structure):
    """yh_request; REQUEST_DTYPE is the same layout as a numpy record."""
    _fields_ = [("arrival", ctypes.c_double), ("rid", ctypes.c_uint64), ("kind", ctypes.c_int),
                ("ring", ctypes.c_int), ("a", ctypes.c_int), ("b", ctypes.c_int), ("msg", ctypes.c_void_p),
                ("result", ctypes.c_void_p), ("output", ctypes.c_void_p), ("done", ctypes.c_double)]


REQUEST_DTYPE = np.dtype({"names": [f[0] for f in ServeRequest._fields_],
                          "formats": ["<f8", "<u8", "<i4", "<i4", "<i4", "<i4", "<u8", "<u8", "<u8", "<f8"],
                          "offsets": [getattr(ServeRequest, f[0]).offset for f in ServeRequest._fields_],
                          "itemsize": ctypes.sizeof(ServeRequest)})
YH_REQ_ADD, YH_REQ_MULKS, YH_REQ_ENCDEC = 0, 1, 2

_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "yh_last_error": (ctypes.c_char_p, []),
    "yh_version": (ctypes.c_int, []),
    "yh_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "yh_ctx_destroy": (ctypes.c_int, [_vp]),
    "yh_ctx_stream": (_vp, [_vp]),
    "yh_ctx_sync": (ctypes.c_int, [_vp]),
    "yh_malloc": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "yh_free": (ctypes.c_int, [_vp, _vp]),
    "yh_memcpy_htod": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "yh_memcpy_dtoh": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "yh_memcpy_dtod": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "yh_memset": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_size_t]),
    "yh_ring_create": (ctypes.c_int, [_vp, ctypes.c_int, _u64p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                      ctypes.POINTER(_vp)]),
    "yh_ring_destroy": (ctypes.c_int, [_vp]),
    "yh_ring_get_info": (ctypes.c_int, [_vp, ctypes.POINTER(RingInfo)]),
    "yh_ring_plan": (ctypes.c_int, [ctypes.c_int, _u64p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                    ctypes.POINTER(RingInfo), _u64p, ctypes.c_int]),
    "yh_ring_get_primes": (ctypes.c_int, [_vp, _u64p, ctypes.c_int]),
    "yh_ring_get_modulus": (ctypes.c_int, [_vp, ctypes.c_int, _u64p, ctypes.c_int]),
    "yh_crt": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]),
    "yh_icrt": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int]),
    "yh_ntt": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "yh_pointwise_mulmod": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "yh_residue_addsub": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int]),
    "yh_residue_scalar": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "yh_crt_bn": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "yh_icrt_bn": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "yh_ring_scalar": (ctypes.c_int, [_vp, _vp, _u64p, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int]),
    "yh_poly_mul_plain": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int)]),
    "yh_poly_mul_plain_length": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "yh_ring_mul": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int]),
    "yh_ring_addsub": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int]),
    "yh_ring_reduce": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int]),
    "yh_yashe_mul": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int]),
    "yh_evk_create": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(_vp)]),
    "yh_evk_destroy": (ctypes.c_int, [_vp]),
    "yh_keyswitch": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int]),
    "yh_yashe_mul_keyswitch": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int]),
    "yh_yashe_mul_keyswitch_static": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int]),
    "yh_yashe_decrypt": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int]),
    "yh_mem_footprint": (ctypes.c_size_t, [_vp, ctypes.c_int]),
    "yh_wire_size": (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "yh_wire_encode": (ctypes.c_int, [ctypes.c_int, _u64p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                      _u64p, ctypes.c_int, _vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "yh_wire_decode": (ctypes.c_int, [ctypes.c_int, _u64p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                      _vp, ctypes.c_size_t, _u64p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "yh_wire_peek": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(WireHeader)]),
    "yh_ctx_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "yh_ctx_set_timing_mask": (ctypes.c_int, [_vp, ctypes.c_uint]),
    "yh_ctx_set_fuse_min_batch": (ctypes.c_int, [_vp, ctypes.c_int]),
    "yh_ctx_kernel_time": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_long)]),
    "yh_ctx_reset_timing": (ctypes.c_int, [_vp]),
    "yh_ring_mul_bcast": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int]),
    "yh_ring_inverse": (ctypes.c_int, [_vp, _vp, _vp]),
    "yh_graph_begin": (ctypes.c_int, [_vp]),
    "yh_graph_end": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "yh_graph_launch": (ctypes.c_int, [_vp]),
    "yh_graph_destroy": (ctypes.c_int, [_vp]),
    "yh_batch_default_config": (ctypes.c_int, [ctypes.POINTER(BatchItem), ctypes.c_int, ctypes.POINTER(BatchConfig)]),
    "yh_batch_create": (ctypes.c_int, [_vp, ctypes.POINTER(BatchItem), ctypes.c_int, ctypes.POINTER(BatchConfig),
                                       ctypes.POINTER(_vp)]),
    "yh_batch_run": (ctypes.c_int, [_vp]),
    "yh_ring_addsub_list": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int]),
    "yh_yashe_mul_keyswitch_list": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int]),
    "yh_yashe_decrypt_list": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int]),
    "yh_ring_use_own_stream": (ctypes.c_int, [_vp]),
    "yh_ring_sync": (ctypes.c_int, [_vp]),
    "yh_event_create": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "yh_event_record": (ctypes.c_int, [_vp, _vp, _vp]),
    "yh_event_query": (ctypes.c_int, [_vp]),
    "yh_event_sync": (ctypes.c_int, [_vp]),
    "yh_event_elapsed_ms": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_float)]),
    "yh_event_destroy": (ctypes.c_int, [_vp]),
    "yh_batch_destroy": (ctypes.c_int, [_vp]),
    "yh_batch_autotune": (ctypes.c_int, [_vp, ctypes.POINTER(BatchItem), ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                         ctypes.POINTER(BatchConfig)]),
    "yh_server_create": (ctypes.c_int, [_vp, ctypes.POINTER(ServeRing), ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "yh_server_set_budget": (ctypes.c_int, [_vp, ctypes.c_size_t]),
    "yh_server_footprint": (ctypes.c_size_t, [_vp, ctypes.c_int, ctypes.c_int]),
    "yh_server_run": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.POINTER(ServeStats)]),
    "yh_server_destroy": (ctypes.c_int, [_vp]),
}


def exported_symbols() -> List[str]:
    return list(_SIGS)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise YashError(f"{LIB_PATH} is not built; run `python -m cuyashe_amd.build` "
                            "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != YH_OK:
        msg = lib().yh_last_error()
        raise YashError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


# ------------------------------------------------------------------ helpers

def int_to_limbs(v: int, L: int) -> List[int]:
    mask = (1 << 64) - 1
    return [(v >> (64 * l)) & mask for l in range(L)]


def _scalar_limbs(v: int):
    """(ctypes limb array, limb count) of a non-negative integer."""
    v = int(v)
    if v < 0:
        raise YashError("scalars are non-negative integers")
    L = max(1, (v.bit_length() + 63) // 64)
    return (ctypes.c_uint64 * L)(*int_to_limbs(v, L)), L


def ints_to_array(values: Sequence[int], L: int) -> np.ndarray:
    """[L][n] uint64 limb-major array."""
    vals = [int(v) for v in values]
    out = np.empty((L, len(vals)), dtype=np.uint64)
    mask = (1 << 64) - 1
    for l in range(L):
        sh = 64 * l
        out[l] = np.fromiter(((v >> sh) & mask for v in vals), dtype=np.uint64, count=len(vals))
    return out


def array_to_ints(arr: np.ndarray) -> List[int]:
    arr = np.asarray(arr, dtype=np.uint64)
    L = arr.shape[0]
    vals = [int(x) for x in arr[L - 1]]
    for l in range(L - 2, -1, -1):
        row = arr[l]
        vals = [(v << 64) | int(x) for v, x in zip(vals, row)]
    return vals


def ring_plan(n: int, q: int, t: int = 0, w_bits: int = 0, keyswitch: bool = False):
    """Host-only parameter planning: (RingInfo, product moduli), or with
    keyswitch=True (RingInfo, product moduli, keyswitch moduli)."""
    q = int(q)
    L = max(1, (q.bit_length() + 63) // 64)
    ql = (ctypes.c_uint64 * L)(*int_to_limbs(q, L))
    info = RingInfo()
    buf = (ctypes.c_uint64 * 512)()
    check(lib().yh_ring_plan(n, ql, L, t, w_bits, ctypes.byref(info), buf, 512), "yh_ring_plan")
    prod = [int(buf[i]) for i in range(info.R)]
    if keyswitch:
        return info, prod, [int(buf[info.R + i]) for i in range(info.R_ks)]
    return info, prod


# ------------------------------------------------------------------ wire format

def wire_encode(n: int, q: int, t: int, w_bits: int, kind: int, polys: np.ndarray) -> bytes:
    """Blob of `polys` ([count][L][n] uint64 limbs, or [L][n] for one) for
    the ring (n, q, t, w_bits); see yh_wire_encode."""
    q = int(q)
    L = max(1, (q.bit_length() + 63) // 64)
    arr = np.ascontiguousarray(polys, dtype=np.uint64)
    if arr.ndim == 2:
        arr = arr[None]
    if arr.shape[1:] != (L, n):
        raise YashError(f"expected [count][{L}][{n}] limbs, got {arr.shape}")
    ql = (ctypes.c_uint64 * L)(*int_to_limbs(q, L))
    size = lib().yh_wire_size(n, L, arr.shape[0])
    out = ctypes.create_string_buffer(size)
    got = ctypes.c_size_t(0)
    check(lib().yh_wire_encode(n, ql, L, t, w_bits, kind, arr.ctypes.data_as(_u64p), arr.shape[0], out, size,
                               ctypes.byref(got)), "yh_wire_encode")
    return out.raw[:got.value]


def wire_peek(data: bytes) -> WireHeader:
    h = WireHeader()
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    check(lib().yh_wire_peek(buf, len(data), ctypes.byref(h)), "yh_wire_peek")
    return h


def wire_blob_size(data: bytes, offset: int = 0) -> int:
    """Size of the blob starting at `offset` (blobs can be concatenated)."""
    h = wire_peek(bytes(data[offset:offset + 96]))
    return int(lib().yh_wire_size(h.n, h.q_limbs, h.count))


def wire_decode(n: int, q: int, t: int, w_bits: int, kind: int, data: bytes) -> np.ndarray:
    """[count][L][n] uint64 limbs from a blob, after every check of
    yh_wire_decode (ring parameters, kind, length, checksum, ranges)."""
    q = int(q)
    L = max(1, (q.bit_length() + 63) // 64)
    ql = (ctypes.c_uint64 * L)(*int_to_limbs(q, L))
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    count = ctypes.c_int(0)
    check(lib().yh_wire_decode(n, ql, L, t, w_bits, kind, buf, len(data), None, 0, ctypes.byref(count)),
          "yh_wire_decode")
    out = np.empty((count.value, L, n), dtype=np.uint64)
    check(lib().yh_wire_decode(n, ql, L, t, w_bits, kind, buf, len(data), out.ctypes.data_as(_u64p), count.value,
                               ctypes.byref(count)), "yh_wire_decode")
    return out


class DeviceBuffer:
    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _vp()
        check(lib().yh_malloc(ctx.h, self.nbytes, ctypes.byref(p)), "yh_malloc")
        self.ptr = p.value

    def free(self):
        if self.ptr:
            check(lib().yh_free(self.ctx.h, self.ptr), "yh_free")
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                lib().yh_free(self.ctx.h, self.ptr)
        except Exception:
            pass

    def view(self, offset: int, nbytes: int = None) -> "BufferView":
        """A non-owning window [offset, offset + nbytes) of this buffer."""
        nbytes = self.nbytes - offset if nbytes is None else int(nbytes)
        if offset < 0 or nbytes < 0 or offset + nbytes > self.nbytes:
            raise YashError(f"view [{offset}, {offset + nbytes}) outside a buffer of {self.nbytes} bytes")
        return BufferView(self, offset, nbytes)

    def upload(self, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        check(lib().yh_memcpy_htod(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes), "yh_memcpy_htod")

    def download(self, shape, dtype=np.uint64) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib().yh_memcpy_dtoh(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes), "yh_memcpy_dtoh")
        return out


class BufferView(DeviceBuffer):
    """Window of a DeviceBuffer; keeps the owner alive, never frees."""

    def __init__(self, owner: DeviceBuffer, offset: int, nbytes: int):
        self.ctx = owner.ctx
        self.owner = owner
        self.nbytes = nbytes
        self.ptr = owner.ptr + offset

    def free(self):
        self.ptr = None

    def __del__(self):
        pass


class Context:
    def __init__(self, device: int = 0):
        h = _vp()
        check(lib().yh_ctx_create(device, ctypes.byref(h)), "yh_ctx_create")
        self.h = h.value

    @property
    def stream(self) -> int:
        return lib().yh_ctx_stream(self.h)

    def sync(self):
        check(lib().yh_ctx_sync(self.h), "yh_ctx_sync")

    def set_timing(self, enable: bool):
        check(lib().yh_ctx_set_timing(self.h, 1 if enable else 0), "yh_ctx_set_timing")

    def set_timing_mask(self, kernel_ids):
        """Time only these kernel ids while timing is on (None: all)."""
        mask = 0xFFFFFFFF if kernel_ids is None else sum(1 << k for k in kernel_ids)
        check(lib().yh_ctx_set_timing_mask(self.h, mask), "yh_ctx_set_timing_mask")

    def reset_timing(self):
        check(lib().yh_ctx_reset_timing(self.h), "yh_ctx_reset_timing")

    def set_fuse_min_batch(self, min_batch: int):
        """Batch size from which mul+keyswitch fuses the scale-and-round into
        the digit transforms (0: the device's CU count)."""
        check(lib().yh_ctx_set_fuse_min_batch(self.h, min_batch), "yh_ctx_set_fuse_min_batch")

    def kernel_time(self, kernel_id: int):
        """(total ms, launches) of one fused kernel since timing was enabled."""
        ms = ctypes.c_double(0)
        cnt = ctypes.c_long(0)
        check(lib().yh_ctx_kernel_time(self.h, kernel_id, ctypes.byref(ms), ctypes.byref(cnt)), "yh_ctx_kernel_time")
        return ms.value, cnt.value

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr: np.ndarray) -> DeviceBuffer:
        a = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, max(a.nbytes, 1))
        b.upload(a)
        return b

    def ntt(self, buf: DeviceBuffer, N: int, count: int, inverse: bool = False):
        check(lib().yh_ntt(self.h, buf.ptr, N, count, YH_INVERSE if inverse else YH_FORWARD), "yh_ntt")

    def pointwise(self, a: DeviceBuffer, b: DeviceBuffer, count: int):
        check(lib().yh_pointwise_mulmod(self.h, a.ptr, b.ptr, count), "yh_pointwise_mulmod")

    def graph_begin(self):
        """Start capturing the context stream's launches into a HIP graph."""
        check(lib().yh_graph_begin(self.h), "yh_graph_begin")

    def graph_end(self) -> "Graph":
        h = _vp()
        check(lib().yh_graph_end(self.h, ctypes.byref(h)), "yh_graph_end")
        return Graph(h.value)

    def close(self):
        """Destroys the context; refused (YashError, the context stays
        usable) while rings, batch plans or graphs created on it are alive."""
        if self.h:
            check(lib().yh_ctx_destroy(self.h), "yh_ctx_destroy")
            self.h = None


class Event:
    """A point in a ring's stream (or the context stream when ring is None)."""

    def __init__(self, ctx: "Context"):
        h = _vp()
        check(lib().yh_event_create(ctx.h, ctypes.byref(h)), "yh_event_create")
        self.h = h.value
        self.ctx = ctx

    def record(self, ring: Optional["Ring"] = None):
        check(lib().yh_event_record(self.h, self.ctx.h, ring.h if ring is not None else None), "yh_event_record")

    def done(self) -> bool:
        rc = lib().yh_event_query(self.h)
        if rc < 0:
            check(rc, "yh_event_query")
        return rc == 1

    def sync(self):
        check(lib().yh_event_sync(self.h), "yh_event_sync")

    def elapsed_ms(self, stop: "Event") -> float:
        ms = ctypes.c_float(0)
        check(lib().yh_event_elapsed_ms(self.h, stop.h, ctypes.byref(ms)), "yh_event_elapsed_ms")
        return ms.value

    def close(self):
        if self.h:
            lib().yh_event_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """An instantiated capture of ring operations, replayed on the context stream."""

    def __init__(self, h):
        self.h = h

    def launch(self):
        check(lib().yh_graph_launch(self.h), "yh_graph_launch")

    def close(self):
        if self.h:
            lib().yh_graph_destroy(self.h)
            self.h = None


class Ring:
    """One ring Z_q[x]/(x^n + 1) with its CRT primes and tables."""

    def __init__(self, ctx: Context, n: int, q: int, t: int = 0, w_bits: int = 0):
        self.ctx = ctx
        self.q = int(q)
        self.n = int(n)
        self.t = int(t)
        L = max(1, (self.q.bit_length() + 63) // 64)
        ql = (ctypes.c_uint64 * L)(*int_to_limbs(self.q, L))
        h = _vp()
        check(lib().yh_ring_create(ctx.h, n, ql, L, t, w_bits, ctypes.byref(h)), "yh_ring_create")
        self.h = h.value
        info = RingInfo()
        check(lib().yh_ring_get_info(self.h, ctypes.byref(info)), "yh_ring_get_info")
        self.info = info
        self.L = info.q_limbs
        self.R = info.R
        self.R_ks = info.R_ks
        self.lw = info.lw
        self.w_bits = info.w_bits
        self.LW = info.work_limbs

    def primes(self) -> List[int]:
        """Moduli of ring products (P first)."""
        buf = (ctypes.c_uint64 * self.R)()
        check(lib().yh_ring_get_primes(self.h, buf, self.R), "yh_ring_get_primes")
        return [int(v) for v in buf]

    def ks_primes(self) -> List[int]:
        """Moduli of the keyswitch multiply-accumulate (P first)."""
        buf = (ctypes.c_uint64 * (self.R + self.R_ks))()
        check(lib().yh_ring_get_primes(self.h, buf, self.R + self.R_ks), "yh_ring_get_primes")
        return [int(v) for v in buf][self.R:]

    def export(self, kind: int, buf: "DeviceBuffer", count: int) -> bytes:
        """Blob (wire format) of `count` device polynomials of this ring."""
        return wire_encode(self.n, self.q, self.t, self.w_bits, kind, buf.download((count, self.L, self.n)))

    def import_(self, kind: int, data: bytes, out: Optional["DeviceBuffer"] = None) -> "DeviceBuffer":
        """Device polynomials from a blob of this ring (checked on the host
        before anything is uploaded)."""
        arr = wire_decode(self.n, self.q, self.t, self.w_bits, kind, data)
        if out is None:
            return self.ctx.to_device(arr)
        out.upload(arr)
        return out

    def modulus(self, ks: bool = False) -> int:
        buf = (ctypes.c_uint64 * self.LW)()
        check(lib().yh_ring_get_modulus(self.h, 1 if ks else 0, buf, self.LW), "yh_ring_get_modulus")
        return sum(int(v) << (64 * i) for i, v in enumerate(buf))

    # -- array helpers: a batch of ring elements is [batch][L][n] uint64
    def pack(self, polys: Sequence[Sequence[int]]) -> np.ndarray:
        return np.stack([ints_to_array(list(p) + [0] * (self.n - len(p)), self.L) for p in polys])

    def unpack(self, arr: np.ndarray) -> List[List[int]]:
        return [array_to_ints(a) for a in arr]

    def footprint(self, op: int) -> int:
        return int(lib().yh_mem_footprint(self.h, op))

    # -- operations on device buffers
    def crt(self, coefs: DeviceBuffer, limbs: int, ncoef: int, spacing: int, out: DeviceBuffer):
        check(lib().yh_crt(self.h, coefs.ptr, limbs, ncoef, spacing, out.ptr), "yh_crt")

    def icrt(self, res: DeviceBuffer, spacing: int, ncoef: int, out: DeviceBuffer, out_limbs: int):
        check(lib().yh_icrt(self.h, res.ptr, spacing, ncoef, out.ptr, out_limbs), "yh_icrt")

    def residue_addsub(self, c: DeviceBuffer, a: DeviceBuffer, b: DeviceBuffer, spacing: int, sub: bool):
        check(lib().yh_residue_addsub(self.h, c.ptr, a.ptr, b.ptr, spacing, YH_SUB if sub else YH_ADD),
              "yh_residue_addsub")

    def residue_scalar(self, c: DeviceBuffer, a: DeviceBuffer, spacing: int, op: int, scalar_res: DeviceBuffer):
        """op (YH_ADD / YH_SUB on coefficient 0, YH_MUL on all) on residues
        [R][spacing], mod each CRT modulus, with the scalar given by its R
        residues (a device buffer of R words)."""
        check(lib().yh_residue_scalar(self.h, c.ptr, a.ptr, spacing, op, scalar_res.ptr), "yh_residue_scalar")

    def crt_bn(self, coefs_bn: DeviceBuffer, ncoef: int, spacing: int, out: DeviceBuffer):
        """CRT of coefficients held as a device array of yh_bn_t."""
        check(lib().yh_crt_bn(self.h, coefs_bn.ptr, ncoef, spacing, out.ptr), "yh_crt_bn")

    def icrt_bn(self, res: DeviceBuffer, spacing: int, ncoef: int, coefs_bn: DeviceBuffer):
        """ICRT into a device array of yh_bn_t."""
        check(lib().yh_icrt_bn(self.h, res.ptr, spacing, ncoef, coefs_bn.ptr), "yh_icrt_bn")

    def scalar(self, a: DeviceBuffer, op: int, scalar: int, out: DeviceBuffer, batch: int = 1):
        """Ring elements: op (YH_ADD / YH_SUB on coefficient 0, YH_MUL on all)
        with scalar mod q."""
        s, L = _scalar_limbs(scalar)
        check(lib().yh_ring_scalar(self.h, a.ptr, s, L, out.ptr, batch, op), "yh_ring_scalar")

    def poly_mul_plain(self, a: DeviceBuffer, deg_a: int, b: DeviceBuffer, deg_b: int, out: DeviceBuffer,
                       out_limbs: int, min_N: int = 1) -> int:
        N = ctypes.c_int(0)
        check(lib().yh_poly_mul_plain(self.h, a.ptr, deg_a, b.ptr, deg_b, min_N, out.ptr, out_limbs, ctypes.byref(N)),
              "yh_poly_mul_plain")
        return N.value

    def ring_mul(self, a: DeviceBuffer, b: DeviceBuffer, out: DeviceBuffer, batch: int = 1):
        check(lib().yh_ring_mul(self.h, a.ptr, b.ptr, out.ptr, batch), "yh_ring_mul")

    # -- pointer-list variants: lists are DeviceBuffers holding `batch` device pointers
    def addsub_list(self, a_list: DeviceBuffer, b_list: DeviceBuffer, out_list: DeviceBuffer, batch: int,
                    sub: bool = False):
        check(lib().yh_ring_addsub_list(self.h, a_list.ptr, b_list.ptr, out_list.ptr, batch, YH_SUB if sub else YH_ADD),
              "yh_ring_addsub_list")

    def mul_keyswitch_list(self, evk: "Evk", c1_list: DeviceBuffer, c2_list: DeviceBuffer, out_list: DeviceBuffer,
                           batch: int):
        check(lib().yh_yashe_mul_keyswitch_list(self.h, evk.h, c1_list.ptr, c2_list.ptr, out_list.ptr, batch),
              "yh_yashe_mul_keyswitch_list")

    def decrypt_list(self, f: DeviceBuffer, c_list: DeviceBuffer, m_list: DeviceBuffer, batch: int):
        check(lib().yh_yashe_decrypt_list(self.h, f.ptr, c_list.ptr, m_list.ptr, batch), "yh_yashe_decrypt_list")

    def use_own_stream(self):
        """Run this ring's work on a stream of its own (overlaps other rings)."""
        check(lib().yh_ring_use_own_stream(self.h), "yh_ring_use_own_stream")

    def sync(self):
        check(lib().yh_ring_sync(self.h), "yh_ring_sync")

    def ring_mul_bcast(self, a: DeviceBuffer, b: DeviceBuffer, out: DeviceBuffer, batch: int = 1):
        """out_k = a * b_k for one a and a batch of b."""
        check(lib().yh_ring_mul_bcast(self.h, a.ptr, b.ptr, out.ptr, batch), "yh_ring_mul_bcast")

    def inverse(self, f: DeviceBuffer, out: DeviceBuffer):
        """out = f^-1 mod (x^n + 1, q); raises if f is not invertible."""
        check(lib().yh_ring_inverse(self.h, f.ptr, out.ptr), "yh_ring_inverse")

    def ring_addsub(self, a: DeviceBuffer, b: DeviceBuffer, out: DeviceBuffer, batch: int = 1, sub: bool = False):
        check(lib().yh_ring_addsub(self.h, a.ptr, b.ptr, out.ptr, batch, YH_SUB if sub else YH_ADD), "yh_ring_addsub")

    def ring_reduce(self, c: DeviceBuffer, in_limbs: int, out: DeviceBuffer, batch: int = 1):
        check(lib().yh_ring_reduce(self.h, c.ptr, in_limbs, out.ptr, batch), "yh_ring_reduce")

    def yashe_mul(self, c1: DeviceBuffer, c2: DeviceBuffer, out: DeviceBuffer, batch: int = 1):
        check(lib().yh_yashe_mul(self.h, c1.ptr, c2.ptr, out.ptr, batch), "yh_yashe_mul")

    def evk(self, gamma: DeviceBuffer) -> "Evk":
        return Evk(self, gamma)

    def keyswitch(self, evk: "Evk", c: DeviceBuffer, out: DeviceBuffer, batch: int = 1):
        check(lib().yh_keyswitch(self.h, evk.h, c.ptr, out.ptr, batch), "yh_keyswitch")

    def mul_keyswitch(self, evk: "Evk", c1: DeviceBuffer, c2: DeviceBuffer, out: DeviceBuffer, batch: int = 1):
        check(lib().yh_yashe_mul_keyswitch(self.h, evk.h, c1.ptr, c2.ptr, out.ptr, batch), "yh_yashe_mul_keyswitch")

    def mul_keyswitch_static(self, evk: "Evk", c1: DeviceBuffer, c2: DeviceBuffer, out: DeviceBuffer,
                             batch: int = 1):
        """mul_keyswitch with the library's static device scratch (no allocation)."""
        check(lib().yh_yashe_mul_keyswitch_static(self.h, evk.h, c1.ptr, c2.ptr, out.ptr, batch),
              "yh_yashe_mul_keyswitch_static")

    def decrypt(self, f: DeviceBuffer, c: DeviceBuffer, m_out: DeviceBuffer, batch: int = 1):
        check(lib().yh_yashe_decrypt(self.h, f.ptr, c.ptr, m_out.ptr, batch), "yh_yashe_decrypt")

    def close(self):
        if self.h:
            check(lib().yh_ring_destroy(self.h), "yh_ring_destroy")
            self.h = None


def _batch_items(items) -> "ctypes.Array":
    arr = (BatchItem * len(items))()
    for k, (ring, a, b, out) in enumerate(items):
        arr[k] = BatchItem(ring.h, a.ptr, b.ptr, out.ptr)
    return arr


def _config(cfg) -> Optional[BatchConfig]:
    if cfg is None:
        return None
    if isinstance(cfg, BatchConfig):
        return cfg
    return BatchConfig(int(cfg["wg_threads"]), int(cfg["wgs_per_cu"]), int(cfg["queues"]), int(cfg.get("waves", 3)))


class Batch:
    """Uneven products c_k = a_k b_k mod (x^n_k + 1, q_k) in one persistent
    launch. items: sequence of (Ring, a, b, out) device buffers."""

    def __init__(self, ctx: Context, items, config=None):
        self.ctx = ctx
        self._keep = list(items)
        arr = _batch_items(items)
        c = _config(config)
        h = _vp()
        check(lib().yh_batch_create(ctx.h, arr, len(items), ctypes.byref(c) if c is not None else None,
                                    ctypes.byref(h)), "yh_batch_create")
        self.h = h.value

    def run(self):
        check(lib().yh_batch_run(self.h), "yh_batch_run")

    def close(self):
        if self.h:
            check(lib().yh_batch_destroy(self.h), "yh_batch_destroy")
            self.h = None

    def __del__(self):
        # never after the context is gone (interpreter shutdown)
        try:
            if self.h and self.ctx.h:
                self.close()
        except Exception:
            pass


def batch_default_config(items) -> dict:
    arr = _batch_items(items)
    c = BatchConfig()
    check(lib().yh_batch_default_config(arr, len(items), ctypes.byref(c)), "yh_batch_default_config")
    return c.as_dict()


def batch_autotune(ctx: Context, items, cache_path: Optional[str] = None, reps: int = 3) -> dict:
    """Fastest launch shape for this batch composition (cached in cache_path)."""
    arr = _batch_items(items)
    c = BatchConfig()
    path = cache_path.encode() if cache_path else None
    check(lib().yh_batch_autotune(ctx.h, arr, len(items), path, reps, ctypes.byref(c)), "yh_batch_autotune")
    return c.as_dict()


class Evk:
    def __init__(self, ring: Ring, gamma: DeviceBuffer):
        h = _vp()
        check(lib().yh_evk_create(ring.h, gamma.ptr, ctypes.byref(h)), "yh_evk_create")
        self.h = h.value
        self.ring = ring

    def close(self):
        if self.h:
            check(lib().yh_evk_destroy(self.h), "yh_evk_destroy")
            self.h = None
