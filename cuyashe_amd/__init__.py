"""MI355X-native (gfx950) YASHE polynomial-ring hot path.

The compute lives in ``lib/libyashe_hip.so`` (hand-written HIP kernels behind
the C ABI of include/yashe_hip.h); this package only binds it for Python
callers (tests, benchmarks, the serving loop).
"""
from ._lib import (  # noqa: F401
    Context, Ring, Evk, Graph, DeviceBuffer, YashError, lib, LIB_PATH,
    Batch, Event, batch_autotune, batch_default_config,
    ints_to_array, array_to_ints, int_to_limbs, ring_plan,
    OP_RING_MUL, OP_YASHE_MUL, OP_KEYSWITCH, OP_MUL_KS, OP_ADD,
    KERNEL_PROD, KERNEL_KEYSWITCH, KERNEL_ICRT, KERNEL_BATCH, KERNEL_KS_DIGITS,
    WIRE_CIPHERTEXT, WIRE_KEYSWITCH_KEY, WIRE_PUBLIC_KEY, WIRE_SECRET_KEY, WIRE_PLAINTEXT,
    wire_encode, wire_decode, wire_peek, wire_blob_size,
    YH_ADD, YH_SUB, YH_MUL,
)
