"""CPU checks of the wire format (yh_wire_* in include/yashe_hip.h): the byte
layout read back with struct, the checksum against zlib's CRC-32, round trips,
and every rejection a corrupted or mismatched blob must meet. No device is
involved: the format works on host memory."""
import random
import struct
import zlib

import numpy as np
import pytest

import cuyashe_amd as ya

Q127 = (1 << 127) - 1
Q255 = (1 << 255) - 19


def rand_polys(rng, count, n, q):
    return np.stack([ya.ints_to_array([rng.randrange(q) for _ in range(n)], max(1, (q.bit_length() + 63) // 64))
                     for _ in range(count)])


@pytest.mark.parametrize("n,q,t,wb,kind,count", [
    (16, Q127, 1024, 32, ya.WIRE_CIPHERTEXT, 3),
    (4096, Q255, 1 << 16, 32, ya.WIRE_KEYSWITCH_KEY, 8),
    (64, 77287149995192912462927307869, 35951, 32, ya.WIRE_PUBLIC_KEY, 1),
    (2, 1171313591017775093490277364417, 0, 0, ya.WIRE_CIPHERTEXT, 0),
])
def test_round_trip_and_layout(n, q, t, wb, kind, count):
    rng = random.Random(n + count)
    L = max(1, (q.bit_length() + 63) // 64)
    polys = rand_polys(rng, count, n, q) if count else np.zeros((0, L, n), np.uint64)
    if count:
        polys[0, :, 0] = ya.ints_to_array([q - 1], L)[:, 0]  # largest coefficient
    blob = ya.wire_encode(n, q, t, wb, kind, polys)
    assert len(blob) == 32 + 8 * L + 8 * L * n * count + 4
    magic, ver, k, n_, L_, t_, wb_, cnt = struct.unpack_from("<4sHHIIQII", blob, 0)
    assert (magic, ver, k, n_, L_, t_, wb_, cnt) == (b"YHW1", 1, kind, n, L, t, wb, count)
    qq = struct.unpack_from(f"<{L}Q", blob, 32)
    assert sum(v << (64 * i) for i, v in enumerate(qq)) == q
    payload = np.frombuffer(blob, dtype="<u8", count=L * n * count, offset=32 + 8 * L)
    assert np.array_equal(payload.reshape(count, L, n), polys)
    assert struct.unpack_from("<I", blob, len(blob) - 4)[0] == zlib.crc32(blob[:-4])
    back = ya.wire_decode(n, q, t, wb, kind, blob)
    assert back.dtype == np.uint64 and np.array_equal(back, polys)
    h = ya.wire_peek(blob)
    assert (h.version, h.kind, h.n, h.q_limbs, h.t, h.w_bits, h.count) == (1, kind, n, L, t, wb, count)
    assert h.modulus() == q
    assert ya.wire_blob_size(blob) == len(blob)


def _blob():
    rng = random.Random(7)
    polys = rand_polys(rng, 2, 16, Q127)
    return polys, ya.wire_encode(16, Q127, 1024, 32, ya.WIRE_CIPHERTEXT, polys)


def test_every_single_bit_flip_is_rejected():
    _, blob = _blob()
    for pos in range(len(blob)):
        for bit in (0, 7):
            bad = bytearray(blob)
            bad[pos] ^= 1 << bit
            with pytest.raises(ya.YashError):
                ya.wire_decode(16, Q127, 1024, 32, ya.WIRE_CIPHERTEXT, bytes(bad))


def test_mismatches_and_truncation_are_rejected():
    polys, blob = _blob()
    dec = lambda *a: ya.wire_decode(*a)  # noqa: E731
    for args in [(32, Q127, 1024, 32), (16, Q255, 1024, 32), (16, Q127 - 2, 1024, 32), (16, Q127, 1023, 32),
                 (16, Q127, 1024, 16)]:
        with pytest.raises(ya.YashError, match="another ring"):
            dec(*args, ya.WIRE_CIPHERTEXT, blob)
    with pytest.raises(ya.YashError, match="another kind"):
        dec(16, Q127, 1024, 32, ya.WIRE_SECRET_KEY, blob)
    assert np.array_equal(dec(16, Q127, 1024, 32, 0, blob), polys)  # kind 0 accepts any
    for cut in (0, 10, 31, 40, len(blob) - 1):
        with pytest.raises(ya.YashError):
            dec(16, Q127, 1024, 32, ya.WIRE_CIPHERTEXT, blob[:cut])
    with pytest.raises(ya.YashError, match="length"):
        dec(16, Q127, 1024, 32, ya.WIRE_CIPHERTEXT, blob + b"\0")
    bad = bytearray(blob)
    struct.pack_into("<H", bad, 4, 2)
    bad[-4:] = struct.pack("<I", zlib.crc32(bytes(bad[:-4])))
    with pytest.raises(ya.YashError, match="version"):
        dec(16, Q127, 1024, 32, ya.WIRE_CIPHERTEXT, bytes(bad))


def test_out_of_range_coefficients_are_rejected():
    rng = random.Random(9)
    polys = rand_polys(rng, 1, 16, Q127)
    polys[0, :, 5] = ya.ints_to_array([Q127], 2)[:, 0]  # exactly q
    with pytest.raises(ya.YashError, match="not below q"):
        ya.wire_encode(16, Q127, 1024, 32, ya.WIRE_CIPHERTEXT, polys)
    # a blob whose checksum is right but whose payload holds q itself
    good = rand_polys(rng, 1, 16, Q127)
    blob = bytearray(ya.wire_encode(16, Q127, 1024, 32, ya.WIRE_CIPHERTEXT, good))
    off = 32 + 16 + 8 * 5
    struct.pack_into("<Q", blob, off, Q127 & ((1 << 64) - 1))
    struct.pack_into("<Q", blob, off + 8 * 16, Q127 >> 64)
    blob[-4:] = struct.pack("<I", zlib.crc32(bytes(blob[:-4])))
    with pytest.raises(ya.YashError, match="not below q"):
        ya.wire_decode(16, Q127, 1024, 32, ya.WIRE_CIPHERTEXT, bytes(blob))
    # plaintexts are bounded by t
    m = ya.ints_to_array([rng.randrange(1024) for _ in range(16)], 2)[None]
    pb = ya.wire_encode(16, Q127, 1024, 32, ya.WIRE_PLAINTEXT, m)
    assert np.array_equal(ya.wire_decode(16, Q127, 1024, 32, ya.WIRE_PLAINTEXT, pb), m)
    m[0, 0, 3] = 1024
    with pytest.raises(ya.YashError, match="below t"):
        ya.wire_encode(16, Q127, 1024, 32, ya.WIRE_PLAINTEXT, m)


def test_concatenated_blobs_walk():
    rng = random.Random(3)
    a = ya.wire_encode(16, Q127, 1024, 32, ya.WIRE_PUBLIC_KEY, rand_polys(rng, 1, 16, Q127))
    b = ya.wire_encode(16, Q127, 1024, 32, ya.WIRE_KEYSWITCH_KEY, rand_polys(rng, 4, 16, Q127))
    data = a + b
    off, kinds = 0, []
    while off < len(data):
        size = ya.wire_blob_size(data, off)
        kinds.append(ya.wire_peek(data[off:off + size]).kind)
        off += size
    assert kinds == [ya.WIRE_PUBLIC_KEY, ya.WIRE_KEYSWITCH_KEY] and off == len(data)
