"""The allocation-free mul+keyswitch (yh_yashe_mul_keyswitch_static)
against the allocating entry point, through the C ABI: two calls in a row on
the shared static scratch, the second with a smaller batch, each bit for bit
equal to yh_yashe_mul_keyswitch, with and without the scale-and-round inside
the digit transforms; a batch beyond the static array is refused."""
import random

import pytest

import cuyashe_amd as ya

pytestmark = pytest.mark.gpu

Q127 = (1 << 127) - 1
Q255 = (1 << 255) - 19


def rand_poly(rng, n, q):
    return [rng.randrange(q) for _ in range(n)]


@pytest.mark.parametrize("n,q,batches", [(4096, Q127, (96, 5)), (8192, Q255, (12, 3)), (64, Q127, (40, 7))])
def test_static_scratch_equals_allocating(gpu_ctx, n, q, batches):
    rng = random.Random(n + len(batches))
    ring = ya.Ring(gpu_ctx, n, q, t=1024, w_bits=32)
    evk = ring.evk(gpu_ctx.to_device(ring.pack([rand_poly(rng, n, q) for _ in range(ring.lw)])))
    for batch in batches:
        C1 = [rand_poly(rng, n, q) for _ in range(batch)]
        C2 = [rand_poly(rng, n, q) for _ in range(batch)]
        d1, d2 = gpu_ctx.to_device(ring.pack(C1)), gpu_ctx.to_device(ring.pack(C2))
        ref = gpu_ctx.alloc(batch * ring.L * n * 8)
        got = gpu_ctx.alloc(batch * ring.L * n * 8)
        try:
            # reference: the allocating entry point with separate kernels;
            # the static one as the default batch rule takes it, then with
            # the scale-and-round inside the digit transforms where the ring
            # allows it
            gpu_ctx.set_fuse_min_batch(1 << 30)
            ring.mul_keyswitch(evk, d1, d2, ref, batch)
            want = ref.download((batch, ring.L, n))
            for min_batch in (0, 1):
                gpu_ctx.set_fuse_min_batch(min_batch)
                ring.mul_keyswitch_static(evk, d1, d2, got, batch)
                gpu_ctx.sync()
                assert (got.download((batch, ring.L, n)) == want).all(), (batch, min_batch)
        finally:
            gpu_ctx.set_fuse_min_batch(0)
    evk.close()
    ring.close()


def test_static_scratch_refuses_oversized_batch(gpu_ctx):
    rng = random.Random(1)
    n, q = 16384, Q255
    ring = ya.Ring(gpu_ctx, n, q, t=1024, w_bits=32)
    evk = ring.evk(gpu_ctx.to_device(ring.pack([rand_poly(rng, n, q) for _ in range(ring.lw)])))
    batch = 64  # ~ 64 x 2.5 MB of scratch, beyond the 64 MiB array
    d = gpu_ctx.alloc(batch * ring.L * n * 8)
    with pytest.raises(ya.YashError, match="static array"):
        ring.mul_keyswitch_static(evk, d, d, d, batch)
    evk.close()
    ring.close()
