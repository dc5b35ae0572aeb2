"""Uneven batches of ring products in one persistent launch (and its autotuner).

Every product of a mixed batch must be bit-identical to the same product run
alone (yh_ring_mul) and to the oracle.
"""
import json

import numpy as np
import pytest

import cuyashe_amd as ya
from cuyashe_amd import workloads
from oracle import cpuref

pytestmark = pytest.mark.gpu

Q127 = (1 << 127) - 1
Q255 = (1 << 255) - 19
Q511 = (1 << 511) - 187


def make_batch(ctx, spec, seed):
    return workloads.make_batch(ctx, spec, seed)


def alone(ctx, rings, bufs):
    """Each ring's products through the per-ring path."""
    out = {}
    for key, (A, B, OUT, ha, hb, idx) in bufs.items():
        ring = rings[key]
        ref = ctx.alloc(ha.nbytes)
        ring.ring_mul(A, B, ref, len(idx))
        out[key] = ref.download(ha.shape)
    return out


def c3_spec(count, seed):
    return workloads.uneven_spec(count, seed)


def test_batch_small_rings_match_oracle(gpu_ctx):
    spec = [(16, Q127), (32, Q255), (64, Q511), (16, Q511), (256, Q127), (1024, Q255), (64, Q127), (128, Q255),
            (512, Q511), (32, Q127), (2048, Q127), (16, Q255)]
    rings, bufs, items = make_batch(gpu_ctx, spec, 1)
    b = ya.Batch(gpu_ctx, items)
    b.run()
    for key, (A, B, OUT, ha, hb, idx) in bufs.items():
        n, q = key
        got = OUT.download(ha.shape)
        for m in range(len(idx)):
            a = ya.array_to_ints(ha[m])
            bb = ya.array_to_ints(hb[m])
            assert ya.array_to_ints(got[m]) == cpuref.ring_mul(a, bb, n, q), (n, q, m)
    b.close()
    for r in rings.values():
        r.close()


def test_batch_c3_equals_alone_and_oracle(gpu_ctx):
    """512 products, n from {1024, 2048, 4096, 8192} weighted .4/.3/.2/.1, q
    127- or 255-bit: one persistent launch, bit-identical to each product
    alone; one product of every n against the oracle."""
    spec = c3_spec(512, 7)
    rings, bufs, items = make_batch(gpu_ctx, spec, 7)
    b = ya.Batch(gpu_ctx, items)
    b.run()
    ref = alone(gpu_ctx, rings, bufs)
    seen = set()
    for key, (A, B, OUT, ha, hb, idx) in bufs.items():
        got = OUT.download(ha.shape)
        assert np.array_equal(got, ref[key]), key
        n, q = key
        if n not in seen:
            seen.add(n)
            want = cpuref.ring_mul(ya.array_to_ints(ha[0]), ya.array_to_ints(hb[0]), n, q)
            assert ya.array_to_ints(got[0]) == want, key
    # a second run over the same plan (counters reset by the launch itself)
    for key, (A, B, OUT, ha, hb, idx) in bufs.items():
        ya.lib().yh_memset(gpu_ctx.h, OUT.ptr, 0, OUT.nbytes)
    b.run()
    for key, (A, B, OUT, ha, hb, idx) in bufs.items():
        assert np.array_equal(OUT.download(ha.shape), ref[key]), key
    b.close()
    for r in rings.values():
        r.close()


def test_batch_with_n16384_rings(gpu_ctx):
    """A mixed batch reaching n = 16384 (q 127-, 255- and 511-bit beside
    smaller rings) runs in one launch of 1024-thread workgroups: every product
    equals the product alone, and each n = 16384 ring matches the oracle."""
    spec = [(16384, Q127), (16384, Q255), (16384, Q511), (8192, Q511), (4096, Q255), (2048, Q127),
            (16384, Q127), (1024, Q511), (64, Q127), (16384, Q255)]
    rings, bufs, items = make_batch(gpu_ctx, spec, 13)
    b = ya.Batch(gpu_ctx, items)
    b.run()
    ref = alone(gpu_ctx, rings, bufs)
    for key, (A, B, OUT, ha, hb, idx) in bufs.items():
        got = OUT.download(ha.shape)
        assert np.array_equal(got, ref[key]), key
        n, q = key
        if n == 16384:
            want = cpuref.ring_mul(ya.array_to_ints(ha[0]), ya.array_to_ints(hb[0]), n, q)
            assert ya.array_to_ints(got[0]) == want, key
    # the same plan again, and a second shape of the wide workgroup
    for key, (A, B, OUT, ha, hb, idx) in bufs.items():
        ya.lib().yh_memset(gpu_ctx.h, OUT.ptr, 0, OUT.nbytes)
    b.run()
    for key, (A, B, OUT, ha, hb, idx) in bufs.items():
        assert np.array_equal(OUT.download(ha.shape), ref[key]), key
    b.close()
    b = ya.Batch(gpu_ctx, items, {"wg_threads": 1024, "wgs_per_cu": 1, "queues": 1, "waves": 4})
    b.run()
    for key, (A, B, OUT, ha, hb, idx) in bufs.items():
        assert np.array_equal(OUT.download(ha.shape), ref[key]), key
    b.close()
    for r in rings.values():
        r.close()


@pytest.mark.parametrize("cfg", [
    {"wg_threads": 512, "wgs_per_cu": 1, "queues": 1, "waves": 3},
    {"wg_threads": 512, "wgs_per_cu": 2, "queues": 8, "waves": 4},
    {"wg_threads": 512, "wgs_per_cu": 4, "queues": 3, "waves": 3},
])
def test_batch_shapes_agree(gpu_ctx, cfg):
    spec = c3_spec(96, 11)
    rings, bufs, items = make_batch(gpu_ctx, spec, 11)
    ref = alone(gpu_ctx, rings, bufs)
    b = ya.Batch(gpu_ctx, items, cfg)
    b.run()
    for key, (A, B, OUT, ha, hb, idx) in bufs.items():
        assert np.array_equal(OUT.download(ha.shape), ref[key]), (cfg, key)
    b.close()
    for r in rings.values():
        r.close()


def test_batch_rejects_bad_shapes(gpu_ctx):
    rings, bufs, items = make_batch(gpu_ctx, [(8192, Q127), (1024, Q127)], 3)
    with pytest.raises(ya.YashError):
        ya.Batch(gpu_ctx, items, {"wg_threads": 256, "wgs_per_cu": 1, "queues": 8})  # n/16 = 512 > 256
    with pytest.raises(ya.YashError):
        ya.Batch(gpu_ctx, items, {"wg_threads": 384, "wgs_per_cu": 1, "queues": 8})
    with pytest.raises(ya.YashError):
        ya.Batch(gpu_ctx, items, {"wg_threads": 512, "wgs_per_cu": 1, "queues": 8, "waves": 5})
    r16k = ya.Ring(gpu_ctx, 16384, Q127)
    buf = gpu_ctx.alloc(r16k.L * 16384 * 8)
    with pytest.raises(ya.YashError, match="wg_threads"):  # n/16 = 1024 > 512
        ya.Batch(gpu_ctx, [(r16k, buf, buf, buf)], {"wg_threads": 512, "wgs_per_cu": 1, "queues": 8, "waves": 4})
    with pytest.raises(ya.YashError, match="four waves"):
        ya.Batch(gpu_ctx, [(r16k, buf, buf, buf)], {"wg_threads": 1024, "wgs_per_cu": 1, "queues": 8, "waves": 3})
    r16k.close()
    for r in rings.values():
        r.close()


def test_batch_autotune_cache_and_graph(gpu_ctx, tmp_path):
    spec = c3_spec(64, 5)
    rings, bufs, items = make_batch(gpu_ctx, spec, 5)
    ref = alone(gpu_ctx, rings, bufs)
    cache = str(tmp_path / "tune.json")
    cfg = ya.batch_autotune(gpu_ctx, items, cache, reps=2)
    data = json.load(open(cache))
    assert len(data) == 1
    (key, entry), = data.items()
    assert {k: entry[k] for k in cfg} == cfg and entry["ms"] > 0
    assert key.split("|")[1].count("x") == len({(n, q) for n, q in spec})  # one histogram entry per ring
    # second call: served from the cache
    assert ya.batch_autotune(gpu_ctx, items, cache, reps=2) == cfg
    b = ya.Batch(gpu_ctx, items, cfg)
    b.run()
    for key, (A, B, OUT, ha, hb, idx) in bufs.items():
        assert np.array_equal(OUT.download(ha.shape), ref[key]), key
    # the tuned batch captured into a graph and replayed
    gpu_ctx.graph_begin()
    b.run()
    g = gpu_ctx.graph_end()
    for rep in range(3):
        for key, (A, B, OUT, ha, hb, idx) in bufs.items():
            ya.lib().yh_memset(gpu_ctx.h, OUT.ptr, 0, OUT.nbytes)
        gpu_ctx.sync()
        g.launch()
        gpu_ctx.sync()
        for key, (A, B, OUT, ha, hb, idx) in bufs.items():
            assert np.array_equal(OUT.download(ha.shape), ref[key]), (rep, key)
    g.close()
    b.close()
    for r in rings.values():
        r.close()
