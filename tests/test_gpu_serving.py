"""The continuous-batching serving loop on the GPU: every decryption checked,
admission by HBM footprint, results identical alone and in the mix."""
import copy

import pytest

import cuyashe_amd as ya
from cuyashe_amd import workloads
from cuyashe_amd.serving import Request, Server, poisson_requests

pytestmark = pytest.mark.gpu

RINGS = [(2048, workloads.Q127), (4096, workloads.Q255), (2048, workloads.Q511)]


@pytest.fixture(scope="module")
def server(gpu_ctx):
    srv = Server(gpu_ctx, RINGS, t=1024, pool=16, slots=48, max_batch=32, seed=5, keep_outputs=True)
    yield srv
    srv.close()


def test_serving_mix_all_decrypts(server):
    reqs = poisson_requests(400, 20000.0, RINGS, 1024, 16, seed=1)
    st = server.run(reqs)
    assert st["all_decrypts_ok"], [r.rid for r in reqs if not r.ok][:10]
    assert all(r.done is not None and r.latency > 0 for r in reqs)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    assert st["mean_batch"] >= 1.0


def test_serving_burst_is_batched(server):
    # a burst: requests arriving while their (ring, kind) has groups in
    # flight wait and go out together
    reqs = poisson_requests(600, 1e6, RINGS, 1024, 16, seed=3)
    st = server.run(reqs)
    assert st["all_decrypts_ok"]
    assert st["mean_batch"] > 2.0, st


def test_serving_holds_rings_and_keys(server):
    rs = next(iter(server.rings.values()))
    with pytest.raises(ya.YashError, match="request loop"):
        rs.ring.close()
    with pytest.raises(ya.YashError, match="request loop"):
        rs.Y.evk.close()


def test_serving_alone_equals_mix(server):
    reqs = poisson_requests(150, 50000.0, RINGS, 1024, 16, seed=2)
    st = server.run(reqs)
    assert st["all_decrypts_ok"]
    picked = [r for r in reqs if r.kind == "mulks"][:6] + [r for r in reqs if r.kind == "add"][:4] + \
        [r for r in reqs if r.kind == "encdec"][:4]
    for r in picked:
        alone = copy.copy(r)
        alone.arrival, alone.done, alone.ok, alone.output = 0.0, None, None, None
        server.run([alone])
        assert alone.ok
        assert (alone.output == r.output).all(), (r.rid, r.kind)


def test_serving_admission_budget(gpu_ctx):
    one = Server(gpu_ctx, RINGS[:1], t=1024, pool=8, slots=16, max_batch=16, seed=9)
    fp = max(one.footprint(Request(0, k, RINGS[0])) for k in ("add", "mulks", "encdec"))
    one.budget = 3 * fp
    reqs = poisson_requests(60, 1e6, RINGS[:1], 1024, 8, seed=4)  # a burst
    st = one.run(reqs)
    assert st["all_decrypts_ok"]
    assert st["peak_reserved_bytes"] <= 3 * fp
    assert st["groups"] >= 12  # only a few requests in flight at a time
    big = max(("add", "mulks", "encdec"), key=lambda k: one.footprint(Request(0, k, RINGS[0])))
    one.budget = fp - 1
    with pytest.raises(ValueError):
        one.run([Request(0, big, RINGS[0], a=0, b=1, msg=reqs[0].msg)])
    one.close()


def test_serving_all_c5_rings(gpu_ctx):
    # every ring of the serving configuration: n in {2048, 4096, 8192, 16384}
    # and q of 127, 255 and 511 bits; each result decrypted and checked, and
    # one multiplication per ring equal alone and in the mix
    rings = workloads_c5()
    srv = Server(gpu_ctx, rings, t=1024, pool=4, slots=8, max_batch=8, seed=11, keep_outputs=True)
    reqs = poisson_requests(180, 20000.0, rings, 1024, 4, seed=12)
    st = srv.run(reqs)
    assert st["all_decrypts_ok"], [r.rid for r in reqs if not r.ok][:10]
    assert {r.ring for r in reqs} == set(rings)
    for key in rings:
        mul = [r for r in reqs if r.ring == key and r.kind == "mulks"][:1]
        for r in mul:
            alone = copy.copy(r)
            alone.arrival, alone.done, alone.ok, alone.output = 0.0, None, None, None
            srv.run([alone])
            assert alone.ok and (alone.output == r.output).all(), key
    srv.close()


def workloads_c5():
    from cuyashe_amd.serving import c5_rings
    rings = c5_rings()
    assert len(rings) == 12
    return rings


def test_serving_refuses_bad_requests(server):
    # the native loop validates every request before running any
    import numpy as np
    key = RINGS[0]
    with pytest.raises(ValueError):
        server.run([Request(1, "nope", key)])
    with pytest.raises(ya.YashError, match="pool index"):
        server.run([Request(2, "add", key, a=0, b=10 ** 6)])
    msg = np.zeros(key[0], dtype=np.int64)
    msg[3] = 1024  # = t
    with pytest.raises(ya.YashError, match="below t"):
        server.run([Request(3, "encdec", key, msg=msg)])
    st = server.run([Request(4, "add", key, a=0, b=1), Request(5, "encdec", key, msg=np.ones(key[0], dtype=np.int64))])
    assert st["all_decrypts_ok"] and st["requests"] == 2
