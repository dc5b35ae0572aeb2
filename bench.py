#!/usr/bin/env python3
"""Benchmark: bit-exact YASHE multiply + keyswitch at n=4096, q=2^127-1,
t=1024, w=2^32 on one MI355X.

A step is one batched multiply+keyswitch over `--batch` resident ciphertext
pairs (default 2048, where throughput saturates: every launch then runs many
rounds of workgroups and the fixed cost per launch is a few per cent; the
`batch_sweep` field reports 512, 1024 and 4096 beside it) (hot path: CRT -> negacyclic Goldilocks NTTs -> pointwise -> INTT ->
ICRT with t/q rounding, then digit decomposition -> NTT -> multiply-
accumulate with the resident keys -> INTT -> ICRT mod q). Inputs sit in HBM
before the timed region. Prints one JSON line.
"""
import argparse
import json
import os
import random
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "bit-exact YASHE mul+keyswitch ops/sec (n=4096,q~2^127); batched vs plain; p99 ms"
N_RING = 4096
Q = (1 << 127) - 1
T = 1024
W_BITS = 32
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--sweep", type=str, default="512,1024,4096",
                    help="other batch sizes timed after the main measurement (empty: none)")
    ap.add_argument("--no-two-streams", dest="two_streams", action="store_false",
                    help="skip the two-stream measurement")
    ap.add_argument("--plain-ops", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-batched", action="store_true",
                    help="run only the warmup + timed batched steps (for rocprofv3 runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a residue add over three 256 MiB buffers (known HBM traffic) for counter calibration")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the uneven-batch and graph-circuit measurements")
    ap.add_argument("--only-uneven", action="store_true",
                    help="run only the uneven batched products (for rocprofv3 runs)")
    ap.add_argument("--only-circuit", action="store_true",
                    help="run only the graph-captured circuit (for rocprofv3 runs)")
    ap.add_argument("--only-ntt", action="store_true", help="run only the yh_ntt primitive measurement")
    ap.add_argument("--serving", action="store_true",
                    help="run only the serving loop over a sweep of arrival rates")
    ap.add_argument("--serving-requests", type=int, default=2000)
    ap.add_argument("--serving-qps", type=str, default="2000,5000,10000,20000")
    return ap.parse_args()


TUNE_CACHE = os.path.join(ROOT, "cuyashe_amd", "tune_cache.json")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
SQ_FILE = os.path.join(ROOT, "profiles", "sq_counters_mul_keyswitch_n4096.md")


def measured_valu(kernel_symbol, batch):
    """SQ_INSTS_VALU per launch of the kernel from the committed rocprofv3
    SQ-counter summary (the count is fixed by the code and the batch), or
    None when the summary was taken at another batch size."""
    import re
    try:
        with open(SQ_FILE) as f:
            m = re.search(r"batch (\d+)", f.readline())
            if not m or int(m.group(1)) != batch:
                return None
            for line in f:
                cells = [c.strip() for c in line.split("|")]
                if len(cells) > 3 and cells[2] == "SQ_INSTS_VALU" and kernel_symbol in cells[1]:
                    return float(cells[3])
    except (OSError, ValueError):
        pass
    return None


def measured_traffic(kernel, batch):
    """HBM bytes per launch of `kernel` at this batch from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, calibrated), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if not line.startswith("traffic "):
                    continue
                kv = dict(tok.split("=", 1) for tok in line.split()[1:])
                if kv.get("kernel") == kernel and int(kv.get("batch", -1)) == batch:
                    return int(float(kv["bytes_per_launch"]))
    except (OSError, ValueError, KeyError):
        pass
    return None


UBENCH_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cuyashe_amd", "lib", "libyashe_ubench.so")


def arith_floor(R, B, n, k_prod_ms, prod_bytes):
    """Arithmetic floor of k_prod from the register-resident microbenchmark
    (cuyashe_amd/csrc/ubench.hip): SIMD cycles per wave of a forward
    radix-16 pass (32 shift butterflies + 15 twiddle products), of the DFT-16
    alone forward and inverse, and of 16 Montgomery products. One k_prod wave
    runs, per lane, two forward transforms (twiddled passes, one pass without
    twiddles, one product more for the folded twist), one inverse (the same
    DFTs, 15 twiddles per twiddled pass plus 16 for the folded untwist) and
    the pointwise product; loads, the CRT, the LDS exchanges and the
    reconstruction output are left out, so this bounds the kernel from
    below. floor_ms = waves * cycles per wave / (SIMDs * clock)."""
    import ctypes
    if not os.path.exists(UBENCH_LIB):
        return None
    lib = ctypes.CDLL(UBENCH_LIB)
    cyc = (ctypes.c_double * 4)()
    info = (ctypes.c_double * 3)()
    if lib.yhu_transform_cycles(0, 2000, cyc, info) != 0:
        return None
    c_pass, c_dif, c_dit, c_mont16 = (float(v) for v in cyc)
    mont = c_mont16 / 16
    logn = n.bit_length() - 1
    nfull = logn // 4
    tw = sum(1 for p in range(nfull) if logn - 4 - 4 * p > 0)
    fwd = tw * c_pass + (nfull - tw) * c_dif + mont
    inv = nfull * c_dit + (15 * tw + 1) * mont
    per_wave = 2 * fwd + inv + c_mont16
    waves = B * R * max(1, n // 16 // 64)
    floor_ms = waves * per_wave / (info[0] * 4 * info[1])
    return {"kernel": "k_prod", "cycles_per_wave": round(per_wave, 1),
            "block_cycles": {"fwd_pass": round(c_pass, 1), "dif16": round(c_dif, 1), "dit16": round(c_dit, 1),
                             "mont_x16": round(c_mont16, 1)},
            "clock_khz": int(info[1]), "floor_ms": round(floor_ms, 4),
            "floor_frac": round(prod_bytes / (floor_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "measured_over_floor": round(k_prod_ms / floor_ms, 3)}


def bench_ntt(ya, ctx, N=4096, count=2048, reps=20):
    """The drop-in transform primitive (yh_ntt, callNTT's replacement): cyclic
    forward and inverse Goldilocks NTTs of `count` vectors of length N in
    natural order, timed with events; algorithmic traffic 16 bytes per
    element (one read, one write)."""
    import numpy as np
    rng = np.random.default_rng(3)
    x = rng.integers(0, (1 << 63), size=(count, N), dtype=np.uint64)
    buf = ctx.to_device(x)
    e0, e1, e2 = ya.Event(ctx), ya.Event(ctx), ya.Event(ctx)
    ctx.ntt(buf, N, count)
    ctx.ntt(buf, N, count, inverse=True)
    e0.record()
    for _ in range(reps):
        ctx.ntt(buf, N, count)
    e1.record()
    for _ in range(reps):
        ctx.ntt(buf, N, count, inverse=True)
    e2.record()
    ctx.sync()
    fwd = e0.elapsed_ms(e1) / reps
    inv = e1.elapsed_ms(e2) / reps
    round_trip = bool(np.array_equal(buf.download((count, N)), x))
    gbytes = 16.0 * N * count / 1e9
    return {"workload": f"cyclic NTT over P, {count} vectors of N={N}, natural order", "forward_ms": round(fwd, 4),
            "inverse_ms": round(inv, 4), "forward_GBps": round(gbytes / (fwd * 1e-3), 1),
            "inverse_GBps": round(gbytes / (inv * 1e-3), 1),
            "butterflies_per_s": round(count * (N // 2) * (N.bit_length() - 1) / (fwd * 1e-3), 1),
            "round_trip_exact": round_trip}


def bench_uneven(ya, ctx, reps=10):
    """512 independent ring products, n in {1024..8192} weighted
    .4/.3/.2/.1, q 127- or 255-bit, operands uniform mod q: one persistent
    launch with the autotuned shape, against one launch pair per ring."""
    from cuyashe_amd import workloads
    spec = workloads.uneven_spec(512, 7)
    rings, groups, items = workloads.make_batch(ctx, spec, 7)
    cfg = ya.batch_autotune(ctx, items, TUNE_CACHE, reps=5)
    b = ya.Batch(ctx, items, cfg)
    b.run()
    ctx.sync()
    ctx.set_timing(True)
    ctx.reset_timing()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run()
    ctx.sync()
    wall = (time.perf_counter() - t0) / reps
    k_ms, k_n = ctx.kernel_time(ya.KERNEL_BATCH)
    ctx.set_timing(False)
    refs = {key: ctx.alloc(g[3].nbytes) for key, g in groups.items()}

    def per_ring():
        for key, (A, B, OUT, ha, hb, idx) in groups.items():
            rings[key].ring_mul(A, B, refs[key], len(idx))

    per_ring()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        per_ring()
    ctx.sync()
    wall_ring = (time.perf_counter() - t0) / reps
    exact = all(bool((g[2].download(g[3].shape) == refs[key].download(g[3].shape)).all())
                for key, g in groups.items())
    # algorithmic bytes of the launch: 3 transforms of 16 n bytes per (product, modulus)
    alg = sum(3 * rings[(n, q)].R * 16 * n for n, q in spec)
    b.close()
    for r in rings.values():
        r.close()
    hist = {}
    for n, q in spec:
        k = f"n{n}/q{q.bit_length()}"
        hist[k] = hist.get(k, 0) + 1
    k_avg = k_ms / max(k_n, 1)
    return {"workload": "512 uneven ring products, n in {1024,2048,4096,8192} (.4/.3/.2/.1), q 2^127-1 or 2^255-19",
            "value": round(512 / wall, 1), "unit": "products/s", "batch_ms": round(wall * 1e3, 4),
            "kernel_avg_ms": round(k_avg, 4), "launches_per_batch": 2,
            "roofline": {"bound": "hbm", "kernel": "k_batch_prod", "bytes_per_launch": alg,
                         "achieved": round(alg / (k_avg * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(alg / (k_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "per_ring_launches_ms": round(wall_ring * 1e3, 4), "speedup_vs_per_ring": round(wall_ring / wall, 3),
            "config": cfg, "rings": hist, "identical_to_per_ring": exact}


def bench_serving(ya, ctx, points):
    """Continuous batching over 12 resident rings (n in {2048..16384}, q
    127/255/511-bit, t=1024): Poisson arrivals, 60% add / 30% mul+keyswitch /
    10% encrypt+decrypt, admission by HBM footprint, every result decrypted
    and checked. points: (offered requests/s, request count) pairs."""
    from cuyashe_amd.serving import Server, c5_rings, poisson_requests
    rings = c5_rings()
    t_setup = time.perf_counter()
    srv = Server(ctx, rings, t=1024, pool=32, slots=64, max_batch=64, seed=11)
    setup_s = time.perf_counter() - t_setup
    # warm-up: a burst sizes every ring's scratch and the loop's staging for full batches
    srv.run(poisson_requests(3000, 1e6, rings, 1024, 32, seed=99))
    out = []
    for k, (qps, count) in enumerate(points):
        st = srv.run(poisson_requests(count, qps, rings, 1024, 32, seed=k))
        st["offered_qps"] = qps
        out.append(st)
    srv.close()
    return {"workload": "serving: 12 rings n in {2048,4096,8192,16384} x q in {2^127-1, 2^255-19, 2^511-187}, t=1024, "
                        "Poisson arrivals, 60% add / 30% mul+ks / 10% enc+dec",
            "setup_s": round(setup_s, 2), "points": out}


def bench_circuit(ya, ctx, reps=5):
    """128 ciphertexts at n=8192, q=2^255-19, t=2^16, w=2^32: depth-3 tree of
    112 mul+keyswitch nodes and 256 additions, captured once into a HIP graph
    and replayed."""
    import numpy as np
    from cuyashe_amd import workloads
    from cuyashe_amd.scheme import Yashe
    n, q, t = 8192, workloads.Q255, 1 << 16
    Y = Yashe(ctx, n, q, t, 32, seed=41).generate_keys()
    x = np.random.default_rng(42).integers(0, t, size=(128, n))
    X = Y.encrypt(x)
    bufs = {k: ctx.alloc(v * Y.elt_bytes) for k, v in workloads.CIRCUIT_BUFFERS.items()}
    workloads.depth3_circuit(Y, X, **bufs)
    ctx.sync()
    t0 = time.perf_counter()
    workloads.depth3_circuit(Y, X, **bufs)
    ctx.sync()
    eager_s = time.perf_counter() - t0
    eager = bufs["Z"].download((16, Y.L, n)).copy()
    ctx.graph_begin()
    workloads.depth3_circuit(Y, X, **bufs)
    g = ctx.graph_end()
    g.launch()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.launch()
    ctx.sync()
    graph_s = (time.perf_counter() - t0) / reps
    same = bool((bufs["Z"].download((16, Y.L, n)) == eager).all())
    g.close()
    Y.evk.close()
    Y.ring.close()
    return {"workload": "depth-3 circuit: 128 ciphertexts, n=8192, q=2^255-19, t=2^16, 112 mul+ks + 256 adds",
            "graph_ms": round(graph_s * 1e3, 3), "eager_ms": round(eager_s * 1e3, 3),
            "mul_ks_per_s": round(112 / graph_s, 1), "replay_equals_eager": same}


def calibrate_copy(ya):
    """yh_residue_addsub reads two [R][S] u64 buffers and writes one: 3*R*S*8
    bytes of HBM traffic per launch, nothing cached between launches."""
    ctx = ya.Context(0)
    ring = ya.Ring(ctx, 64, Q)
    S = (256 << 20) // 8 // ring.R
    nbytes = ring.R * S * 8
    a, b, c = ctx.alloc(nbytes), ctx.alloc(nbytes), ctx.alloc(nbytes)
    ctx.sync()
    for _ in range(5):
        ring.residue_addsub(c, a, b, S, False)
    ctx.sync()
    print(json.dumps({"calibration": "k_residue_addsub", "read_bytes_per_launch": 2 * nbytes,
                      "write_bytes_per_launch": nbytes, "launches": 5}))
    ring.close()
    ctx.close()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(full, one_out, A1, A2, gammas, uniq, B, tiled_ok, seconds):
    """The oracle (oracle/cpuref.c: radix-2 Stockham NTT with s_mul as
    uint128 % P, CRT, ICRT, fold, scale-and-round and keyswitch all in C,
    -O3) on the same ciphertext pairs and keys, timed on one thread and on
    every core of the host's share (threads: the C calls release the
    interpreter lock). Every op it computes is checked bit for bit against the
    GPU's row for that pair; all distinct pairs are covered. Returns
    (cpu_baseline, every output checked equal)."""
    import concurrent.futures
    import numpy as np
    from oracle import cpuref

    n = A1.shape[2]

    def op(u):
        g = cpuref.yashe_mul_limbs(A1[u], A2[u], n, Q, T)
        return u, cpuref.keyswitch_limbs(g, gammas, n, Q, W_BITS)

    match = tiled_ok and bool(np.array_equal(one_out[0], full[0]))
    # one thread: ops of the batch in order until the time budget is spent
    done, c0 = 0, time.perf_counter()
    while done < min(B, uniq) or time.perf_counter() - c0 < seconds / 2:
        u, want = op(done % uniq)
        match &= bool(np.array_equal(want, full[u]))
        done += 1
    t1 = time.perf_counter() - c0
    # every core of the host share (at most 16 threads: one GPU's share)
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    threads = max(1, min(16, avail))
    done_mt, c0 = 0, time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as ex:
        while done_mt == 0 or time.perf_counter() - c0 < seconds / 2:
            for u, want in ex.map(op, [(done_mt + k) % uniq for k in range(threads)]):
                match &= bool(np.array_equal(want, full[u]))
            done_mt += threads
    tm = time.perf_counter() - c0
    cpu = {"value": round(done_mt / tm, 4), "unit": "ops/s", "cores": threads, "kind": "port",
           "sample": f"{done_mt} mul+keyswitch ops at n=4096, q=2^127-1 through the oracle (oracle/cpuref.c, "
                     f"-O3 -march=x86-64-v3, all steps in C) on {threads} threads in {tm:.1f} s; single thread: "
                     f"{done} ops in {t1:.1f} s",
           "single_thread": {"value": round(done / t1, 4), "unit": "ops/s", "cores": 1},
           "cpu_model": cpu_model(), "nproc": os.cpu_count(), "threads_available": avail}
    return cpu, bool(match)


def uniform_poly(rng, n, q):
    return [rng.randrange(q) for _ in range(n)]


def percentile(xs, p):
    xs = sorted(xs)
    k = (len(xs) - 1) * p / 100.0
    lo = int(k)
    hi = min(lo + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the path shards by request, "
              "with no collective; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import cuyashe_amd as ya

    if args.calibrate_copy:
        calibrate_copy(ya)
        return
    if args.only_uneven or args.only_circuit or args.serving or args.only_ntt:
        ctx = ya.Context(0)
        if args.only_ntt:
            out = bench_ntt(ya, ctx)
        elif args.serving:
            out = bench_serving(ya, ctx, [(float(v), args.serving_requests) for v in args.serving_qps.split(",")])
        else:
            out = bench_uneven(ya, ctx) if args.only_uneven else bench_circuit(ya, ctx)
        print(json.dumps(out))
        ctx.close()
        return

    rng = random.Random(20240613)
    ctx = ya.Context(0)
    ring = ya.Ring(ctx, N_RING, Q, t=T, w_bits=W_BITS)
    n, L, B = ring.n, ring.L, args.batch

    # Synthetic resident data: keyswitch keys and ciphertext pairs uniform mod q.
    gammas = [uniform_poly(rng, n, Q) for _ in range(ring.lw)]
    evk = ring.evk(ctx.to_device(ring.pack(gammas)))
    uniq = 16  # distinct pairs, tiled over the batch
    C1 = [uniform_poly(rng, n, Q) for _ in range(uniq)]
    C2 = [uniform_poly(rng, n, Q) for _ in range(uniq)]
    import numpy as np
    A1 = ring.pack(C1)
    A2 = ring.pack(C2)
    reps = (B + uniq - 1) // uniq
    d1 = ctx.to_device(np.ascontiguousarray(np.tile(A1, (reps, 1, 1))[:B]))
    d2 = ctx.to_device(np.ascontiguousarray(np.tile(A2, (reps, 1, 1))[:B]))
    out = ctx.alloc(B * L * n * 8)

    # Algorithmic bytes per launch: 16*n per length-n transform of one residue
    # vector (8n read + 8n write); a fused kernel is charged every transform
    # it runs. k_prod: 2 forward + 1 inverse per (pair, modulus);
    # k_ks_scale_digits: the product's R + 1 rows of 4 bytes per coefficient
    # read once, then one forward per (ciphertext, digit); k_keyswitch (k_ks_mac):
    # one inverse per (ciphertext, keyswitch modulus); k_icrt: the keyswitch
    # reconstruction mod q.
    R, Rks, lw = ring.R, ring.R_ks, ring.lw
    kernels = {
        "k_prod": {"id": ya.KERNEL_PROD, "bytes_per_launch": 3 * R * B * 16 * n},
        "k_ks_scale_digits": {"id": ya.KERNEL_KS_DIGITS, "bytes_per_launch": (lw * 16 + 4 * (R + 1)) * B * n},
        "k_keyswitch": {"id": ya.KERNEL_KEYSWITCH, "bytes_per_launch": Rks * B * 16 * n},
        "k_icrt": {"id": ya.KERNEL_ICRT, "bytes_per_launch": None},
    }

    # ---- warmup, then a short pass with HIP events around every launch for
    # the per-kernel table (not the timed region)
    for _ in range(args.warmup):
        ring.mul_keyswitch(evk, d1, d2, out, B)
    ctx.sync()
    ctx.set_timing(True)
    ctx.reset_timing()
    for _ in range(5):
        ring.mul_keyswitch(evk, d1, d2, out, B)
    ctx.sync()
    for k in kernels.values():
        ms, cnt = ctx.kernel_time(k["id"])
        k["avg_ms"] = ms / cnt if cnt else None
        k["launches_per_step"] = cnt // 5
    dom_name = max(("k_prod", "k_ks_scale_digits", "k_keyswitch"),
                   key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches_per_step"])

    # ---- timed region: exactly K steps between two synchronisations, with
    # events only around the dominant kernel's launches
    ctx.set_timing_mask([kernels[dom_name]["id"]])
    ctx.reset_timing()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ring.mul_keyswitch(evk, d1, d2, out, B)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    dom_ms, dom_n = ctx.kernel_time(kernels[dom_name]["id"])
    ctx.set_timing(False)
    ctx.set_timing_mask(None)
    ms_per_step = 1000.0 * elapsed / args.steps
    value = B * args.steps / elapsed
    if args.profile_batched:
        full = out.download((B, L, n))
        print(json.dumps({"profile": "batched mul+keyswitch", "batch": B, "steps": args.steps,
                          "tiled_outputs_equal": bool(np.array_equal(full, full[np.arange(B) % uniq])),
                          "value": round(value, 1), "dominant": dom_name,
                          "dominant_avg_ms": dom_ms / max(dom_n, 1),
                          "kernels_avg_ms": {k: v["avg_ms"] for k, v in kernels.items()}}))
        return

    # ---- throughput at other batch sizes (same data tiled; 3 warm-up and 20 timed
    # steps each), right after the main timed region so the clocks are the same
    sweep = []
    for bs in [int(v) for v in args.sweep.split(",") if v.strip()]:
        reps_s = (bs + uniq - 1) // uniq
        s1 = ctx.to_device(np.ascontiguousarray(np.tile(A1, (reps_s, 1, 1))[:bs]))
        s2 = ctx.to_device(np.ascontiguousarray(np.tile(A2, (reps_s, 1, 1))[:bs]))
        so = ctx.alloc(bs * L * n * 8)
        for _ in range(3):
            ring.mul_keyswitch(evk, s1, s2, so, bs)
        ctx.sync()
        ts = time.perf_counter()
        for _ in range(20):
            ring.mul_keyswitch(evk, s1, s2, so, bs)
        ctx.sync()
        el = time.perf_counter() - ts
        sweep.append({"batch": bs, "ops_per_s": round(bs * 20 / el, 1), "ms_per_step": round(1000.0 * el / 20, 4)})
        del s1, s2, so

    # ---- two independent batches in flight: steps alternate between two
    # rings of the same parameters on streams of their own (default
    # priority), so one batch's latency-bound reconstruction overlaps the
    # other's transforms. Reported beside the single-stream value (whose
    # per-kernel event times stay those of unshared launches).
    two = None
    if args.two_streams:
        rings2 = [ya.Ring(ctx, N_RING, Q, t=T, w_bits=W_BITS) for _ in range(2)]
        evks2 = [r.evk(ctx.to_device(r.pack(gammas))) for r in rings2]
        outs2 = [ctx.alloc(B * L * n * 8) for _ in rings2]
        for r in rings2:
            r.use_own_stream()
        for _ in range(2):
            for k2, r in enumerate(rings2):
                r.mul_keyswitch(evks2[k2], d1, d2, outs2[k2], B)
        for r in rings2:
            r.sync()
        ts = time.perf_counter()
        for st in range(args.steps):
            k2 = st % 2
            rings2[k2].mul_keyswitch(evks2[k2], d1, d2, outs2[k2], B)
        for r in rings2:
            r.sync()
        el = time.perf_counter() - ts
        ref = out.download((B, L, n))
        same = all(bool((o.download((B, L, n)) == ref).all()) for o in outs2)
        two = {"value": round(B * args.steps / el, 1), "unit": "ops/s", "ms_per_step": round(1000.0 * el / args.steps, 4),
               "steps": args.steps, "identical_to_single_stream": same}
        for e in evks2:
            e.close()
        for r in rings2:
            r.close()
        del outs2

    # per-step latency (each step synchronised)
    lat = []
    for _ in range(max(10, args.steps)):
        s = time.perf_counter()
        ring.mul_keyswitch(evk, d1, d2, out, B)
        ctx.sync()
        lat.append(1000.0 * (time.perf_counter() - s))

    # ---- plain: one op at a time, synchronised after each (reference style)
    one = ctx.alloc(L * n * 8)
    for _ in range(3):
        ring.mul_keyswitch(evk, d1, d2, one, 1)
    ctx.sync()
    plain_lat = []
    tp0 = time.perf_counter()
    for _ in range(args.plain_ops):
        s = time.perf_counter()
        ring.mul_keyswitch(evk, d1, d2, one, 1)
        ctx.sync()
        plain_lat.append(1000.0 * (time.perf_counter() - s))
    plain_value = args.plain_ops / (time.perf_counter() - tp0)

    # ---- roofline of the dominant kernel, from its events in the timed region
    dom = kernels[dom_name]
    dom_avg = dom_ms / dom_n
    achieved = dom["bytes_per_launch"] / (dom_avg * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom_name, B),
                "kernel": dom_name, "avg_ms": round(dom_avg, 5), "launches": dom_n}
    symbols = {"k_prod": "k_prod<12, true>", "k_ks_scale_digits": "k_ks_scale_digits<12, 2>", "k_keyswitch": "k_ks_mac<12>"}
    # VALU instruction rate of the dominant kernel (instruction count from the
    # committed SQ counters, time from the events above). Reported without a
    # fraction: on gfx950 the register-resident microbenchmark retires this
    # code's mix faster than one wave64 instruction per 4 cycles per SIMD, so
    # that figure is no ceiling; `arith_floor` is the measured one.
    valu = measured_valu(symbols[dom_name], B)
    valu_rate = None
    if valu:
        valu_rate = {"kernel": dom_name, "instructions_per_launch": valu,
                     "achieved": round(valu / (dom_avg * 1e-3) / 1e9, 1), "unit": "G wave-instructions/s"}
    floor = arith_floor(R, B, n, dom_avg if dom_name == "k_prod" else kernels["k_prod"]["avg_ms"],
                        kernels["k_prod"]["bytes_per_launch"])
    ntt_k = [k for k in kernels.values() if k["bytes_per_launch"]]
    ntt_bytes = sum(k["bytes_per_launch"] * k["launches_per_step"] for k in ntt_k)
    ntt_ms = sum(k["avg_ms"] * k["launches_per_step"] for k in ntt_k)

    # ---- CPU baseline: the oracle on a bounded sample of the same workload,
    # which also checks those GPU results bit for bit.
    cpu = None
    exact = None
    # every output of the timed batch: the batch tiles `uniq` distinct pairs,
    # so row k must equal row k mod uniq (checked on the whole download), and
    # those rows (and the plain path's output, pair 0) are checked against
    # the oracle below
    full = out.download((B, L, n))
    tiled_ok = bool(np.array_equal(full, full[np.arange(B) % uniq]))
    one_out = one.download((1, L, n))
    if not args.no_cpu:
        cpu, exact = cpu_baseline(full, one_out, A1, A2, ring.pack(gammas), uniq, B, tiled_ok, args.cpu_seconds)

    secondary = None
    if not args.no_secondary:
        secondary = {"uneven_products": bench_uneven(ya, ctx), "graph_circuit": bench_circuit(ya, ctx),
                     "ntt_primitive": bench_ntt(ya, ctx),
                     # below saturation (latency) and far above it (the loop's ceiling)
                     "serving": bench_serving(ya, ctx, [(1000.0, 1500), (1e6, 30000)])}

    res = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "ops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: seeded ciphertext pairs and keyswitch keys uniform mod q, resident in HBM",
        "config": {"workload": "YASHE mul+keyswitch, n=4096, q=2^127-1, t=1024, w=2^32 (4 digits)",
                   "batch": B, "crt_primes": R, "keyswitch_primes": Rks, "digits": lw},
        "p50_ms": round(percentile(lat, 50), 4),
        "p99_ms": round(percentile(lat, 99), 4),
        "plain": {"value": round(plain_value, 1), "unit": "ops/s",
                  "p50_ms": round(percentile(plain_lat, 50), 4), "p99_ms": round(percentile(plain_lat, 99), 4)},
        "batched_vs_plain": round(value / plain_value, 2),
        "batch_sweep": sweep,
        "two_streams": two,
        "ntt_stages": {"algorithmic_GBps": round(ntt_bytes / (ntt_ms * 1e-3) / 1e9, 1) if ntt_ms else None,
                       "frac_of_hbm": round(ntt_bytes / (ntt_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if ntt_ms else None},
        "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items() if kk != "id"}
                    for k, v in kernels.items()},
        "roofline": roofline,
        "valu_rate": valu_rate,
        "arith_floor": floor,
        "cpu_baseline": cpu,
        "bit_exact_sample": exact,
        "outputs_checked": {"timed_batch": B, "plain": 1, "distinct_pairs_vs_oracle": min(B, uniq),
                            "tiled_rows_equal": tiled_ok},
        "secondary": secondary,
    }
    evk.close()
    ring.close()
    ctx.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
