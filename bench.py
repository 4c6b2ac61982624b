#!/usr/bin/env python3
"""Benchmark: 2048-bit modular exponentiation throughput on one MI355X.

A step is one xmpIntegersPowm over 65,536 instances with 2048-bit odd moduli
(top bit set, uniform random limbs), bases uniform below the modulus and
uniform random 2048-bit exponents -- inputs resident in HBM.  The JSON line
also carries the kernel's roofline, GMP's throughput on this host's cores,
the elementwise add/sub/mul and div/mod rates, the mixed-size batched-vs-plain
comparison and the serving loop's latency percentiles.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The serving loop runs each batch slot on two streams, four with reserved CUs
# (6 slots: 12 or 24 streams).  HIP maps streams onto this many hardware
# queues per process (default 4), so raise it to 16 for the streams to run
# concurrently; it must be set before the first HIP call.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

import xmp_amd  # noqa: E402
from xmp_amd import hip, inputs  # noqa: E402
from xmp_amd.ext import Batch, Montgomery  # noqa: E402

HBM_PEAK_GBS = 8000.0
# one v_mad_u64_u32 (32x32+64 -> 64) per lane per clock on 256 CUs x 64 lanes
# at 2.4 GHz = the published 78.6 TFLOP/s FP64 vector rate of the same datapath;
# a multiply-add counts as 2 integer operations
VALU_PEAK_TOPS = 78.6
# HBM bytes per launch of powm_kernel<2,37,28> on 65,536 x 2048-bit instances
# by window width, from --pmc runs of `bench.py --only powm`: reads are
# 32/64/128 x TCC_EA0_RDREQ_{32B,64B,128B} (exact on a copy kernel of known
# size), writes are WRITE_SIZE (exact on the same copy)
POWM_TRAFFIC_BYTES = {(2, 37, 28, 6): 9053542464 + 1444922912,
                      (2, 37, 28, 7): 7917544640 + 2808747328}


def canonical_powm_macs(bits, exp_bits):
    """(b + ceil(b/w)) * (2n^2 + n) 32x32 multiply-adds per powm, w from the
    reference's windowBitsForPrecision (src/include/powm_operators.h:15-31)."""
    n = bits // 32
    w = 4 if exp_bits < 400 else 5 if exp_bits < 800 else 6 if exp_bits < 1600 else 7 if exp_bits <= 4096 else 8
    return (exp_bits + math.ceil(exp_bits / w)) * (2 * n * n + n)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_async(fn, stream, reps):
    """Average ms per call of fn (enqueued on `stream`), measured with HIP events."""
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.sync()
    return e0.elapsed_ms(e1) / reps


def lib_powm_async(h, out, base, exp, mod, count):
    rc = xmp_amd.lib().xmpIntegersPowmAsync(h.h, out.x, base.x, exp.x, mod.x, count)
    if rc != 0:
        raise xmp_amd.XmpError(rc, "xmpIntegersPowmAsync")


def cpu_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def powm_workload(h, n, bits, seed):
    limbs = bits // 32
    m = inputs.odd_moduli(seed, n, limbs)
    b = inputs.below(seed + 1, m)
    e = inputs.uniform(seed + 2, n, limbs)
    return m, b, e, h.from_array(m), h.from_array(b), h.from_array(e)


MIXED_SIZES = inputs.MIXED_SIZES
mixed_counts = inputs.mixed_counts


def run_mixed(h, stream, total, seed, reps=2):
    """Mixed-size batch: one persistent graph launch vs one plain launch per class."""
    counts = mixed_counts(total, seed)
    host, outs, bases, exps, mods = [], [], [], [], []
    for k, (bits, n) in enumerate(zip(MIXED_SIZES, counts)):
        m, b, e, M, B, E = powm_workload(h, n, bits, seed + 10 * k)
        host.append((m, b, e))
        outs.append(h.integers(bits, n))
        bases.append(B)
        exps.append(E)
        mods.append(M)
    batch = Batch(h, outs, bases, exps, mods, counts)
    batch.launch_async()
    hip.stream_sync(stream)
    batched_ms = time_async(batch.launch_async, stream, reps)
    # a sample of every class's batched results, checked in the CPU leg
    samples = []
    for k, bits in enumerate(MIXED_SIZES):
        m, b, e = host[k]
        idx = np.linspace(0, counts[k] - 1, num=min(8, counts[k])).astype(int)
        samples.append((bits, b[idx], e[idx], m[idx], outs[k].get()[idx]))
    batch.launch_plain_async()
    hip.stream_sync(stream)
    plain_ms = time_async(batch.launch_plain_async, stream, reps)
    batch.close()
    work = sum(c * canonical_powm_macs(bits, bits) for bits, c in zip(MIXED_SIZES, counts))
    eq2048 = work / canonical_powm_macs(2048, 2048)   # 2048-bit-powm equivalents
    return {"instances": total, "class_counts": dict(zip([str(x) for x in MIXED_SIZES], counts)),
            "batched_ms": round(batched_ms, 3), "plain_ms": round(plain_ms, 3),
            "batched_2048eq_powm_per_s": round(eq2048 / batched_ms * 1e3, 1),
            "plain_2048eq_powm_per_s": round(eq2048 / plain_ms * 1e3, 1),
            "batched_over_plain": round(plain_ms / batched_ms, 3)}, samples


# 1,200,000 requests per rate point: the last batches' drain (20-40 ms), which
# the achieved rate counts in the run's wall time, is then 2-3 % of a run at
# 800 K-1 M req/s (it was 4-5 % with 600,000).  Six slots of
# 8,192 (the latency shape) and of 12,288 (the throughput shape), each batch on
# the layout for its own size: measured on one MI355X against 4-8 slots of
# 6,144-16,384 and the layouts chosen for all slots in flight, the first held
# 97 % of the offered rate at 80 % of its 1.31 M req/s saturation, the second
# saturated highest (1.59 M req/s, p99 24 ms at 800 K req/s).
SERVE_SLOTS, SERVE_BATCH, SERVE_REQUESTS = 6, 8192, 1200000
SERVE_TP_BATCH = 12288


def run_serving(h, nkeys, requests, seed, reserved_cus=0, slots=SERVE_SLOTS, batch=SERVE_BATCH, layout_scale=1):
    """Serving loop over `nkeys` resident 2048-bit keys.  Saturation is the
    throughput achieved under Poisson arrivals offered far above capacity
    (3 M req/s); then Poisson arrivals at fixed 400 K and 800 K req/s and at
    50% and 80% of the saturation, each with p50/p99 latency."""
    from xmp_amd.ext import Server, make_keys
    keys = make_keys(nkeys, 2048, seed)
    rng = np.random.default_rng(seed)
    kinds = (rng.random(requests) < 0.5).astype(np.uint32)
    key_idx = rng.integers(0, nkeys, requests).astype(np.uint32)
    bases = inputs.uniform(seed + 1, requests, 64).copy()
    nmat = keys["n"][key_idx]
    bases[:, -1] = (bases[:, -1].astype(np.uint64) % nmat[:, -1].astype(np.uint64)).astype(np.uint32)
    exps = inputs.uniform(seed + 2, requests, 64)
    srv = Server(h, keys, 2048, streams=slots, max_batch=batch, reserved_cus=reserved_cus, layout_scale=layout_scale)
    out = {"keys": nkeys, "requests": requests, "mix": "50% 2048-bit powm, 50% CRT pairs of 1024-bit moduli",
           "streams": slots, "max_batch": batch, "reserved_cus": reserved_cus, "layout_scale": layout_scale}

    def point(rate, seed_):
        res, lat, wall = srv.run_poisson(rate, kinds, key_idx, bases, exps, seed=seed_)
        return res, {"offered_req_per_s": round(rate, 1), "achieved_req_per_s": round(requests / wall, 1),
                     "achieved_frac": round(requests / wall / rate, 4),
                     "p50_ms": round(float(np.percentile(lat, 50)) / 1e3, 3),
                     "p99_ms": round(float(np.percentile(lat, 99)) / 1e3, 3)}
    try:
        srv.run_poisson(1e9, kinds[:4096], key_idx[:4096], bases[:4096], exps[:4096], seed=seed)   # warm-up
        res, sat = point(3e6, seed)
        out["saturation_req_per_s"] = sat["achieved_req_per_s"]
        out["saturation_run"] = sat
        for rate in (400000.0, 800000.0):
            res, out["rate_%dK" % int(rate / 1000)] = point(rate, seed + 1)
        for frac in (0.5, 0.8):
            res, out["rate_%d%%" % int(frac * 100)] = point(frac * out["saturation_req_per_s"], seed + 1)
        out["stats"] = {k: v for k, v in srv.stats().items() if k in ("batches", "peak_batch", "peak_inflight")}
    finally:
        srv.close()
    sample = np.linspace(0, requests - 1, num=64).astype(int)
    return out, (keys, kinds[sample], key_idx[sample], bases[sample], exps[sample], res[sample])


def run_only(args):
    """One kernel, `warmup` + `steps` launches, for counter runs; prints ms per launch."""
    h = xmp_amd.Handle()
    stream = hip.stream_create()
    h.set_stream(stream)
    lib = xmp_amd.lib()
    if args.only == "serve":
        res, _ = run_serving(h, 1024, SERVE_REQUESTS, 99, reserved_cus=args.serve_reserved_cus, slots=args.serve_slots,
                             batch=args.serve_batch, layout_scale=args.serve_layout_scale)
        print(json.dumps(res), flush=True)
        h.close()
        return
    if args.only == "mixed":
        res, _ = run_mixed(h, stream, 262144, 777, reps=max(1, args.steps))
        print(json.dumps(res), flush=True)
        h.close()
        return
    if args.only == "powm":
        m, b, e, M, B, E = powm_workload(h, 65536, 2048, 2048)
        out = h.integers(2048, 65536)
        ctx = Montgomery(h, M, 2048, 65536)
        fn = lambda: ctx.powm_async(out, B, E, 65536)  # noqa: E731
        what = "powm_kernel 65,536 x 2048-bit (algorithmic bytes: 4 x 256 B in/out per instance)"
    elif args.only == "powm1024":
        m1, b1, e1, M1, B1, E1 = powm_workload(h, 4096, 1024, 1024)
        out1 = h.integers(1024, 4096)
        fn = lambda: h.powm(out1, B1, E1, M1, 4096, asynchronous=True)  # noqa: E731
        what = "powm 4,096 x 1024-bit (R^2 mod m and the exponentiation, as xmpIntegersPowm)"
    elif args.only == "mul32768":
        a5, b5 = inputs.uniform(5151, 4096, 1024), inputs.uniform(5152, 4096, 1024)
        A5, B5, P5 = h.from_array(a5), h.from_array(b5), h.integers(65536, 4096)
        fn = lambda: lib.xmpIntegersMulAsync(h.h, P5.x, A5.x, B5.x, 4096)  # noqa: E731
        what = "mul 4,096 x 32768x32768 -> 65536-bit (Karatsuba over the row kernels)"
    elif args.only == "divwide":
        a6 = inputs.uniform(6161, 16384, 512)
        d6 = inputs.uniform(6162, 16384, 256)
        d6[:, -1] |= np.uint32(0x80000000)
        A6, D6 = h.from_array(a6), h.from_array(d6)
        Q6, R6 = h.integers(16384, 16384), h.integers(8192, 16384)
        fn = lambda: lib.xmpIntegersDivModAsync(h.h, Q6.x, R6.x, A6.x, D6.x, 16384)  # noqa: E731
        what = "divmod_wave_kernel 16,384 x 16384-bit / 8192-bit (one wave per division)"
    elif args.only == "mul":
        n2 = 1 << 20
        a2, b2 = inputs.ordered_pair(4242, n2, 64)
        A2, B2 = h.from_array(a2), h.from_array(b2)
        P2 = h.integers(4096, n2)
        fn = lambda: lib.xmpIntegersMulAsync(h.h, P2.x, A2.x, B2.x, n2)  # noqa: E731
        what = "mul_cols_kernel<74,74,28> 1,048,576 x 2048x2048 -> 4096-bit (4096 canonical MACs and 1024 B per product)"
    elif args.only == "sqr":
        n2 = 1 << 20
        a2, _ = inputs.ordered_pair(4242, n2, 64)
        A2 = h.from_array(a2)
        P2 = h.integers(4096, n2)
        fn = lambda: lib.xmpIntegersMulAsync(h.h, P2.x, A2.x, A2.x, n2)  # noqa: E731
        what = "mul_cols_kernel<74,74,28,square> 1,048,576 x 2048-bit squares -> 4096-bit"
    elif args.only == "div":
        a3 = inputs.uniform(4343, 65536, 128)
        d3 = inputs.uniform(4344, 65536, 64)
        d3[:, -1] |= np.uint32(0x80000000)
        A3, D3 = h.from_array(a3), h.from_array(d3)
        Q3 = h.integers(4096, 65536)
        fn = lambda: lib.xmpIntegersDivAsync(h.h, Q3.x, A3.x, D3.x, 65536)  # noqa: E731
        what = "divmod_pair_kernel<32> 65,536 x 4096-bit / 2048-bit (n(m-n) = 4096 canonical MACs per division)"
    elif args.only == "add":
        n2 = 1 << 20
        a2, b2 = inputs.ordered_pair(4242, n2, 64)
        A2, B2 = h.from_array(a2), h.from_array(b2)
        S2 = h.integers(2048, n2)
        fn = lambda: lib.xmpIntegersAddAsync(h.h, S2.x, A2.x, B2.x, n2)  # noqa: E731
        what = "addsub_vec4_kernel 1,048,576 x 2048-bit (algorithmic bytes: 768 B per pair, 805,306,368 B)"
    else:
        n2 = 1 << 20
        src = hip.malloc(n2 * 256, fill=7)   # device-resident compact source (256 MiB)
        dst = h.integers(2048, n2)
        fn = lambda: lib.xmpIntegersImportAsync(h.h, dst.x, 64, -1, 4, 0, 0, src, n2)  # noqa: E731
        what = "import_fast_kernel copy of 268,435,456 B read + 268,435,456 B written"
    for _ in range(args.warmup):
        fn()
    hip.stream_sync(stream)
    ms = time_async(fn, stream, max(1, args.steps))
    print(json.dumps({"only": args.only, "what": what, "ms_per_launch": round(ms, 4)}), flush=True)
    h.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-reserved-cus", type=int, default=0,
                    help="CUs the serving loop keeps for its short kernels (0 = none)")
    ap.add_argument("--serve-slots", type=int, default=SERVE_SLOTS, help="serving batch slots")
    ap.add_argument("--serve-batch", type=int, default=SERVE_BATCH, help="serving requests per slot")
    ap.add_argument("--serve-layout-scale", type=int, default=1,
                    help="serving layout chosen as for a batch this many times larger (0 = slots in flight)")
    ap.add_argument("--only", choices=["powm", "powm1024", "add", "mul", "mul32768", "sqr", "div", "divwide", "copy",
                                       "mixed", "serve"],
                    help="run just one workload (for profiling): the 2048-bit powm kernel with resident "
                         "constants, 4,096 x 1024-bit powm, the 2048-bit add/mul/square, 32768-bit products, "
                         "4096/2048 and 16384/8192 division, a plain 256 MiB import copy for counter calibration, "
                         "the mixed batch or the serving loop")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the instances are independent, so N GPUs would each run this "
              "same per-GPU workload on their own share; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.only:
        run_only(args)
        return

    N, BITS = 65536, 2048
    h = xmp_amd.Handle()
    stream = hip.stream_create()
    h.set_stream(stream)
    m, b, e, M, B, E = powm_workload(h, N, BITS, 2048)
    OUT = h.integers(BITS, N)

    def step():
        h.powm(OUT, B, E, M, N, asynchronous=True)

    log("warmup %d steps" % args.warmup)
    for _ in range(args.warmup):
        step()
    hip.stream_sync(stream)
    hip.device_sync()
    # HIP events on the library's stream bracket the timed steps, so the
    # kernels' own time (the exponentiation and its R^2 mod m set-up) is
    # measured on exactly the launches the wall clock covers
    ev = [hip.Event() for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for _ in range(args.steps):
        step()
    ev[1].record(stream)
    hip.stream_sync(stream)
    hip.device_sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = N * args.steps / (t1 - t0)
    kern_ms = ev[0].elapsed_ms(ev[1]) / args.steps
    log("headline: %.1f powm/s (%.2f ms/step, %.3f ms of kernels per step)" % (value, ms_per_step, kern_ms))

    got = OUT.get()   # checked against GMP in the CPU leg

    # roofline of the exponentiation kernel: canonical multiply-adds over the
    # event-timed steps (which also hold the R^2 mod m kernel, about 0.1 % of a
    # step, so the fraction is a slight underestimate)
    ctx = Montgomery(h, M, BITS, N)   # the layout the plain call chose for this shape
    lanes, digits, radix, wbits = ctx.layout()
    ctx.close()
    macs = canonical_powm_macs(BITS, BITS) * N
    achieved = 2.0 * macs / (kern_ms * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                "frac": round(achieved / VALU_PEAK_TOPS, 4), "traffic": None,
                "kernel": "powm_kernel<%d,%d,%d> window %d" % (lanes, digits, radix, wbits),
                "kernel_ms": round(kern_ms, 4), "kernel_ms_source": "HIP events around the timed steps"}
    # the issue rate a pure v_mad_u64_u32 stream reaches at 2 and 8 waves per SIMD,
    # and the kernel's fraction of that measured ceiling at its own occupancy (2)
    from xmp_amd.ext import probe_mad_rate
    probe2, probe8 = probe_mad_rate(h, 2), probe_mad_rate(h, 8)
    roofline["measured_mad_per_clk_per_cu"] = {"2_waves": round(probe2, 2), "8_waves": round(probe8, 2),
                                               "full_rate": 64}
    roofline["frac_of_probe_2_waves"] = round(achieved / VALU_PEAK_TOPS * 64.0 / probe2, 4)
    roofline["traffic"] = POWM_TRAFFIC_BYTES.get((lanes, digits, radix, wbits))
    log("powm kernel %.3f ms -> %.2f TOP/s (%.1f%% of %.1f)" % (kern_ms, achieved, 100 * achieved / VALU_PEAK_TOPS,
                                                            VALU_PEAK_TOPS))

    extra = {}
    # GMP counterparts of the secondary figures, timed in the CPU leg:
    # (extra key, GMP op, count, out limbs, a, b, c)
    cpu_work = []
    if not args.quick:
        # ---- the same 65,536 x 2048-bit work as one persistent graph launch vs the plain call
        OUT2 = h.integers(BITS, N)
        batch = Batch(h, [OUT2], [B], [E], [M], [N])
        batch.launch_async()
        hip.stream_sync(stream)
        b_ms = time_async(batch.launch_async, stream, max(2, args.steps // 2))
        p_ms = time_async(lambda: lib_powm_async(h, OUT2, B, E, M, N), stream, max(2, args.steps // 2))
        same = bool(np.array_equal(OUT2.get(), got))
        batch.close()
        extra["batched_2048"] = {"instances": N, "batched_ms": round(b_ms, 3), "plain_ms": round(p_ms, 3),
                                 "batched_powm_per_s": round(N / b_ms * 1e3, 1),
                                 "batched_over_plain": round(p_ms / b_ms, 3), "same_results": same}
        log("2048-bit batched %.3f ms vs plain %.3f ms" % (b_ms, p_ms))
        del OUT2

        # ---- the same job from host buffers to host buffers (pageable numpy memory):
        # import base, exponent and modulus, powm, export -- the PCIe-inclusive rate
        Mh, Bh, Eh, Oh = (h.integers(BITS, N) for _ in range(4))

        def host_job():
            Mh.load(m)
            Bh.load(b)
            Eh.load(e)
            h.powm(Oh, Bh, Eh, Mh, N)
            return Oh.get()
        host_job()
        t_a = time.perf_counter()
        res_h = host_job()
        t_b = time.perf_counter()
        Mh.load(m)
        Bh.load(b)
        Eh.load(e)
        t_c = time.perf_counter()
        Oh.get()
        t_d = time.perf_counter()
        extra["powm_2048_host_buffers"] = {
            "instances": N, "ms": round((t_b - t_a) * 1e3, 3), "powm_per_s": round(N / (t_b - t_a), 1),
            "import_ms": round((t_c - t_b) * 1e3, 3), "export_ms": round((t_d - t_c) * 1e3, 3),
            "import_GB_per_s": round(3 * N * BITS / 8 / (t_c - t_b) / 1e9, 2),
            "export_GB_per_s": round(N * BITS / 8 / (t_d - t_c) / 1e9, 2),
            "same_results": bool(np.array_equal(res_h, got))}
        log("2048-bit host buffers to host buffers: %.3f ms" % ((t_b - t_a) * 1e3))
        for x in (Mh, Bh, Eh, Oh):
            x.close()

        # ---- elementwise add/sub/mul on 1,048,576 pairs of 2048-bit integers
        n2 = 1 << 20
        a2, b2 = inputs.ordered_pair(4242, n2, 64)
        A2, B2 = h.from_array(a2), h.from_array(b2)
        S2, P2 = h.integers(2048, n2), h.integers(4096, n2)
        lib = xmp_amd.lib()
        add = lambda: lib.xmpIntegersAddAsync(h.h, S2.x, A2.x, B2.x, n2)  # noqa: E731
        sub = lambda: lib.xmpIntegersSubAsync(h.h, S2.x, A2.x, B2.x, n2)  # noqa: E731
        mul = lambda: lib.xmpIntegersMulAsync(h.h, P2.x, A2.x, B2.x, n2)  # noqa: E731
        sqr = lambda: lib.xmpIntegersMulAsync(h.h, P2.x, A2.x, A2.x, n2)  # noqa: E731
        xor = lambda: lib.xmpIntegersXorAsync(h.h, S2.x, A2.x, B2.x, n2)  # noqa: E731
        add()
        hip.stream_sync(stream)
        add_ms = time_async(add, stream, 20)
        sub_ms = time_async(sub, stream, 20)
        xor()
        hip.stream_sync(stream)
        xor_ms = time_async(xor, stream, 20)
        mul()
        hip.stream_sync(stream)
        mul_ms = time_async(mul, stream, 10)
        sqr()
        hip.stream_sync(stream)
        sqr_ms = time_async(sqr, stream, 10)
        extra["add_2048"] = {"pairs": n2, "ms": round(add_ms, 4), "GB_per_s": round(768 * n2 / add_ms / 1e6, 1),
                             "hbm_frac": round(768 * n2 / add_ms / 1e6 / HBM_PEAK_GBS, 4)}
        extra["sub_2048"] = {"pairs": n2, "ms": round(sub_ms, 4), "GB_per_s": round(768 * n2 / sub_ms / 1e6, 1),
                             "hbm_frac": round(768 * n2 / sub_ms / 1e6 / HBM_PEAK_GBS, 4)}
        extra["xor_2048"] = {"pairs": n2, "ms": round(xor_ms, 4), "GB_per_s": round(768 * n2 / xor_ms / 1e6, 1),
                             "hbm_frac": round(768 * n2 / xor_ms / 1e6 / HBM_PEAK_GBS, 4)}
        extra["mul_2048x2048"] = {"pairs": n2, "ms": round(mul_ms, 4), "mul_per_s": round(n2 / mul_ms * 1e3, 1),
                                  "GB_per_s": round(1024 * n2 / mul_ms / 1e6, 1),
                                  "TOP_per_s": round(2 * 4096 * n2 / mul_ms / 1e9, 3)}
        extra["sqr_2048"] = {"count": n2, "ms": round(sqr_ms, 4), "sqr_per_s": round(n2 / sqr_ms * 1e3, 1),
                             "GB_per_s": round(768 * n2 / sqr_ms / 1e6, 1),
                             "TOP_per_s_as_products": round(2 * 4096 * n2 / sqr_ms / 1e9, 3)}
        log("add %.3f ms  sub %.3f ms  xor %.3f ms  mul %.3f ms  sqr %.3f ms" % (add_ms, sub_ms, xor_ms, mul_ms,
                                                                             sqr_ms))
        cpu_work += [("add_2048", "GO_ADD", n2, 64, a2, b2, None), ("sub_2048", "GO_SUB", n2, 64, a2, b2, None),
                     ("mul_2048x2048", "GO_MUL", n2, 128, a2, b2, None)]
        del A2, B2, S2, P2

        # ---- mul beyond the row kernels: 4,096 x 32768x32768 bits (Karatsuba)
        n5 = 4096
        a5, b5 = inputs.uniform(5151, n5, 1024), inputs.uniform(5152, n5, 1024)
        A5, B5, P5 = h.from_array(a5), h.from_array(b5), h.integers(65536, n5)
        mul5 = lambda: lib.xmpIntegersMulAsync(h.h, P5.x, A5.x, B5.x, n5)  # noqa: E731
        mul5()
        hip.stream_sync(stream)
        mul5_ms = time_async(mul5, stream, 5)
        extra["mul_32768x32768"] = {"pairs": n5, "ms": round(mul5_ms, 4), "mul_per_s": round(n5 / mul5_ms * 1e3, 1)}
        log("mul 32768x32768 x4096: %.3f ms" % mul5_ms)
        cpu_work += [("mul_32768x32768", "GO_MUL", n5, 2048, a5, b5, None)]
        del A5, B5, P5

        # ---- div/mod: 65,536 x (4096-bit / 2048-bit, divisor top bit set)
        a3 = inputs.uniform(4343, N, 128)
        d3 = inputs.uniform(4344, N, 64)
        d3[:, -1] |= np.uint32(0x80000000)
        A3, D3 = h.from_array(a3), h.from_array(d3)
        Q3, R3 = h.integers(4096, N), h.integers(2048, N)
        div = lambda: lib.xmpIntegersDivAsync(h.h, Q3.x, A3.x, D3.x, N)  # noqa: E731
        mod = lambda: lib.xmpIntegersModAsync(h.h, R3.x, A3.x, D3.x, N)  # noqa: E731
        div()
        hip.stream_sync(stream)
        div_ms = time_async(div, stream, 10)
        mod_ms = time_async(mod, stream, 10)
        extra["div_4096_by_2048"] = {"count": N, "ms": round(div_ms, 4), "div_per_s": round(N / div_ms * 1e3, 1)}
        extra["mod_4096_by_2048"] = {"count": N, "ms": round(mod_ms, 4), "mod_per_s": round(N / mod_ms * 1e3, 1)}
        log("div %.3f ms  mod %.3f ms" % (div_ms, mod_ms))
        cpu_work += [("div_4096_by_2048", "GO_DIV", N, 128, a3, d3, None),
                     ("mod_4096_by_2048", "GO_MOD", N, 64, a3, d3, None)]

        # ---- 4,096 x 1024-bit powm (the reference's CPU-runnable case)
        m1, b1, e1, M1, B1, E1 = powm_workload(h, 4096, 1024, 1024)
        O1 = h.integers(1024, 4096)
        p1 = lambda: h.powm(O1, B1, E1, M1, 4096, asynchronous=True)  # noqa: E731
        p1()
        hip.stream_sync(stream)
        p1_ms = time_async(p1, stream, 5)
        extra["powm_1024_x4096"] = {"ms": round(p1_ms, 4), "powm_per_s": round(4096 / p1_ms * 1e3, 1)}
        log("powm 1024 x4096: %.3f ms" % p1_ms)
        cpu_work.append(("powm_1024_x4096", "GO_POWM", 4096, 32, b1, e1, m1))

        # ---- mixed-size batch: 262,144 instances, one persistent launch vs plain per-class launches
        mixed, mixed_samples = run_mixed(h, stream, 262144, 777)
        extra["mixed_batch"] = mixed
        log("mixed batch: batched %.1f ms, plain %.1f ms (x%.2f)" % (
            mixed["batched_ms"], mixed["plain_ms"], mixed["batched_over_plain"]))
        # ---- serving: 1024 resident keys, Poisson arrivals
        # the latency-oriented shape (6 slots of 8,192) gives the p50/p99
        # figures; the throughput-oriented shape (6 slots of 12,288) the
        # highest saturation
        serving, serve_sample = run_serving(h, 1024, SERVE_REQUESTS, 99, reserved_cus=args.serve_reserved_cus,
                                            slots=args.serve_slots, batch=args.serve_batch,
                                            layout_scale=args.serve_layout_scale)
        extra["serving"] = serving
        log("serving: %s" % json.dumps(serving))
        serving_tp, _ = run_serving(h, 1024, SERVE_REQUESTS, 99, slots=SERVE_SLOTS, batch=SERVE_TP_BATCH,
                                    layout_scale=1)
        extra["serving_throughput_shape"] = serving_tp
        log("serving (throughput shape): %s" % json.dumps(serving_tp))

    # ---- CPU leg: GMP (the reference test suite's oracle) on this host's cores.
    # It times mpz_powm on a bounded sample of the same instances (the CPU
    # baseline) and, as the checker, compares the GPU results it has sampled.
    cpu, parity = None, {}
    if not args.no_cpu:
        from oracle import oracle as O
        if O.gmp_available():
            cores = cpu_cores()
            S = 4096 if args.quick else 8192
            want, _, secs = O.gmp_batch(O.GO_POWM, S, 64, b[:S], e[:S], m[:S], threads=cores)
            cpu = {"value": round(S / secs, 1), "unit": "powm/s", "cores": cores, "kind": "port",
                   "sample": "GMP %s mpz_powm (the reference test suite's oracle, unit_tests/test.cu:1214) on the "
                             "first %d of the same 2048-bit instances, %d pthreads, %.1f s wall" % (
                                 O.gmp_version(), S, cores, secs)}
            # mpz_powm_sec, the constant-time routine the reference's perf harness
            # times (perf_tests/main.cu:205), on a quarter of that sample
            S2 = S // 4
            want2, _, secs2 = O.gmp_batch(O.GO_POWM_SEC, S2, 64, b[:S2], e[:S2], m[:S2], threads=cores)
            cpu["powm_sec_value"] = round(S2 / secs2, 1)
            cpu["host_cpu"] = "%s, %d cores visible" % (cpu_model(), os.cpu_count() or 0)
            log("cpu baseline: %.1f powm/s (mpz_powm_sec %.1f) on %d cores" % (S / secs, S2 / secs2, cores))
            if not np.array_equal(want[:S2], want2):
                raise SystemExit("mpz_powm and mpz_powm_sec disagree on the CPU sample")
            parity["headline"] = bool(np.array_equal(got[:S], want))
            if not args.quick:
                for key, opname, n, ol, xa, xb, xc in cpu_work:
                    _, _, t = O.gmp_batch(getattr(O, opname), n, ol, xa, xb, xc, threads=cores)
                    extra[key]["cpu_gmp_per_s"] = round(n / t, 1)
                    log("cpu %s: %.1f /s" % (key, n / t))
                # every class's sampled results; the samples (one instance per thread)
                # also give GMP's per-core rate per class, from which the whole mixed
                # draw's time on the same cores is estimated
                ok, est_s = True, 0.0
                for bits, bb, ee, mm, gg in mixed_samples:
                    k = bb.shape[0]
                    w, _, t = O.gmp_batch(O.GO_POWM, k, bits // 32, bb, ee, mm, threads=min(cores, k))
                    ok &= bool(np.array_equal(gg, w))
                    est_s += mixed["class_counts"][str(bits)] * t / cores
                parity["mixed_batch_sample"] = ok
                extra["mixed_batch"]["cpu_gmp_estimated_ms"] = round(est_s * 1e3, 1)
                keys, kk, ki, bb, ee, gg = serve_sample
                parity["serving_sample"] = bool(np.array_equal(
                    gg, O.serving_expected(keys, kk, ki, bb, ee, threads=cores)))
            log("parity: %s" % parity)

    line = {
        "metric": "2048-bit powm/sec on 1 MI355X, batched vs plain launch; p50/p99 latency",
        "value": round(value, 1),
        "unit": "powm/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "powm 2048-bit: 65,536 instances, odd 2048-bit moduli (top bit set), bases below m, "
                               "2048-bit exponents; one xmpIntegersPowm per step",
                   "instances": N, "bits": BITS, "exp_bits": BITS, "parallelism": "none (1 GPU)"},
        "parity": {"bit_exact_vs_gmp": parity},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    h.close()


if __name__ == "__main__":
    main()
