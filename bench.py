"""Benchmark: hafnians/s on the mixed n in [8, 40] complex128 request stream, one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--requests R]

Workload (the metric's): the serving stream of the benchmark configurations -- even n
uniform in [8, 40], half of them loop hafnians, entries complex N(0, 1/n) symmetrised
(``thewalrus_amd.workloads.config5``, seed 141).  A step is one pass of the sieve over a batch
of R requests of that stream with every request's reordered matrix already resident in HBM;
``value`` = R * K / (wall time of K steps).  The same JSON line also carries:
  * per_call: the same requests through the drop-in ``hafnian()`` one call at a time
    (synchronous, host buffers), and the batched / per-call throughput ratio;
  * serving: the continuous-batching server fed by Poisson arrivals at 0.5x and 0.9x of the
    measured batched capacity: p50 / p99 latency (submit to result) and throughput;
  * small_stream: batched vs per-call on the same stream restricted to n <= 24, where
    per-call dispatch cost matters;
  * configs: the other benchmark configurations (1: n = 20 per call; 2: 4096 real 24 x 24 in
    one launch; 3: n = 40 loop single solve; 4: 1024 hafnian_repeated requests);
  * next_rows: the GBS samplers' batched loop hafnians (loop_hafnian_batch, _gamma) and the
    montrealer at 16 and 20 modes, per call, with the CPU oracle beside them;
  * roofline: FP64 product flops of the dominant kernel / its HIP-event time (same stream);
  * cpu_baseline: the CPU oracle (C++ restatement of the reference algorithm, OpenMP over
    terms, all host cores) on a bounded sample of the stream, scaled to hafnians/s by
    algorithmic work.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from thewalrus_amd import workloads as W  # noqa: E402

FP64_PEAK = 78.6e12  # MI355X FP64 (matrix and vector), FLOP/s
METRIC = "hafnians/s + p50/p99 ms, mixed n in [8,40] c128 stream vs per-call GPU"


def timed_runs(batch, k, torch, s):
    """Run a DeviceBatch k times on stream s between synchronisations: (HIP-event s, wall s)."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(k):
        batch.run(s)
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / 1e3, wall


def per_call(reqs, tw):
    t0 = time.perf_counter()
    vals = [tw.hafnian(A, loop=l) for A, l in reqs]
    return len(reqs) / (time.perf_counter() - t0), vals


def serve_run(reqs, qps, Server):
    arrivals = W.arrivals(len(reqs), qps)
    with Server(slots=4, slot_pool_mb=16) as srv:
        t0 = time.perf_counter()
        for (A, l), ta in zip(reqs, arrivals):
            wait = ta - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            srv.submit(A, loop=l)
        srv.drain()
        s = srv.stats.summary()
    s["offered_qps"] = qps
    return s


def serve_closed(reqs, Server, reps=3):
    """Every request submitted at once to a server, after one untimed pass (first launch of
    each kernel class), ``reps`` times: (median hafnians/s from the first submission to the
    last result, the values of the last pass)."""
    rates = []
    with Server(slots=4, slot_pool_mb=16) as srv:
        for f in [srv.submit(A, loop=l) for A, l in reqs]:
            f.result()
        for _ in range(reps):
            t0 = time.perf_counter()
            futs = [srv.submit(A, loop=l) for A, l in reqs]
            vals = [f.result() for f in futs]
            rates.append(len(reqs) / (time.perf_counter() - t0))
    return float(np.median(rates)), vals


def cpu_baseline(reqs, budget_s=12.0):
    """Oracle on a sample of the stream (largest n <= 32 first), parallel over terms (OpenMP,
    one request at a time) and over requests (one single-threaded sieve per thread); hafnians/s
    of the whole stream from the faster measured algorithmic flop rate."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O

    cores = len(os.sched_getaffinity(0))
    sample = sorted([r for r in reqs if r[0].shape[0] <= 32], key=lambda r: -r[0].shape[0])
    done_w, spent, used = 0.0, 0.0, []
    O.hafnian_sieve(np.ones((8, 8)), nthreads=cores)
    for A, loop in sample:
        if spent > budget_s:
            break
        t0 = time.perf_counter()
        O.hafnian_sieve(A, loop=loop, nthreads=cores)
        spent += time.perf_counter() - t0
        done_w += W.w_glynn(A.shape[0])
        used.append((A, loop))
    rate_terms = done_w / spent
    with ThreadPoolExecutor(cores) as ex:
        t0 = time.perf_counter()
        list(ex.map(lambda r: O.hafnian_sieve(r[0], loop=r[1], nthreads=1), used))
        spent_mats = time.perf_counter() - t0
    rate_mats = done_w / spent_mats
    rate = max(rate_terms, rate_mats)
    total_w = sum(W.w_glynn(A.shape[0]) for A, _ in reqs)
    ns = [A.shape[0] for A, _ in used]
    return {
        "value": len(reqs) / (total_w / rate),
        "unit": "hafnians/s",
        "cores": cores,
        "kind": "port",
        "sample": "CPU oracle (C++ restatement of thewalrus _calc_hafnian/_calc_loop_hafnian) on %d stream "
        "requests with n in [%d, %d]: OpenMP over terms, %d threads, %.1f s (%.3g flop/s of the reference "
        "algorithm's products); over requests, %d threads, %.1f s (%.3g flop/s); the faster rate scaled to "
        "the full stream by those flops"
        % (len(used), min(ns), max(ns), cores, spent, rate_terms, cores, spent_mats, rate_mats),
    }


def configs_run(torch, tw, B, stream, with_cpu=True):
    """Configurations 1-4, one measurement each (inputs resident for the batched ones); with
    ``with_cpu`` the n = 24 batch is also timed on the CPU oracle over a bounded sample."""
    out = {}
    A1 = W.config1()
    tw.hafnian(A1)
    t0 = time.perf_counter()
    for _ in range(5):
        tw.hafnian(A1)
    out["config1_n20_per_call_ms"] = (time.perf_counter() - t0) / 5 * 1e3
    if with_cpu:
        from oracle import oracle as O

        cores = len(os.sched_getaffinity(0))
        O.hafnian_sieve(A1, nthreads=cores)
        t0 = time.perf_counter()
        for _ in range(5):
            O.hafnian_sieve(A1, nthreads=cores)
        out["config1_n20_cpu_oracle_ms"] = (time.perf_counter() - t0) / 5 * 1e3

    mats = W.config2()
    b2 = B.DeviceBatch([B.prepare(M) for M in mats])
    b2.run(stream)
    ev, wall = timed_runs(b2, 3, torch, stream)
    r2 = b2.results()
    out["config2_real24_batch"] = {
        "hafnians_per_s": 3 * len(mats) / wall,
        "kernel_tflops": b2.flops / (ev / 3) / 1e12,
        "all_ones_exact": bool(all(round(v.real) == 316234143225 for v in r2[-16:])),
    }
    b2.close()
    if with_cpu:
        # the same real n = 24 matrices on the CPU oracle (all host threads), 32 of them
        from oracle import oracle as O

        from concurrent.futures import ThreadPoolExecutor

        cores = len(os.sched_getaffinity(0))
        O.hafnian_sieve(mats[0], nthreads=cores)
        t0 = time.perf_counter()
        for M in mats[:32]:  # OpenMP over terms, one matrix at a time
            O.hafnian_sieve(M, nthreads=cores)
        rate_terms = 32 / (time.perf_counter() - t0)
        with ThreadPoolExecutor(cores) as ex:  # over matrices: one single-threaded sieve per thread
            list(ex.map(lambda M: O.hafnian_sieve(M, nthreads=1), mats[:cores]))
            t0 = time.perf_counter()
            list(ex.map(lambda M: O.hafnian_sieve(M, nthreads=1), mats[:8 * cores]))
            rate_mats = 8 * cores / (time.perf_counter() - t0)
        cpu_rate = max(rate_terms, rate_mats)
        row = out["config2_real24_batch"]
        row["cpu_oracle_hafnians_per_s"] = cpu_rate
        row["cpu_oracle_over_terms"] = rate_terms
        row["cpu_oracle_over_matrices"] = rate_mats
        row["cpu_cores"] = cores
        row["gpu_over_cpu"] = row["hafnians_per_s"] / cpu_rate

    A3 = W.config3()
    tw.hafnian(A3, loop=True)
    b3 = B.DeviceBatch([B.prepare(A3, loop=True)])
    b3.run(stream)
    ev, wall = timed_runs(b3, 3, torch, stream)
    t0 = time.perf_counter()
    tw.hafnian(A3, loop=True)
    out["config3_n40_loop"] = {
        "per_call_s": time.perf_counter() - t0,
        "kernel_s": ev / 3,
        "kernel_tflops": b3.flops / (ev / 3) / 1e12,
        "frac_fp64_peak": b3.flops / (ev / 3) / FP64_PEAK,
        "reference_algorithm_equivalent_tflops": W.w_glynn(40) / (ev / 3) / 1e12,
    }
    b3.close()

    reqs4 = W.config4()
    b4 = B.DeviceBatch([B.prepare_repeated(A, rpt, mu=mu, loop=True) for A, rpt, mu in reqs4])
    b4.run(stream)
    ev, wall = timed_runs(b4, 3, torch, stream)
    out["config4_repeated_batch"] = {"requests_per_s": 3 * len(reqs4) / wall, "kernel_ms": ev / 3 * 1e3}
    b4.close()
    if with_cpu:
        # the same requests on the CPU oracle: over terms (all threads per request) on 128 of them,
        # over requests (one single-threaded sieve per thread) on all; the faster kept
        from concurrent.futures import ThreadPoolExecutor

        from oracle import oracle as O

        cores = len(os.sched_getaffinity(0))

        def one(r, nt):
            A, rpt, mu = r
            return O.hafnian_sieve(A, loop=True, rpt=rpt, mu=mu, nthreads=nt)

        t0 = time.perf_counter()
        for r in reqs4[:128]:
            one(r, cores)
        rate_terms = 128 / (time.perf_counter() - t0)
        with ThreadPoolExecutor(cores) as ex:
            t0 = time.perf_counter()
            list(ex.map(lambda r: one(r, 1), reqs4))
            rate_reqs = len(reqs4) / (time.perf_counter() - t0)
        row = out["config4_repeated_batch"]
        row["cpu_oracle_requests_per_s"] = max(rate_terms, rate_reqs)
        row["cpu_oracle_over_terms"] = rate_terms
        row["cpu_oracle_over_requests"] = rate_reqs
        row["gpu_over_cpu"] = row["requests_per_s"] / max(rate_terms, rate_reqs)
    return out


def next_rows_run(with_cpu):
    """The sampler's batched loop hafnian and the montrealer: device time per call (host
    buffers in and out, as the drop-ins take them) and, with with_cpu, the CPU oracle on the
    same inputs."""
    from thewalrus_amd import lmtl, mtl
    from thewalrus_amd.loop_hafnian_batch import loop_hafnian_batch
    from thewalrus_amd.loop_hafnian_batch_gamma import loop_hafnian_batch_gamma

    out = {}
    rng = np.random.default_rng(2024)

    def per_call_ms(fn, reps):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) / reps * 1e3

    # hafnian sampler step: 16 modes, 15 detected with 0-3 photons each, cutoff 12 on the last
    n = 16
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A = (X + X.T) * 0.1
    D = rng.normal(size=n) + 1j * rng.normal(size=n)
    fixed = list(rng.integers(0, 4, size=n - 1))
    row = {"workload": "loop_hafnian_batch, %d modes, fixed photons %d, N_cutoff 12" % (n, sum(fixed)),
           "gpu_ms": per_call_ms(lambda: loop_hafnian_batch(A, D, fixed, 12), 5)}
    if with_cpu:
        from oracle import oracle as O

        row["cpu_oracle_ms"] = per_call_ms(lambda: O.loop_hafnian_batch(A, D, fixed, 12), 1)
    out["loop_hafnian_batch"] = row
    # threshold sampler step: fanout 10 diagonals, cutoff 1
    Dg = rng.normal(size=(10, n)) + 1j * rng.normal(size=(10, n))
    fixed1 = [int(f > 0) for f in fixed]
    row = {"workload": "loop_hafnian_batch_gamma, %d modes, %d clicks, fanout 10, N_cutoff 1" % (n, sum(fixed1)),
           "gpu_ms": per_call_ms(lambda: loop_hafnian_batch_gamma(A, Dg, fixed1, 1), 5)}
    if with_cpu:
        row["cpu_oracle_ms"] = per_call_ms(lambda: O.loop_hafnian_batch_gamma(A, Dg, fixed1, 1), 1)
    out["loop_hafnian_batch_gamma"] = row

    # montrealer: 2^n subsets of 2n x 2n products
    for nm in (16, 20):
        S = rng.normal(size=(2 * nm, 2 * nm)) + 1j * rng.normal(size=(2 * nm, 2 * nm))
        S = (S + S.T) / (2 * np.sqrt(2 * nm))
        z = rng.normal(size=2 * nm) + 1j * rng.normal(size=2 * nm)
        flops = 2.0 ** nm * 8 * (2 * nm) ** 3 * ((nm + 1) // 2 - 1)
        ms = per_call_ms(lambda: mtl(S), 2)
        row = {"gpu_ms": ms, "gpu_ms_loop": per_call_ms(lambda: lmtl(S, z), 2),
               "product_tflops": flops / (ms / 1e3) / 1e12}
        if with_cpu and nm == 16:
            row["cpu_oracle_ms"] = per_call_ms(lambda: O.mtl(S), 1)
        out["mtl_n%d" % nm] = row
    return out


# HBM bytes per launch of the dominant kernel on the n in [34, 40] loop sub-batch of the
# default stream: FETCH_SIZE x 2 + WRITE_SIZE from rocprofv3, scaled by the copy calibration
TRAFFIC_PM1S_T5_LOOP = 4.41e6


def mfma_flops_executed(n):
    """FP64 flops the symmetric-power kernel issues on the matrix cores for one n x n request:
    per product T(T+1)/2 block groups of 8 x 8, each 4 real 4x4x4 (four-block) MFMAs per k chunk
    of 4, 2T chunks (n_pad = 8T); ceil(h/2) - 1 products per term, 2^(h-1) terms."""
    h = n // 2
    T = (n + 7) // 8
    per_product = T * (T + 1) // 2 * 4 * 512 * 2 * T
    return 2.0 ** (h - 1) * ((h + 1) // 2 - 1) * per_product


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--requests", type=int, default=128)
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--nmin", type=int, default=8, help="drop stream requests with n < nmin")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-percall", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--no-next", action="store_true", help="skip the batched loop hafnian / montrealer lines")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import torch

    import thewalrus_amd as tw
    from thewalrus_amd import batch as B
    from thewalrus_amd.serve import Server

    torch.cuda.init()
    if args.calibrate_copy_mb:
        src = torch.ones(args.calibrate_copy_mb << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    stream_all = W.config5(count=max(args.requests, args.serve_requests, 512))
    reqs = [r for r in stream_all[: args.requests] if r[0].shape[0] >= args.nmin]
    wl = "mixed stream: even n uniform in [8,40], 50% loop, complex128 CN(0,1/n) symmetrised, seed 141"
    if args.nmin > 8:
        wl += " (only n >= %d)" % args.nmin
    preps = [B.prepare(A, loop=l) for A, l in reqs]
    batch = B.DeviceBatch(preps)
    stream = torch.cuda.Stream()  # default priority; kernels and events share it

    for _ in range(args.warmup):
        batch.run(stream)
    torch.cuda.synchronize()
    ev_s, wall = timed_runs(batch, args.steps, torch, stream)
    value = len(reqs) * args.steps / wall
    res = batch.results()

    # dominant kernel: the largest tile class (n in [34, 40]); plain and loop launches timed
    # separately on the stream they run on, the slower one reported
    roof = None
    best = None
    for want_loop in (False, True):
        idx = [i for i, (A, l) in enumerate(reqs) if A.shape[0] >= 34 and l == want_loop]
        if not idx:
            continue
        sub = B.DeviceBatch([preps[i] for i in idx])
        sub.run(stream)
        t, _ = timed_runs(sub, 1, torch, stream)
        wref = sum(W.w_glynn(reqs[i][0].shape[0]) for i in idx)
        wexe = sum(mfma_flops_executed(reqs[i][0].shape[0]) for i in idx)
        if best is None or t > best[0]:
            best = (t, sub.flops, want_loop, len(idx), wref, wexe)
        sub.close()
    if best:
        t, fl, lp, cnt, wref, wexe = best
        roof = {
            "bound": "mfma",
            "achieved": fl / t / 1e12,
            "peak": FP64_PEAK / 1e12,
            "unit": "TFLOP/s",
            "frac": fl / t / FP64_PEAK,
            # HBM bytes per launch of wh_sieve_pm1s<5, loop> on this sub-batch from rocprofv3
            # FETCH_SIZE x 2 + WRITE_SIZE (copy-calibrated), profiles/pm1s_T5_loop_counters.md
            "traffic": TRAFFIC_PM1S_T5_LOOP,
            "kernel": "wh_sieve_pm1s<5,%s> (FP64 matrix cores, symmetric powers, n in [34,40], %d requests)"
            % ("loop" if lp else "plain", cnt),
            "flops_formula": "sum over requests of 2^(h-1) * 8 * n(n+1)/2 * n * (ceil(h/2)-1): the upper "
            "triangle of the symmetric complex n x n products per term (powers up to ceil(h/2)); "
            "lower-order trace and loop terms not counted",
            "mfma_executed_tflops": wexe / t / 1e12,
            "kernel_time_s": t,
            "reference_algorithm_equivalent_tflops": wref / t / 1e12,
        }

    pc = None
    if not args.no_percall:
        rate, vals = per_call(reqs, tw)
        pc = {
            "value": rate,
            "unit": "hafnians/s",
            "max_rel_diff_vs_batched": float(max(abs(a - b) / max(abs(b), 1e-300) for a, b in zip(vals, res))),
        }

    serving = None
    if not args.no_serve:
        sreqs = stream_all[: args.serve_requests]
        serving = {"requests": len(sreqs)}
        for load in (0.5, 0.9):
            serving["load_%.1f" % load] = serve_run(sreqs, load * value, Server)

    small = None
    if not args.no_percall:
        sm = [r for r in stream_all if r[0].shape[0] <= 24][:256]
        sb = B.DeviceBatch([B.prepare(A, loop=l) for A, l in sm])
        sb.run(stream)
        _, w = timed_runs(sb, 3, torch, stream)
        r_eager = sb.results()
        sb.capture(stream)  # the same plan as one HIP graph launch
        sb.run(stream)
        _, wg = timed_runs(sb, 3, torch, stream)
        rate, _ = per_call(sm, tw)
        served, served_vals = serve_closed(sm, Server)
        small = {"requests": len(sm), "batched": 3 * len(sm) / w, "per_call": rate,
                 "batched_over_per_call": 3 * len(sm) / w / rate,
                 "batched_hip_graph": 3 * len(sm) / wg,
                 "graph_equals_eager": bool(np.array_equal(sb.results(), r_eager)),
                 "served": served, "served_over_per_call": served / rate,
                 "served_max_rel_diff_vs_batched": float(max(
                     abs(a - b) / max(abs(b), 1e-300) for a, b in zip(served_vals, r_eager)))}
        sb.close()

    configs = None if args.no_configs else configs_run(torch, tw, B, stream, with_cpu=not args.no_cpu)
    next_rows = None if args.no_next else next_rows_run(not args.no_cpu)
    cpu = None if args.no_cpu else cpu_baseline(reqs)
    head = serving["load_0.9"] if serving else {}
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "hafnians/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "ms_per_step_hip_events": ev_s / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl, "requests_per_step": len(reqs), "parallelism": "single GPU"},
        "p50_ms": head.get("p50_ms"),
        "p99_ms": head.get("p99_ms"),
        "per_call": pc,
        "batched_over_per_call": (value / pc["value"]) if pc else None,
        "serving": serving,
        "small_stream": small,
        "configs": configs,
        "next_rows": next_rows,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
