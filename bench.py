"""Benchmark: hafnians/s and p50/p99 latency on the mixed n in [8, 40] complex128 request stream
served on one MI355X, beside per-call GPU dispatch.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--step-requests R] [--load L]

Workload (the metric's): the serving stream of the benchmark configurations -- Poisson
arrivals, even n uniform in [8, 40], half of them loop hafnians, entries complex N(0, 1/n)
symmetrised (``thewalrus_amd.workloads.config5``, seed 141).  A step is R = 500 consecutive
requests of that stream; K = 20 steps serve its 10^4 requests once.  Before the timed region the
same K R requests run as K resident device batches of R (the batched capacity C).  The timed
region then runs the K steps as one continuous stream: each request is submitted to the
continuous-batching server at its Poisson arrival time (offered rate ``qps`` = L x C, L = 1 by
default), host preparation (validation, pairing, reorder) and submission inside the region, and
the region ends when the last result is back.  ``value`` = K R / (wall time of the region):
served hafnians/s.  p50 / p99 are the latencies (submit to result) of those same K R requests.
Warm-up: the W steps of a warm-up stream (seed 142) run batched and served first.

The same JSON line also carries:
  * batched: the batched capacity C (one plan per step, inputs already in HBM) and served / C;
  * serving_half_load / serving_0.9_load: latencies of the first two steps served at 0.5 C and
    at 0.9 C;
  * warmup_step_served_closed_loop / warmup_step_batched: one warm-up step submitted to the
    server at once, and the same step as one resident batch;
  * per_call: the same K R requests through the drop-in ``hafnian()`` one call at a time
    (synchronous, host buffers), served / per-call and batched / per-call;
  * small_stream: batched vs per-call on the stream restricted to n <= 24, where per-call
    dispatch cost matters;
  * configs: the other benchmark configurations (1: n = 20 per call; 2: 4096 real 24 x 24 in
    one launch; 3: n = 40 loop single solve; 4: 1024 hafnian_repeated requests);
  * next_rows: the GBS samplers' batched loop hafnians (loop_hafnian_batch, _gamma) and the
    montrealer at 16 and 20 modes, per call, with the CPU oracle beside them;
  * roofline: FP64 product flops of the dominant kernel / its HIP-event time (same stream); its
    HBM traffic per launch from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) of
    ``--roofline-only`` run as child processes, copy-calibrated (the committed counter summary
    when rocprofv3 is absent or with --no-live-traffic);
  * cpu_baseline: the CPU restatement of the reference algorithm on host cores, timed per n on
    a bounded sample of terms, summed over the stream's requests.
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


def serve_closed(reqs, Server, reps=7):
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


def cpu_class_seconds(A, loop, cores, target_s=0.25):
    """Seconds per hafnian of one (n, loop) class on the timed CPU build (all host threads):
    n >= 24, over terms on a bounded term range of one request scaled by its term count (every
    term of an all-ones request costs the same); n < 24, over requests (one sieve per thread)."""
    from oracle import oracle as O

    n = A.shape[0]
    h = n // 2
    steps = 2 ** (h - 1)
    Ax, Dx = O.ct_prepare(A, loop)
    if n >= 24:
        k = min(steps, 64 * cores)
        O.ct_sieve(Ax, Dx, 0, k, cores)
        t0 = time.perf_counter()
        O.ct_sieve(Ax, Dx, 0, k, cores)
        dt = time.perf_counter() - t0
        k2 = int(min(steps, max(k, k * target_s / max(dt, 1e-6))))
        t0 = time.perf_counter()
        O.ct_sieve(Ax, Dx, 0, k2, cores)
        return (time.perf_counter() - t0) / k2 * steps, "%d of %d terms" % (k2, steps)
    count = 8 * cores
    O.ct_sieve_many([Ax] * cores, None if Dx is None else [Dx] * cores, cores)
    t0 = time.perf_counter()
    O.ct_sieve_many([Ax] * count, None if Dx is None else [Dx] * count, cores)
    dt = time.perf_counter() - t0
    reps = int(max(1, target_s / max(dt, 1e-6)))
    if reps > 1:
        t0 = time.perf_counter()
        for _ in range(reps):
            O.ct_sieve_many([Ax] * count, None if Dx is None else [Dx] * count, cores)
        dt = (time.perf_counter() - t0) / reps
    return dt / count, "%d requests x %d" % (count, reps)


def cpu_baseline(reqs):
    """The reference's CPU path, timed: the performance restatement (oracle/cpu_timing.cpp,
    blocked complex products with FMA, OpenMP over terms as numba's prange, or over requests for
    small n) per (n, loop) class of the stream on every host thread, each class on a bounded
    sample; hafnians/s of the whole stream = requests / the sum of their class times."""
    from oracle import oracle as O

    cores = len(os.sched_getaffinity(0))
    first = {}
    for A, loop in reqs:
        first.setdefault((A.shape[0], loop), A)
    per = {}
    notes = []
    t_all = time.perf_counter()
    for (n, loop) in sorted(first):
        sec, how = cpu_class_seconds(first[(n, loop)], loop, cores)
        per[(n, loop)] = sec
        notes.append("n=%d%s %.4g ms (%s)" % (n, " loop" if loop else "", sec * 1e3, how))
    spent = time.perf_counter() - t_all
    total = sum(per[(A.shape[0], loop)] for A, loop in reqs)
    _, isa = O.cpu_timing_lib()
    return {
        "value": len(reqs) / total,
        "unit": "hafnians/s",
        "cores": cores,
        "kind": "port",
        "sample": "CPU restatement of the reference algorithm's CPU path (oracle/cpu_timing.cpp: per-term "
        "products as register-blocked complex products with FMA, -O3 -march=%s, OpenMP over terms as numba's "
        "prange; over requests for n < 24) on %d host threads; each (n, loop) class of the %d-request stream "
        "timed on one request (a bounded range of its terms, scaled by the term count) in %.1f s in all; "
        "stream time = sum over requests of their class time = %.0f s. Per class: %s"
        % (isa, cores, len(reqs), spent, total, "; ".join(notes)),
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
        Ax1, _ = O.ct_prepare(A1, False)
        O.ct_sieve(Ax1, None, 0, 512, cores)
        t0 = time.perf_counter()
        for _ in range(5):
            O.ct_sieve(Ax1, None, 0, 512, cores)
        out["config1_n20_cpu_ms"] = (time.perf_counter() - t0) / 5 * 1e3

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
        # the same real n = 24 matrices on the timed CPU build (complex arithmetic, as the
        # reference casts them), all host threads: over terms (one matrix at a time) and over
        # matrices (one sieve per thread), the faster kept
        from oracle import oracle as O

        cores = len(os.sched_getaffinity(0))
        prepped = [O.ct_prepare(M, False)[0] for M in mats[:8 * cores]]
        O.ct_sieve(prepped[0], None, 0, 2048, cores)
        t0 = time.perf_counter()
        for Ax in prepped[:16]:
            O.ct_sieve(Ax, None, 0, 2048, cores)
        rate_terms = 16 / (time.perf_counter() - t0)
        O.ct_sieve_many(prepped[:cores], None, cores)
        t0 = time.perf_counter()
        O.ct_sieve_many(prepped, None, cores)
        rate_mats = len(prepped) / (time.perf_counter() - t0)
        cpu_rate = max(rate_terms, rate_mats)
        row = out["config2_real24_batch"]
        row["cpu_hafnians_per_s"] = cpu_rate
        row["cpu_over_terms"] = rate_terms
        row["cpu_over_matrices"] = rate_mats
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
    p4 = [B.prepare_repeated(A, rpt, mu=mu, loop=True) for A, rpt, mu in reqs4]
    b4 = B.DeviceBatch(p4)
    b4.run(stream)
    ev, wall = timed_runs(b4, 3, torch, stream)
    wref = sum(w_rep(p.job) for p in p4 if p.on_device)
    out["config4_repeated_batch"] = {
        "requests_per_s": 3 * len(reqs4) / wall, "kernel_ms": ev / 3 * 1e3,
        "kernel_tflops": b4.flops / (ev / 3) / 1e12, "frac_fp64_peak": b4.flops / (ev / 3) / FP64_PEAK,
        "flops_formula": "per term ceil(h/2) - 1 products of the padded 2E x 2E term matrix (8 flops per "
        "complex multiply-add), h = N/2",
        "reference_algorithm_equivalent_tflops": wref / (ev / 3) / 1e12,
        "reference_formula": "per term 8 m^3 max(min(h+1, m) - 2, 0), m = 2 nnz(delta) (charpoly.py:316-318)"}
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
    """The next rows: the samplers' batched loop hafnians, the montrealer, hafnian_batched (the
    Hermite recursion) and a probability tensor: device time per call (host buffers in and out,
    as the drop-ins take them) and, with with_cpu, the CPU oracle on the same inputs."""
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

    # hafnian_batched: the Hermite recursion's whole tensor, 6 modes, cutoff 10 (10^6 entries)
    from thewalrus_amd._hermite_multidimensional import hafnian_batched

    nh = 6
    Xh = rng.normal(size=(nh, nh)) + 1j * rng.normal(size=(nh, nh))
    Ah = (Xh + Xh.T) * 0.15
    muh = 0.3 * (rng.normal(size=nh) + 1j * rng.normal(size=nh))
    row = {"workload": "hafnian_batched, %d modes, cutoff 10, loop weights (the whole tensor)" % nh,
           "gpu_ms": per_call_ms(lambda: hafnian_batched(Ah, 10, mu=muh), 3)}
    if with_cpu:
        row["cpu_oracle_ms"] = per_call_ms(lambda: O.hafnian_batched(Ah, 10, mu=muh), 1)
    out["hafnian_batched"] = row

    # probabilities of a 4-mode Gaussian state up to 6 photons per mode (1296 density-matrix
    # diagonal elements: repeated-row loop hafnians, one device batch)
    from thewalrus_amd import quantum as Q

    nq = 4
    Sg = rng.normal(size=(2 * nq, 2 * nq))
    cov = np.identity(2 * nq) + 0.2 * (Sg @ Sg.T) / (2 * nq)
    muq = 0.3 * rng.normal(size=2 * nq)
    out["probabilities"] = {"workload": "probabilities, %d modes, cutoff 6, displaced mixed state" % nq,
                            "gpu_ms": per_call_ms(lambda: Q.probabilities(muq, cov, 6), 3)}
    return out


def counter_traffic(name):
    """HBM bytes per launch of a kernel from its committed counter summary
    (profiles/<name>_counters.md, line ``traffic_bytes_per_launch: X``: FETCH_SIZE and
    WRITE_SIZE of two rocprofv3 --pmc runs of this benchmark, scaled by a copy kernel's
    calibration), or None when there is none."""
    path = os.path.join(ROOT, "profiles", name + "_counters.md")
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch:"):
                    return float(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def live_traffic(loop, timeout_s=120):
    """HBM bytes per launch of the dominant kernel measured in this run: two rocprofv3 --pmc
    child runs of ``bench.py --roofline-only --calibrate-copy-mb 256`` (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch, averaged over the kernel's launches), each scaled by the
    reading of the same counter for the child's 256 MiB device-to-device copy (the bytes the copy
    reads, resp. writes, over its counter).  None when rocprofv3 is absent or a pass fails."""
    import csv
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    want = "wh_sieve_pm1s<5, %s" % ("true" if loop else "false")
    copy_bytes = 256 << 20
    got = {}
    env = dict(os.environ, TMPDIR="/tmp")
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            od = os.path.join(d, ctr.lower())
            cmd = [exe, "--pmc", ctr, "--kernel-trace", "--output-format", "csv", "-d", od, "-o", "p", "--",
                   sys.executable, os.path.abspath(__file__), "--roofline-only", "--calibrate-copy-mb", "256"]
            try:
                r = subprocess.run(cmd, cwd=d, env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                                   timeout=timeout_s)
            except (subprocess.TimeoutExpired, OSError):
                return None
            path = os.path.join(od, "p_counter_collection.csv")
            if not os.path.exists(path):
                return None
            per = {}
            with open(path, newline="") as f:
                for row in csv.DictReader(f):
                    if row.get("Counter_Name") != ctr:
                        continue
                    key = (row["Dispatch_Id"], row["Kernel_Name"])
                    per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
            copies = [v for (_, name), v in per.items() if "copyBuffer" in name]
            kern = [v for (_, name), v in per.items() if name.startswith("void " + want)]
            if r.returncode not in (0, -11, 139) or not copies or not kern or max(copies) <= 0:
                return None
            got[ctr] = (sum(kern) / len(kern), max(copies))
    fetch, fcopy = got["FETCH_SIZE"]
    write, wcopy = got["WRITE_SIZE"]
    ff, wf = copy_bytes / (1024.0 * fcopy), copy_bytes / (1024.0 * wcopy)
    return {"bytes_per_launch": fetch * 1024 * ff + write * 1024 * wf, "fetch_kib": fetch, "write_kib": write,
            "fetch_scale": ff, "write_scale": wf}


def w_rep(job):
    """The reference's product flops for one prepared repeated-row request: over its terms
    (mixed-radix digits, _hafnian.py:162-180), 8 m^3 max(min(h+1, m) - 2, 0) with m = 2 nnz(delta)
    (charpoly.py:316-318)."""
    e = np.asarray(job.edge_reps, dtype=np.int64)
    h = (2 * int(e.sum()) + (job.oddloop is not None)) // 2
    first = (e[0] + 2) // 2 if job.glynn and job.oddloop is None else e[0] + 1
    grids = np.meshgrid(*([np.arange(first)] + [np.arange(x + 1) for x in e[1:]]), indexing="ij")
    k = np.stack([g.ravel() for g in grids])
    delta = 2 * k - e[:, None] if job.glynn else k
    m = 2 * (delta != 0).sum(axis=0)
    return float((8.0 * m ** 3 * np.maximum(np.minimum(h + 1, m) - 2, 0)).sum())


def mfma_flops_executed(n, loop):
    """FP64 flops the symmetric-power kernel issues on the matrix cores for one complex n x n
    request: per product T(T+1)/2 block groups of 8 x 8, each 3 real 4x4x4 (four-block) MFMAs
    per k chunk of 4 (Gauss's three-multiplication form; 4 for the T = 6 loop and T >= 7
    instances), 2T chunks (n_pad = 8T); ceil(h/2) - 1 products per term, 2^(h-1) terms."""
    h = n // 2
    T = (n + 7) // 8
    mults = 3 if T <= 5 or (T == 6 and not loop) else 4
    per_product = T * (T + 1) // 2 * mults * 512 * 2 * T
    return 2.0 ** (h - 1) * ((h + 1) // 2 - 1) * per_product


def serve_stream(reqs, qps, Server, seed):
    """Submit ``reqs`` to a fresh server at Poisson arrival times of rate ``qps`` (seeded) and
    wait for every result: (wall seconds from the first submission to the last result, the
    server's stats summary, the values in request order)."""
    arrivals = W.arrivals(len(reqs), qps, seed=seed)
    with Server(slots=4, slot_pool_mb=16) as srv:
        futs = []
        t0 = time.perf_counter()
        for (A, l), ta in zip(reqs, arrivals):
            wait = ta - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            futs.append(srv.submit(A, loop=l))
        vals = [f.result() for f in futs]
        wall = time.perf_counter() - t0
        stats = srv.stats.summary()
    stats["offered_qps"] = qps
    return wall, stats, np.array(vals)


def batched_steps(reqs, R, B, torch, stream, reps=1):
    """The requests as consecutive resident batches of R (one plan each), run in turn on one
    stream: (total wall seconds of the runs, values).  Packing and the H2D copy happen before
    the timed runs."""
    batches = [B.DeviceBatch([B.prepare(A, loop=l) for A, l in reqs[k:k + R]]) for k in range(0, len(reqs), R)]
    for b in batches:
        b.run(stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for b in batches:
            b.run(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    vals = np.concatenate([b.results() for b in batches])
    for b in batches:
        b.close()
    return wall, vals


def dominant_kernel(step0, B, torch, stream):
    """Roofline of the dominant kernel: the largest tile class (n in [34, 40]) of a step's
    requests, plain and loop launches timed separately (HIP events on the stream they run on),
    the slower one reported."""
    roof = None
    best = None
    for want_loop in (False, True):
        idx = [i for i, (A, l) in enumerate(step0) if A.shape[0] >= 34 and l == want_loop]
        if not idx:
            continue
        sub = B.DeviceBatch([B.prepare(*step0[i]) for i in idx])
        sub.run(stream)
        t, _ = timed_runs(sub, 1, torch, stream)
        wref = sum(W.w_glynn(step0[i][0].shape[0]) for i in idx)
        wexe = sum(mfma_flops_executed(step0[i][0].shape[0], step0[i][1]) for i in idx)
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
            "traffic": counter_traffic("pm1s_T5_%s" % ("loop" if lp else "plain")),
            "kernel": "wh_sieve_pm1s<5,%s> (FP64 matrix cores, symmetric powers, n in [34,40], %d requests of "
            "the first step)" % ("loop" if lp else "plain", cnt),
            "flops_formula": "sum over requests of 2^(h-1) * 8 * n(n+1)/2 * n * (ceil(h/2)-1): the upper "
            "triangle of the symmetric complex n x n products per term (powers up to ceil(h/2)), 8 flops "
            "per complex multiply-add as in the standard product (the kernel forms it with 3 real "
            "multiplications: mfma_executed_tflops); lower-order trace and loop terms not counted",
            "mfma_executed_tflops": wexe / t / 1e12,
            "kernel_time_s": t,
            "reference_algorithm_equivalent_tflops": wref / t / 1e12,
        }

    return roof


def kernel_probe(what, B, torch, stream, reps=5):
    """Launch sequences for kernel traces: real n = 32 and n = 40 batches at the default waves
    per workgroup and at 3 (real_waves), the real n = 24 batch of configuration 2 (config2), the
    repeated-row batch of configuration 4 on the
    default route (tile sides 3 and 1), with every request at tile side 2 (2E <= 16, one term per
    wave) and on the general kernel (config4), or the first 1000 requests of the stream as
    batches of 500 with wh_tune_persistent 0, 1 and 2 (persistent); wall seconds per pass."""
    from thewalrus_amd import _native

    out = {}
    if what == "real_waves":
        rng = np.random.default_rng(7)
        for n, loop, cnt in ((32, False, 16), (32, True, 16), (40, False, 1), (40, True, 1)):
            mats = [rng.uniform(-1, 1, size=(n, n)) for _ in range(cnt)]
            preps = [B.prepare((M + M.T) / 2, loop=loop) for M in mats]
            row = {}
            for w in (0, 3):
                b = B.DeviceBatch(preps, waves=w)
                b.run(stream)
                ev, wall = timed_runs(b, reps, torch, stream)
                row["waves_%d" % w] = {"kernel_ms": ev / reps * 1e3, "blocks": b.blocks}
                b.close()
            out["real_n%d_%s_x%d" % (n, "loop" if loop else "plain", cnt)] = row
        return out
    if what == "config2":
        b = B.DeviceBatch([B.prepare(M) for M in W.config2()])
        b.run(stream)
        ev, wall = timed_runs(b, reps, torch, stream)
        vals = b.results()
        out["default"] = {"kernel_ms": ev / reps * 1e3, "hafnians_per_s": reps * len(vals) / wall,
                          "symmetric_tflops": b.flops / (ev / reps) / 1e12}
        out["ones_exact"] = bool(round(vals[-1].real) == 316234143225)
        b.close()
        return out
    if what == "config4":
        preps = [B.prepare_repeated(A, rpt, mu=mu, loop=True) for A, rpt, mu in W.config4()]
        vals = {}
        for name, kw in (("default", {}), ("tile_side_2", {"tile": 2}), ("general_kernel", {"force_generic": True})):
            b = B.DeviceBatch(preps, **kw)
            b.run(stream)
            ev, wall = timed_runs(b, reps, torch, stream)
            out[name] = {"kernel_ms": ev / reps * 1e3, "requests_per_s": reps * len(preps) / wall}
            vals[name] = b.results()
            b.close()
        out["max_rel_diff"] = max_rel(vals["default"], vals["general_kernel"])
        out["max_rel_diff_side_2"] = max_rel(vals["default"], vals["tile_side_2"])
        return out
    L = _native.load()
    prev = L.wh_tune_persistent(-1)
    reqs = W.config5(count=1000, seed=141)
    try:
        sums = {}
        for mode in (0, 1, 2):
            L.wh_tune_persistent(mode)
            wall, vals = batched_steps(reqs, 500, B, torch, stream, reps=2)
            out["mode_%d_s_per_1000" % mode] = wall
            sums[mode] = vals
        out["bit_identical"] = bool(np.array_equal(sums[0], sums[1]) and np.array_equal(sums[0], sums[2]))
    finally:
        L.wh_tune_persistent(prev)
    return out


def max_rel(a, b):
    return float(max(abs(x - y) / max(abs(y), 1e-300) for x, y in zip(a, b)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--step-requests", type=int, default=500, help="requests per step")
    ap.add_argument("--load", type=float, default=1.0,
                    help="offered rate / the batched capacity measured on the same requests")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-percall", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--no-next", action="store_true", help="skip the batched loop hafnian / montrealer lines")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the dominant kernel's launches (for counter runs under rocprofv3)")
    ap.add_argument("--only", choices=["config2", "config4", "persistent", "real_waves"],
                    help="for kernel traces under rocprofv3: real_waves = real n = 32 and 40 batches "
                    "with the default waves per workgroup and with 3; config2 = the real n = 24 batch on the "
                    "default route; config4 = the repeated-row batch on the "
                    "default route and on the general kernel; persistent = 1000 stream requests as "
                    "batches of 500 with the persistent launch off, on (default rule) and forced")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (counter calibration)")
    ap.add_argument("--no-live-traffic", action="store_true",
                    help="take the roofline's traffic from the committed counter summary instead of two "
                    "rocprofv3 --pmc passes in this run")
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
    R, K = args.step_requests, args.steps
    stream = torch.cuda.Stream()  # default priority; kernels and events share it
    if args.roofline_only:
        print(json.dumps({"roofline": dominant_kernel(W.config5(count=R, seed=141), B, torch, stream)}))
        return
    if args.only:
        print(json.dumps(kernel_probe(args.only, B, torch, stream)))
        return
    warm = W.config5(count=max(args.warmup, 1) * R, seed=142)
    reqs = W.config5(count=K * R, seed=141)

    # warm-up: the warm-up stream as resident batches of R, then served at that capacity; then
    # one warm-up step submitted at once (closed loop: the served capacity)
    wwall, _ = batched_steps(warm, R, B, torch, stream)
    if args.warmup:
        serve_stream(warm, len(warm) / wwall, Server, seed=142)
    with Server(slots=4, slot_pool_mb=16) as srv:
        t0 = time.perf_counter()
        for f in [srv.submit(A, loop=l) for A, l in warm[:R]]:
            f.result()
        served_cap = R / (time.perf_counter() - t0)
    wstep, _ = batched_steps(warm[:R], R, B, torch, stream)

    # the batched capacity of the timed requests (K resident batches of R): sets the offered rate
    bwall, bvals = batched_steps(reqs, R, B, torch, stream)
    capacity = len(reqs) / bwall
    qps = args.load * capacity

    torch.cuda.synchronize()
    wall, stats, served_vals = serve_stream(reqs, qps, Server, seed=141)
    torch.cuda.synchronize()
    value = len(reqs) / wall
    batched = {"value": capacity, "unit": "hafnians/s", "ms_per_step": bwall / K * 1e3,
               "served_over_batched": value / capacity,
               "served_max_rel_diff_vs_batched": max_rel(served_vals, bvals)}
    # latency at half and at 0.9 of the capacity on the first two steps
    half = reqs[:2 * R]
    _, stats_half, _ = serve_stream(half, 0.5 * capacity, Server, seed=143)
    _, stats_09, _ = serve_stream(half, 0.9 * capacity, Server, seed=144)

    roof = dominant_kernel(reqs[:R], B, torch, stream)
    if roof:
        roof["traffic_source"] = "profiles/%s_counters.md" % ("pm1s_T5_loop" if "loop" in roof["kernel"] else
                                                               "pm1s_T5_plain")
        lt = None if args.no_live_traffic else live_traffic("loop" in roof["kernel"])
        if lt:
            roof["traffic"] = lt["bytes_per_launch"]
            roof["traffic_source"] = "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes of this run, copy-calibrated"
            roof["traffic_counters"] = lt

    pc = None
    if not args.no_percall:
        rate, vals = per_call(reqs, tw)
        pc = {"value": rate, "unit": "hafnians/s", "served_over_per_call": value / rate,
              "max_rel_diff_vs_served": max_rel(vals, served_vals)}
        if batched:
            batched["batched_over_per_call"] = batched["value"] / rate

    small = None
    if not args.no_percall:
        sm = [r for r in warm + reqs if r[0].shape[0] <= 24][:256]
        sb = B.DeviceBatch([B.prepare(A, loop=l) for A, l in sm])
        sb.run(stream)
        _, w = timed_runs(sb, 3, torch, stream)
        r_eager = sb.results()
        sb.capture(stream)  # the same plan as one HIP graph launch
        sb.run(stream)
        _, wg = timed_runs(sb, 3, torch, stream)
        rate, _ = per_call(sm, tw)
        served, served_sm = serve_closed(sm, Server)
        small = {"requests": len(sm), "batched": 3 * len(sm) / w, "per_call": rate,
                 "batched_over_per_call": 3 * len(sm) / w / rate,
                 "batched_hip_graph": 3 * len(sm) / wg,
                 "graph_equals_eager": bool(np.array_equal(sb.results(), r_eager)),
                 "served": served, "served_over_per_call": served / rate,
                 "served_max_rel_diff_vs_batched": max_rel(served_sm, r_eager)}
        sb.close()

    configs = None if args.no_configs else configs_run(torch, tw, B, stream, with_cpu=not args.no_cpu)
    next_rows = None if args.no_next else next_rows_run(not args.no_cpu)
    cpu = None if args.no_cpu else cpu_baseline(reqs)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "hafnians/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "serving stream: %d requests (%d steps of %d), Poisson arrivals at %.1f QPS (%.2f x "
                   "the batched capacity %.1f/s of the same requests, measured first), even n uniform in [8,40], "
                   "50%% loop, complex128 CN(0,1/n) symmetrised, seed 141; host preparation and submission "
                   "inside the timed region" % (len(reqs), K, R, qps, args.load, capacity),
                   "requests": len(reqs), "step_requests": R, "qps": qps, "parallelism": "single GPU"},
        "p50_ms": stats.get("p50_ms"),
        "p99_ms": stats.get("p99_ms"),
        "serving": stats,
        "serving_half_load_first_%d" % len(half): stats_half,
        "serving_0.9_load_first_%d" % len(half): stats_09,
        "warmup_step_served_closed_loop": served_cap,
        "warmup_step_batched": R / wstep,
        "batched": batched,
        "per_call": pc,
        "small_stream": small,
        "configs": configs,
        "next_rows": next_rows,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
