"""GPU tests of the continuous-batching server: results are the same alone and in the mix."""
import threading
import time

import numpy as np
import pytest

from tests.conftest import random_symmetric

import thewalrus_amd as tw
from thewalrus_amd import workloads as W
from thewalrus_amd.serve import Server

pytestmark = pytest.mark.gpu


def _requests(seed=7, count=60):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        kind = k % 4
        if kind == 0:
            n = 2 * int(rng.integers(3, 13))
            out.append(dict(A=random_symmetric(rng, n, scale=1 / np.sqrt(n)), loop=False))
        elif kind == 1:
            n = int(rng.integers(5, 25))
            out.append(dict(A=random_symmetric(rng, n, scale=1 / np.sqrt(n)), loop=True))
        elif kind == 2:
            A, rpt, mu = W.config4(count=1, seed=int(rng.integers(1 << 30)))[0]
            out.append(dict(A=A, rpt=rpt, mu=mu, loop=True))
        else:
            n = 2 * int(rng.integers(2, 4))  # n <= 6: includes host shortcut sizes
            out.append(dict(A=random_symmetric(rng, n), loop=bool(rng.random() < 0.5)))
    return out


def _alone(r):
    if "rpt" in r:
        return tw.hafnian_repeated(r["A"], r["rpt"], mu=r["mu"], loop=r["loop"])
    return tw.hafnian(r["A"], loop=r["loop"])


def test_served_equals_alone_bitwise():
    reqs = _requests()
    with Server(slots=3, slot_pool_mb=2) as srv:
        futs = [None] * len(reqs)

        def submitter(idx):
            rng = np.random.default_rng(idx)
            for i in idx:
                time.sleep(float(rng.uniform(0, 0.002)))
                futs[i] = srv.submit(**reqs[i])

        parts = [list(range(k, len(reqs), 3)) for k in range(3)]
        threads = [threading.Thread(target=submitter, args=(p,)) for p in parts]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert srv.drain(timeout=120)
        for r, f in zip(reqs, futs):
            got = f.result(timeout=60)
            want = _alone(r)
            assert complex(got) == complex(want), (r["A"].shape[0], r.get("rpt"), got, want)
        s = srv.stats.summary()
        assert s["completed"] > 0 and s["p99_ms"] >= s["p50_ms"] > 0


def test_admission_by_footprint_queues_and_completes():
    """A budget that admits about one n = 20 request at a time still completes everything."""
    rng = np.random.default_rng(3)
    mats = [random_symmetric(rng, 20, scale=0.2) for _ in range(6)]
    with Server(slots=4, slot_pool_mb=1, hbm_budget_mb=0.02) as srv:
        futs = [srv.submit(A) for A in mats]
        assert srv.drain(timeout=120)
        for A, f in zip(mats, futs):
            assert complex(f.result()) == complex(tw.hafnian(A))
        assert srv.stats.batches >= 3


def test_oversized_request_rejected():
    with Server(slots=1, slot_pool_mb=0.01) as srv:
        with pytest.raises(ValueError):
            srv.submit(np.ones((30, 30)))


def test_queue_deadline_and_cancel():
    """With the only slot busy on an n = 40 request, a queued request past its deadline fails
    with TimeoutError, a cancelled one is dropped, and the others complete with their values
    (deadlines are host-side: nothing is interrupted on the device)."""
    rng = np.random.default_rng(91)
    big = random_symmetric(rng, 40, scale=1 / np.sqrt(40))
    small = [random_symmetric(rng, 10, scale=1 / np.sqrt(10)) for _ in range(3)]
    with Server(slots=1, slot_pool_mb=2) as srv:
        f_big = srv.submit(big, loop=True)
        t0 = time.perf_counter()
        while srv.stats.batches < 1 and time.perf_counter() - t0 < 30:
            time.sleep(0.0002)
        assert srv.stats.batches == 1  # the big request holds the only slot
        f_late = srv.submit(small[0], deadline_s=0.0)
        f_cancel = srv.submit(small[1])
        assert f_cancel.cancel()
        f_ok = srv.submit(small[2], deadline_s=120.0)
        assert srv.drain(timeout=120)
        with pytest.raises(TimeoutError):
            f_late.result()
        assert f_cancel.cancelled()
        assert f_ok.result() == tw.hafnian(small[2])  # same request, same kernel and chunking
        assert f_big.result() == tw.hafnian(big, loop=True)
        assert srv.stats.expired == 1 and srv.stats.cancelled == 1


def test_close_runs_queued_requests():
    """close() with requests still queued behind a busy slot: every future gets its value, the
    counters stay readable, and a later submit is refused."""
    rng = np.random.default_rng(17)
    big = random_symmetric(rng, 36, scale=1 / np.sqrt(36))
    small = [random_symmetric(rng, 2 * int(rng.integers(4, 10)), scale=0.3) for _ in range(40)]
    srv = Server(slots=1, slot_pool_mb=2)
    try:
        f_big = srv.submit(big)
        futs = [srv.submit(A, loop=bool(k % 2)) for k, A in enumerate(small)]
    finally:
        srv.close()
    assert f_big.done() and all(f.done() for f in futs)
    assert complex(f_big.result()) == complex(tw.hafnian(big))
    for k, (A, f) in enumerate(zip(small, futs)):
        assert complex(f.result()) == complex(tw.hafnian(A, loop=bool(k % 2)))
    assert srv.stats.launched == 41 and srv.stats.batches >= 2
    with pytest.raises(RuntimeError):
        srv.submit(small[0])


def test_injected_launch_failures_fail_the_batch_and_the_server_keeps_serving():
    """wh_debug_fail_launches makes the next batch launches fail before any kernel runs: a
    resident batch and a per-call hafnian raise RuntimeError and then run again with the same
    values; in the server the requests of the failed batch get RuntimeError and the later ones
    their values, from the same server."""
    from thewalrus_amd import _native
    from thewalrus_amd import batch as B

    L = _native.load()
    rng = np.random.default_rng(23)
    mats = [random_symmetric(rng, 2 * int(rng.integers(4, 12)), scale=0.3) for _ in range(12)]
    want = [tw.hafnian(A) for A in mats]
    try:
        db = B.DeviceBatch([B.prepare(A) for A in mats])
        assert L.wh_debug_fail_launches(1) == 0
        with pytest.raises(RuntimeError, match="launch failure"):
            db.run()
        assert L.wh_debug_fail_launches(-1) == 0
        db.run()
        assert np.array_equal(db.results(), np.array(want))
        db.close()
        L.wh_debug_fail_launches(1)
        with pytest.raises(RuntimeError):
            tw.hafnian(mats[0])
        assert tw.hafnian(mats[0]) == want[0]

        with Server(slots=1, slot_pool_mb=2) as srv:
            L.wh_debug_fail_launches(1)
            first = srv.submit(mats[0])
            with pytest.raises(RuntimeError, match="launch failure"):
                first.result(timeout=60)
            futs = [srv.submit(A) for A in mats]
            assert [f.result(timeout=60) for f in futs] == want
    finally:
        L.wh_debug_fail_launches(0)
