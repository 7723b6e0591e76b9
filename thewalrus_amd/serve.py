"""Continuous-batching server for hafnian, loop-hafnian and repeated-row requests on one GPU.

Requests are prepared on the submitting thread (validation, shortcuts, mode pairing: the same
host code as the drop-in frontends) and packed into their device form.  The library's serving
engine (``wh_server_*``) then owns them: a native dispatcher thread forms batches from whatever
is queued whenever a slot is free and launches each batch on that slot's own stream, so batches
that arrive while others run overlap on the device.  Admission is by HBM footprint: each slot
owns fixed device arenas (matrix pool, job table, partial sums, results) allocated once at
start-up, a batch takes queued requests in arrival order while their footprint fits the slot,
and the total admitted footprint never exceeds ``hbm_budget``.  Nothing is allocated, freed or
synchronised on the device while serving, and each request's work is laid out from the request
alone, so its result is bit-identical alone and in any mix.  A Python thread collects the
engine's completions (waiting inside the library, without the interpreter lock) and resolves
the futures.
"""
import ctypes
import itertools
import threading
import time
import weakref
from concurrent.futures import Future

import numpy as np

from . import _hafnian as H
from . import _native
from .batch import pack_one

__all__ = ["Server", "ServerStats"]

_WAIT_MAX = 1024  # completions collected per library call
# After a first completion, wait this long for more while other requests are in flight: the
# completion thread then takes the interpreter lock once per group, not once per request.
_LINGER_S = 200e-6


class ServerStats:
    """Per-request timings of completed requests (seconds, perf_counter clock) and counters."""

    def __init__(self, counters=None):
        self.submit, self.done, self.launch = [], [], []
        self.cancelled = 0  # dropped from the queue after Future.cancel()
        self.expired = 0    # failed with TimeoutError: not launched within their deadline
        self._counters = counters  # the engine's live (batches, launched), or None once closed
        self._cached = (0, 0)

    def _read(self):
        c = self._counters() if self._counters is not None else None
        if c is not None:
            self._cached = c
        return self._cached

    @property
    def batches(self):
        """Batches launched so far."""
        return self._read()[0]

    @property
    def launched(self):
        """Requests launched on the device so far."""
        return self._read()[1]

    def latencies_ms(self):
        return (np.asarray(self.done) - np.asarray(self.submit)) * 1e3

    def summary(self):
        lat = self.latencies_ms()
        if lat.size == 0:
            return {"completed": 0}
        span = max(self.done) - min(self.submit)
        return {
            "completed": int(lat.size),
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()),
            "throughput": float(lat.size / span) if span > 0 else None,
            "batches": self.batches,
            "mean_batch": float(self.launched / self.batches) if self.batches else 0.0,
        }


class _ServedFuture(Future):
    """A request's future: ``cancel()`` removes the request from the engine's queue, and fails
    (returns False) once the request has been launched."""

    def __init__(self, server, tag):
        super().__init__()
        self._server, self._tag = server, tag

    def cancel(self):
        if self.done():
            return super().cancel()
        if not self._server._cancel(self._tag):
            return False
        return super().cancel()


_ONE_PASS_DTYPES = (np.dtype(np.float64), np.dtype(np.complex128))


def _one_pass(A, method, rtol, atol):
    """Whether the library's one-pass preparation may be tried (it declines, returning 0, what
    it cannot decide exactly as the host path)."""
    return (type(A) is np.ndarray and A.ndim == 2 and 6 <= A.shape[0] <= 64 and A.shape[0] == A.shape[1]
            and A.shape[0] % 2 == 0 and A.dtype in _ONE_PASS_DTYPES and A.flags.c_contiguous
            and method in ("glynn", "inclexcl") and isinstance(rtol, (float, int)) and isinstance(atol, (float, int)))


def _stop_engine(lib, srv, collector):
    """Runs what is queued, stops the engine and the completion thread, frees the engine (at
    ``close()``, or at interpreter exit for a server never closed)."""
    lib.wh_server_shutdown(srv)
    collector.join()
    lib.wh_server_destroy(srv)


class Server:
    """Continuous-batching hafnian server on one HIP device.

    Args:
        device (int): HIP device index
        slots (int): batches in flight at once (each on its own stream)
        slot_pool_mb (float): matrix-pool arena per slot (MiB of complex128)
        slot_partials (int): partial-sum slots per batch
        slot_requests (int): requests per batch
        hbm_budget_mb (float): cap on the admitted footprint of all in-flight batches
        chunk_terms (int): terms per workgroup (0 = automatic, request-determined)
    """

    def __init__(self, device=0, slots=4, slot_pool_mb=16.0, slot_partials=1 << 17, slot_requests=2048,
                 hbm_budget_mb=None, chunk_terms=0):
        self._lib = _native.load()
        self._host = _native.load_host()  # submission and cancel: the interpreter lock stays held
        self.device = int(device)
        self.chunk_terms = int(chunk_terms)
        pool_elems = int(slot_pool_mb * (1 << 20) // 16)
        self.hbm_budget = int((hbm_budget_mb if hbm_budget_mb is not None else slots * slot_pool_mb) * (1 << 20))
        srv = ctypes.c_void_p()
        _native.check(self._lib.wh_server_create(self.device, int(slots), pool_elems, int(slot_partials),
                                                 int(slot_requests), self.chunk_terms, self.hbm_budget,
                                                 ctypes.byref(srv)))
        self._srv = srv
        self.stats = ServerStats(self._engine_counters)
        self._lock = threading.Lock()
        self._futures = {}  # tag -> (future, t_enqueue, t_submit)
        self.pending = 0    # submitted to the engine and not yet resolved
        self._tags = itertools.count(1)
        self._closed = False
        self._collector = threading.Thread(target=self._collect, name="hafnian-completions", daemon=True)
        self._collector.start()
        self._finalizer = weakref.finalize(self, _stop_engine, self._lib, srv, self._collector)

    # ---------------------------------------------------------------- submission
    def submit(self, A, loop=False, rpt=None, mu=None, glynn=True, method="glynn", rtol=1e-05, atol=1e-08,
               deadline_s=None):
        """Queue one request; returns a Future with the value (numpy complex128, or the
        frontend's exact value for shortcut cases).  Validation errors raise immediately.
        Latency is measured from entry to this call (host preparation included).

        ``deadline_s``: a request still queued (not launched) that long after submission fails
        with ``TimeoutError`` instead; a launched request always completes.  A request whose
        future is cancelled (``Future.cancel()``) before launch is dropped from the queue."""
        t_submit = time.perf_counter()
        if rpt is None and _one_pass(A, method, rtol, atol):
            # the common request (n even, 6..64, exactly symmetric): validated, paired, reordered
            # and queued by the library in one call; 0 = it needs the host path below
            tag = next(self._tags)
            fut = _ServedFuture(self, tag)
            with self._lock:
                if self._closed:
                    raise RuntimeError("server is closed")
                t_enq = time.perf_counter()
                rc = self._host.wh_server_submit_matrix(
                    self._srv, A.ctypes.data, A.shape[0], int(A.dtype == np.complex128), int(bool(loop)),
                    int(method == "glynn"), float(rtol), float(atol), tag,
                    -1.0 if deadline_s is None else float(deadline_s))
                if rc == 1:
                    self._futures[tag] = (fut, t_enq, t_submit)
                    self.pending += 1
            if rc == 1:
                return fut
            if rc < 0:
                _native.check(int(rc))
        if rpt is None:
            r = H.prepare_hafnian(A, loop=loop, rtol=rtol, atol=atol, method=method)
        else:
            r = H.prepare_hafnian_repeated(A, rpt, mu=mu, loop=loop, rtol=rtol, atol=atol, glynn=glynn)
        if not isinstance(r, H.SieveJob):
            fut = Future()
            fut.set_result(r)
            return fut
        packed = pack_one(r)
        tag = next(self._tags)
        fut = _ServedFuture(self, tag)
        deadline = -1.0 if deadline_s is None else float(deadline_s)
        with self._lock:  # registered before the collector can see its completion; not after close()
            if self._closed:
                raise RuntimeError("server is closed")
            t_enq = time.perf_counter()
            rc = self._host.wh_server_submit(self._srv, ctypes.byref(packed.req), packed.payload.ctypes.data,
                                             packed.elems, tag, deadline)
            if rc == 0:
                self._futures[tag] = (fut, t_enq, t_submit)
                self.pending += 1
        _native.check(rc)
        return fut

    def _cancel(self, tag):
        with self._lock:  # the engine is not freed meanwhile (close() clears _srv under the lock)
            if self._srv is None or self._host.wh_server_cancel(self._srv, tag) != 1:
                return False
            self._futures.pop(tag, None)
            self.pending -= 1
            self.stats.cancelled += 1
        return True

    # ---------------------------------------------------------------- completions
    def _collect(self):
        tags = np.zeros(_WAIT_MAX, dtype=np.uint64)
        vals = np.zeros(2 * _WAIT_MAX, dtype=np.float64)
        status = np.zeros(_WAIT_MAX, dtype=np.int32)
        times = np.zeros(2 * _WAIT_MAX, dtype=np.float64)
        ptrs = (tags.ctypes.data, vals.ctypes.data, status.ctypes.data, times.ctypes.data)
        while True:
            n = self._lib.wh_server_wait(self._srv, *ptrs, _WAIT_MAX, 0.05, _LINGER_S)
            if n < 0:
                return
            if n == 0:
                continue
            now = time.perf_counter()
            err = self._engine_error() if (status[:n] < 0).any() else None
            done = []
            with self._lock:
                for k in range(n):
                    rec = self._futures.pop(int(tags[k]), None)
                    if rec is None:
                        continue
                    fut, t_enq, t_sub = rec
                    st = int(status[k])
                    if st == 0:
                        self.stats.submit.append(t_sub)
                        self.stats.launch.append(t_enq + float(times[2 * k]))
                        self.stats.done.append(now)
                    elif st == 1:
                        self.stats.expired += 1
                    done.append((fut, st, k))
                self.pending -= len(done)
            for fut, st, k in done:
                if st == 0:
                    fut.set_result(np.complex128(complex(vals[2 * k], vals[2 * k + 1])))
                elif st == 1:
                    fut.set_exception(TimeoutError("request not launched within its deadline"))
                else:
                    fut.set_exception(RuntimeError("batch failed (%d): %s" % (st, err)))

    def _engine_error(self):
        c = (ctypes.c_int64 * 6)()
        buf = ctypes.create_string_buffer(512)
        self._lib.wh_server_stats(self._srv, c, buf, 512)
        return buf.value.decode(errors="replace")

    def _engine_counters(self):
        srv = self._srv
        if srv is None:
            return None
        c = (ctypes.c_int64 * 6)()
        self._lib.wh_server_stats(srv, c, None, 0)
        return int(c[0]), int(c[1])

    # ---------------------------------------------------------------- control
    def drain(self, timeout=None):
        """Wait until every queued and in-flight request has completed."""
        t0 = time.perf_counter()
        while True:
            with self._lock:
                if self.pending == 0:
                    break
            if timeout is not None and time.perf_counter() - t0 > timeout:
                return False
            time.sleep(0.0005)
        return True

    def close(self):
        """Runs everything queued, stops the engine and frees its arenas."""
        with self._lock:
            if self._closed:
                return
            self._closed = True
        self._lib.wh_server_shutdown(self._srv)
        self._collector.join()
        self.stats._read()  # the final counters stay readable
        with self._lock:
            self._srv = None
        self._finalizer()  # frees the engine (once)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
