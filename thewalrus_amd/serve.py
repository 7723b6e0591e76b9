"""Continuous-batching server for hafnian, loop-hafnian and repeated-row requests on one GPU.

Requests are prepared on the submitting thread (validation, shortcuts, mode pairing: the same
host code as the drop-in frontends) and queued.  A dispatcher thread forms batches from
whatever is queued whenever a slot is free and launches each batch on that slot's own stream,
so batches that arrive while others run overlap on the device.  Admission is by HBM
footprint: each slot owns fixed device arenas (matrix pool, job table, partial sums, results)
allocated once at start-up, a batch takes queued requests in arrival order while their
footprint fits the slot, and the total admitted footprint never exceeds ``hbm_budget``.
Nothing is allocated, freed or synchronised on the device while serving, and each request's
work is laid out from the request alone, so its result is bit-identical alone and in any mix.
"""
import collections
import ctypes
import threading
import time
from concurrent.futures import Future

import numpy as np

from . import _hafnian as H
from . import _native
from .batch import layout_packed, pack_one

__all__ = ["Server", "ServerStats"]

_JOB_BYTES = 256  # upper bound on sizeof(wh_job_dev), rounded


def _partial_slots(lib, packed, chunk_terms):
    """Partial-sum slots (term chunks) of one packed request, as the library lays it out; the
    layout depends on the request alone, so this is also its share of any batch."""
    ps = ctypes.c_int64()
    _native.check(lib.wh_request_partials(ctypes.byref(packed.req), chunk_terms, ctypes.byref(ps)))
    return int(ps.value)


class _Item:
    __slots__ = ("packed", "future", "t_submit", "t_launch", "t_done", "elems", "slots", "value", "t_deadline")

    def __init__(self, packed, future, slots, t_submit, t_deadline=None):
        self.packed, self.future = packed, future
        self.elems, self.slots = packed.elems, slots
        self.t_submit = t_submit
        self.t_launch = self.t_done = None
        self.value = None
        self.t_deadline = t_deadline


class ServerStats:
    """Per-request timings of completed requests (seconds, perf_counter clock)."""

    def __init__(self):
        self.submit, self.done, self.launch = [], [], []
        self.batches = 0
        self.batch_sizes = []
        self.cancelled = 0  # dropped from the queue after Future.cancel()
        self.expired = 0    # failed with TimeoutError: not launched within their deadline

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
            "mean_batch": float(np.mean(self.batch_sizes)) if self.batch_sizes else 0.0,
        }


class _Slot:
    def __init__(self, torch, device, pool_elems, partial_slots, max_requests):
        self.stream = torch.cuda.Stream(device=device)  # default priority
        self.event = torch.cuda.Event()
        f64 = torch.float64
        self.pool_elems = pool_elems
        self.partial_slots = partial_slots
        self.max_requests = max_requests
        self.h_pool = torch.empty(2 * pool_elems, dtype=f64, pin_memory=True)
        self.d_pool = torch.empty(2 * pool_elems, dtype=f64, device=device)
        self.h_jobs = torch.empty(max_requests * _JOB_BYTES, dtype=torch.uint8, pin_memory=True)
        self.d_jobs = torch.empty(max_requests * _JOB_BYTES, dtype=torch.uint8, device=device)
        self.d_partials = torch.empty(4 * partial_slots, dtype=f64, device=device)
        self.d_out = torch.empty(2 * max_requests, dtype=f64, device=device)
        self.h_out = torch.empty(2 * max_requests, dtype=f64, pin_memory=True)
        self.items = None
        self.batch = None
        self.footprint = 0
        self.pool_view = self.h_pool.numpy().view(np.complex128)
        self.out_view = self.h_out.numpy().view(np.complex128)
        # raw addresses for the library calls of the serving loop
        self.ptrs = (self.h_jobs.data_ptr(), self.d_jobs.data_ptr(), self.h_jobs.numel(), self.h_pool.data_ptr(),
                     self.d_pool.data_ptr(), self.d_partials.data_ptr(), partial_slots, self.d_out.data_ptr(),
                     self.h_out.data_ptr(), max_requests, self.stream.cuda_stream)


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
        import torch

        self.torch = torch
        self.device = torch.device("cuda", device)
        self.chunk_terms = int(chunk_terms)
        pool_elems = int(slot_pool_mb * (1 << 20) // 16)
        self.hbm_budget = int((hbm_budget_mb if hbm_budget_mb is not None else slots * slot_pool_mb) * (1 << 20))
        with torch.cuda.device(self.device):
            self.slots = [_Slot(torch, self.device, pool_elems, slot_partials, slot_requests) for _ in range(slots)]
        self.free = collections.deque(self.slots)
        self.busy = []
        self.queue = collections.deque()
        self.admitted_bytes = 0
        self.pending = 0  # submitted and not yet completed
        self.queued_deadlines = 0  # queued requests with a deadline (the queue is scanned while > 0)
        self.stats = ServerStats()
        self._lock = threading.Lock()
        self._wake = threading.Condition(self._lock)
        self._stop = False
        self._lib = _native.load()
        self._host = _native.load_host()  # host-only layout calls: the interpreter lock stays held
        self._thread = threading.Thread(target=self._run, name="hafnian-dispatcher", daemon=True)
        self._thread.start()

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
        if rpt is None:
            r = H.prepare_hafnian(A, loop=loop, rtol=rtol, atol=atol, method=method)
        else:
            r = H.prepare_hafnian_repeated(A, rpt, mu=mu, loop=loop, rtol=rtol, atol=atol, glynn=glynn)
        fut = Future()
        if not isinstance(r, H.SieveJob):
            fut.set_result(r)
            return fut
        packed = pack_one(r)
        item = _Item(packed, fut, _partial_slots(self._host, packed, self.chunk_terms), t_submit,
                     None if deadline_s is None else t_submit + float(deadline_s))
        slot0 = self.slots[0]
        if item.elems > slot0.pool_elems or item.slots > slot0.partial_slots:
            raise ValueError("request exceeds the per-slot arena (%d elements, %d slots)" % (item.elems, item.slots))
        if self._footprint(item) > self.hbm_budget:
            raise ValueError("request footprint %d B exceeds the HBM budget %d B" % (self._footprint(item), self.hbm_budget))
        with self._wake:
            self.queue.append(item)
            self.pending += 1
            if item.t_deadline is not None:
                self.queued_deadlines += 1
            self._wake.notify()
        return fut

    def _footprint(self, item):
        return 16 * item.elems + 32 * item.slots + 2 * _JOB_BYTES

    # ---------------------------------------------------------------- dispatcher
    def _launch(self, slot, items):
        reqs, used = layout_packed([it.packed for it in items], slot.pool_view)
        b = ctypes.c_void_p()
        _native.check(self._host.wh_batch_create(reqs, len(items), self.chunk_terms, ctypes.byref(b)))
        # job table, pool copy, kernels and the results' copy back: one library call
        hj, dj, jcap, hp, dp, dpart, pcap, dout, hout, ocap, strm = slot.ptrs
        rc = self._lib.wh_batch_submit(b, hj, dj, jcap, hp, dp, used.shape[0], dpart, pcap, dout, hout, ocap, strm)
        if rc != 0:
            self._host.wh_batch_destroy(b)
            _native.check(rc)
        slot.event.record(slot.stream)
        now = time.perf_counter()
        for it in items:
            it.t_launch = now
        slot.items = items
        slot.batch = b
        self.stats.batches += 1
        self.stats.batch_sizes.append(len(items))

    def _complete(self, slot):
        out = slot.out_view[: len(slot.items)].copy()
        now = time.perf_counter()
        self._host.wh_batch_destroy(slot.batch)
        for k, it in enumerate(slot.items):
            it.t_done = now
            self.stats.submit.append(it.t_submit)
            self.stats.launch.append(it.t_launch)
            self.stats.done.append(now)
            it.future.set_result(np.complex128(out[k]))
        with self._lock:
            self.admitted_bytes -= slot.footprint
            self.pending -= len(slot.items)
        slot.items = slot.batch = None
        slot.footprint = 0

    def _expire(self):
        """Drops queued requests whose queue deadline has passed, and cancelled ones on the way
        (the caller holds the lock).  Without deadlines in the queue there is nothing to scan:
        a cancelled request is dropped when it reaches the head (``_take``)."""
        if self.queued_deadlines == 0:
            return
        now = time.perf_counter()
        keep = collections.deque()
        for it in self.queue:
            if it.future.cancelled():
                self.pending -= 1
                self.stats.cancelled += 1
            elif it.t_deadline is not None and now > it.t_deadline:
                self.pending -= 1
                self.stats.expired += 1
                it.future.set_exception(TimeoutError("request not launched within its deadline"))
            else:
                keep.append(it)
                continue
            if it.t_deadline is not None:
                self.queued_deadlines -= 1
        self.queue = keep

    def _take(self, slot):
        """Requests for one batch: arrival order, within the slot arena and the HBM budget."""
        items, elems, parts, fp = [], 0, 0, 0
        while self.queue and len(items) < slot.max_requests:
            it = self.queue[0]
            f = self._footprint(it)
            if (elems + it.elems > slot.pool_elems or parts + it.slots > slot.partial_slots
                    or self.admitted_bytes + fp + f > self.hbm_budget):
                break
            self.queue.popleft()
            if it.t_deadline is not None:
                self.queued_deadlines -= 1
            # RUNNING from here on (no longer cancellable); False: cancelled while queued
            if not it.future.set_running_or_notify_cancel():
                self.pending -= 1
                self.stats.cancelled += 1
                continue
            items.append(it)
            elems += it.elems
            parts += it.slots
            fp += f
        return items, fp

    def _run(self):
        torch = self.torch
        torch.cuda.set_device(self.device)
        while True:
            progressed = False
            for slot in list(self.busy):
                if slot.event.query():
                    self._complete(slot)
                    self.busy.remove(slot)
                    self.free.append(slot)
                    progressed = True
            with self._wake:
                if self._stop and self.pending == 0:
                    return
                launch = []
                if self.queue:
                    self._expire()
                while self.free and self.queue:
                    slot = self.free[0]
                    items, fp = self._take(slot)
                    if not items:
                        break
                    self.free.popleft()
                    slot.footprint = fp
                    self.admitted_bytes += fp
                    launch.append((slot, items))
                if not launch and not progressed:
                    self._wake.wait(timeout=0.0002 if (self.busy or self.queue) else 0.05)
            for slot, items in launch:
                try:
                    self._launch(slot, items)
                    self.busy.append(slot)
                except Exception as exc:  # pylint: disable=broad-except
                    for it in items:
                        it.future.set_exception(exc)
                    with self._lock:
                        self.admitted_bytes -= slot.footprint
                        self.pending -= len(items)
                    slot.footprint = 0
                    self.free.append(slot)

    # ---------------------------------------------------------------- control
    def drain(self, timeout=None):
        """Wait until every queued and in-flight request has completed."""
        t0 = time.perf_counter()
        while True:
            with self._lock:
                if self.pending == 0:
                    return True
            if timeout is not None and time.perf_counter() - t0 > timeout:
                return False
            time.sleep(0.0005)

    def close(self):
        with self._wake:
            self._stop = True
            self._wake.notify()
        self._thread.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
