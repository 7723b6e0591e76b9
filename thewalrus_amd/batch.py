"""Batched hafnians: many prepared requests in one launch sequence over device-resident data.

A request is prepared on the host exactly as the drop-in frontends do it (validation, the
closed-form shortcuts, ``matched_reps`` pairing and reordering: ``_hafnian.py:470-631,
718-936``).  Requests that need the sieve are packed into one complex128 pool that lives in
HBM (a torch tensor on the HIP device); ``wh_plan_create`` builds the job table once and
``wh_plan_run`` enqueues the sieve kernels and the fold on a stream with no host work, so a
plan can be replayed (and captured in a HIP graph) as often as needed.
"""
import ctypes

import numpy as np

from . import _hafnian as H
from . import _native

__all__ = ["Prepared", "prepare", "prepare_repeated", "DeviceBatch", "hafnian_batch"]


class Prepared:
    """One request after host preparation: either a host ``value`` (the frontend resolved it
    without the sieve) or a :class:`~thewalrus_amd._hafnian.SieveJob` for the device."""

    __slots__ = ("value", "job", "n")

    def __init__(self, result, n):
        self.job = result if isinstance(result, H.SieveJob) else None
        self.value = None if self.job is not None else result
        self.n = n

    @property
    def on_device(self):
        return self.job is not None


def prepare(A, loop=False, rtol=1e-05, atol=1e-08, method="glynn"):
    """Host preparation of ``hafnian(A, loop, rtol, atol, method=method)``."""
    return Prepared(H.prepare_hafnian(A, loop=loop, rtol=rtol, atol=atol, method=method), A.shape[0])


def prepare_repeated(A, rpt, mu=None, loop=False, rtol=1e-05, atol=1e-08, glynn=True):
    """Host preparation of ``hafnian_repeated(A, rpt, mu, loop, rtol, atol, glynn)``."""
    return Prepared(H.prepare_hafnian_repeated(A, rpt, mu=mu, loop=loop, rtol=rtol, atol=atol, glynn=glynn),
                    A.shape[0])


def _is_real(*arrays):
    """No array has a nonzero imaginary part (None entries are skipped)."""
    for a in arrays:
        if a is None or not np.iscomplexobj(a) or a.size == 0:
            continue
        if a.flat[a.size // 2].imag != 0 or a.imag.any():  # one entry first: most inputs stop there
            return False
    return True


def job_elements(job):
    """complex128 elements a sieve job occupies in a device pool."""
    twoE = job.Ax.shape[0]
    return twoE * twoE + (twoE if job.loop else 0) + (twoE if job.oddloop is not None else 0)


class PackedJob:
    """A sieve job in its device form: one complex128 ``payload`` ([Ax | Dx | oddV], the parts
    the job has) and its ``wh_request`` with offsets relative to the payload's start."""

    __slots__ = ("payload", "req")

    def __init__(self, payload, req):
        self.payload, self.req = payload, req

    @property
    def elems(self):
        return self.payload.shape[0]


def pack_one(job, options=0):
    """The device form of one sieve job (see :class:`PackedJob`)."""
    twoE = job.Ax.shape[0]
    E = twoE // 2
    if E > _native.MAX_EDGES:
        raise ValueError("%d edges exceed the device limit %d" % (E, _native.MAX_EDGES))
    r = _native.WhRequest()
    r.E = E
    r.glynn = int(bool(job.glynn))
    r.options = options
    r.edge_reps[:E] = np.asarray(job.edge_reps, dtype=np.int64).tolist()
    r.a_off = 0
    r.d_off = r.v_off = -1
    parts = [job.Ax.ravel()]
    if job.loop:
        r.d_off = twoE * twoE
        parts.append(job.Dx)
        if job.oddloop is not None:
            r.v_off = twoE * twoE + twoE
            parts.append(job.oddV)
            r.oddloop[0] = complex(job.oddloop).real
            r.oddloop[1] = complex(job.oddloop).imag
    r.real = int(job.oddloop is None and _is_real(job.Ax, job.Dx))
    payload = np.concatenate(parts).astype(np.complex128, copy=False) if len(parts) > 1 else \
        np.ascontiguousarray(parts[0], dtype=np.complex128)
    return PackedJob(payload, r)


def layout_packed(packed, pool):
    """Lays packed jobs out one after another in ``pool`` (a writable complex128 array).

    Returns (requests, used): a ctypes ``wh_request`` array with offsets into ``pool`` and the
    used part of the pool."""
    total = sum(p.elems for p in packed)
    if pool.shape[0] < total:
        raise ValueError("pool holds %d elements, %d needed" % (pool.shape[0], total))
    reqs = (_native.WhRequest * max(len(packed), 1))()
    off = 0
    for k, p in enumerate(packed):
        reqs[k] = p.req
        r = reqs[k]
        r.a_off = off
        if r.d_off >= 0:
            r.d_off += off
        if r.v_off >= 0:
            r.v_off += off
        n = p.elems
        pool[off:off + n] = p.payload
        off += n
    return reqs, pool[:max(total, 1)]


def pack_jobs(jobs, pool=None, options=0):
    """Lays sieve jobs out in one complex128 pool.

    Returns (requests, pool): a ctypes ``wh_request`` array with offsets into ``pool``.  When
    ``pool`` (a writable complex128 array, e.g. a view of pinned memory) is given, the data
    is written into it; otherwise a new array is allocated.
    """
    packed = [pack_one(j, options) for j in jobs]
    if pool is None:
        pool = np.empty(max(sum(p.elems for p in packed), 1), dtype=np.complex128)
    return layout_packed(packed, pool)


class DeviceBatch:
    """A set of prepared requests resident on one HIP device, ready to run as one plan.

    Args:
        preps (list[Prepared]): prepared requests
        device (int): HIP device index
        chunk_terms (int): terms per workgroup (0 = automatic)
        force_generic (bool): run every request on the general kernel
        full_products (bool): on the all-ones kernel, form every power as the reference does
            (h - 1 products per term) instead of the half-products formulation
        valu_products (bool): on the all-ones kernel with complex inputs, run the products on
            the vector FMA units instead of the FP64 matrix cores
        full_powers (bool): on the matrix cores, form the full powers B^a instead of the
            symmetric powers B^a X diag(d) (upper triangle only)
        charpoly (bool): on the general kernel, take the traces above the term matrix's side
            from the characteristic polynomial as the reference does, instead of forming those
            powers directly
        tile (int): run the all-ones and repeated-row kernels at this tile side (n_pad = 8 tile)
            instead of the smallest that holds each matrix (0 = automatic)
        waves (int): waves per workgroup of those kernels (0 = automatic)
    """

    def __init__(self, preps, device=0, chunk_terms=0, force_generic=False, full_products=False,
                 valu_products=False, full_powers=False, charpoly=False, tile=0, waves=0):
        import torch

        self.torch = torch
        self.device = torch.device("cuda", device)
        self.preps = list(preps)
        self.n = len(self.preps)
        self.host_values = {}
        self.dev_index = []  # position in the device output for each request (or -1)
        jobs = []
        for i, p in enumerate(self.preps):
            if not p.on_device:
                self.host_values[i] = p.value
                self.dev_index.append(-1)
                continue
            self.dev_index.append(len(jobs))
            jobs.append(p.job)
        self.nreq = len(jobs)
        self.plan = None
        self.graph = None
        if self.nreq == 0:
            return
        options = ((1 if force_generic else 0) | (2 if full_products else 0) | (8 if valu_products else 0)
                   | (16 if full_powers else 0) | (32 if charpoly else 0) | (int(tile) << 8) | (int(waves) << 12))
        reqs, pool = pack_jobs(jobs, options=options)
        with torch.cuda.device(self.device):
            self.pool = torch.from_numpy(pool.view(np.float64).copy()).to(self.device)
            self.out = torch.zeros(2 * self.nreq, dtype=torch.float64, device=self.device)
            L = _native.load()
            self._reqs = reqs
            plan = ctypes.c_void_p()
            _native.check(L.wh_plan_create(reqs, self.nreq, int(chunk_terms), ctypes.byref(plan)))
            self.plan = plan
        terms = ctypes.c_int64()
        blocks = ctypes.c_int64()
        flops = ctypes.c_double()
        _native.check(L.wh_plan_info(self.plan, ctypes.byref(terms), ctypes.byref(blocks), ctypes.byref(flops)))
        self.terms, self.blocks, self.flops = terms.value, blocks.value, flops.value

    def run(self, stream=None):
        """Enqueue the plan on ``stream`` (torch stream; default: the current stream), as one
        graph launch after ``capture``."""
        if self.plan is None:
            return
        torch = self.torch
        with torch.cuda.device(self.device):
            s = stream if stream is not None else torch.cuda.current_stream(self.device)
            if self.graph is not None:
                _native.check(_native.load().wh_graph_launch(self.graph, ctypes.c_void_p(s.cuda_stream)))
                return
            _native.check(_native.load().wh_plan_run(self.plan, self.pool.data_ptr(), self.out.data_ptr(),
                                                     ctypes.c_void_p(s.cuda_stream)))

    def capture(self, stream):
        """Capture the plan's launches on ``stream`` (a torch stream other than the default
        stream) into a HIP graph; later ``run`` calls replay it with one launch.  Nothing runs
        during capture."""
        if self.plan is None or self.graph is not None:
            return
        if stream is None or stream.cuda_stream == 0:
            raise ValueError("graph capture needs a stream other than the default stream")
        with self.torch.cuda.device(self.device):
            g = ctypes.c_void_p()
            _native.check(_native.load().wh_plan_graph_create(self.plan, self.pool.data_ptr(), self.out.data_ptr(),
                                                              ctypes.c_void_p(stream.cuda_stream), ctypes.byref(g)))
            self.graph = g

    def results(self):
        """Synchronise and return all results (numpy complex128, request order)."""
        out = np.zeros(self.n, dtype=np.complex128)
        if self.plan is not None:
            self.torch.cuda.synchronize(self.device)
            dev = self.out.cpu().numpy().view(np.complex128)
        for i in range(self.n):
            j = self.dev_index[i]
            out[i] = complex(self.host_values[i]) if j < 0 else dev[j]
        return out

    def close(self):
        if self.graph is not None:
            _native.load().wh_graph_destroy(self.graph)
            self.graph = None
        if self.plan is not None:
            _native.load().wh_plan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # pylint: disable=broad-except
            pass


def hafnian_batch(matrices, loop=False, rtol=1e-05, atol=1e-08, method="glynn", device=0):
    """Hafnians of many matrices (mixed sizes allowed) in one plan; returns complex128[n]."""
    if isinstance(loop, (bool, np.bool_)):
        loop = [bool(loop)] * len(matrices)
    preps = [prepare(A, loop=l, rtol=rtol, atol=atol, method=method) for A, l in zip(matrices, loop)]
    b = DeviceBatch(preps, device=device)
    b.run()
    res = b.results()
    b.close()
    return res
