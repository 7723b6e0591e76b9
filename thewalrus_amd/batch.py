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
    return all(a is None or not np.iscomplexobj(a) or not np.any(np.imag(a)) for a in arrays)


class DeviceBatch:
    """A set of prepared requests resident on one HIP device, ready to run as one plan.

    Args:
        preps (list[Prepared]): prepared requests
        device (int): HIP device index
        chunk_terms (int): terms per workgroup (0 = automatic)
        force_generic (bool): run every request on the general kernel
        full_products (bool): on the all-ones kernel, form every power as the reference does
            (h - 1 products per term) instead of the half-products formulation
    """

    def __init__(self, preps, device=0, chunk_terms=0, force_generic=False, full_products=False):
        import torch

        self.torch = torch
        self.device = torch.device("cuda", device)
        self.preps = list(preps)
        self.n = len(self.preps)
        self.host_values = {}
        reqs = []
        chunks = []
        off = 0
        self.dev_index = []  # position in the device output for each request (or -1)
        for i, p in enumerate(self.preps):
            if not p.on_device:
                self.host_values[i] = p.value
                self.dev_index.append(-1)
                continue
            r = _native.WhRequest()
            p = p.job
            twoE = p.Ax.shape[0]
            E = twoE // 2
            if E > _native.MAX_EDGES:
                raise ValueError("request %d: %d edges exceed the device limit %d" % (i, E, _native.MAX_EDGES))
            r.E = E
            r.glynn = int(bool(p.glynn))
            r.options = (1 if force_generic else 0) | (2 if full_products else 0)
            for k in range(E):
                r.edge_reps[k] = int(p.edge_reps[k])
            r.a_off = off
            chunks.append(np.ascontiguousarray(p.Ax, dtype=np.complex128).ravel())
            off += twoE * twoE
            r.d_off = r.v_off = -1
            if p.loop:
                r.d_off = off
                chunks.append(np.ascontiguousarray(p.Dx, dtype=np.complex128))
                off += twoE
                if p.oddloop is not None:
                    r.v_off = off
                    chunks.append(np.ascontiguousarray(p.oddV, dtype=np.complex128))
                    off += twoE
                    r.oddloop[0] = complex(p.oddloop).real
                    r.oddloop[1] = complex(p.oddloop).imag
            r.real = int(p.oddloop is None and _is_real(p.Ax, p.Dx))
            self.dev_index.append(len(reqs))
            reqs.append(r)
        self.nreq = len(reqs)
        self.plan = None
        if self.nreq == 0:
            return
        pool = np.concatenate(chunks) if chunks else np.zeros(1, dtype=np.complex128)
        with torch.cuda.device(self.device):
            self.pool = torch.from_numpy(pool.view(np.float64).copy()).to(self.device)
            self.out = torch.zeros(2 * self.nreq, dtype=torch.float64, device=self.device)
            L = _native.load()
            arr = (_native.WhRequest * self.nreq)(*reqs)
            self._reqs = arr
            plan = ctypes.c_void_p()
            _native.check(L.wh_plan_create(arr, self.nreq, int(chunk_terms), ctypes.byref(plan)))
            self.plan = plan
        terms = ctypes.c_int64()
        blocks = ctypes.c_int64()
        flops = ctypes.c_double()
        _native.check(L.wh_plan_info(self.plan, ctypes.byref(terms), ctypes.byref(blocks), ctypes.byref(flops)))
        self.terms, self.blocks, self.flops = terms.value, blocks.value, flops.value

    def run(self, stream=None):
        """Enqueue the plan on ``stream`` (torch stream; default: the current stream)."""
        if self.plan is None:
            return
        torch = self.torch
        with torch.cuda.device(self.device):
            s = stream if stream is not None else torch.cuda.current_stream(self.device)
            _native.check(_native.load().wh_plan_run(self.plan, self.pool.data_ptr(), self.out.data_ptr(),
                                                     ctypes.c_void_p(s.cuda_stream)))

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
