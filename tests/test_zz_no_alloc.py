"""wh_batch_launch_noalloc (job table and partial sums in the library's static scratch array)
against wh_batch_launch (the same buffers passed in by the caller), through the C ABI.

Both launch the same kernels with the same chunking, so the results must agree bit for bit.
Two calls in a row, the second with a smaller batch, check that the second call does not
depend on what the first left in the scratch array; a batch larger than the scratch array is
refused before anything is enqueued."""
import ctypes

import numpy as np
import pytest

from tests.conftest import random_symmetric
from thewalrus_amd import _hafnian as H
from thewalrus_amd import _native
from thewalrus_amd import batch as B

pytestmark = pytest.mark.gpu


def _requests(seed, sizes):
    rng = np.random.default_rng(seed)
    jobs = []
    for k, n in enumerate(sizes):
        A = random_symmetric(rng, n, complex_=k % 3 != 2, scale=1 / np.sqrt(2 * n))
        jobs.append(H.prepare_hafnian(A, loop=k % 2 == 1))
    reps = [2, 1, 3, 0, 1, 2, 1, 1]  # one repeated-row request: the general kernel
    A = random_symmetric(rng, 8, scale=0.3)
    jobs.append(H.prepare_hafnian_repeated(A, reps, mu=np.diag(A).copy(), loop=True))
    return B.pack_jobs(jobs)


def _launch_both(L, torch, reqs, pool, n):
    b = ctypes.c_void_p()
    assert L.wh_batch_create(reqs, n, 0, ctypes.byref(b)) == 0, L.wh_last_error()
    try:
        jb = ctypes.c_size_t()
        slots = ctypes.c_int64()
        assert L.wh_batch_info(b, ctypes.byref(jb), ctypes.byref(slots), None, None) == 0
        h_jobs = torch.empty(jb.value, dtype=torch.uint8, pin_memory=True)
        assert L.wh_batch_write_jobs(b, h_jobs.data_ptr()) == 0
        dev = torch.device("cuda", 0)
        d_pool = torch.from_numpy(pool.view(np.float64).copy()).to(dev)
        d_jobs = torch.empty(jb.value, dtype=torch.uint8, device=dev)
        d_parts = torch.empty(4 * slots.value, dtype=torch.float64, device=dev)
        out_ref = torch.zeros(2 * n, dtype=torch.float64, device=dev)
        out_static = torch.full((2 * n,), float("nan"), dtype=torch.float64, device=dev)
        s = torch.cuda.current_stream(dev).cuda_stream
        rc = L.wh_batch_launch(b, h_jobs.data_ptr(), d_jobs.data_ptr(), d_pool.data_ptr(), d_parts.data_ptr(),
                               out_ref.data_ptr(), s)
        assert rc == 0, L.wh_last_error()
        rc = L.wh_batch_launch_noalloc(b, h_jobs.data_ptr(), d_pool.data_ptr(), out_static.data_ptr(), s)
        assert rc == 0, L.wh_last_error()
        torch.cuda.synchronize(dev)
        return out_ref.cpu().numpy().view(np.complex128), out_static.cpu().numpy().view(np.complex128)
    finally:
        L.wh_batch_destroy(b)


def test_noalloc_equals_batch_launch_twice_then_smaller():
    import torch

    L = _native.load()
    assert L.wh_init(0) == 0, L.wh_last_error()
    big = [8, 40, 12, 30, 16, 24, 20, 36, 10, 26, 38, 14]
    small = [10, 22, 8]
    for seed, sizes in ((1, big), (2, small)):
        reqs, pool = _requests(seed, sizes)
        n = len(sizes) + 1
        ref, got = _launch_both(L, torch, reqs, pool, n)
        print("no-alloc launch: %d requests, max |diff| %.3g" % (n, np.max(np.abs(ref - got))), flush=True)
        assert np.all(np.isfinite(got))
        assert np.array_equal(ref.view(np.float64), got.view(np.float64)), (sizes, ref, got)


def test_noalloc_refuses_batches_beyond_the_scratch():
    """One n = 40 request with one term per chunk needs 2^19 partial slots, more than the scratch
    array holds: refused with WH_EINVAL before any copy or launch."""
    L = _native.load()
    X = np.random.default_rng(3).normal(size=(40, 40))
    pk = B.pack_one(H.prepare_hafnian((X + X.T) / 2))
    b = ctypes.c_void_p()
    assert L.wh_batch_create(ctypes.byref(pk.req), 1, 1, ctypes.byref(b)) == 0
    try:
        fake = 1 << 12  # never dereferenced: the size check fails first
        rc = L.wh_batch_launch_noalloc(b, fake, fake, fake, None)
        assert rc == -1 and b"static scratch" in L.wh_last_error()
    finally:
        L.wh_batch_destroy(b)
