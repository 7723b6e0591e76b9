"""Batched loop hafnians (thewalrus_amd.loop_hafnian_batch / loop_hafnian_batch_gamma).

Host logic (edge bookkeeping, argument checks, the empty term range) runs on the CPU.  The
GPU tests compare every output with the oracle's double-double restatement of
loop_hafnian_batch.py, itself pinned by Kan's formula, at the bar of the other sieve tests:
relative error <= 1e-10, or <= 2x the reference algorithm's own FP64 error where that is larger.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import random_symmetric
from thewalrus_amd import _native
from thewalrus_amd.loop_hafnian_batch import add_batch_edges_even, add_batch_edges_odd, loop_hafnian_batch
from thewalrus_amd.loop_hafnian_batch_gamma import loop_hafnian_batch_gamma


def test_batch_edges_match_oracle():
    for fixed in ([], [0, 1], [2, 0, 1, 3], [4, 1, 2, 0, 3, 5]):
        assert list(add_batch_edges_even(np.array(fixed, dtype=int))) == list(O._batch_edges_even(fixed))
        for odd in (0, 3, 7):
            assert list(add_batch_edges_odd(np.array(fixed, dtype=int), odd)) == list(O._batch_edges_odd(fixed, odd))
    assert list(add_batch_edges_even(np.array([0, 1]))) == [2, 0, 2, 1]
    assert list(add_batch_edges_odd(np.array([], dtype=int), 0)) == [1, 0, 1, 1]


def test_shape_checks():
    A = np.eye(3)
    with pytest.raises(AssertionError):
        loop_hafnian_batch(A, np.ones(2), [1, 1], 2)
    with pytest.raises(AssertionError):
        loop_hafnian_batch(A, np.ones(3), [1], 2)
    with pytest.raises(AssertionError):
        loop_hafnian_batch_gamma(A, np.ones((2, 2)), [1, 1], 2)


def test_odd_zero_cutoff_is_empty_range():
    """N_cutoff = 0 with an unpaired fixed mode: batch_max = -1, no term, [0] as in the reference."""
    A = random_symmetric(np.random.default_rng(1), 3)
    H = loop_hafnian_batch(A, np.ones(3, dtype=complex), [1, 0], 0)
    assert H.shape == (1,) and H[0] == 0
    G = loop_hafnian_batch_gamma(A, np.ones((2, 3), dtype=complex), [1, 0], 0)
    assert G.shape == (2, 1) and not G.any()


def test_limits_rejected_before_device():
    """Argument errors come back from the C ABI before any device call."""
    rng = np.random.default_rng(2)
    Ax = random_symmetric(rng, 4)
    Dx = np.ones(4, dtype=complex)
    with pytest.raises(RuntimeError, match="exceed 128"):
        _native.loop_hafnian_batch(Ax, Dx, [1], 64, 0, 0)      # 129 outputs
    with pytest.raises(RuntimeError, match="exceed 128"):
        _native.loop_hafnian_batch(Ax, np.ones((20, 4)), [1], 3, 0, 0)  # 20 x 7 outputs
    with pytest.raises(ValueError, match="matrix side"):
        _native.loop_hafnian_batch(Ax, Dx, [1, 1], 2, 0, 0)    # side 4 != 2 (2 + 1)
    with pytest.raises(ValueError, match="cutoff"):
        _native.loop_hafnian_batch(Ax, Dx, [1], 2, 2, 0)


def _check(dev, A, D, fixed, N_cutoff, glynn=True):
    truth = O.loop_hafnian_batch(A, D, fixed, N_cutoff, glynn=glynn, precision="dd")
    ref64 = O.loop_hafnian_batch(A, D, fixed, N_cutoff, glynn=glynn, precision="double")
    assert dev.shape == truth.shape
    for q in range(len(truth)):
        if truth[q] == 0:
            assert dev[q] == 0, q
            continue
        err = abs(dev[q] - truth[q]) / abs(truth[q])
        bar = max(1e-10, 2 * abs(ref64[q] - truth[q]) / abs(truth[q]))
        assert err <= bar, (q, err, bar)


GPU_CASES = [
    (4, [1, 2, 0], 5),
    (4, [1, 1, 1], 4),
    (3, [0, 0], 3),
    (2, [1], 4),
    (2, [0], 7),
    (6, [3, 1, 0, 2, 1], 7),
    (8, [1, 0, 2, 1, 1, 0, 3], 10),      # sampler shape: 8 modes, cutoff 10
    (10, [2, 1, 1, 0, 3, 1, 2, 1, 1], 12),
    (12, [1] * 11, 9),                   # 11 single photons: odd variant, 2E = 14
]


@pytest.mark.gpu
@pytest.mark.parametrize("n,fixed,N_cutoff", GPU_CASES)
@pytest.mark.parametrize("glynn", [True, False])
def test_gpu_loop_hafnian_batch(n, fixed, N_cutoff, glynn):
    rng = np.random.default_rng(4000 + 13 * n + N_cutoff)
    A = random_symmetric(rng, n, scale=0.5)
    D = rng.normal(size=n) + 1j * rng.normal(size=n)
    print("lhb n=%d fixed=%s N_cutoff=%d glynn=%s" % (n, fixed, N_cutoff, glynn), flush=True)
    _check(loop_hafnian_batch(A, D, fixed, N_cutoff, glynn=glynn), A, D, fixed, N_cutoff, glynn)


@pytest.mark.gpu
def test_gpu_loop_hafnian_batch_real_inputs():
    """Real B and gamma as the hafnian sampler passes them for a real covariance."""
    rng = np.random.default_rng(8)
    n = 7
    A = random_symmetric(rng, n, complex_=False, scale=0.4)
    D = rng.normal(size=n)
    fixed = [2, 0, 1, 1, 0, 2]
    _check(loop_hafnian_batch(A, D, fixed, 8), A, D, fixed, 8)


@pytest.mark.gpu
@pytest.mark.parametrize("n,fixed,N_cutoff,nD", [(5, [1, 0, 2, 1], 1, 10), (4, [1, 1, 1], 1, 10), (6, [2, 1, 0, 1, 1], 6, 4)])
def test_gpu_loop_hafnian_batch_gamma(n, fixed, N_cutoff, nD):
    """Row k equals the single-diagonal batch with D[k], bit for bit, and the oracle to the bar.
    These calls have fewer than 1024 term chunks, so each diagonal gets its own workgroups and
    the same per-row arithmetic as the single-diagonal call."""
    rng = np.random.default_rng(900 + n + nD)
    A = random_symmetric(rng, n, scale=0.5)
    D = rng.normal(size=(nD, n)) + 1j * rng.normal(size=(nD, n))
    print("lhb gamma n=%d fixed=%s N_cutoff=%d nD=%d" % (n, fixed, N_cutoff, nD), flush=True)
    G = loop_hafnian_batch_gamma(A, D, fixed, N_cutoff)
    assert G.shape == (nD, N_cutoff + 1)
    for k in range(nD):
        row = loop_hafnian_batch(A, D[k], fixed, N_cutoff)
        assert np.array_equal(G[k], row), k
        _check(G[k], A, D[k], fixed, N_cutoff)


@pytest.mark.gpu
def test_gpu_loop_hafnian_batch_many_chunks():
    """Enough terms for several workgroups (chunked partials folded in order)."""
    rng = np.random.default_rng(31)
    n = 9
    A = random_symmetric(rng, n, scale=0.35)
    D = rng.normal(size=n) + 1j * rng.normal(size=n)
    fixed = [3, 2, 2, 1, 3, 2, 1, 2]   # 2 x 16 = 32 fixed photons
    print("lhb many chunks", flush=True)
    _check(loop_hafnian_batch(A, D, fixed, 10), A, D, fixed, 10)


@pytest.mark.gpu
def test_gpu_loop_hafnian_batch_gamma_many_chunks():
    """More than 1024 term chunks: each workgroup then serves all nD diagonals of its chunk
    (the row and column chains per diagonal), a different path from the few-chunk split; every
    row is checked against the oracle's single-diagonal batch to the bar."""
    rng = np.random.default_rng(32)
    n = 10
    A = random_symmetric(rng, n, scale=0.35)
    nD = 3
    D = rng.normal(size=(nD, n)) + 1j * rng.normal(size=(nD, n))
    fixed = [3, 3, 2, 2, 3, 2, 2, 2, 1]  # edge multiplicities [3, 2, 2, 2, 1]: 6 x 216 = 1296 terms
    print("lhb gamma many chunks", flush=True)
    G = loop_hafnian_batch_gamma(A, D, fixed, 10)
    assert G.shape == (nD, 11)
    for k in range(nD):
        _check(G[k], A, D[k], fixed, 10)
