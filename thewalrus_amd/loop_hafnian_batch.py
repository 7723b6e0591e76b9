"""Batched loop hafnians for Gaussian boson sampling, on the GPU.

Drop-in for ``thewalrus.loop_hafnian_batch`` (``loop_hafnian_batch.py``): the same names,
arguments and results.  ``loop_hafnian_batch(A, D, fixed_reps, N_cutoff)`` returns, in one
sieve, the loop hafnians of ``A`` (diagonal ``D``) with repetitions ``fixed_reps`` on the
first ``n - 1`` modes and ``0..N_cutoff`` on the last, the chain-rule step of the GBS sampler.
The edge bookkeeping runs on the host exactly as in the reference; every sieve runs in the
``wh_sieve_lhb`` kernel through ``libwalrus_hip.so`` (no CPU fallback).
"""
import numpy as np

from . import _native
from ._hafnian import matched_reps

__all__ = [
    "_calc_loop_hafnian_batch_even",
    "_calc_loop_hafnian_batch_odd",
    "add_batch_edges_even",
    "add_batch_edges_odd",
    "loop_hafnian_batch",
]


def _calc_loop_hafnian_batch_even(A, D, fixed_edge_reps, batch_max, odd_cutoff, glynn=True):
    """Loop hafnians for 0..2 batch_max + odd_cutoff photons in the batched mode when the fixed
    modes pair up (``loop_hafnian_batch.py:52-123``).  ``A`` / ``D`` are ordered with the
    batched edge first; returns ``2 batch_max + odd_cutoff + 1`` complex values."""
    return _native.loop_hafnian_batch(A, D, fixed_edge_reps, batch_max, odd_cutoff, 0, glynn)


def _calc_loop_hafnian_batch_odd(A, D, fixed_edge_reps, batch_max, even_cutoff, glynn=True):
    """As :func:`_calc_loop_hafnian_batch_even` when one fixed mode is left unpaired
    (``loop_hafnian_batch.py:128-208``); edge 1 joins that mode to the batched mode.  Returns
    ``2 batch_max + even_cutoff + 2`` complex values."""
    return _native.loop_hafnian_batch(A, D, fixed_edge_reps, batch_max, even_cutoff, 1, glynn)


def add_batch_edges_even(fixed_edges):
    """Vertex order with the batched mode (index ``max + 1``) paired with itself as edge 0
    (``loop_hafnian_batch.py:211-231``)."""
    fixed_edges = np.asarray(fixed_edges, dtype=int)
    if len(fixed_edges) == 0:
        return np.array([0, 0], dtype=int)
    half = len(fixed_edges) // 2
    new = max(fixed_edges) + 1
    return np.concatenate(([new], fixed_edges[:half], [new], fixed_edges[half:])).astype(int)


def add_batch_edges_odd(fixed_edges, oddmode):
    """Vertex order with edge 0 = (batched, batched) and edge 1 = (oddmode, batched)
    (``loop_hafnian_batch.py:234-257``)."""
    fixed_edges = np.asarray(fixed_edges, dtype=int)
    if len(fixed_edges) == 0:
        return np.array([1, oddmode, 1, 1], dtype=int)
    half = len(fixed_edges) // 2
    new = max(max(fixed_edges), oddmode) + 1
    return np.concatenate(([new, oddmode], fixed_edges[:half], [new, new], fixed_edges[half:])).astype(int)


def loop_hafnian_batch(A, D, fixed_reps, N_cutoff, glynn=True):
    """Loop hafnians of ``A`` (diagonal ``D``) with ``fixed_reps`` repetitions on modes
    ``0..n-2`` and ``0..N_cutoff`` on mode ``n-1`` (``loop_hafnian_batch.py:260-304``).

    Returns:
        array: ``N_cutoff + 1`` complex loop hafnians, one per photon number of the last mode.
    """
    n = A.shape[0]
    assert A.shape[1] == n
    assert D.shape == (n,)
    assert len(fixed_reps) == n - 1

    nz = np.nonzero(list(fixed_reps) + [1])[0]
    Anz = A[np.ix_(nz, nz)]
    Dnz = D[nz]
    fixed_reps_nz = np.asarray(fixed_reps)[nz[:-1]]
    fixed_edges, fixed_m_reps, oddmode = matched_reps(fixed_reps_nz)

    if oddmode is None:
        batch_max = N_cutoff // 2
        odd_cutoff = N_cutoff % 2
        edges = add_batch_edges_even(fixed_edges)
        Ax = Anz[np.ix_(edges, edges)].astype(np.complex128)
        Dx = Dnz[edges].astype(np.complex128)
        return _calc_loop_hafnian_batch_even(Ax, Dx, fixed_m_reps, batch_max, odd_cutoff, glynn=glynn)
    edges = add_batch_edges_odd(fixed_edges, oddmode)
    Ax = Anz[np.ix_(edges, edges)].astype(np.complex128)
    Dx = Dnz[edges].astype(np.complex128)
    batch_max = (N_cutoff - 1) // 2
    even_cutoff = 1 - (N_cutoff % 2)
    return _calc_loop_hafnian_batch_odd(Ax, Dx, fixed_m_reps, batch_max, even_cutoff, glynn=glynn)
