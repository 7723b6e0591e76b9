"""Batched loop hafnians for threshold-detector (torontonian) sampling, on the GPU.

Drop-in for ``thewalrus.loop_hafnian_batch_gamma`` (``loop_hafnian_batch_gamma.py``): the same
names, arguments and results.  ``D`` holds one diagonal per row (the sampler's fanout rows);
row ``k`` of the result is :func:`loop_hafnian_batch` with diagonal ``D[k]``, and all rows come
from one sieve that builds each term's matrix and power traces once (``wh_sieve_lhb`` with
``nD`` diagonals).
"""
import numpy as np

from . import _native
from ._hafnian import matched_reps
from .loop_hafnian_batch import add_batch_edges_even, add_batch_edges_odd

__all__ = [
    "_calc_loop_hafnian_batch_gamma_even",
    "_calc_loop_hafnian_batch_gamma_odd",
    "loop_hafnian_batch_gamma",
]


def _calc_loop_hafnian_batch_gamma_even(A, D, fixed_edge_reps, batch_max, odd_cutoff, glynn=True):
    """``loop_hafnian_batch_gamma.py:53-129``: an (nD, 2 batch_max + odd_cutoff + 1) array."""
    return _native.loop_hafnian_batch(A, np.atleast_2d(D), fixed_edge_reps, batch_max, odd_cutoff, 0, glynn)


def _calc_loop_hafnian_batch_gamma_odd(A, D, fixed_edge_reps, batch_max, even_cutoff, glynn=True):
    """``loop_hafnian_batch_gamma.py:132-219``: an (nD, 2 batch_max + even_cutoff + 2) array."""
    return _native.loop_hafnian_batch(A, np.atleast_2d(D), fixed_edge_reps, batch_max, even_cutoff, 1, glynn)


def loop_hafnian_batch_gamma(A, D, fixed_reps, N_cutoff, glynn=True):
    """Loop hafnians of ``A`` for every diagonal ``D[k]`` (rows), with ``fixed_reps`` on modes
    ``0..n-2`` and ``0..N_cutoff`` photons on mode ``n-1`` (``loop_hafnian_batch_gamma.py:222-269``).

    Returns:
        array: ``(D.shape[0], N_cutoff + 1)`` complex loop hafnians.
    """
    n = A.shape[0]
    assert A.shape[1] == n
    assert D.shape[1] == n
    assert len(fixed_reps) == n - 1

    nz = np.nonzero(list(fixed_reps) + [1])[0]
    Anz = A[np.ix_(nz, nz)]
    Dnz = D[:, nz]
    fixed_reps_nz = np.asarray(fixed_reps)[nz[:-1]]
    fixed_edges, fixed_m_reps, oddmode = matched_reps(fixed_reps_nz)

    if oddmode is None:
        batch_max = N_cutoff // 2
        odd_cutoff = N_cutoff % 2
        edges = add_batch_edges_even(fixed_edges)
        Ax = Anz[np.ix_(edges, edges)].astype(np.complex128)
        Dx = Dnz[:, edges].astype(np.complex128)
        return _calc_loop_hafnian_batch_gamma_even(Ax, Dx, fixed_m_reps, batch_max, odd_cutoff, glynn=glynn)
    edges = add_batch_edges_odd(fixed_edges, oddmode)
    Ax = Anz[np.ix_(edges, edges)].astype(np.complex128)
    Dx = Dnz[:, edges].astype(np.complex128)
    batch_max = (N_cutoff - 1) // 2
    even_cutoff = 1 - (N_cutoff % 2)
    return _calc_loop_hafnian_batch_gamma_odd(Ax, Dx, fixed_m_reps, batch_max, even_cutoff, glynn=glynn)
