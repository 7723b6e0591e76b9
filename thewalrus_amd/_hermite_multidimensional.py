"""Multidimensional Hermite polynomials and hafnian_batched on the GPU.

Drop-in for ``thewalrus._hermite_multidimensional``'s ``hermite_multidimensional`` (:26) and
``hafnian_batched`` (:173): the same arguments, checks, dtype rules and results.
``hafnian_batched(A, cutoff, mu)`` returns the (loop) hafnians of ``reduction(A, k)`` for every
``k`` below the cutoff as one tensor; the recursion that fills it runs in the ``hermite.hip``
kernels (one contiguous slice of the tensor per step, every entry of a slice in parallel).
"""
from typing import Iterable

import numpy as np

from . import _native
from ._hafnian import input_validation

__all__ = ["hermite_multidimensional", "hafnian_batched"]


def hermite_multidimensional(
    R, cutoff, y=None, C=1, renorm=False, make_tensor=True, modified=False, rtol=1e-05, atol=1e-08
):
    """Multidimensional Hermite polynomials H_k^(R)(y) for every 0 <= k_j < cutoff
    (``_hermite_multidimensional.py:26-111``).  ``modified=False`` evaluates the standard
    polynomials as the modified ones at ``R @ y``; ``renorm`` divides by sqrt(prod k_j!)."""
    input_validation(R, atol=atol, rtol=rtol)
    n, _ = R.shape

    if (modified is False) and (y is not None):
        m = y.shape[0]
        if m == n:
            ym = R @ y
            return hermite_multidimensional(
                R, cutoff, y=ym, C=C, renorm=renorm, make_tensor=make_tensor, modified=True
            )

    if y is None:
        y = np.zeros([n], dtype=complex)

    m = y.shape[0]
    if m != n:
        raise ValueError("The matrix R and vector y have incompatible dimensions")

    num_indices = len(y)
    if isinstance(cutoff, np.ndarray) and (cutoff.ndim == 0 or len(cutoff) == 1):
        cutoff = int(cutoff)
    if isinstance(cutoff, Iterable):
        cutoffs = tuple(cutoff)
    else:
        cutoffs = tuple([cutoff]) * num_indices

    Rt = np.real_if_close(R)
    yt = np.real_if_close(y)
    dtype = np.result_type(Rt, yt, C)
    values = _native.hermite_recursion(Rt, yt, cutoffs, C=complex(C), renorm=renorm)
    if not np.issubdtype(dtype, np.complexfloating):
        values = values.real
    values = values.astype(dtype, copy=False)

    if not make_tensor:
        values = values.flatten()
    return values


def hafnian_batched(A, cutoff, mu=None, rtol=1e-05, atol=1e-08, renorm=False, make_tensor=True):
    """Hafnians of ``reduction(A, k)`` (loop hafnians with diagonal ``reduction(mu, k)`` when
    ``mu`` is given) for every ``k`` below ``cutoff`` (``_hermite_multidimensional.py:173-216``)."""
    input_validation(A, atol=atol, rtol=rtol)
    n, _ = A.shape

    if mu is None:
        mu = np.zeros([n], dtype=complex)
    # (p, q) ordering of the Hermite recursion: R = -A
    return hermite_multidimensional(-A, cutoff, y=mu, renorm=renorm, make_tensor=make_tensor, modified=True)
