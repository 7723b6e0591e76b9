"""Montrealer and loop montrealer on the GPU.

Drop-in for ``thewalrus._montrealer`` (``mtl`` :121, ``lmtl`` :104, ``montrealer`` :37,
``lmontrealer`` :79, ``power_loop`` :60, ``dec2bin`` :17): the same names, arguments, checks
and results.  The inclusion-exclusion over the 2^n mode subsets runs in the all-ones sieve
kernel: a subset S scales the columns of X Sigma X by 1 (kept) or 0 (dropped), so each term's
tr(Sigma_S^n) and conj(zeta_S) Sigma_S^(n-1) zeta_S come from the same register-tiled powers
as the hafnian's terms.  Up to 24 modes.
"""
import numpy as np

from . import _native

__all__ = ["dec2bin", "montrealer", "power_loop", "lmontrealer", "lmtl", "mtl"]


def _xmat(n):
    I = np.identity(n)
    O = np.zeros_like(I)
    return np.block([[O, I], [I, O]])


def tor_input_checks(A, loops=None):
    """Square, even-sized NumPy matrix (and matching vector) (``_torontonian.py:23-44``)."""
    if not isinstance(A, np.ndarray):
        raise TypeError("Input matrix must be a NumPy array.")
    matshape = A.shape
    if matshape[0] != matshape[1]:
        raise ValueError("Input matrix must be square.")
    if matshape[0] % 2 != 0:
        raise ValueError("matrix dimension must be even")
    if loops is not None:
        if not isinstance(loops, np.ndarray):
            raise TypeError("Input matrix must be a NumPy array.")
        if matshape[0] != len(loops):
            raise ValueError("gamma must be a vector matching the dimension of A")


def dec2bin(num, n):
    """Subset of ``n`` objects labelled by ``num``: bit ``n-1-i`` selects element ``i``
    (``_montrealer.py:17-33``)."""
    digits = [(int(num) >> (n - 1 - i)) & 1 for i in range(n)]
    return np.nonzero(digits)[0]


def power_loop(Sigma, zeta, n):
    """``conj(zeta) @ Sigma^(n-1) @ zeta`` (``_montrealer.py:60-76``); a host helper."""
    vec = zeta
    for _ in range(n - 1):
        vec = Sigma @ vec
    return np.conj(zeta) @ vec


def montrealer(Sigma):
    """Montrealer of the zero-displacement Gaussian state with Glauber covariance ``Sigma``
    (``_montrealer.py:37-57``)."""
    return _native.montrealer(Sigma)


def lmontrealer(Sigma, zeta):
    """Loop montrealer with displacement vector ``zeta`` (``_montrealer.py:79-101``)."""
    return _native.montrealer(Sigma, zeta)


def lmtl(A, zeta):
    """Loop montrealer of ``A`` and ``zeta``: ``lmontrealer(X A, zeta)`` (``_montrealer.py:104-118``)."""
    tor_input_checks(A, zeta)
    n = len(A) // 2
    Sigma = _xmat(n) @ A
    return lmontrealer(Sigma, zeta)


def mtl(A):
    """Montrealer of ``A``: ``montrealer(X A)`` (``_montrealer.py:121-135``)."""
    tor_input_checks(A)
    n = len(A) // 2
    Sigma = _xmat(n) @ A
    return montrealer(Sigma)
