"""Drop-in hafnian frontends over the MI355X sieve kernels.

Same names, signatures, defaults, exceptions and return types as ``thewalrus._hafnian``
(``hafnian`` :718, ``hafnian_repeated`` :864, ``loop_hafnian`` :581, ``_haf`` :470,
``matched_reps`` :80, ``input_validation`` :634, ``reduction`` :697).  Validation, the
closed-form shortcuts for n <= 4 and the mode pairing run on the host exactly as in the
reference; every sieve evaluation runs on the GPU through ``libwalrus_hip.so``.
"""
import functools
import warnings

import numpy as np

from . import _native


def input_validation(A, rtol=1e-05, atol=1e-08):
    """Checks that ``A`` is a square, NaN-free, symmetric NumPy array (``_hafnian.py:634-665``).

    Raises:
        TypeError: ``A`` is not a NumPy array.
        ValueError: ``A`` is not square, contains NaNs, or is not symmetric.
    """
    if not isinstance(A, np.ndarray):
        raise TypeError("Input matrix must be a NumPy array.")
    shape = A.shape
    if shape[0] != shape[1]:
        raise ValueError("Input matrix must be square.")
    if np.isnan(A).any():
        raise ValueError("Input matrix must not contain NaNs.")
    # exactly equal entries are close under any tolerance (numpy.isclose: ... or x == y), so
    # the tolerance test only runs for a matrix that is not exactly symmetric
    exact = A.dtype.kind in "biufc" and bool((A == A.T).all())
    if not exact and not np.allclose(A, A.T, rtol=rtol, atol=atol):
        raise ValueError("Input matrix must be symmetric.")
    return True


def _surely_not_diagonal(A, rtol, atol):
    """True when the off-diagonal entry A[0, 1] is far from close to 0 under numpy.isclose's rule
    (|0 - y| <= atol + rtol |y|, y finite): then ``np.allclose(diag(diag(A)), A)`` is False.  The
    margin covers the last-bit differences between numpy's array and scalar complex moduli; a
    False here decides nothing (the full test runs)."""
    if A.shape[0] < 2 or A.dtype.kind not in "fc" or np.ndim(rtol) or np.ndim(atol):
        return False
    ay = abs(complex(A[0, 1]))
    if not ay < float("inf"):
        return False
    return ay * (1.0 - 1e-9) > float(atol) + float(rtol) * ay * (1.0 + 1e-9)


def _is_diagonal(A, rtol, atol):
    """``np.allclose(np.diag(np.diag(A)), A)`` (``_hafnian.py:766``)."""
    if _surely_not_diagonal(A, rtol, atol):
        return False
    return np.allclose(np.diag(np.diag(A)), A, rtol=rtol, atol=atol)


def reduction(A, rpt):
    """Repeats row/column ``i`` of ``A`` ``rpt[i]`` times (``_hafnian.py:697-714``).

    Works on matrices and on vectors (1-D arrays).
    """
    rows = [idx for idx, count in enumerate(rpt) for _ in range(count)]
    if A.ndim == 1:
        return A[rows]
    return A[:, rows][rows]


def matched_reps(reps):
    """Pairs repeated modes into edges of a perfect matching (``_hafnian.py:80-159``).

    Greedy rule: with modes sorted by (repetitions, index) descending, a mode whose count is
    more than twice the next one is paired with itself (``r // 2`` times, leaving 1 behind if
    ``r`` is odd); otherwise the two leading modes are paired ``r_2`` times.

    Returns:
        tuple[array, array, int]: vertex order ``x`` (length ``2E``; index ``i`` is matched
        with ``i + E``), edge multiplicities (length ``E``), and the unpaired mode (or None).
    """
    if sum(reps) == 0:
        return np.array([], dtype=int), np.array([], dtype=int), None

    # (count, mode) pairs, largest first; zero counts dropped
    live = [(int(r), i) for r, i in sorted(zip(reps, range(len(reps))), reverse=True) if r > 0]
    first, second, mult = [], [], []
    while len(live) > 1 or (len(live) == 1 and live[0][0] > 1):
        live.sort(reverse=True)
        (r0, i0) = live[0]
        if len(live) == 1 or r0 > 2 * live[1][0]:
            first.append(i0)
            second.append(i0)
            mult.append(r0 // 2)
            live = live[1:] if r0 % 2 == 0 else [(1, i0)] + live[1:]
        else:
            (r1, i1) = live[1]
            first.append(i0)
            second.append(i1)
            mult.append(r1)
            live = ([(r0 - r1, i0)] if r0 > r1 else []) + live[2:]

    oddmode = live[0][1] if len(live) == 1 else None
    x = np.asarray(first + second, dtype=np.int64)
    return x, np.asarray(mult, dtype=np.int64), oddmode


class SieveJob:
    """Sieve inputs at the reference kernel seam (``_calc_hafnian`` / ``_calc_loop_hafnian``).

    ``Ax``: reordered matrix (complex128, 2E x 2E); ``edge_reps``: multiplicities;
    ``Dx``/``oddloop``/``oddV``: loop weights (``Dx is None`` for the plain hafnian).
    """

    __slots__ = ("Ax", "edge_reps", "Dx", "oddloop", "oddV", "glynn")

    def __init__(self, Ax, edge_reps, Dx=None, oddloop=None, oddV=None, glynn=True):
        self.Ax, self.edge_reps, self.Dx = Ax, edge_reps, Dx
        self.oddloop, self.oddV, self.glynn = oddloop, oddV, glynn

    @property
    def loop(self):
        return self.Dx is not None

    def run(self):
        """Evaluate on the GPU (synchronous, host buffers)."""
        if self.Dx is None:
            return _native.calc_hafnian(self.Ax, self.edge_reps, self.glynn)
        return _native.calc_loop_hafnian(self.Ax, self.Dx, self.edge_reps, self.oddloop, self.oddV,
                                         self.glynn)


@functools.lru_cache(maxsize=None)
def _matched_ones(n):
    """``matched_reps([1] * n)``, computed once per n."""
    x, e, odd = matched_reps([1] * n)
    x.flags.writeable = False
    e.flags.writeable = False
    return x, e, odd


def _pairing(reps, ones):
    """``matched_reps(reps)``; every multiplicity 1 (``ones``) comes from the per-n cache."""
    if not ones:
        return matched_reps(reps)
    x, e, odd = _matched_ones(len(reps))
    return x, e.copy(), odd


def _reorder(A, x):
    """``A[np.ix_(x, x)].astype(np.complex128)`` (``_hafnian.py:505, :627``), as two gathers: a new
    C-contiguous complex128 array."""
    return A.take(x, axis=0).take(x, axis=1).astype(np.complex128, copy=False)


def _prepare_haf(A, reps=None, glynn=True):
    """Host part of ``_haf`` (``_hafnian.py:486-505``): a value, or the sieve job."""
    n = A.shape[0]
    ones = reps is None
    if ones:
        reps = [1] * n
    N = sum(reps)
    if N == 0:
        return 1.0
    if N % 2 == 1:
        return 0.0
    assert n == len(reps)
    x, edge_reps, _ = _pairing(reps, ones)
    return SieveJob(_reorder(A, x), edge_reps, glynn=glynn)


def _prepare_loop(A, D=None, reps=None, glynn=True):
    """Host part of ``loop_hafnian`` (``_hafnian.py:598-628``): a value, or the sieve job."""
    n = A.shape[0]
    ones = reps is None
    if ones:
        reps = [1] * n
    if D is None:
        D = A.diagonal()
    N = sum(reps)
    if N == 0:
        return 1.0
    if N == 1:
        return D[np.where(np.array(reps) == 1)[0][0]]
    assert n == len(reps)
    assert D.shape[0] == n
    x, edge_reps, oddmode = _pairing(reps, ones)
    if oddmode is not None:
        oddloop = D[oddmode].astype(np.complex128)
        oddV = A[oddmode, x].astype(np.complex128)
    else:
        oddloop = None
        oddV = None
    Dx = D[x].astype(np.complex128)
    return SieveJob(_reorder(A, x), edge_reps, Dx, oddloop, oddV, glynn)


def _finish(r):
    return r.run() if isinstance(r, SieveJob) else r


def _haf(A, reps=None, glynn=True):
    """Hafnian with optional repeated rows/columns (``_hafnian.py:470-508``).

    Pairs the modes, reorders ``A`` and evaluates the sieve on the GPU.
    """
    return _finish(_prepare_haf(A, reps, glynn))


def loop_hafnian(A, D=None, reps=None, glynn=True):
    """Loop hafnian with optional repeated rows/columns (``_hafnian.py:581-631``).

    ``D`` replaces the diagonal (defaults to ``A``'s diagonal).
    """
    return _finish(_prepare_loop(A, D, reps, glynn))


def hafnian(A, loop=False, rtol=1e-05, atol=1e-08, approx=False, num_samples=1000, method="glynn"):
    """Returns the (loop) hafnian of ``A`` (``_hafnian.py:718-818``).

    Args:
        A (array): square, symmetric array
        loop (bool): return the loop hafnian
        rtol, atol (float): tolerances of the symmetry check
        approx (bool): Barvinok estimator -- not part of this library (raises)
        num_samples (int): samples of the estimator
        method (str): ``"glynn"`` (finite-difference sieve) or ``"inclexcl"``
            (inclusion-exclusion); ``"recursive"`` is not part of this library

    Returns:
        int or float or complex: the hafnian (numpy complex128 from the sieve)
    """
    return _finish(prepare_hafnian(A, loop, rtol, atol, approx, num_samples, method))


def prepare_hafnian(A, loop=False, rtol=1e-05, atol=1e-08, approx=False, num_samples=1000,
                    method="glynn"):
    """Host part of :func:`hafnian`: validation and shortcuts; a value or a :class:`SieveJob`.

    Args:
        A (array): square, symmetric array
        loop (bool): return the loop hafnian
        rtol, atol (float): tolerances of the symmetry check
        approx (bool): Barvinok estimator -- not part of this library (raises)
        num_samples (int): samples of the estimator
        method (str): ``"glynn"`` (finite-difference sieve) or ``"inclexcl"``
            (inclusion-exclusion); ``"recursive"`` is not part of this library

    Returns:
        int or float or complex: the hafnian (numpy complex128 from the sieve)
    """
    # pylint: disable=too-many-return-statements,too-many-branches
    input_validation(A, rtol=rtol, atol=atol)
    shape = A.shape
    if method == "glynn":
        glynn = True
    if method == "inclexcl":
        glynn = False
    if shape == (0, 0):
        return 1
    if shape[0] % 2 != 0 and not loop:
        return 0.0
    if _is_diagonal(A, rtol, atol):
        if loop:
            return np.prod(np.diag(A))
        return 0
    if shape[0] == 2:
        if loop:
            return A[0, 1] + A[0, 0] * A[1, 1]
        return A[0][1]
    if shape[0] == 3 and loop:
        return A[0, 0] * A[1, 2] + A[1, 1] * A[0, 2] + A[2, 2] * A[0, 1] + A[0, 0] * A[1, 1] * A[2, 2]
    if shape[0] == 4:
        pairs = A[0, 1] * A[2, 3] + A[0, 2] * A[1, 3] + A[0, 3] * A[1, 2]
        if loop:
            return (
                pairs
                + A[0, 0] * A[1, 1] * A[2, 3]
                + A[0, 1] * A[2, 2] * A[3, 3]
                + A[0, 2] * A[1, 1] * A[3, 3]
                + A[0, 0] * A[2, 2] * A[1, 3]
                + A[0, 0] * A[3, 3] * A[1, 2]
                + A[0, 3] * A[1, 1] * A[2, 2]
                + A[0, 0] * A[1, 1] * A[2, 2] * A[3, 3]
            )
        return pairs
    if approx:
        raise NotImplementedError(
            "approx=True (Barvinok estimator) is outside this library's power-trace path"
        )
    if loop:
        if method == "recursive":
            warnings.warn("Recursive algorithm does not support the loop hafnian")
        return _prepare_loop(A, D=None, reps=None, glynn=True)
    if method == "recursive":
        raise NotImplementedError(
            "method='recursive' is outside this library's power-trace path; use 'glynn'"
        )
    return _prepare_haf(A, reps=None, glynn=glynn)


def hafnian_repeated(A, rpt, mu=None, loop=False, rtol=1e-05, atol=1e-08, glynn=True):
    """Hafnian of ``A`` with row/column ``i`` repeated ``rpt[i]`` times (``_hafnian.py:864-936``).

    Equals ``hafnian(reduction(A, rpt))``; with ``loop=True`` the diagonal is ``mu``
    (default: ``A``'s diagonal).
    """
    return _finish(prepare_hafnian_repeated(A, rpt, mu, loop, rtol, atol, glynn))


def prepare_hafnian_repeated(A, rpt, mu=None, loop=False, rtol=1e-05, atol=1e-08, glynn=True):
    """Host part of :func:`hafnian_repeated`: a value or a :class:`SieveJob`."""
    # pylint: disable=too-many-return-statements,too-many-branches
    input_validation(A, atol=atol, rtol=rtol)
    if len(rpt) != len(A):
        raise ValueError("the rpt argument must be 1-dimensional sequence of length len(A).")
    nud = np.array(rpt, dtype=np.int32)
    if not np.all(np.mod(rpt, 1) == 0) or np.any(nud < 0):
        raise ValueError("the rpt argument must contain non-negative integers.")
    if np.all(nud == 0):
        return 1.0
    if np.sum(nud) % 2 != 0 and not loop:
        return 0.0
    if mu is None:
        mu = A.diagonal().copy()
    if np.allclose(A, 0, rtol=rtol, atol=atol):
        if loop:
            return np.prod(mu**rpt)
        return 0
    if len(mu) != len(A):
        raise ValueError("Length of means vector must be the same length as the matrix A.")
    if loop:
        return _prepare_loop(A, D=mu, reps=rpt, glynn=glynn)
    return _prepare_haf(A, reps=rpt, glynn=glynn)
