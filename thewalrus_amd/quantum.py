"""Gaussian-state probability requests over the GPU hafnians.

The reference's callers of ``hafnian`` / ``hafnian_repeated`` on Gaussian states
(thewalrus/quantum/fock_tensors.py): ``density_matrix_element`` (:191-232),
``pure_state_amplitude`` (:45-98) and ``probabilities`` (:392-430), with the conversions they
use (thewalrus/quantum/conversions.py: ``Xmat`` :55, ``Qmat`` :70, ``Covmat`` :99, ``Amat``
:124, ``complex_to_real_displacements`` :153, ``real_to_complex_displacements`` :172), the
purity check (thewalrus/quantum/gaussian_checks.py:12-77) and the prefactor (fock_tensors.py
:566-581).  Same names, arguments and results.

Each element reduces to one (loop) hafnian -- of the reduced matrix when the repetitions are
few (geometric mean of rpt + 1 below 3) or ``hafnian_repeated`` otherwise, the reference's own
rule (fock_tensors.py:80, :214).  ``probabilities`` prepares every Fock index on the host and
evaluates all their hafnians in one device batch (the reference loops over them one call at a
time, or farms them out to dask processes).
"""
from itertools import product

import numpy as np
from scipy.special import factorial as fac

from . import _hafnian as H

__all__ = [
    "Xmat",
    "Qmat",
    "Covmat",
    "Amat",
    "complex_to_real_displacements",
    "real_to_complex_displacements",
    "sympmat",
    "is_valid_cov",
    "is_pure_cov",
    "density_matrix_element",
    "pure_state_amplitude",
    "probabilities",
]


def _swap_halves(N):
    """The 2N x 2N permutation that exchanges the first and second N coordinates."""
    P = np.zeros((2 * N, 2 * N))
    idx = np.arange(2 * N)
    P[idx, (idx + N) % (2 * N)] = 1.0
    return P


def _ladder_basis(N):
    """W with (a, a^dagger) = W (x, p) / sqrt(hbar) for a = (x + i p) / sqrt(2 hbar): the unitary
    change from quadratures to ladder operators, W = [[I, iI], [I, -iI]] / sqrt(2)."""
    W = np.zeros((2 * N, 2 * N), dtype=np.complex128)
    r = np.arange(N)
    W[r, r] = W[r + N, r] = 1.0
    W[r, r + N] = 1j
    W[r + N, r + N] = -1j
    return W / np.sqrt(2.0)


def Xmat(N):
    """[[0, I_N], [I_N, 0]] (conversions.py:55-67)."""
    return _swap_halves(N)


def sympmat(N, dtype=np.float64):
    """The symplectic form Omega = [[0, I_N], [-I_N, 0]]."""
    Om = np.zeros((2 * N, 2 * N), dtype=dtype)
    r = np.arange(N)
    Om[r, r + N] = 1
    Om[r + N, r] = -1
    return Om


def Qmat(cov, hbar=2):
    """Husimi Q matrix of a Gaussian state from its xp-covariance (conversions.py:70-96).

    In the basis W* (the conjugate of the ladder basis), Q = W* cov W^T / hbar + I/2: its
    top-left block is <a_i^dagger a_j> + delta_ij, its bottom-left block <a_i a_j>, and the
    vacuum (cov = hbar/2 I) gives Q = I."""
    cov = np.asarray(cov)
    Wc = _ladder_basis(len(cov) // 2).conj()
    return Wc @ cov @ Wc.conj().T / hbar + 0.5 * np.identity(len(cov))


def Covmat(Q, hbar=2):
    """xp-covariance from the Husimi Q matrix, the inverse map of :func:`Qmat`
    (conversions.py:99-121): cov = hbar W^T (Q - I/2) W*, real for a physical Q."""
    Q = np.asarray(Q)
    Wc = _ladder_basis(len(Q) // 2).conj()
    return (hbar * (Wc.conj().T @ (Q - 0.5 * np.identity(len(Q))) @ Wc)).real


def Amat(cov, hbar=2, cov_is_qmat=False):
    """A = X (I - Q^-1)^*, whose hafnians give the photon-number statistics (conversions.py:124-150);
    X only exchanges the two row blocks."""
    Q = np.asarray(cov) if cov_is_qmat else Qmat(cov, hbar=hbar)
    N = len(Q) // 2
    M = np.conj(np.identity(2 * N) - np.linalg.inv(Q))
    return np.concatenate([M[N:], M[:N]], axis=0)


def complex_to_real_displacements(mu, hbar=2):
    """(alpha, alpha^*) from the quadrature means mu = (x, p), alpha = (x + i p) / sqrt(2 hbar)
    (conversions.py:153-169)."""
    mu = np.asarray(mu)
    return _ladder_basis(len(mu) // 2) @ mu / np.sqrt(hbar)


def real_to_complex_displacements(beta, hbar=2):
    """Quadrature means from (alpha, alpha^*), the inverse map of
    :func:`complex_to_real_displacements` (conversions.py:172-187); only alpha is read."""
    beta = np.asarray(beta)
    N = len(beta) // 2
    full = np.concatenate([beta[:N], np.conj(beta[:N])])
    return np.sqrt(hbar) * (_ladder_basis(N).conj().T @ full).real


def is_valid_cov(cov, hbar=2, rtol=1e-05, atol=1e-08):
    """Square, symmetric, even-sized, and obeying the uncertainty relation
    cov + i hbar/2 Omega >= 0 up to ``atol`` (gaussian_checks.py:12-56)."""
    rows, cols = cov.shape
    if rows != cols or rows % 2:
        return False
    if not np.allclose(cov, cov.T, rtol=rtol, atol=atol):
        return False
    lam = np.linalg.eigvalsh(cov + 0.5j * hbar * sympmat(rows // 2))
    return bool(np.all((lam > -atol) | (np.abs(lam) < atol)))


def is_pure_cov(cov, hbar=2, rtol=1e-05, atol=1e-08):
    """Valid and of unit purity (gaussian_checks.py:59-77).  The purity of a Gaussian state is
    det(2 cov / hbar)^(-1/2)."""
    if not is_valid_cov(cov, hbar=hbar, rtol=rtol, atol=atol):
        return False
    purity = np.linalg.det(2.0 * np.asarray(cov) / hbar) ** -0.5
    return bool(np.allclose(purity, 1.0, rtol=rtol, atol=atol))


def _prefactor(mu, cov, hbar=2):
    """<0|D(-alpha) rho D(alpha)|0>-type normalisation of a Gaussian element
    (fock_tensors.py:566-581): exp(-beta^T Q^-1 beta^* / 2) / sqrt(det Q) with beta = (alpha,
    alpha^*); the quadratic form by a linear solve."""
    Q = Qmat(cov, hbar=hbar)
    beta = complex_to_real_displacements(mu, hbar=hbar)
    quad = beta @ np.linalg.solve(Q, np.conj(beta))
    return np.exp(-0.5 * quad) / np.sqrt(np.linalg.det(Q))


def _few_repetitions(rpt):
    return np.prod([k + 1 for k in rpt]) ** (1 / len(rpt)) < 3


def _element_job(Amatrix, rpt, gamma):
    """The (loop) hafnian of one element, prepared by the reference's rule: a value, or a
    sieve job (fock_tensors.py:80-92, :212-227)."""
    if gamma is None:
        if _few_repetitions(rpt):
            return H.prepare_hafnian(H.reduction(Amatrix, rpt))
        return H.prepare_hafnian_repeated(Amatrix, rpt)
    if _few_repetitions(rpt):
        A_rpt = H.reduction(Amatrix, rpt)
        np.fill_diagonal(A_rpt, H.reduction(gamma, rpt))
        return H.prepare_hafnian(A_rpt, loop=True)
    return H.prepare_hafnian_repeated(Amatrix, rpt, mu=gamma, loop=True)


def _mixed_setup(mu, cov, tol, hbar):
    beta = complex_to_real_displacements(mu, hbar=hbar)
    A = Amat(cov, hbar=hbar)
    gamma = None if np.linalg.norm(beta) < tol else beta.conj() - A @ beta
    return A, gamma


def density_matrix_element(mu, cov, i, j, include_prefactor=True, tol=1e-10, hbar=2):
    """<i|rho|j> of a Gaussian state (fock_tensors.py:191-232)."""
    rpt = list(i) + list(j)
    A, gamma = _mixed_setup(mu, cov, tol, hbar)
    haf = H._finish(_element_job(A, rpt, gamma))
    if include_prefactor:
        haf *= _prefactor(mu, cov, hbar=hbar)
    return haf / np.sqrt(np.prod(fac(rpt)))


def _pure_setup(mu, cov, tol, hbar):
    beta = complex_to_real_displacements(mu, hbar=hbar)
    Q = Qmat(cov, hbar=hbar)
    A = Amat(cov, hbar=hbar)
    N = cov.shape[0] // 2
    B = A[0:N, 0:N].conj()
    alpha = beta[0:N]
    gamma = None if np.linalg.norm(alpha) < tol else alpha - B @ np.conj(alpha)
    return Q, B, alpha, gamma


def pure_state_amplitude(mu, cov, i, include_prefactor=True, tol=1e-10, hbar=2, check_purity=True):
    """<i|psi> of a pure Gaussian state (fock_tensors.py:45-98)."""
    if check_purity and not is_pure_cov(cov, hbar=hbar, rtol=1e-05, atol=1e-08):
        raise ValueError("The covariance matrix does not correspond to a pure state")
    rpt = list(i)
    Q, B, alpha, gamma = _pure_setup(mu, cov, tol, hbar)
    haf = H._finish(_element_job(B, rpt, gamma))
    if include_prefactor:
        haf *= np.exp(-0.5 * (np.linalg.norm(alpha) ** 2 - alpha.conj() @ B @ alpha.conj()))
    return haf / np.sqrt(np.prod(fac(rpt)) * np.sqrt(np.linalg.det(Q)))


def _evaluate(jobs, device=0):
    """Values of prepared elements: host values as they are, sieve jobs in one device batch."""
    from .batch import DeviceBatch, Prepared

    preps = [Prepared(j, 0) for j in jobs]
    if not any(p.on_device for p in preps):
        return np.array([complex(p.value) for p in preps])
    b = DeviceBatch(preps, device=device)
    b.run()
    vals = b.results()
    b.close()
    return vals


def probabilities(mu, cov, cutoff, parallel=False, hbar=2.0, rtol=1e-05, atol=1e-08, device=0):
    """Fock-space probabilities up to ``cutoff`` per mode (fock_tensors.py:392-430).

    Every index's hafnian is prepared on the host and all of them run in one device batch;
    ``parallel`` is accepted for signature compatibility (the batch is already parallel).
    """
    num_modes = len(mu) // 2
    idx = list(product(range(cutoff), repeat=num_modes))
    if is_pure_cov(cov, hbar=hbar, rtol=rtol, atol=atol):
        Q, B, alpha, gamma = _pure_setup(mu, cov, 1e-10, hbar)
        vals = _evaluate([_element_job(B, list(i), gamma) for i in idx], device)
        pref = np.exp(-0.5 * (np.linalg.norm(alpha) ** 2 - alpha.conj() @ B @ alpha.conj()))
        denom = np.sqrt(np.sqrt(np.linalg.det(Q)))
        amps = np.array([v * pref / (np.sqrt(np.prod(fac(list(i)))) * denom) for v, i in zip(vals, idx)])
        return (np.abs(amps) ** 2).reshape([cutoff] * num_modes)
    A, gamma = _mixed_setup(mu, cov, 1e-10, hbar)
    vals = _evaluate([_element_job(A, list(i) + list(i), gamma) for i in idx], device)
    pref = _prefactor(mu, cov, hbar=hbar)
    probs = np.array([np.maximum(0.0, np.real_if_close(v * pref / np.sqrt(np.prod(fac(list(i) + list(i))))))
                      for v, i in zip(vals, idx)], dtype=np.float64)
    return probs.reshape([cutoff] * num_modes)
