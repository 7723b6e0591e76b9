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


def Xmat(N):
    """[[0, I_N], [I_N, 0]] (conversions.py:55-67)."""
    I = np.identity(N)
    O = np.zeros_like(I)
    return np.block([[O, I], [I, O]])


def sympmat(N, dtype=np.float64):
    """The symplectic form [[0, I_N], [-I_N, 0]]."""
    I = np.identity(N, dtype=dtype)
    O = np.zeros_like(I, dtype=dtype)
    return np.block([[O, I], [-I, O]])


def Qmat(cov, hbar=2):
    """Husimi Q matrix of a Gaussian state from its xp-covariance (conversions.py:70-96)."""
    N = len(cov) // 2
    I = np.identity(N)
    x = cov[:N, :N] * 2 / hbar
    xp = cov[:N, N:] * 2 / hbar
    p = cov[N:, N:] * 2 / hbar
    aidaj = (x + p + 1j * (xp - xp.T) - 2 * I) / 4   # <a_i^dagger a_j>
    aiaj = (x - p + 1j * (xp + xp.T)) / 4            # <a_i a_j>
    return np.block([[aidaj, aiaj.conj()], [aiaj, aidaj.conj()]]) + np.identity(2 * N)


def Covmat(Q, hbar=2):
    """xp-covariance from the Husimi Q matrix; inverse of :func:`Qmat` (conversions.py:99-121)."""
    n = len(Q) // 2
    I = np.identity(n)
    N = Q[0:n, 0:n] - I
    M = Q[n:2 * n, 0:n]
    mm11a = 2 * (N.real + M.real) + np.identity(n)
    mm22a = 2 * (N.real - M.real) + np.identity(n)
    mm12a = 2 * (M.imag + N.imag)
    return (hbar / 2) * np.block([[mm11a, mm12a], [mm12a.T, mm22a]])


def Amat(cov, hbar=2, cov_is_qmat=False):
    """A = X (I - Q^-1)^*, whose hafnians give the photon-number statistics (conversions.py:124-150)."""
    N = len(cov) // 2
    Q = cov if cov_is_qmat else Qmat(cov, hbar=hbar)
    return Xmat(N) @ (np.identity(2 * N) - np.linalg.inv(Q)).conj()


def complex_to_real_displacements(mu, hbar=2):
    """(alpha, alpha^*) from the quadrature means (conversions.py:153-169)."""
    N = len(mu) // 2
    alpha = (mu[:N] + 1j * mu[N:]) / np.sqrt(2 * hbar)
    return np.concatenate([alpha, alpha.conj()])


def real_to_complex_displacements(beta, hbar=2):
    """Quadrature means from (alpha, alpha^*) (conversions.py:172-187)."""
    N = len(beta) // 2
    alpha = beta[0:N]
    return np.sqrt(2 * hbar) * np.concatenate([alpha.real, alpha.imag])


def is_valid_cov(cov, hbar=2, rtol=1e-05, atol=1e-08):
    """Square, symmetric, even, and cov + i hbar/2 Omega >= 0 (gaussian_checks.py:12-56)."""
    (n, m) = cov.shape
    if n != m or n % 2 != 0:
        return False
    if not np.allclose(cov, np.transpose(cov), rtol=rtol, atol=atol):
        return False
    vals = np.linalg.eigvalsh(cov + 0.5j * hbar * sympmat(n // 2))
    vals[np.abs(vals) < atol] = 0.0
    return bool(np.all(vals >= 0))


def is_pure_cov(cov, hbar=2, rtol=1e-05, atol=1e-08):
    """Valid and of unit purity (gaussian_checks.py:59-77)."""
    if is_valid_cov(cov, hbar=hbar, rtol=rtol, atol=atol):
        purity = 1 / np.sqrt(np.linalg.det(2 * cov / hbar))
        if np.allclose(purity, 1.0, rtol=rtol, atol=atol):
            return True
    return False


def _prefactor(mu, cov, hbar=2):
    """exp(-beta Q^-1 beta^* / 2) / sqrt(det Q) (fock_tensors.py:566-581)."""
    Q = Qmat(cov, hbar=hbar)
    beta = complex_to_real_displacements(mu, hbar=hbar)
    return np.exp(-0.5 * beta @ np.linalg.inv(Q) @ beta.conj()) / np.sqrt(np.linalg.det(Q))


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
