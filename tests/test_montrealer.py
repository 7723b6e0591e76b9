"""Montrealer / loop montrealer (thewalrus_amd._montrealer) against the oracle's restatement of
_montrealer.py, which is pinned here by the closed-walk enumeration (the combinatorial
definition, no power traces), the all-ones closed forms of test_montrealer.py:24-39 and the
identities of test_montrealer.py:126-188.  GPU tests: relative error <= 1e-10, or <= 2x the
reference algorithm's own FP64 error where that is larger."""
import numpy as np
import pytest
from scipy.special import factorial2

from oracle import oracle as O
from tests.conftest import random_symmetric, relerr
from thewalrus_amd import _montrealer as M
from thewalrus_amd import lmtl, mtl


def aad(rng, n):
    """X (Q - I) of a random mixed Gaussian state: V = S diag(nu) S^T with S a product of a
    random two-mode-mixing orthogonal symplectic and single-mode squeezers."""
    from thewalrus_amd import quantum as q

    r = rng.uniform(0.1, 0.6, size=n)
    Ssq = np.diag(np.concatenate([np.exp(-r), np.exp(r)]))
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    U, _ = np.linalg.qr(X)
    O_ = np.block([[U.real, -U.imag], [U.imag, U.real]])
    nu = rng.uniform(1.0, 1.5, size=n)
    V = O_ @ Ssq @ np.diag(np.concatenate([nu, nu])) @ Ssq @ O_.T
    return O.xmat(n) @ (q.Qmat(V) - np.identity(2 * n))


# ---- oracle pins (CPU)
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_oracle_vs_walk_enumeration(n):
    rng = np.random.default_rng(50 + n)
    A = random_symmetric(rng, 2 * n)
    z = rng.normal(size=2 * n) + 1j * rng.normal(size=2 * n)
    B = rng.normal(size=(2 * n, 2 * n))  # not symmetric: the identity holds for any matrix
    for prec in ("double", "dd"):
        assert relerr(O.mtl(A, precision=prec), O.mtl_walks(A)) < 1e-13
        assert relerr(O.mtl(A, z, precision=prec), O.mtl_walks(A, z)) < 1e-13
        assert relerr(O.mtl(B, precision=prec), O.mtl_walks(B)) < 1e-13


@pytest.mark.parametrize("n", range(1, 8))
def test_oracle_all_ones(n):
    """(2n-2)!! and (n+1)(2n-2)!! (test_montrealer.py:24-39)."""
    J = np.ones((2 * n, 2 * n))
    assert abs(O.mtl(J) - factorial2(2 * n - 2)) < 1e-12 * factorial2(2 * n - 2)
    assert abs(O.mtl(J, np.diag(J)) - (n + 1) * factorial2(2 * n - 2)) < 1e-12 * (n + 1) * factorial2(2 * n - 2)


@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_oracle_identities(n):
    """Permutation invariance, diagonal scaling (Eq. 41), lmtl(zeta = 0) = mtl.  In double-double:
    the montrealer's alternating subset sum cancels, and the FP64 value's last digits depend on
    the order in which the OpenMP threads add their partial sums."""
    rng = np.random.default_rng(70 + n)
    A = aad(rng, n)
    perm = rng.permutation(n)
    perm = np.concatenate((perm, perm + n))
    ref = O.mtl(A, precision="dd")
    assert relerr(O.mtl(A[np.ix_(perm, perm)], precision="dd"), ref) < 1e-13
    g = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    G = np.diag(np.concatenate((g, g.conj())))
    assert relerr(O.mtl(G @ A @ G, precision="dd"), np.prod(np.abs(g) ** 2) * ref) < 1e-13
    assert relerr(O.mtl(A, np.zeros(2 * n), precision="dd"), ref) < 1e-13


# ---- host logic (CPU)
def test_input_checks():
    with pytest.raises(TypeError):
        mtl([[1, 2], [3, 4]])
    with pytest.raises(ValueError, match="square"):
        mtl(np.ones((2, 4)))
    with pytest.raises(ValueError, match="even"):
        mtl(np.ones((3, 3)))
    with pytest.raises(TypeError):
        lmtl(np.ones((2, 2)), [1, 1])
    with pytest.raises(ValueError, match="gamma"):
        lmtl(np.ones((2, 2)), np.ones(4))


def test_dec2bin_and_power_loop():
    assert list(M.dec2bin(5, 4)) == [1, 3]
    assert list(M.dec2bin(0, 3)) == []
    assert list(M.dec2bin(7, 3)) == [0, 1, 2]
    rng = np.random.default_rng(3)
    S = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    z = rng.normal(size=4) + 1j * rng.normal(size=4)
    assert np.isclose(M.power_loop(S, z, 3), np.conj(z) @ S @ S @ z)


def test_size_limit_rejected():
    with pytest.raises(RuntimeError, match="exceeds 24"):
        mtl(np.ones((50, 50)))


# ---- device (GPU)
def _bar(dev, A, z=None):
    truth = O.mtl(A, z, precision="dd")
    ref64 = O.mtl(A, z, precision="double")
    err = relerr(dev, truth)
    assert err <= max(1e-10, 2 * relerr(ref64, truth)), (err, relerr(ref64, truth))


@pytest.mark.gpu
@pytest.mark.parametrize("n", range(1, 8))
def test_gpu_all_ones(n):
    J = np.ones([2 * n, 2 * n])
    print("mtl all ones n=%d" % n, flush=True)
    assert np.isclose(mtl(J), factorial2(2 * n - 2), rtol=1e-12)
    assert np.isclose(lmtl(J, zeta=np.diag(J)), (n + 1) * factorial2(2 * n - 2), rtol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 11, 14])
def test_gpu_vs_oracle(n):
    rng = np.random.default_rng(300 + n)
    A = random_symmetric(rng, 2 * n, scale=0.5)
    z = rng.normal(size=2 * n) + 1j * rng.normal(size=2 * n)
    print("mtl random n=%d" % n, flush=True)
    _bar(mtl(A), A)
    _bar(lmtl(A, z), A, z)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [4, 9])
def test_gpu_physical_and_real(n):
    """Aad = X (Q - I) of a mixed state with zeta = conj(diag) (test_montrealer.py:99-105), and a
    real symmetric matrix with a real zeta (the real-arithmetic kernel)."""
    rng = np.random.default_rng(400 + n)
    A = aad(rng, n)
    z = np.diag(A).conj()
    _bar(mtl(A), A)
    _bar(lmtl(A, z), A, z)
    R = random_symmetric(rng, 2 * n, complex_=False, scale=0.5)
    zr = rng.normal(size=2 * n)
    _bar(mtl(R), R)
    _bar(lmtl(R, zr), R, zr)


@pytest.mark.gpu
def test_gpu_identities():
    """Permutation invariance and Eq. 41 at n = 10: the device meets the bar on the permuted and
    rescaled inputs, whose double-double values satisfy the identities."""
    rng = np.random.default_rng(12)
    n = 10
    A = aad(rng, n)
    perm = rng.permutation(n)
    perm = np.concatenate((perm, perm + n))
    P = A[np.ix_(perm, perm)]
    g = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    G = np.diag(np.concatenate((g, g.conj())))
    S = G @ A @ G
    t = O.mtl(A, precision="dd")
    assert relerr(O.mtl(P, precision="dd"), t) < 1e-12
    assert relerr(O.mtl(S, precision="dd"), np.prod(np.abs(g) ** 2) * t) < 1e-12
    _bar(mtl(A), A)
    _bar(mtl(P), P)
    _bar(mtl(S), S)
