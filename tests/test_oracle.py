"""Pins of the CPU oracle (CPU only).

Every oracle function is checked against something independent of its own formula: library
routines (numpy matrix powers, numpy.poly), mpmath Taylor coefficients of the generating
functions (docs/algorithms.rst:105-118), the reference's own known-answer values
(test_labudde.py, test_hafnian.py, test_hafnian_repeated.py), closed forms, brute force over
perfect matchings, and Kan's repeated-moment formula.
"""
import math

import mpmath
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import random_symmetric, relerr


def rand_c(rng, m, scale=1.0):
    return (rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))) * scale


# ---------------------------------------------------------------------------- powertrace
@pytest.mark.parametrize("m,n", [(6, 4), (6, 6), (8, 5), (5, 9), (4, 12), (2, 7), (7, 15)])
def test_powertrace_vs_matrix_power(m, n):
    """charpoly.py:302: p_k = tr(H^k), through the product loop and the La Budde extension."""
    rng = np.random.default_rng(m * 100 + n)
    H = rand_c(rng, m, 0.4)
    p = O.powertrace(H, n)
    want = [m] + [np.trace(np.linalg.matrix_power(H, k)) for k in range(1, n)]
    for k in range(n):
        assert abs(p[k] - want[k]) <= 1e-11 * max(1.0, abs(want[k])), (k, p[k], want[k])


@pytest.mark.parametrize("n,val", [(1, 2.09951), (2, 2.40794), (3, 2.95599)])
def test_powertrace_labudde_2x2_known(n, val):
    """test_labudde.py:35-53 known values (n + 1 > m exercises the La Budde extension)."""
    phi = 0.1 * math.pi
    mat = np.array([[math.cosh(phi), math.sinh(phi)], [math.sinh(phi), math.cosh(phi)]])
    assert np.allclose(O.powertrace(mat, n + 1)[-1], val)


@pytest.mark.parametrize("n,val", [(1, 15.0995), (2, 301.18), (4, 81466.1)])
def test_powertrace_labudde_4x4_known(n, val):
    """test_labudde.py:56-74 known values."""
    mat = np.array(
        [
            [1.04975523, 0.31935254, 1, 2],
            [0.31935254, 1.04975523, 3, 4],
            [0.31635254, 2.444, 5, 6],
            [21.31935254, 3.14975523, 7, 8],
        ]
    )
    assert np.allclose(O.powertrace(mat, n + 1)[-1], val)


# ---------------------------------------------------------------------------- charpoly
@pytest.mark.parametrize("phi", [0.1, 0.2, 0.3])
def test_charpoly_2x2_known(phi):
    """test_labudde.py:21-31: coefficients of lambda^2 - 2 cosh(phi) lambda + 1."""
    mat = np.array([[math.cosh(phi), math.sinh(phi)], [math.sinh(phi), math.cosh(phi)]])
    c = O.charpoly(mat)
    assert np.allclose(c[0], -2 * math.cosh(phi))
    assert np.allclose(c[1], 1)


@pytest.mark.parametrize("m", [2, 3, 5, 8, 12])
def test_charpoly_vs_numpy_poly(m):
    rng = np.random.default_rng(m)
    H = rand_c(rng, m, 0.5)
    c = O.charpoly(H)
    want = np.poly(H)[1:]
    assert np.allclose(c, want, rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------------------- f, f_loop, f_loop_odd
mpmath.mp.dps = 40


def _taylor(fun, order):
    return [complex(c) for c in mpmath.taylor(fun, 0, order)]


def _mpmat(X):
    return mpmath.matrix([[mpmath.mpc(complex(x)) for x in row] for row in X])


@pytest.mark.parametrize("m,N", [(4, 4), (6, 6), (4, 8), (6, 10)])
def test_f_vs_generating_function(m, N):
    """f(C) = [eta^h] det(I - eta C)^(-1/2) (docs/algorithms.rst:114-118)."""
    rng = np.random.default_rng(m + N)
    C = rand_c(rng, m, 0.5)
    got = O.fcoeffs(C, N, mode=0)
    Cm = _mpmat(C)
    I = mpmath.eye(m)
    want = _taylor(lambda z: mpmath.det(I - z * Cm) ** (-0.5), N // 2)
    for k in range(N // 2 + 1):
        assert abs(got[k] - want[k]) < 1e-11 * max(1, abs(want[k])), (k, got[k], want[k])


@pytest.mark.parametrize("m,N", [(4, 4), (6, 8)])
def test_f_loop_vs_generating_function(m, N):
    """f_loop = [eta^h] det(I - eta C)^(-1/2) exp(eta/2 * XD (I - eta C)^(-1) D)."""
    rng = np.random.default_rng(10 + m + N)
    C = rand_c(rng, m, 0.5)
    XD = rng.normal(size=m) + 1j * rng.normal(size=m)
    D = rng.normal(size=m) + 1j * rng.normal(size=m)
    got = O.fcoeffs(C, N, mode=1, XD=XD, D=D)
    Cm, I = _mpmat(C), mpmath.eye(m)
    xd = mpmath.matrix([[mpmath.mpc(complex(v)) for v in XD]])
    dv = mpmath.matrix([mpmath.mpc(complex(v)) for v in D])

    def gen(z):
        M = I - z * Cm
        quad = (xd * mpmath.inverse(M) * dv)[0]
        return mpmath.det(M) ** (-0.5) * mpmath.exp(z * quad / 2)

    want = _taylor(gen, N // 2)
    for k in range(N // 2 + 1):
        assert abs(got[k] - want[k]) < 1e-11 * max(1, abs(want[k])), (k, got[k], want[k])


@pytest.mark.parametrize("m,N", [(4, 5), (6, 7)])
def test_f_loop_odd_vs_generating_function(m, N):
    """f_loop_odd = [z^N] det(I - z^2 C)^(-1/2) exp(z^2/2 XD R D + oddloop z + z^3 oddVX R D),
    R = (I - z^2 C)^(-1)  (_hafnian.py:268-275 restated as a series)."""
    rng = np.random.default_rng(20 + m + N)
    C = rand_c(rng, m, 0.5)
    XD = rng.normal(size=m) + 1j * rng.normal(size=m)
    D = rng.normal(size=m) + 1j * rng.normal(size=m)
    OV = rng.normal(size=m) + 1j * rng.normal(size=m)
    ol = complex(rng.normal(), rng.normal())
    got = O.fcoeffs(C, N, mode=2, XD=XD, D=D, oddloop=ol, oddVX=OV)
    Cm, I = _mpmat(C), mpmath.eye(m)
    xd = mpmath.matrix([[mpmath.mpc(complex(v)) for v in XD]])
    ov = mpmath.matrix([[mpmath.mpc(complex(v)) for v in OV]])
    dv = mpmath.matrix([mpmath.mpc(complex(v)) for v in D])

    def gen(z):
        M = I - z * z * Cm
        R = mpmath.inverse(M)
        return mpmath.det(M) ** (-0.5) * mpmath.exp(
            z * z * (xd * R * dv)[0] / 2 + mpmath.mpc(ol) * z + z ** 3 * (ov * R * dv)[0]
        )

    want = _taylor(gen, N)
    for k in range(N + 1):
        assert abs(got[k] - want[k]) < 1e-10 * max(1, abs(want[k])), (k, got[k], want[k])


# ---------------------------------------------------------------------------- whole sieve
@pytest.mark.parametrize("n", [4, 6, 8, 12, 16, 20, 24])
def test_all_ones(n):
    """haf(J_n) = (n-1)!! (test_hafnian.py:211-216)."""
    A = np.ones((n, n))
    for prec in ("double", "dd"):
        v = O.hafnian_sieve(A, precision=prec)
        assert round(v.real) == O.double_factorial(n - 1) and abs(v.real - O.double_factorial(n - 1)) < 0.5


@pytest.mark.parametrize("n", [6, 8])
def test_block_ones(n):
    """haf([[0, J], [J, 0]]) = n! (test_hafnian.py:220-228), both sieves."""
    Z, J = np.zeros((n, n)), np.ones((n, n))
    A = np.block([[Z, J], [J, Z]])
    for glynn in (True, False):
        assert abs(O.hafnian_sieve(A, glynn=glynn) - math.factorial(n)) < 1e-8 * math.factorial(n)


@pytest.mark.parametrize("n", [6, 7, 8, 9, 10])
def test_loop_telephone(n):
    """lhaf(J_n) = T(n) (test_hafnian.py:285-290; T from :30)."""
    assert abs(O.hafnian_sieve(np.ones((n, n)), loop=True) - O.TELEPHONE[n]) < 1e-9 * O.TELEPHONE[n]


@pytest.mark.parametrize("n", [6, 8])
def test_identity(n):
    """haf(I) = 0, lhaf(I) = 1 (test_hafnian.py:203-207, :277-281)."""
    assert abs(O.hafnian_sieve(np.identity(n))) < 1e-12
    assert abs(O.hafnian_sieve(np.identity(n), loop=True) - 1) < 1e-12


@pytest.mark.parametrize("n", [4, 6, 8, 10, 12, 14])
@pytest.mark.parametrize("loop", [False, True])
def test_sieve_vs_brute_force(n, loop):
    rng = np.random.default_rng(300 + n + loop)
    A = random_symmetric(rng, n, scale=1 / np.sqrt(n))
    bf = O.brute_force(A, loop=loop)
    for prec in ("double", "dd"):
        for glynn in (True, False):
            assert relerr(O.hafnian_sieve(A, loop=loop, glynn=glynn, precision=prec), bf) < 1e-11
    assert relerr(O.hafnian_sieve(A, loop=loop, precision="dd", variant=1), bf) < 1e-13


@pytest.mark.parametrize("n", [5, 7, 9, 11])
def test_odd_loop_vs_brute_force(n):
    rng = np.random.default_rng(400 + n)
    A = random_symmetric(rng, n, scale=1 / np.sqrt(n))
    bf = O.brute_force(A, loop=True)
    assert relerr(O.hafnian_sieve(A, loop=True), bf) < 1e-11
    assert relerr(O.hafnian_sieve(A, loop=True, precision="dd", variant=1), bf) < 1e-13


def test_brute_force_closed_forms():
    """The brute force itself: (n-1)!!, telephone numbers, rank one x x^T -> (n-1)!! prod(x)."""
    for n in (2, 4, 6, 8, 10):
        assert abs(O.brute_force(np.ones((n, n))) - O.double_factorial(n - 1)) < 1e-9
    for n in range(1, 11):
        assert abs(O.brute_force(np.ones((n, n)), loop=True) - O.TELEPHONE[n]) < 1e-9
    rng = np.random.default_rng(3)
    x = rng.normal(size=8) + 1j * rng.normal(size=8)
    assert relerr(O.brute_force(np.outer(x, x)), O.double_factorial(7) * np.prod(x)) < 1e-13


def test_brute_force_tiny_by_enumeration():
    """n = 4 and 5 spelled out term by term (test_hafnian.py:182-199 formulas)."""
    rng = np.random.default_rng(4)
    A = random_symmetric(rng, 4)
    pairs = A[0, 1] * A[2, 3] + A[0, 2] * A[1, 3] + A[0, 3] * A[1, 2]
    assert relerr(O.brute_force(A), pairs) < 1e-14
    loops = (
        pairs
        + A[0, 0] * A[1, 1] * A[2, 3]
        + A[0, 1] * A[2, 2] * A[3, 3]
        + A[0, 2] * A[1, 1] * A[3, 3]
        + A[0, 0] * A[2, 2] * A[1, 3]
        + A[0, 0] * A[3, 3] * A[1, 2]
        + A[0, 3] * A[1, 1] * A[2, 2]
        + A[0, 0] * A[1, 1] * A[2, 2] * A[3, 3]
    )
    assert relerr(O.brute_force(A, loop=True), loops) < 1e-14


@pytest.mark.parametrize("rpt", [[1, 2, 3, 0], [2, 2, 2, 2], [5, 1, 0, 2], [3, 3, 1, 2], [6, 1, 1, 1], [4, 0, 4, 0]])
@pytest.mark.parametrize("loop", [False, True])
def test_repeated_vs_reduction_and_kan(rpt, loop):
    """hafnian_repeated(A, rpt) == hafnian(reduction(A, rpt)) (_hafnian.py:873-878) and Kan."""
    if not loop and sum(rpt) % 2:
        pytest.skip("odd photon number without loops is zero")
    rng = np.random.default_rng(sum(rpt) + 10 * loop)
    A = random_symmetric(rng, 4, scale=0.6)
    mu = rng.normal(size=4) + 1j * rng.normal(size=4)
    got = O.hafnian_sieve(A, loop=loop, rpt=rpt, mu=mu, precision="dd")
    Ar = O.reduction(A, rpt).astype(np.complex128)
    np.fill_diagonal(Ar, O.reduction(mu, rpt))
    bf = O.brute_force(Ar, loop=loop)
    assert relerr(got, bf) < 1e-13
    assert relerr(O.kan(A, rpt, mu=mu, loop=loop), bf) < 1e-13
    assert relerr(O.hafnian_sieve(A, loop=loop, rpt=rpt, mu=mu), bf) < 1e-10


@pytest.mark.parametrize("n", [3, 5, 6, 8])
def test_repeated_closed_forms(n):
    """[[1]] with rpt [2n] -> (2n-1)!!; [[0,1],[1,0]] with [n,n] -> n! (test_hafnian_repeated.py:193-222)."""
    v = O.hafnian_sieve(np.array([[1.0]]), rpt=[2 * n])
    assert abs(v - O.double_factorial(2 * n - 1)) < 1e-8 * O.double_factorial(2 * n - 1)
    v = O.hafnian_sieve(np.array([[0.0, 1.0], [1.0, 0.0]]), rpt=[n, n])
    assert abs(v - math.factorial(n)) < 1e-8 * math.factorial(n)


@pytest.mark.parametrize("n", [6, 10, 15, 20])
def test_permanent_repeated_deep(n):
    """permanent_repeated([[1]], [n]) = n!: B = [[0,1],[1,0]], rpt [n, n] (test_permanent.py:185-190)."""
    v = O.hafnian_sieve(np.array([[0.0, 1.0], [1.0, 0.0]]), rpt=[n, n], precision="dd")
    assert abs(v - math.factorial(n)) < 1e-10 * math.factorial(n)


def test_pairing_properties():
    """pair_modes: every mode's count is covered exactly (2 per self-pair, 1 per side)."""
    rng = np.random.default_rng(9)
    for _ in range(200):
        rpt = list(rng.integers(0, 7, size=int(rng.integers(1, 8))))
        x, e, odd = O.pair_modes(rpt)
        E = len(e)
        cover = np.zeros(len(rpt), dtype=int)
        for i in range(E):
            cover[x[i]] += e[i]
            cover[x[i + E]] += e[i]
        if odd is not None:
            cover[odd] += 1
        assert list(cover) == [int(r) for r in rpt]


def test_thread_count_invariance():
    """OpenMP partition changes only the rounding of the final sum."""
    rng = np.random.default_rng(12)
    A = random_symmetric(rng, 16, scale=0.25)
    v1 = O.hafnian_sieve(A, nthreads=1)
    v8 = O.hafnian_sieve(A, nthreads=8)
    assert relerr(v1, v8) < 1e-12


# ---- loop_hafnian_batch (GBS sampler core): every output is one loop hafnian, pinned by Kan
LHB_CASES = [
    (4, [1, 2, 0], 5),   # paired fixed modes (even variant), a zero repetition dropped
    (4, [1, 1, 1], 4),   # one mode left unpaired (odd variant)
    (3, [0, 0], 3),      # no fixed photons: edges [0, 0]
    (2, [1], 4),         # odd variant with no fixed edge: edges [1, odd, 1, 1]
    (5, [2, 0, 1, 1], 6),
    (6, [3, 1, 0, 2, 1], 7),
]


@pytest.mark.parametrize("n,fixed,N_cutoff", LHB_CASES)
@pytest.mark.parametrize("glynn", [True, False])
@pytest.mark.parametrize("precision", ["double", "dd"])
def test_loop_hafnian_batch_vs_kan(n, fixed, N_cutoff, glynn, precision):
    """Output q is lhaf of A with repetitions fixed + [q] and diagonal D (Kan's formula)."""
    rng = np.random.default_rng(1000 + 7 * n + N_cutoff)
    A = random_symmetric(rng, n, scale=0.6)
    D = rng.normal(size=n) + 1j * rng.normal(size=n)
    H = O.loop_hafnian_batch(A, D, fixed, N_cutoff, glynn=glynn, precision=precision)
    assert H.shape == (N_cutoff + 1,)
    for q in range(N_cutoff + 1):
        k = O.kan(A, list(fixed) + [q], mu=D, loop=True)
        assert relerr(H[q], k) < (1e-11 if precision == "double" else 1e-14), q


def test_loop_hafnian_batch_odd_zero_cutoff():
    """Odd variant at N_cutoff = 0: batch_max = -1 leaves no term, the result is [0]
    (loop_hafnian_batch.py:298-304 with steps = prod(edge_reps + 1) = 0)."""
    rng = np.random.default_rng(5)
    A = random_symmetric(rng, 3)
    D = rng.normal(size=3) + 0j
    H = O.loop_hafnian_batch(A, D, [1, 0], 0)
    assert H.shape == (1,) and H[0] == 0


def test_loop_hafnian_batch_gamma_rows_vs_kan():
    rng = np.random.default_rng(77)
    A = random_symmetric(rng, 4, scale=0.5)
    D = rng.normal(size=(3, 4)) + 1j * rng.normal(size=(3, 4))
    H = O.loop_hafnian_batch_gamma(A, D, [1, 0, 2], 3, precision="dd")
    for k in range(3):
        for q in range(4):
            assert relerr(H[k, q], O.kan(A, [1, 0, 2, q], mu=D[k], loop=True)) < 1e-14


@pytest.mark.parametrize("n", [6, 8, 12, 14])
@pytest.mark.parametrize("loop", [False, True])
def test_timed_cpu_baseline_matches_the_oracle(n, loop):
    """The timed CPU baseline (cpu_timing.cpp: the reference's per-term products, FMA contraction
    on) gives the oracle's hafnian to rounding, over all terms at once and over a split range of
    terms (the bounded samples bench.py times); J_n gives (n - 1)!!."""
    rng = np.random.default_rng(300 + n + loop)
    A = random_symmetric(rng, n, scale=1 / np.sqrt(n))
    Ax, Dx = O.ct_prepare(A, loop)
    got = O.ct_sieve_many([Ax, Ax], None if Dx is None else [Dx, Dx], nthreads=2)
    want = O.hafnian_sieve(A, loop=loop, precision="dd")
    assert relerr(got[0], want) < 1e-11 and got[0] == got[1]
    steps = 1 << (n // 2 - 1)
    half = steps // 2
    split = O.ct_sieve(Ax, Dx, 0, half, nthreads=2) + O.ct_sieve(Ax, Dx, half, steps, nthreads=2)
    assert relerr(split * 2.0 ** -(n // 2 - 1), want) < 1e-11
    J = np.ones((n, n))
    Jx, JD = O.ct_prepare(J, False)
    assert O.ct_sieve_many([Jx], None)[0] == math.prod(range(n - 1, 0, -2))
