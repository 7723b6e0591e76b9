"""GPU parity of the HIP sieve (through the C ABI) against the CPU oracle and closed forms.

Tolerance: relative error <= 1e-10 against double-double truth (the oracle's high-precision
restatement), the bar the reference's FP64 path is held to; exact integers where the matrix is
integral (|Re - truth| < 0.5, round(Re) == truth).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import random_symmetric, relerr

import thewalrus_amd as tw
from thewalrus_amd import _native

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.mark.parametrize("n", [6, 8, 10, 12, 16, 20, 24])
def test_all_ones_exact(n):
    """haf(J_n) = (n-1)!! exactly after rounding (test_hafnian.py:211-216 generalised)."""
    A = np.ones((n, n))
    v = tw.hafnian(A)
    want = O.double_factorial(n - 1)
    assert abs(v.real - want) < 0.5 and round(v.real) == want and abs(v.imag) < 1e-6


@pytest.mark.parametrize("n", [6, 8])
def test_block_ones(n):
    """haf([[0, J], [J, 0]]) = n! (test_hafnian.py:220-228)."""
    import math

    O0, B = np.zeros((n, n)), np.ones((n, n))
    A = np.block([[O0, B], [B, O0]])
    assert abs(tw.hafnian(A) - math.factorial(n)) < 1e-9 * math.factorial(n)


@pytest.mark.parametrize("n", [6, 7, 8, 9, 10])
def test_loop_ones_telephone(n):
    """lhaf(J_n) = T(n) (test_hafnian.py:285-290)."""
    A = np.ones((n, n))
    assert abs(tw.hafnian(A, loop=True) - O.TELEPHONE[n]) < 1e-9 * O.TELEPHONE[n]


@pytest.mark.parametrize("n", [6, 8, 10, 14, 20])
@pytest.mark.parametrize("loop", [False, True])
def test_random_complex_vs_truth(n, loop):
    rng = np.random.default_rng(1000 + n + 50 * loop)
    A = random_symmetric(rng, n, scale=1.0 / np.sqrt(n))
    got = tw.hafnian(A, loop=loop)
    truth = O.hafnian_sieve(A, loop=loop, precision="dd")
    ref64 = O.hafnian_sieve(A, loop=loop, precision="double")
    assert relerr(got, truth) < TOL, (got, truth, ref64)


@pytest.mark.parametrize("n", [5, 7, 9, 11, 13, 17, 19, 21])
def test_odd_loop_vs_truth(n):
    rng = np.random.default_rng(2000 + n)
    A = random_symmetric(rng, n, scale=1.0 / np.sqrt(n))
    got = tw.hafnian(A, loop=True)
    truth = O.hafnian_sieve(A, loop=True, precision="dd")
    assert relerr(got, truth) < TOL


@pytest.mark.parametrize("rpt", [[1] * 17 + [2], [2, 1, 1, 3, 1, 1, 2, 1, 1, 1, 1, 2, 1, 1]])
def test_odd_repeated_loop_vs_truth(rpt):
    """Odd photon numbers whose reordered side 2E exceeds 16: the general kernel's path that
    forms powers out of place, with the odd-N loop scalars as bilinear sums over them."""
    n = len(rpt)
    rng = np.random.default_rng(2100 + n)
    A = random_symmetric(rng, n, scale=1.0 / np.sqrt(2 * n))
    mu = rng.normal(size=n) + 1j * rng.normal(size=n)
    assert sum(rpt) % 2 == 1
    print("odd repeated loop rpt=%s" % rpt, flush=True)
    got = tw.hafnian_repeated(A, rpt, mu=mu, loop=True)
    truth = O.hafnian_sieve(A, loop=True, rpt=rpt, mu=mu, precision="dd")
    ref64 = O.hafnian_sieve(A, loop=True, rpt=rpt, mu=mu, precision="double")
    bar = max(TOL, 2 * relerr(ref64, truth))
    assert relerr(got, truth) < bar, (got, truth, ref64)


@pytest.mark.parametrize("n", [6, 8, 12])
@pytest.mark.parametrize("loop", [False, True])
def test_inclexcl_vs_truth(n, loop):
    rng = np.random.default_rng(3000 + n)
    A = random_symmetric(rng, n, scale=1.0 / np.sqrt(n))
    truth = O.hafnian_sieve(A, loop=loop, precision="dd")
    if loop:
        got = tw.loop_hafnian(A, glynn=False)
    else:
        got = tw.hafnian(A, method="inclexcl")
    assert relerr(got, truth) < TOL


@pytest.mark.parametrize(
    "rpt",
    [[1, 2, 3, 0], [2, 2, 2, 2], [5, 1, 0, 2], [3, 3, 1, 2], [6, 1, 1, 1], [4, 4, 4, 4], [10, 2, 1, 1]],
)
@pytest.mark.parametrize("loop", [False, True])
def test_repeated_vs_truth(rpt, loop):
    """hafnian_repeated against the oracle (La Budde path whenever h + 1 > m) and Kan's formula."""
    if not loop and sum(rpt) % 2:
        pytest.skip("odd photon number: zero by definition")
    rng = np.random.default_rng(4000 + sum(rpt))
    A = random_symmetric(rng, 4, scale=0.5)
    mu = rng.normal(size=4) + 1j * rng.normal(size=4)
    got = tw.hafnian_repeated(A, rpt, mu=mu, loop=loop)
    truth = O.hafnian_sieve(A, loop=loop, rpt=rpt, mu=mu, precision="dd")
    kan = O.kan(A, rpt, mu=mu, loop=loop)
    assert relerr(truth, kan) < 1e-12
    assert relerr(got, truth) < TOL


def test_repeated_config_shape():
    """8x8 complex, rpt ~ Multinomial(32, uniform), loop weights (the batch-of-requests shape)."""
    rng = np.random.default_rng(140)
    for _ in range(3):
        A = random_symmetric(rng, 8, scale=np.sqrt(1 / 16))
        rpt = rng.multinomial(32, [1 / 8] * 8)
        mu = (rng.normal(size=8) + 1j * rng.normal(size=8)) / np.sqrt(2)
        got = tw.hafnian_repeated(A, rpt, mu=mu, loop=True)
        truth = O.hafnian_sieve(A, loop=True, rpt=rpt, mu=mu, precision="dd")
        assert relerr(got, truth) < TOL


@pytest.mark.parametrize("n", [6, 8])
def test_permanent_identity_deep_labudde(n):
    """hafnian_repeated([[0,1],[1,0]], [n, n]) = n! (test_hafnian_repeated.py:219-222)."""
    import math

    A = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    got = tw.hafnian_repeated(A, [n, n])
    assert abs(got - math.factorial(n)) < 1e-8 * math.factorial(n)


def test_seam_matches_oracle_seam():
    """Same prepared inputs at the kernel seam: device vs oracle double restatement."""
    rng = np.random.default_rng(7)
    A = random_symmetric(rng, 12, scale=0.3)
    x, e, _ = tw.matched_reps([1] * 12)
    Ax = A[np.ix_(x, x)].astype(np.complex128)
    got = _native.calc_hafnian(Ax, e, True)
    ref = O.calc_sieve(Ax, e, glynn=True, precision="double")
    assert relerr(got, ref) < 1e-12


def _involutions(n):
    """Telephone numbers T(n): perfect matchings with loops of K_n, the loop hafnian of J_n."""
    a, b = 1, 1
    for k in range(2, n + 1):
        a, b = b, b + (k - 1) * a
    return b if n >= 1 else 1


def _dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


@pytest.mark.parametrize("n", [46, 48, 50])
@pytest.mark.parametrize("loop", [False, True])
def test_rank_one_closed_forms(n, loop):
    """n = 46, 48 (tile side 6) and 50 (tile side 7, past the earlier limit of 48):
    haf(x x^T) = (n-1)!! prod x and lhaf(x x^T with diagonal x) = T(n) prod x, exact closed forms;
    the sieve barely cancels on these (sum |term| / |sum| about 3 at n = 28)."""
    rng = np.random.default_rng(5000 + n + loop)
    x = np.exp(1j * rng.uniform(0, 2 * np.pi, n)) * rng.uniform(0.9, 1.1, n)
    A = np.outer(x, x)
    if loop:
        A[np.diag_indices(n)] = x
    print("n=%d rank one loop=%s" % (n, loop), flush=True)
    got = tw.hafnian(A, loop=loop)
    want = complex(np.prod(x)) * float(_involutions(n) if loop else _dfact(n - 1))
    assert relerr(got, want) < TOL, (got, want)


def test_n50_direct_sum_of_rank_ones():
    """n = 50 as the direct sum of rank-one blocks of 24 and 26 modes, permuted: the product of
    the two closed forms (the term matrices' powers have rank two)."""
    rng = np.random.default_rng(5002)
    x1 = np.exp(1j * rng.uniform(0, 2 * np.pi, 24)) * rng.uniform(0.9, 1.1, 24)
    x2 = np.exp(1j * rng.uniform(0, 2 * np.pi, 26)) * rng.uniform(0.9, 1.1, 26)
    A = np.zeros((50, 50), dtype=np.complex128)
    A[:24, :24] = np.outer(x1, x1)
    A[24:, 24:] = np.outer(x2, x2)
    p = rng.permutation(50)
    A = A[np.ix_(p, p)]
    print("n=50 direct sum of rank ones", flush=True)
    got = tw.hafnian(A)
    want = complex(np.prod(x1) * np.prod(x2)) * float(_dfact(23) * _dfact(25))
    assert relerr(got, want) < TOL, (got, want)
