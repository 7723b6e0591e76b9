"""hafnian_batched / hermite_multidimensional (thewalrus_amd._hermite_multidimensional).

The oracle restates the recursion of _hermite_multidimensional.py and is pinned here by Kan's
formula for every entry (the hafnian_batched identity of test_hermite_multidimensional.py:69-129),
the one-dimensional Hermite polynomials of scipy and the renormalised closed form of
test_hermite_multidimensional.py:33-40.  GPU tests: every entry within max(1e-10 relative,
2x the reference algorithm's FP64 error) of the oracle's double-double tensor."""
from itertools import product

import numpy as np
import pytest
from scipy.special import eval_hermite, eval_hermitenorm

from oracle import oracle as O
from thewalrus_amd import hafnian_repeated
from thewalrus_amd._hermite_multidimensional import hafnian_batched, hermite_multidimensional


def rand_sym(rng, n):
    A = rng.random((n, n)) + 1j * rng.random((n, n))
    return A + A.T


# ---- oracle pins (CPU)
@pytest.mark.parametrize("loop", [False, True])
def test_oracle_vs_kan(loop):
    rng = np.random.default_rng(11)
    n, c = 4, 5
    A = rand_sym(rng, n)
    mu = rng.random(n) + 1j * rng.random(n) if loop else None
    G = O.hafnian_batched(A, c, mu=mu)
    for q in product(range(c), repeat=n):
        k = O.kan(A, list(q), mu=mu, loop=loop)
        if k == 0:
            assert G[q] == 0
        else:
            assert abs(G[q] - k) <= 1e-12 * abs(k), q


def test_oracle_one_dimensional_polynomials():
    for x in np.arange(-1, 1, 0.1):
        v = O.hermite_recursion(np.array([[2.0]]), np.array([2 * x]), [6])
        assert np.allclose(v, [eval_hermite(k, x) for k in range(6)], rtol=1e-13, atol=1e-13)
        v = O.hermite_recursion(np.array([[1.0]]), np.array([x]), [6])
        assert np.allclose(v, [eval_hermitenorm(k, x) for k in range(6)], rtol=1e-13, atol=1e-13)


def test_oracle_renorm_closed_form():
    B = np.sqrt(0.5) * np.array([[0, 1], [1, 0]]) + 0j
    v = O.hermite_recursion(-B, np.zeros(2), [10, 10], renorm=True, precision="dd")
    assert np.abs(v - np.diag(0.5 ** (np.arange(10) / 2))).max() < 1e-15


# ---- host logic (CPU)
def test_dimension_mismatch():
    with pytest.raises(ValueError, match="incompatible"):
        hermite_multidimensional(np.eye(2), 3, y=np.ones(3), modified=True)
    with pytest.raises(ValueError, match="symmetric"):
        hafnian_batched(np.array([[0.0, 1.0], [0.0, 0.0]]), 3)


# ---- device (GPU)
def _check(dev, truth, ref64):
    assert dev.shape == truth.shape
    t, d, r = truth.ravel(), dev.ravel(), ref64.ravel()
    zero = t == 0
    assert np.all(d[zero] == 0)
    err = np.abs(d[~zero] - t[~zero]) / np.abs(t[~zero])
    bar = np.maximum(1e-10, 2 * np.abs(r[~zero] - t[~zero]) / np.abs(t[~zero]))
    assert np.all(err <= bar), (err.max(), np.argmax(err / bar))


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["plain", "loop", "loop_no_edges", "zero"])
def test_gpu_hafnian_batched_vs_repeated(case):
    """test_hermite_multidimensional.py:69-129: every entry equals hafnian_repeated."""
    rng = np.random.default_rng(21)
    n, c = 4, 5
    A = np.zeros((n, n), dtype=complex) if case in ("loop_no_edges", "zero") else rand_sym(rng, n)
    mu = rng.random(n) + 1j * rng.random(n) if case in ("loop", "loop_no_edges") else None
    print("hafnian_batched %s" % case, flush=True)
    G = hafnian_batched(A, c, mu=mu, make_tensor=False)
    loop = mu is not None or case == "zero"
    v1 = np.array([hafnian_repeated(A, q, mu=mu, loop=loop) for q in product(np.arange(c), repeat=n)])
    assert np.allclose(G, v1)
    _check(G.reshape([c] * n), O.hafnian_batched(A, c, mu=mu, precision="dd"), O.hafnian_batched(A, c, mu=mu))


@pytest.mark.gpu
def test_gpu_polynomials_and_renorm():
    x = np.arange(-1, 1, 0.1)
    A = np.ones([1, 1], dtype=complex)
    vals = np.array([hermite_multidimensional(2 * A, 5, y=np.array([x0], dtype=complex)) for x0 in x]).T
    assert np.allclose(vals, np.array([eval_hermite(i, x) for i in range(5)]))
    vals = np.array([hermite_multidimensional(A, 5, y=np.array([x0], dtype=complex)) for x0 in x]).T
    assert np.allclose(vals, np.array([eval_hermitenorm(i, x) for i in range(5)]))
    B = np.sqrt(0.5) * np.array([[0, 1], [1, 0]]) + 0 * 1j
    assert np.allclose(hermite_multidimensional(-B, 10, renorm=True), np.diag(0.5 ** (np.arange(0, 10) / 2)))


@pytest.mark.gpu
@pytest.mark.parametrize("n,c,renorm", [(6, 8, False), (6, 8, True), (3, 30, True), (9, 4, False)])
def test_gpu_larger_tensors(n, c, renorm):
    """Tensors of 2.6e5 (6 modes, cutoff 8) to 2.7e4 entries, per-mode slices large enough for
    the one-launch-per-step kernel, against the oracle."""
    rng = np.random.default_rng(100 + n + c)
    A = rand_sym(rng, n) * 0.3
    mu = (rng.random(n) + 1j * rng.random(n)) * 0.5
    print("hafnian_batched n=%d c=%d renorm=%s" % (n, c, renorm), flush=True)
    G = hafnian_batched(A, c, mu=mu, renorm=renorm)
    _check(G, O.hafnian_batched(A, c, mu=mu, renorm=renorm, precision="dd"), O.hafnian_batched(A, c, mu=mu, renorm=renorm))


@pytest.mark.gpu
def test_gpu_real_dtype_and_cutoff_forms():
    """Real R and y give a float64 tensor; per-mode cutoffs; make_tensor=False flattens."""
    rng = np.random.default_rng(5)
    R = rng.random((3, 3))
    R = R + R.T
    y = rng.random(3)
    G = hermite_multidimensional(R, [3, 4, 5], y=y, modified=True)
    assert G.dtype == np.float64 and G.shape == (3, 4, 5)
    T = O.hermite_recursion(R, y, [3, 4, 5], precision="dd")
    assert np.allclose(G, T.real, rtol=1e-12, atol=1e-12)
    F = hermite_multidimensional(R, np.array(4), y=y, modified=True, make_tensor=False)
    assert F.shape == (64,)
