"""Host-side logic of the drop-in frontends (CPU only; no sieve evaluation).

Mirrors the reference's wrapper tests (test_hafnian.py:74-176, test_hafnian_repeated.py:29-125):
validation errors, the closed-form shortcuts, pairing and reduction, and the prepared sieve
inputs handed to the device.
"""
import numpy as np
import pytest

import thewalrus_amd as tw
from thewalrus_amd import _hafnian as H
from thewalrus_amd import batch as B
from oracle import oracle as O
from tests.conftest import random_symmetric


class TestValidation:
    def test_array_exception(self):
        with pytest.raises(TypeError):
            tw.hafnian(1)

    def test_square_exception(self):
        with pytest.raises(ValueError):
            tw.hafnian(np.zeros([2, 3]))

    def test_non_symmetric_exception(self):
        A = np.ones([4, 4])
        A[0, 1] = 0.0
        with pytest.raises(ValueError):
            tw.hafnian(A)

    def test_nan(self):
        with pytest.raises(ValueError):
            tw.hafnian(np.array([[2, 1], [1, np.nan]]))

    def test_repeated_validation(self):
        A = np.array([[2, 1], [1, 3]])
        with pytest.raises(TypeError):
            tw.hafnian_repeated(1, [1])
        with pytest.raises(ValueError):
            tw.hafnian_repeated(A, [1])
        with pytest.raises(ValueError):
            tw.hafnian_repeated(A, [1, -1])
        with pytest.raises(ValueError):
            tw.hafnian_repeated(A, [1.1, 1])
        with pytest.raises(ValueError):
            tw.hafnian_repeated(np.ones((2, 2)), [1, 1], mu=np.ones(3), loop=True)

    def test_approx_and_recursive_are_not_offered(self):
        A = np.ones((6, 6))
        with pytest.raises(NotImplementedError):
            tw.hafnian(A, approx=True)
        with pytest.raises(NotImplementedError):
            tw.hafnian(A, method="recursive")

    def test_unknown_method_quirk(self):
        """The reference leaves ``glynn`` unbound for an unknown method (_hafnian.py:754-818)."""
        with pytest.raises(UnboundLocalError):
            tw.hafnian(np.ones((6, 6)), method="nonsense")


class TestShortcuts:
    def test_empty(self):
        assert tw.hafnian(np.ndarray((0, 0))) == 1

    def test_odd_plain(self):
        assert tw.hafnian(np.ones([3, 3])) == 0

    def test_diagonal(self):
        v = np.array([1.5, 2.0, -3.0, 0.5])
        assert tw.hafnian(np.diag(v)) == 0
        assert np.isclose(tw.hafnian(np.diag(v), loop=True), np.prod(v))

    def test_2x2_3x3_4x4(self, rng):
        A = random_symmetric(rng, 2)
        assert tw.hafnian(A) == A[0, 1]
        assert tw.hafnian(A, loop=True) == A[0, 1] + A[0, 0] * A[1, 1]
        assert tw.hafnian(np.ones([3, 3]), loop=True) == 4.0
        A = random_symmetric(rng, 4)
        assert np.isclose(tw.hafnian(A), O.brute_force(A))
        assert np.isclose(tw.hafnian(A, loop=True), O.brute_force(A, loop=True))

    def test_2x2_dtype_preserved(self):
        A = np.array([[1, 7], [7, 2]], dtype=np.int64)
        r = tw.hafnian(A)
        assert r == 7 and isinstance(r, np.int64)

    def test_repeated_shortcuts(self):
        A = np.array([[2, 1], [1, 3]])
        assert tw.hafnian_repeated(A, [0, 0]) == 1.0
        assert tw.hafnian_repeated(np.ones((3, 3)), [1, 1, 1]) == 0.0
        Z = np.zeros((3, 3))
        mu = np.array([2.0, 3.0, 0.5])
        assert tw.hafnian_repeated(Z, [1, 1, 1]) == 0
        assert np.isclose(tw.hafnian_repeated(Z, [1, 2, 3], mu=mu, loop=True), np.prod(mu ** np.array([1, 2, 3])))

    def test_loop_single_photon(self):
        A = np.array([[2.0, 1.0], [1.0, 3.0]])
        assert tw.loop_hafnian(A, reps=[0, 1]) == 3.0
        assert tw.loop_hafnian(A, reps=[0, 0]) == 1.0
        assert H._haf(A, reps=[1, 2]) == 0.0  # test_hafnian.py:361-364


class TestPairingAndReduction:
    @pytest.mark.parametrize("n", [6, 8])
    def test_reduction(self, n):
        """test_hafnian.py:45-57."""
        res = tw.reduction(np.array([[0, 1], [1, 0]]), [n, n])
        Z, J = np.zeros([n, n]), np.ones([n, n])
        assert np.all(res == np.vstack([np.hstack([Z, J]), np.hstack([J, Z])]))

    @pytest.mark.parametrize("n", [6, 8])
    def test_reduction_vector(self, n):
        res = tw.reduction(np.array([0, 1]), [n, n])
        assert np.all(res == np.hstack([np.zeros([n, n]), np.ones([n, n])]))

    def test_empty_matched_reps(self):
        """test_hafnian.py:367-373."""
        x, y, z = tw.matched_reps([0, 0, 0, 0])
        assert x.shape == (0,) and y.shape == (0,) and z is None

    def test_matched_reps_plain(self):
        """rpt all ones: x = [n-1, n-3, ..., 1, n-2, ..., 0], e = 1 (_hafnian.py:102-159)."""
        x, e, odd = tw.matched_reps([1] * 8)
        assert list(x) == [7, 5, 3, 1, 6, 4, 2, 0] and list(e) == [1] * 4 and odd is None

    def test_matched_reps_equals_oracle_restatement(self):
        """Bit-exact agreement with the oracle's independent restatement of the pairing."""
        rng = np.random.default_rng(77)
        for _ in range(500):
            rpt = list(rng.integers(0, 9, size=int(rng.integers(1, 10))))
            x, e, odd = tw.matched_reps(rpt)
            xo, eo, oddo = O.pair_modes(rpt)
            assert list(x) == list(xo) and list(e) == list(eo) and odd == oddo, rpt

    def test_matched_reps_self_pair(self):
        x, e, odd = tw.matched_reps([7, 2])
        # 7 > 2*2: self-pair 3 times, keep 1; then (2,1) pair once ...
        xo, eo, oddo = O.pair_modes([7, 2])
        assert list(x) == list(xo) and list(e) == list(eo) and odd == oddo


class TestPreparedJobs:
    def test_plain_job(self, rng):
        A = random_symmetric(rng, 10)
        job = H.prepare_hafnian(A)
        assert isinstance(job, H.SieveJob) and not job.loop
        x, e, _ = tw.matched_reps([1] * 10)
        assert np.array_equal(job.Ax, A[np.ix_(x, x)]) and list(job.edge_reps) == [1] * 5

    def test_loop_job_odd(self, rng):
        A = random_symmetric(rng, 7)
        job = H.prepare_hafnian(A, loop=True)
        assert job.loop and job.oddloop is not None and job.oddV.shape == (6,)

    def test_repeated_job(self, rng):
        A = random_symmetric(rng, 4)
        mu = np.arange(4) + 1j
        job = H.prepare_hafnian_repeated(A, [3, 1, 1, 0], mu=mu, loop=True)
        x, e, odd = tw.matched_reps([3, 1, 1, 0])
        assert odd is not None
        assert np.array_equal(job.Dx, mu[x]) and list(job.edge_reps) == list(e)
        assert job.oddloop == mu[odd]

    def test_batch_prepare_host_values(self):
        p = B.prepare(np.ones((4, 4)))
        assert not p.on_device and p.value == 3.0
        p = B.prepare(np.ones((6, 6)))
        assert p.on_device
