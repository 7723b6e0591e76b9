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


class TestHostFastPaths:
    """The shortcuts of the host checks decide exactly as numpy.allclose does."""

    @staticmethod
    def _cases(rng):
        out = []
        for n in (1, 2, 3, 6, 9):
            S = random_symmetric(rng, n)
            out.append(S)
            for scale in (1e-12, 1e-9, 1e-8, 1e-7, 1e-5, 1e-3):  # around atol and rtol
                P = S.copy()
                if n > 1:
                    P[0, 1] += scale
                out.append(P)
            D = np.diag(rng.normal(size=n)).astype(np.complex128)
            for off in (0.0, 5e-9, 1e-8, 1.0000001e-8, 2e-8, 1e-3):
                Q = D.copy()
                if n > 1:
                    Q[0, 1] = Q[1, 0] = off
                out.append(Q)
                out.append(Q.real.copy())
            out.append(np.round(S.real * 4).astype(np.int64))
            out.append(S.astype(np.complex64))
            out.append(S.real.astype(np.float32))
            out.append(np.eye(n, dtype=bool))
        I = np.diag([1.0, 2.0, 3.0])
        I[0, 1] = np.inf
        out.append(I)
        I2 = I.copy()
        I2[1, 0] = np.inf
        out.append(I2)
        return out

    @pytest.mark.parametrize("tol", [(1e-05, 1e-08), (0.0, 0.0), (0.5, 1e-3), (2.0, 0.0)])
    def test_symmetry_and_diagonal_decisions(self, rng, tol):
        rtol, atol = tol
        for A in self._cases(rng):
            want = np.allclose(A, A.T, rtol=rtol, atol=atol)
            try:
                H.input_validation(A, rtol=rtol, atol=atol)
                got = True
            except ValueError:
                got = False
            assert got == want, (A, tol)
            assert H._is_diagonal(A, rtol, atol) == np.allclose(np.diag(np.diag(A)), A, rtol=rtol, atol=atol)

    def test_all_ones_pairing_cache(self):
        for n in range(0, 70):
            a, b = tw.matched_reps([1] * n), H._pairing([1] * n, True)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
        x, e, _ = H._pairing([1] * 8, True)
        e[0] = 7  # a caller's copy: the cache is unchanged
        assert list(H._pairing([1] * 8, True)[1]) == [1] * 4

    def test_reorder_equals_fancy_index(self, rng):
        for n in (2, 5, 8):
            x = rng.permutation(n)
            for A in (random_symmetric(rng, n), rng.normal(size=(n, n)), rng.integers(-5, 5, size=(n, n))):
                R = H._reorder(A, x)
                assert R.dtype == np.complex128 and R.flags.c_contiguous
                assert np.array_equal(R, A[np.ix_(x, x)].astype(np.complex128))
